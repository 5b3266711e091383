// Test-only driver around the reference transform engine.
//
// TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load the library built from this file,
// and only as the checker / the timed CPU baseline.
//
// The engine itself is rocket_fft/_pocketfft_hdronly.h, compiled in place from
// the reference checkout (see oracle/Makefile) with the reference's own flags
// (setup.py:38-46: -std=c++11 -O3, POCKETFFT_NO_SANITYCHECK,
// POCKETFFT_CACHE_SIZE=16, POCKETFFT_PTHREADS).  This file restates the thin
// dispatch of rocket_fft/_pocketfft_numba.cpp:25-223 with plain pointers
// instead of Numba's arystruct_t, so the engine can be called from ctypes
// without Numba.  It additionally exposes the engine's private plan facts
// (factor lists, Bluestein choice, twiddles) so that the product's host plan
// layer can be pinned bit-exactly against them.
#include <numpy/npy_common.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <queue>
#include <stdexcept>
#include <thread>
#include <vector>
#include <pthread.h>

// Plan facts (factor lists, Bluestein n2) are private members of the engine's
// plan classes; the standard headers above are already expanded, so this only
// affects the engine header.
#define private public
#include "_pocketfft_hdronly.h"
#undef private

using pocketfft::axes_view_t;
using pocketfft::shape_view_t;
using pocketfft::stride_view_t;
using pocketfft::detail::ndarr;
using pocketfft::detail::rev_iter;
using pocketfft::detail::util;

namespace {

struct Views {
  shape_view_t shape;
  stride_view_t sin, sout;
  axes_view_t axes;
  Views(uint64_t ndim, const npy_intp *shp, const npy_intp *si,
        const npy_intp *so, const uint64_t *ax, uint64_t nax)
      : shape(shp, ndim), sin(si, ndim), sout(so, ndim), axes(ax, nax) {}
};

template <typename T>
void mirror_fill(const Views &v, std::complex<T> *out) {
  // Hermitian completion of numba_c2c_sym (_pocketfft_numba.cpp:111-143).
  ndarr<std::complex<T>> ares(out, v.shape, v.sout);
  rev_iter iter(ares, v.axes);
  while (iter.remaining() > 0) {
    auto val = ares[iter.ofs()];
    ares[iter.rev_ofs()] = std::conj(val);
    iter.advance();
  }
}

}  // namespace

extern "C" {

uint64_t ref_good_size(uint64_t n, int real) {
  return real ? util::good_size_real(n) : util::good_size_cmplx(n);
}

uint64_t ref_largest_prime_factor(uint64_t n) { return util::largest_prime_factor(n); }

double ref_cost_guess(uint64_t n) { return util::cost_guess(n); }

// Complex FFTPACK factor list in execution order (cfftp::factorize).
int ref_cfftp_factors(uint64_t n, uint64_t *out, int cap) {
  pocketfft::detail::cfftp<double> p(n);
  int k = 0;
  for (auto &f : p.fact) {
    if (k < cap) out[k] = f.fct;
    ++k;
  }
  return k;
}

// Real FFTPACK factor list (rfftp::factorize).
int ref_rfftp_factors(uint64_t n, uint64_t *out, int cap) {
  pocketfft::detail::rfftp<double> p(n);
  int k = 0;
  for (auto &f : p.fact) {
    if (k < cap) out[k] = f.fct;
    ++k;
  }
  return k;
}

// 1 = Bluestein (n2 written), 0 = FFTPACK; complex plan choice.
int ref_plan_c(uint64_t n, uint64_t *n2) {
  pocketfft::detail::pocketfft_c<double> p(n);
  if (p.blueplan) {
    *n2 = p.blueplan->n2;
    return 1;
  }
  *n2 = 0;
  return 0;
}

// Same for the real plan.
int ref_plan_r(uint64_t n, uint64_t *n2) {
  pocketfft::detail::pocketfft_r<double> p(n);
  if (p.blueplan) {
    *n2 = p.blueplan->n2;
    return 1;
  }
  *n2 = 0;
  return 0;
}

// e^{+2 pi i k / n} for k < count, as produced by sincos_2pibyn<T>.
void ref_twiddle(uint64_t n, uint64_t count, int dbl, double *out) {
  if (dbl) {
    pocketfft::detail::sincos_2pibyn<double> tw(n);
    for (uint64_t k = 0; k < count; ++k) {
      auto c = tw[k];
      out[2 * k] = c.r;
      out[2 * k + 1] = c.i;
    }
  } else {
    pocketfft::detail::sincos_2pibyn<float> tw(n);
    for (uint64_t k = 0; k < count; ++k) {
      auto c = tw[k];
      out[2 * k] = c.r;
      out[2 * k + 1] = c.i;
    }
  }
}

#define REF_VIEWS Views v(ndim, shape, sin, sout, axes, naxes)

int ref_c2c(int dbl, uint64_t ndim, const npy_intp *shape, const npy_intp *sin,
            const npy_intp *sout, const uint64_t *axes, uint64_t naxes, int forward,
            double fct, uint64_t nthreads, const void *in, void *out) {
  REF_VIEWS;
  try {
    if (dbl)
      pocketfft::c2c<double>(v.shape, v.sin, v.sout, v.axes, forward != 0,
                             (const std::complex<double> *)in, (std::complex<double> *)out,
                             fct, nthreads);
    else
      pocketfft::c2c<float>(v.shape, v.sin, v.sout, v.axes, forward != 0,
                            (const std::complex<float> *)in, (std::complex<float> *)out,
                            (float)fct, nthreads);
  } catch (...) {
    return -1;
  }
  return 0;
}

int ref_r2c(int dbl, uint64_t ndim, const npy_intp *shape, const npy_intp *sin,
            const npy_intp *sout, const uint64_t *axes, uint64_t naxes, int forward,
            double fct, uint64_t nthreads, const void *in, void *out) {
  REF_VIEWS;
  try {
    if (dbl)
      pocketfft::r2c<double>(v.shape, v.sin, v.sout, v.axes, forward != 0,
                             (const double *)in, (std::complex<double> *)out, fct, nthreads);
    else
      pocketfft::r2c<float>(v.shape, v.sin, v.sout, v.axes, forward != 0, (const float *)in,
                            (std::complex<float> *)out, (float)fct, nthreads);
  } catch (...) {
    return -1;
  }
  return 0;
}

// shape is the REAL output shape (numba_c2r reads it from aout).
int ref_c2r(int dbl, uint64_t ndim, const npy_intp *shape, const npy_intp *sin,
            const npy_intp *sout, const uint64_t *axes, uint64_t naxes, int forward,
            double fct, uint64_t nthreads, const void *in, void *out) {
  REF_VIEWS;
  try {
    if (dbl)
      pocketfft::c2r<double>(v.shape, v.sin, v.sout, v.axes, forward != 0,
                             (const std::complex<double> *)in, (double *)out, fct, nthreads);
    else
      pocketfft::c2r<float>(v.shape, v.sin, v.sout, v.axes, forward != 0,
                            (const std::complex<float> *)in, (float *)out, (float)fct,
                            nthreads);
  } catch (...) {
    return -1;
  }
  return 0;
}

int ref_c2c_sym(int dbl, uint64_t ndim, const npy_intp *shape, const npy_intp *sin,
                const npy_intp *sout, const uint64_t *axes, uint64_t naxes, int forward,
                double fct, uint64_t nthreads, const void *in, void *out) {
  REF_VIEWS;
  try {
    if (dbl) {
      pocketfft::r2c<double>(v.shape, v.sin, v.sout, v.axes, forward != 0,
                             (const double *)in, (std::complex<double> *)out, fct, nthreads);
      mirror_fill<double>(v, (std::complex<double> *)out);
    } else {
      pocketfft::r2c<float>(v.shape, v.sin, v.sout, v.axes, forward != 0, (const float *)in,
                            (std::complex<float> *)out, (float)fct, nthreads);
      mirror_fill<float>(v, (std::complex<float> *)out);
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

int ref_dct(int dbl, uint64_t ndim, const npy_intp *shape, const npy_intp *sin,
            const npy_intp *sout, const uint64_t *axes, uint64_t naxes, uint64_t type,
            double fct, int ortho, uint64_t nthreads, const void *in, void *out) {
  REF_VIEWS;
  try {
    if (dbl)
      pocketfft::dct<double>(v.shape, v.sin, v.sout, v.axes, (int)type, (const double *)in,
                             (double *)out, fct, ortho != 0, nthreads);
    else
      pocketfft::dct<float>(v.shape, v.sin, v.sout, v.axes, (int)type, (const float *)in,
                            (float *)out, (float)fct, ortho != 0, nthreads);
  } catch (...) {
    return -1;
  }
  return 0;
}

int ref_dst(int dbl, uint64_t ndim, const npy_intp *shape, const npy_intp *sin,
            const npy_intp *sout, const uint64_t *axes, uint64_t naxes, uint64_t type,
            double fct, int ortho, uint64_t nthreads, const void *in, void *out) {
  REF_VIEWS;
  try {
    if (dbl)
      pocketfft::dst<double>(v.shape, v.sin, v.sout, v.axes, (int)type, (const double *)in,
                             (double *)out, fct, ortho != 0, nthreads);
    else
      pocketfft::dst<float>(v.shape, v.sin, v.sout, v.axes, (int)type, (const float *)in,
                            (float *)out, (float)fct, ortho != 0, nthreads);
  } catch (...) {
    return -1;
  }
  return 0;
}

int ref_r2r_fftpack(int dbl, uint64_t ndim, const npy_intp *shape, const npy_intp *sin,
                    const npy_intp *sout, const uint64_t *axes, uint64_t naxes,
                    int real2hermitian, int forward, double fct, uint64_t nthreads,
                    const void *in, void *out) {
  REF_VIEWS;
  try {
    if (dbl)
      pocketfft::r2r_fftpack<double>(v.shape, v.sin, v.sout, v.axes, real2hermitian != 0,
                                     forward != 0, (const double *)in, (double *)out, fct,
                                     nthreads);
    else
      pocketfft::r2r_fftpack<float>(v.shape, v.sin, v.sout, v.axes, real2hermitian != 0,
                                    forward != 0, (const float *)in, (float *)out,
                                    (float)fct, nthreads);
  } catch (...) {
    return -1;
  }
  return 0;
}

int ref_hartley(int dbl, int genuine, uint64_t ndim, const npy_intp *shape,
                const npy_intp *sin, const npy_intp *sout, const uint64_t *axes,
                uint64_t naxes, double fct, uint64_t nthreads, const void *in, void *out) {
  REF_VIEWS;
  try {
    if (genuine) {
      if (dbl)
        pocketfft::r2r_genuine_hartley<double>(v.shape, v.sin, v.sout, v.axes,
                                               (const double *)in, (double *)out, fct,
                                               nthreads);
      else
        pocketfft::r2r_genuine_hartley<float>(v.shape, v.sin, v.sout, v.axes,
                                              (const float *)in, (float *)out, (float)fct,
                                              nthreads);
    } else {
      if (dbl)
        pocketfft::r2r_separable_hartley<double>(v.shape, v.sin, v.sout, v.axes,
                                                 (const double *)in, (double *)out, fct,
                                                 nthreads);
      else
        pocketfft::r2r_separable_hartley<float>(v.shape, v.sin, v.sout, v.axes,
                                                (const float *)in, (float *)out,
                                                (float)fct, nthreads);
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

}  // extern "C"
