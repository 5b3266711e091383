// plan.hip -- host plan layer of the gfx950 transform engine.
//
// * Integer plan facts of the reference engine, restated bit-exactly:
//   good_size (rocket_fft/_pocketfft_hdronly.h:588-650), largest prime
//   factor and cost model (:548-586), FFTPACK factor orders (:1750-1778,
//   :2629-2652) and the FFTPACK-vs-Bluestein decision (:2843-2861,
//   :2885-2903).  The GPU follows the reference's Bluestein decision and its
//   Bluestein length n2 = good_size_cmplx(2n-1) (:2765-2767).
// * The GPU's own radix schedule (radix 16 first, then 8/4/2, then 9 for
//   each pair of threes, then odd primes), independent of the reference's
//   FFTPACK order: both compute the same DFT, the schedule only changes
//   rounding.
// * Twiddle, chirp and Bluestein-kernel tables generated in FP64 on the host
//   (rounded to FP32 for complex64 plans) and kept resident in HBM for the
//   life of the process, in a thread-safe cache keyed by (device, length,
//   precision, LDS budget, interleaved lines).  There is no LRU cap: the tables are O(n) and HBM is large.
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <type_traits>

#include "rfh.hpp"

namespace rfh {

DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

void DevBuf::upload(const void *host, size_t nbytes) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
  bytes = nbytes;
  if (!nbytes) return;
  HIP_CHECK(hipMalloc(&p, nbytes));
  HIP_CHECK(hipMemcpy(p, host, nbytes, hipMemcpyHostToDevice));
}

// ------------------------------------------------------- reference facts

uint64_t largest_prime_factor(uint64_t n) {
  uint64_t res = 1;
  while ((n & 1) == 0) {
    res = 2;
    n >>= 1;
  }
  for (uint64_t x = 3; x * x <= n; x += 2)
    while (n % x == 0) {
      res = x;
      n /= x;
    }
  return n > 1 ? n : res;
}

double cost_guess(uint64_t n) {
  const double lfp = 1.1;  // penalty of the reference for primes above 5
  const uint64_t ni = n;
  double result = 0.;
  while ((n & 1) == 0) {
    result += 2;
    n >>= 1;
  }
  for (uint64_t x = 3; x * x <= n; x += 2)
    while (n % x == 0) {
      result += (x <= 5) ? double(x) : lfp * double(x);
      n /= x;
    }
  if (n > 1) result += (n <= 5) ? double(n) : lfp * double(n);
  return result * double(ni);
}

// Smallest 2^a 3^b 5^c 7^d 11^e >= n (n itself up to 12): enumerate the
// 11/7/5 part, complete it with powers of two, then trade factors of 2 for
// factors of 3 while the candidate stays even.
uint64_t good_size_cmplx(uint64_t n) {
  if (n <= 12) return n;
  uint64_t best = 2 * n;
  for (uint64_t f11 = 1; f11 < best; f11 *= 11)
    for (uint64_t f7 = f11; f7 < best; f7 *= 7)
      for (uint64_t f5 = f7; f5 < best; f5 *= 5) {
        uint64_t x = f5;
        while (x < n) x *= 2;
        for (;;) {
          if (x < n) {
            x *= 3;
          } else if (x > n) {
            if (x < best) best = x;
            if (x & 1) break;
            x >>= 1;
          } else {
            return n;
          }
        }
      }
  return best;
}

// Same for 2^a 3^b 5^c (n itself up to 6).
uint64_t good_size_real(uint64_t n) {
  if (n <= 6) return n;
  uint64_t best = 2 * n;
  for (uint64_t f5 = 1; f5 < best; f5 *= 5) {
    uint64_t x = f5;
    while (x < n) x *= 2;
    for (;;) {
      if (x < n) {
        x *= 3;
      } else if (x > n) {
        if (x < best) best = x;
        if (x & 1) break;
        x >>= 1;
      } else {
        return n;
      }
    }
  }
  return best;
}

std::vector<uint64_t> cfftp_factors(uint64_t n) {
  std::vector<uint64_t> f;
  if (n <= 1) return f;
  uint64_t len = n;
  while ((len & 7) == 0) {
    f.push_back(8);
    len >>= 3;
  }
  while ((len & 3) == 0) {
    f.push_back(4);
    len >>= 2;
  }
  if ((len & 1) == 0) {
    len >>= 1;
    f.push_back(2);
    std::swap(f.front(), f.back());
  }
  for (uint64_t d = 3; d * d <= len; d += 2)
    while (len % d == 0) {
      f.push_back(d);
      len /= d;
    }
  if (len > 1) f.push_back(len);
  return f;
}

std::vector<uint64_t> rfftp_factors(uint64_t n) {
  std::vector<uint64_t> f;
  if (n <= 1) return f;
  uint64_t len = n;
  while ((len & 3) == 0) {
    f.push_back(4);
    len >>= 2;
  }
  if ((len & 1) == 0) {
    len >>= 1;
    f.push_back(2);
    std::swap(f.front(), f.back());
  }
  for (uint64_t d = 3; d * d <= len; d += 2)
    while (len % d == 0) {
      f.push_back(d);
      len /= d;
    }
  if (len > 1) f.push_back(len);
  return f;
}

bool plan_bluestein(uint64_t n, bool real, uint64_t *n2) {
  *n2 = 0;
  if (n == 0) return false;
  const uint64_t lpf = (n < 50) ? 0 : largest_prime_factor(n);
  if (lpf * lpf <= n) return false;
  const double comp1 = (real ? 0.5 : 1.0) * cost_guess(n);
  const double comp2 = 1.5 * (2 * cost_guess(good_size_cmplx(2 * n - 1)));
  if (comp2 < comp1) {
    *n2 = good_size_cmplx(2 * n - 1);
    return true;
  }
  return false;
}

// ------------------------------------------------------------- GPU schedule

std::vector<int> gpu_radices(uint64_t n) {
  std::vector<int> r;
  if (n <= 1) return r;
  int twos = 0;
  while ((n & 1) == 0) {
    n >>= 1;
    ++twos;
  }
  while (twos >= 4) {
    r.push_back(16);
    twos -= 4;
  }
  if (twos == 3) r.push_back(8);
  if (twos == 2) r.push_back(4);
  if (twos == 1) r.push_back(2);
  int threes = 0;
  while (n % 3 == 0) {
    n /= 3;
    ++threes;
  }
  for (; threes >= 2; threes -= 2) r.push_back(9);
  if (threes) r.push_back(3);
  for (uint64_t d = 5; d * d <= n; d += 2)
    while (n % d == 0) {
      r.push_back(int(d));
      n /= d;
    }
  if (n > 1) r.push_back(int(n));
  return r;
}

size_t lds_limit() {
  static size_t lim = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return size_t(65536);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess ||
        v <= 0)
      return size_t(65536);
    return size_t(v);
  }();
  return lim;
}

// Static LDS of the line kernels (per-line offsets) is reserved first.
static size_t lds_dynamic_budget() { return lds_limit() - kLineStaticLds; }

uint64_t direct_max(bool dbl, int lines, size_t budget) {
  const size_t elem = dbl ? 16 : 8;
  return (budget ? budget : lds_dynamic_budget()) / (2 * elem * size_t(lines));
}

// --------------------------------------------------------------- twiddles

// e^{-2 pi i e / m} in FP64 with exact integer octant reduction.
static std::complex<double> wfwd(uint64_t e, uint64_t m) {
  e %= m;
  const unsigned __int128 x = (unsigned __int128)e * 8u;
  const int oct = int(x / m);
  const uint64_t rem = uint64_t(x - (unsigned __int128)oct * m);
  const double quarter_pi = 0.78539816339744830961566084581987572;
  const double a = (oct & 1) ? double(m - rem) / double(m) * quarter_pi
                             : double(rem) / double(m) * quarter_pi;
  const double ca = std::cos(a), sa = std::sin(a);
  double c, s;
  switch (oct) {
    case 0: c = ca; s = sa; break;
    case 1: c = sa; s = ca; break;
    case 2: c = -sa; s = ca; break;
    case 3: c = -ca; s = sa; break;
    case 4: c = -ca; s = -sa; break;
    case 5: c = -sa; s = -ca; break;
    case 6: c = sa; s = -ca; break;
    default: c = ca; s = -sa; break;
  }
  return {c, -s};  // the angle above is +2 pi e/m; forward sign flips it
}

template <typename T>
static void upload_cplx(DevBuf &b, const std::vector<std::complex<double>> &v) {
  std::vector<T> h(2 * v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    h[2 * i] = T(v[i].real());
    h[2 * i + 1] = T(v[i].imag());
  }
  b.upload(h.data(), h.size() * sizeof(T));
}

static void upload_any(DevBuf &b, const std::vector<std::complex<double>> &v, bool dbl) {
  if (dbl)
    upload_cplx<double>(b, v);
  else
    upload_cplx<float>(b, v);
}

static bool has_fixed_kernel(int r) {
  return r == 2 || r == 3 || r == 4 || r == 5 || r == 8 || r == 9 || r == 16;
}

// ---------------------------------------------------------- plan records
// A plan is decided and its tables generated on the host (FP64) into a
// record; the device plan is built from the record.  Records are what the
// on-disk plan cache stores.

struct DirectRec {
  FftDesc desc{};
  std::vector<std::complex<double>> tw, gen;
};

struct PlanRec {
  int kind = kDirect;
  uint64_t n = 0;
  bool dbl = false;
  size_t budget = 0;
  bool inter = false;
  uint64_t n1 = 0, n2 = 0, bn2 = 0;
  int shift = 0;
  DirectRec main, pa, pb;
  std::vector<std::complex<double>> big_lo, big_hi, bk, bkf;
  uint64_t sub_n = 0;  // sub-plan (same precision and budget), 0: none
  bool sub_inter = false;
};

struct RealRec {
  uint64_t n = 0;
  bool dbl = false;
  size_t budget = 0;
  bool fused = false, half = false;
  DirectRec d;
  std::vector<std::complex<double>> rtw;
};

static void make_direct(DirectRec &d, uint64_t n) {
  std::vector<int> rad = gpu_radices(n);
  if (rad.size() > size_t(kMaxStages)) RFH_THROW(3, "too many FFT stages");
  d.desc = FftDesc{};
  d.desc.n = int(n);
  d.desc.nst = int(rad.size());
  d.tw.clear();
  d.gen.clear();
  int ns = 1;
  for (size_t s = 0; s < rad.size(); ++s) {
    const int R = rad[s];
    Stage &st = d.desc.st[s];
    st.radix = R;
    st.ns = ns;
    st.twoff = int(d.tw.size());
    st.genoff = int(d.gen.size());
    if (ns > 1)
      for (int r = 1; r < R; ++r)
        for (int k = 0; k < ns; ++k) d.tw.push_back(wfwd(uint64_t(k) * r, uint64_t(ns) * R));
    if (!has_fixed_kernel(R))
      for (int q = 0; q < R; ++q) d.gen.push_back(std::conj(wfwd(q, R)));  // (cos, +sin)
    ns *= R;
  }
  if (d.tw.empty()) d.tw.push_back({1, 0});
  if (d.gen.empty()) d.gen.push_back({1, 0});
}

static void upload_direct(Direct &d, const DirectRec &r, bool dbl) {
  d.desc = r.desc;
  upload_any(d.tw, r.tw, dbl);
  upload_any(d.gen, r.gen, dbl);
}

// Plain recursive mixed-radix DFT in FP64 (plan time only: the Bluestein
// kernel).  Forward sign.
static void host_dft(const std::complex<double> *x, size_t xs, std::complex<double> *y,
                     uint64_t n, const std::vector<int> &fac, size_t fi,
                     const std::vector<std::complex<double>> &w, uint64_t wn,
                     std::vector<std::complex<double>> &scratch) {
  if (n == 1) {
    y[0] = x[0];
    return;
  }
  const uint64_t p = uint64_t(fac[fi]), m = n / p, st = wn / n;
  for (uint64_t r = 0; r < p; ++r)
    host_dft(x + r * xs, xs * p, y + r * m, m, fac, fi + 1, w, wn, scratch);
  std::vector<std::complex<double>> t(n);
  for (uint64_t k = 0; k < m; ++k)
    for (uint64_t q = 0; q < p; ++q) {
      const uint64_t kk = k + m * q;
      std::complex<double> acc = y[k];
      for (uint64_t r = 1; r < p; ++r) acc += w[((r * kk) % n) * st] * y[r * m + k];
      t[kk] = acc;
    }
  std::copy(t.begin(), t.end(), y);
}

static void make_bluestein(PlanRec &p, uint64_t n, uint64_t n2) {
  // chirp e^{+i pi m^2 / n}; the exponent m^2 mod 2n is exact integer math
  // (the reference's recurrence coeff += 2m-1, _pocketfft_hdronly.h:2773-2780).
  std::vector<std::complex<double>> &bk = p.bk;
  bk.assign(n, {0, 0});
  uint64_t coeff = 0;
  bk[0] = {1, 0};
  for (uint64_t m = 1; m < n; ++m) {
    coeff += 2 * m - 1;
    if (coeff >= 2 * n) coeff -= 2 * n;
    bk[m] = std::conj(wfwd(coeff, 2 * n));
  }
  std::vector<std::complex<double>> bz(n2, {0, 0});
  p.bkf.assign(n2, {0, 0});
  const double xn2 = 1.0 / double(n2);
  bz[0] = bk[0] * xn2;
  for (uint64_t m = 1; m < n; ++m) bz[m] = bz[n2 - m] = bk[m] * xn2;
  std::vector<std::complex<double>> w(n2);
  for (uint64_t e = 0; e < n2; ++e) w[e] = wfwd(e, n2);
  std::vector<int> fac = gpu_radices(n2);
  std::vector<std::complex<double>> scratch;
  host_dft(bz.data(), 1, p.bkf.data(), n2, fac, 0, w, n2, scratch);
  p.bn2 = n2;
}

// Two-pass split N = n1 * n2 with both passes in LDS at `lines` lines per
// group; the most balanced split wins.  0 when none exists.
static uint64_t pick_split(uint64_t N, bool dbl, int lines, size_t budget) {
  const uint64_t lim = direct_max(dbl, lines, budget);
  uint64_t best = 0;
  double bestd = 1e300;
  for (uint64_t d = 2; d <= lim && d <= N / 2; ++d) {
    if (N % d) continue;
    const uint64_t e = N / d;
    if (e > lim || e < 2) continue;
    const double dd = std::fabs(std::log(double(d)) - std::log(double(e)));
    if (dd < bestd) {
      bestd = dd;
      best = d;
    }
  }
  return best;
}

static int two_pass_lines(bool dbl) { return dbl ? 4 : 8; }  // 64-byte row segments

static void make_big_twiddle(PlanRec &p, uint64_t N) {
  int shift = 0;
  while ((uint64_t(1) << (2 * shift)) < N) ++shift;
  p.shift = shift;
  const uint64_t lo_n = uint64_t(1) << shift, hi_n = (N + lo_n - 1) / lo_n;
  p.big_lo.resize(lo_n);
  p.big_hi.resize(hi_n);
  for (uint64_t i = 0; i < lo_n; ++i) p.big_lo[i] = wfwd(i, N);
  for (uint64_t i = 0; i < hi_n; ++i) p.big_hi[i] = wfwd(i * lo_n, N);
}

static void make_two_pass(PlanRec &p, uint64_t N, size_t budget, uint64_t n1_given = 0) {
  const uint64_t n1 = n1_given ? n1_given : pick_split(N, p.dbl, two_pass_lines(p.dbl), budget);
  if (!n1) RFH_THROW(3, "no two-pass split for length " + std::to_string(N));
  p.n1 = n1;
  p.n2 = N / n1;
  make_direct(p.pa, p.n1);
  make_direct(p.pb, p.n2);
  make_big_twiddle(p, N);
}

std::shared_ptr<C2CPlan> get_c2c_plan_any(uint64_t n, bool dbl, size_t budget, bool inter);

static bool all_factors_at_most(uint64_t n, uint64_t lim) {
  return largest_prime_factor(n) <= lim;
}

// Column length n1 of a multi-pass split N = n1 * M (n1 <= lim, all prime
// factors of M <= lim): M should take two passes of its own (or more when N
// exceeds lim^3), so n1 near the cube root of N is preferred.
static uint64_t pick_multi(uint64_t N, bool dbl, int lines, size_t budget) {
  const uint64_t lim = direct_max(dbl, lines, budget);
  const double target = std::cbrt(double(N));
  uint64_t best = 0;
  double bestd = 1e300;
  for (uint64_t d = 2; d <= lim && d <= N / 2; ++d) {
    if (N % d) continue;
    const uint64_t M = N / d;
    if (M <= lim) continue;  // the two-pass split covers it
    const bool two = pick_split(M, dbl, lines, budget) != 0;
    if (!two && M <= lim * lim) continue;
    if (!all_factors_at_most(M, lim)) continue;
    const double dd = std::fabs(std::log(double(d)) - std::log(target)) + (two ? 0.0 : 100.0);
    if (dd < bestd) {
      bestd = dd;
      best = d;
    }
  }
  return best;
}

static void make_multi(PlanRec &p, uint64_t N, uint64_t n1) {
  p.n1 = n1;
  p.n2 = N / n1;
  make_direct(p.pa, n1);
  make_big_twiddle(p, N);
  // the rows are written with the column index as their fastest output stride
  p.sub_n = p.n2;
  p.sub_inter = true;
}

// GPU Bluestein length.  The reference pads to good_size(2n - 1)
// (_pocketfft_hdronly.h:2765-2767), whose balanced two-pass split rarely has
// a power-of-two side (131220 = 324 x 405 for n = 65537).  Any length
// N >= 2n - 2 gives the same transform: the circular convolution of the
// chirped input (indices 0 .. n-1) with the kernel b[m] = e^{i pi m^2 / n}
// (m = -(n-1) .. n-1) only needs each difference k - j in -(n-1) .. n-1 at a
// distinct index mod N, except that +(n-1) and -(n-1) may share index n-1
// because b is even (b[m] = b[-m]); N = 2n - 2 is the least such length.
// So when the padded length needs two passes the GPU uses n1 * m with n1 a
// power of two (the column passes on the specialised kernel) and
// m = good_size(ceil((2n - 2) / n1)), if that is at most 6% longer than the
// reference's length; among equal lengths the larger n1 wins (shorter rows on
// the generic kernel).  For n = 2^k + 1 (65537) this is exactly 2^(k+1).
// Returns n1 (0: keep the reference length).
static uint64_t pick_blue_pow2(uint64_t n, uint64_t ref_n2, bool dbl, size_t budget,
                               uint64_t *gpu_n2) {
  const uint64_t lim = direct_max(dbl, two_pass_lines(dbl), budget);
  const uint64_t need = 2 * n - 2;
  uint64_t best_n1 = 0, best = 0;
  for (uint64_t n1 : {uint64_t(1024), uint64_t(512), uint64_t(256)}) {
    if (n1 > lim || !fast_interleaved_ok(n1, dbl)) continue;
    const uint64_t m = good_size_cmplx((need + n1 - 1) / n1);
    if (m < 2 || m > lim) continue;
    if (!best || n1 * m < best) {
      best = n1 * m;
      best_n1 = n1;
    }
  }
  if (!best || double(best) > 1.06 * double(ref_n2)) return 0;
  *gpu_n2 = best;
  return best_n1;
}

static PlanRec build_c2c_rec(uint64_t n, bool dbl, size_t budget, bool inter) {
  PlanRec p;
  p.n = n;
  p.dbl = dbl;
  p.budget = budget;
  p.inter = inter;
  const int lines = two_pass_lines(dbl);
  // a length fits one LDS pass: any line group for contiguous lines; for
  // interleaved lines, a group of adjacent lines wide enough for full row
  // segments (the specialised power-of-two kernel, or `lines` lines)
  auto direct_ok = [&](uint64_t N) {
    if (!inter) return N <= direct_max(dbl, 1, budget);
    return fast_interleaved_ok(N, dbl) || N <= direct_max(dbl, lines, budget);
  };
  uint64_t bn2 = 0;
  bool blue = n > 1 && plan_bluestein(n, false, &bn2);
  uint64_t n1 = 0;
  if (!blue) {
    if (direct_ok(n)) {
      p.kind = kDirect;
      make_direct(p.main, n);
      return p;
    }
    if (pick_split(n, dbl, lines, budget)) {
      p.kind = kTwoPass;
      make_two_pass(p, n, budget);
      return p;
    }
    if ((n1 = pick_multi(n, dbl, lines, budget)) != 0) {
      p.kind = kMultiPass;
      make_multi(p, n, n1);
      return p;
    }
    // a long length with a prime factor too large for any LDS pass
    blue = true;
    bn2 = good_size_cmplx(2 * n - 1);
  }
  uint64_t gbn2 = 0, gn1 = 0;
  if (direct_ok(bn2)) {
    p.kind = kBlueDirect;
    make_direct(p.main, bn2);
  } else if ((gn1 = pick_blue_pow2(n, bn2, dbl, budget, &gbn2)) != 0) {
    p.kind = kBlueTwoPass;
    bn2 = gbn2;
    make_two_pass(p, bn2, budget, gn1);
  } else if (pick_split(bn2, dbl, lines, budget)) {
    p.kind = kBlueTwoPass;
    make_two_pass(p, bn2, budget);
  } else {
    p.kind = kBlueMulti;
    p.sub_n = bn2;
    p.sub_inter = false;
  }
  make_bluestein(p, n, bn2);
  return p;
}

static RealRec build_real_rec(uint64_t n, bool dbl, size_t budget) {
  RealRec p;
  p.n = n;
  p.dbl = dbl;
  p.budget = budget;
  const uint64_t m = (n % 2 == 0) ? n / 2 : n;
  uint64_t bn2;
  // The fused pass runs the complex FFT of m as one LDS-resident mixed-radix
  // plan.  When the Bluestein rule picks Bluestein for m (a large prime
  // factor, whose generic radix stage would cost O(m p) in one workgroup), or
  // m does not fit LDS, even lengths run the FFT of the n/2 pairs through the
  // complex plan of n/2 with the split twiddle in the pre/post kernels, and
  // odd lengths the complex plan of n.
  const bool blue = m > 1 && plan_bluestein(m, false, &bn2);
  if (n >= 2 && !blue && m <= direct_max(dbl, 1, budget)) {
    p.fused = true;
    make_direct(p.d, m);
    if (n % 2 == 0) {
      p.rtw.resize(n / 2 + 1);
      for (uint64_t k = 0; k <= n / 2; ++k) p.rtw[k] = wfwd(k, n);
    }
  } else {
    p.half = n >= 2 && n % 2 == 0;
  }
  return p;
}

static std::shared_ptr<C2CPlan> materialize(const PlanRec &r) {
  auto p = std::make_shared<C2CPlan>();
  p->n = r.n;
  p->dbl = r.dbl;
  p->budget = r.budget;
  p->kind = r.kind;
  p->n1 = r.n1;
  p->n2 = r.n2;
  p->shift = r.shift;
  if (r.kind == kDirect || r.kind == kBlueDirect) upload_direct(p->main, r.main, r.dbl);
  if (r.kind == kTwoPass || r.kind == kBlueTwoPass || r.kind == kMultiPass)
    upload_direct(p->pa, r.pa, r.dbl);
  if (r.kind == kTwoPass || r.kind == kBlueTwoPass) upload_direct(p->pb, r.pb, r.dbl);
  if (!r.big_lo.empty()) {
    upload_any(p->big_lo, r.big_lo, r.dbl);
    upload_any(p->big_hi, r.big_hi, r.dbl);
  }
  if (r.bn2) {
    upload_any(p->bk, r.bk, r.dbl);
    upload_any(p->bkf, r.bkf, r.dbl);
    p->bn2 = r.bn2;
  }
  if (r.sub_n) p->sub = get_c2c_plan_any(r.sub_n, r.dbl, r.budget, r.sub_inter);
  return p;
}

static std::shared_ptr<RealPlan> materialize(const RealRec &r) {
  auto p = std::make_shared<RealPlan>();
  p->n = r.n;
  p->dbl = r.dbl;
  p->fused = r.fused;
  p->half = r.half;
  if (r.fused) upload_direct(p->d, r.d, r.dbl);
  if (!r.rtw.empty()) upload_any(p->rtw, r.rtw, r.dbl);
  return p;
}

// ------------------------------------------------------------------ cache
//
// In memory: one plan per (device, kind, n, precision, LDS budget,
// interleaved lines), for the life of the process.  On disk (opt-in,
// rfh_plan_cache): the plan records, one file per (kind, n, precision, LDS
// budget, stride class) and GPU architecture, so that a later process skips
// the FP64 table generation (the Bluestein kernel's host DFT above all):
//   <dir>/plan_<arch>_<c2c|real>_<n>_<c64|c128>_b<budget>_<i|c>.bin
// Format: "RFHPLAN1", the key fields, the record, an FNV-1a checksum of all
// that; a file that does not parse, checksum or match its key is ignored
// (and rewritten).  Files are written to a temporary name and renamed.

namespace {

struct Writer {
  std::string b;
  template <typename V>
  void put(const V &v) {
    b.append(reinterpret_cast<const char *>(&v), sizeof(V));
  }
  void vec(const std::vector<std::complex<double>> &v) {
    put(uint64_t(v.size()));
    b.append(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(v[0]));
  }
  void direct(const DirectRec &d) {
    put(d.desc);
    vec(d.tw);
    vec(d.gen);
  }
};

struct Reader {
  const std::string &b;
  size_t at = 0;
  bool ok = true;
  template <typename V>
  void get(V &v) {
    if (!ok || at + sizeof(V) > b.size()) {
      ok = false;
      return;
    }
    std::memcpy(&v, b.data() + at, sizeof(V));
    at += sizeof(V);
  }
  void vec(std::vector<std::complex<double>> &v) {
    uint64_t n = 0;
    get(n);
    if (!ok || n > (b.size() - at) / sizeof(std::complex<double>)) {
      ok = false;
      return;
    }
    v.resize(n);
    std::memcpy(v.data(), b.data() + at, n * sizeof(std::complex<double>));
    at += n * sizeof(std::complex<double>);
  }
  void direct(DirectRec &d) {
    get(d.desc);
    vec(d.tw);
    vec(d.gen);
    if (ok && (d.desc.nst < 0 || d.desc.nst > kMaxStages)) ok = false;
  }
};

uint64_t fnv1a(const char *p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ uint64_t(uint8_t(p[i]))) * 1099511628211ull;
  return h;
}

const char kMagic[8] = {'R', 'F', 'H', 'P', 'L', 'A', 'N', '1'};

std::string seal(const Writer &w) {
  std::string out(kMagic, 8);
  out += w.b;
  const uint64_t h = fnv1a(out.data(), out.size());
  out.append(reinterpret_cast<const char *>(&h), sizeof h);
  return out;
}

// the payload of a sealed file, or empty when it does not check
std::string unseal(const std::string &f) {
  if (f.size() < 16 || std::memcmp(f.data(), kMagic, 8) != 0) return {};
  uint64_t h = 0;
  std::memcpy(&h, f.data() + f.size() - 8, 8);
  if (h != fnv1a(f.data(), f.size() - 8)) return {};
  return f.substr(8, f.size() - 16);
}

std::string encode(const PlanRec &r) {
  Writer w;
  w.put(uint8_t(0));  // c2c
  w.put(r.n);
  w.put(uint8_t(r.dbl));
  w.put(uint64_t(r.budget));
  w.put(uint8_t(r.inter));
  w.put(int32_t(r.kind));
  w.put(r.n1);
  w.put(r.n2);
  w.put(r.bn2);
  w.put(int32_t(r.shift));
  w.direct(r.main);
  w.direct(r.pa);
  w.direct(r.pb);
  w.vec(r.big_lo);
  w.vec(r.big_hi);
  w.vec(r.bk);
  w.vec(r.bkf);
  w.put(r.sub_n);
  w.put(uint8_t(r.sub_inter));
  return seal(w);
}

std::string encode(const RealRec &r) {
  Writer w;
  w.put(uint8_t(1));  // real
  w.put(r.n);
  w.put(uint8_t(r.dbl));
  w.put(uint64_t(r.budget));
  w.put(uint8_t(0));
  w.put(uint8_t(r.fused));
  w.put(uint8_t(r.half));
  w.direct(r.d);
  w.vec(r.rtw);
  return seal(w);
}

// key fields of a record file; false when they do not match
bool read_key(Reader &rd, uint8_t kind, uint64_t n, bool dbl, size_t budget, bool inter) {
  uint8_t k = 255, d = 255, it = 255;
  uint64_t nn = 0, bb = 0;
  rd.get(k);
  rd.get(nn);
  rd.get(d);
  rd.get(bb);
  rd.get(it);
  return rd.ok && k == kind && nn == n && d == uint8_t(dbl) && bb == budget &&
         it == uint8_t(inter);
}

bool decode(const std::string &f, uint64_t n, bool dbl, size_t budget, bool inter, PlanRec *r) {
  const std::string pay = unseal(f);
  if (pay.empty()) return false;
  Reader rd{pay};
  if (!read_key(rd, 0, n, dbl, budget, inter)) return false;
  r->n = n;
  r->dbl = dbl;
  r->budget = budget;
  r->inter = inter;
  int32_t kind = -1, shift = 0;
  uint8_t si = 0;
  rd.get(kind);
  rd.get(r->n1);
  rd.get(r->n2);
  rd.get(r->bn2);
  rd.get(shift);
  rd.direct(r->main);
  rd.direct(r->pa);
  rd.direct(r->pb);
  rd.vec(r->big_lo);
  rd.vec(r->big_hi);
  rd.vec(r->bk);
  rd.vec(r->bkf);
  rd.get(r->sub_n);
  rd.get(si);
  r->kind = kind;
  r->shift = shift;
  r->sub_inter = si != 0;
  return rd.ok && rd.at == pay.size() && kind >= kDirect && kind <= kBlueMulti;
}

bool decode(const std::string &f, uint64_t n, bool dbl, size_t budget, RealRec *r) {
  const std::string pay = unseal(f);
  if (pay.empty()) return false;
  Reader rd{pay};
  if (!read_key(rd, 1, n, dbl, budget, false)) return false;
  r->n = n;
  r->dbl = dbl;
  r->budget = budget;
  uint8_t fu = 0, ha = 0;
  rd.get(fu);
  rd.get(ha);
  rd.direct(r->d);
  rd.vec(r->rtw);
  r->fused = fu != 0;
  r->half = ha != 0;
  return rd.ok && rd.at == pay.size();
}

std::mutex g_cache_mu;
typedef std::tuple<int, uint64_t, bool, size_t, bool> PlanKey;
std::map<PlanKey, std::shared_ptr<C2CPlan>> g_c2c;
std::map<PlanKey, std::shared_ptr<RealPlan>> g_real;
std::vector<std::shared_ptr<const void>> g_retired;  // dropped from the maps
std::string g_plan_dir;              // empty: no disk cache
uint64_t g_disk_hits = 0, g_disk_writes = 0;

std::string device_arch() {
  int dev = 0;
  std::string arch = "gpu";
  hipDeviceProp_t pr;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&pr, dev) == hipSuccess) {
    arch.clear();
    for (const char *c = pr.gcnArchName; *c && *c != ':'; ++c)
      arch += std::isalnum(static_cast<unsigned char>(*c)) ? *c : '_';
  }
  return arch;
}

std::string plan_file(const std::string &dir, const std::string &arch, const char *kind,
                      uint64_t n, bool dbl, size_t budget, bool inter) {
  return dir + "/plan_" + arch + "_" + kind + "_" + std::to_string(n) + (dbl ? "_c128" : "_c64") +
         "_b" + std::to_string(budget) + (inter ? "_i" : "_c") + ".bin";
}

bool read_file(const std::string &path, std::string *out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  *out = ss.str();
  return true;
}

void write_file(const std::string &path, const std::string &data) {
  const std::string tmp = path + ".tmp" + std::to_string(::getpid());
  {
    std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
    if (!f) return;  // an unwritable directory only costs the cache
    f.write(data.data(), std::streamsize(data.size()));
    if (!f) {
      std::remove(tmp.c_str());
      return;
    }
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

// The record of a plan: from the disk cache when a valid file is there,
// else built (and written when a cache directory is set).
template <typename Rec, typename B>
Rec cached_record(const char *kind, uint64_t n, bool dbl, size_t budget, bool inter, B &&build) {
  std::string dir;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    dir = g_plan_dir;
  }
  std::string path;
  if (!dir.empty()) {
    path = plan_file(dir, device_arch(), kind, n, dbl, budget, inter);
    std::string f;
    Rec r;
    bool hit;
    if constexpr (std::is_same<Rec, PlanRec>::value)
      hit = read_file(path, &f) && decode(f, n, dbl, budget, inter, &r);
    else
      hit = read_file(path, &f) && decode(f, n, dbl, budget, &r);
    if (hit) {
      std::lock_guard<std::mutex> lk(g_cache_mu);
      ++g_disk_hits;
      return r;
    }
  }
  Rec r = build();
  if (!path.empty()) {
    write_file(path, encode(r));
    std::lock_guard<std::mutex> lk(g_cache_mu);
    ++g_disk_writes;
  }
  return r;
}

template <typename P, typename B>
std::shared_ptr<P> cached(std::map<PlanKey, std::shared_ptr<P>> &m, uint64_t n, bool dbl,
                          size_t budget, bool inter, B &&build) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  const size_t b = budget ? budget : lds_dynamic_budget();
  const PlanKey key = std::make_tuple(dev, n, dbl, b, inter);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = m.find(key);
    if (it != m.end()) return it->second;
  }
  auto p = build(n, dbl, b, inter);  // outside the lock (sub-plans recurse)
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = m.find(key);
  if (it != m.end()) return it->second;
  m[key] = p;
  return p;
}
}  // namespace

std::shared_ptr<C2CPlan> get_c2c_plan_any(uint64_t n, bool dbl, size_t budget, bool inter) {
  return cached(g_c2c, n, dbl, budget, inter, [](uint64_t n_, bool d_, size_t b_, bool i_) {
    return materialize(cached_record<PlanRec>("c2c", n_, d_, b_, i_,
                                              [&] { return build_c2c_rec(n_, d_, b_, i_); }));
  });
}

std::shared_ptr<C2CPlan> get_c2c_plan(uint64_t n, bool dbl, size_t budget, bool interleaved) {
  return get_c2c_plan_any(n, dbl, budget, interleaved);
}

std::shared_ptr<C2CPlan> find_c2c_plan(uint64_t n, bool dbl, size_t budget, bool interleaved) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  const PlanKey key =
      std::make_tuple(dev, n, dbl, budget ? budget : lds_dynamic_budget(), interleaved);
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_c2c.find(key);
  return it == g_c2c.end() ? nullptr : it->second;
}

std::shared_ptr<RealPlan> get_real_plan(uint64_t n, bool dbl, size_t budget) {
  return cached(g_real, n, dbl, budget, false, [](uint64_t n_, bool d_, size_t b_, bool) {
    return materialize(cached_record<RealRec>("real", n_, d_, b_, false,
                                              [&] { return build_real_rec(n_, d_, b_); }));
  });
}

void plan_cache_dir(const std::string &dir, bool clear_memory) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_plan_dir = dir;
  while (g_plan_dir.size() > 1 && g_plan_dir.back() == '/') g_plan_dir.pop_back();
  if (clear_memory) {
    // dropped plans stay alive: kernels already enqueued, or captured in
    // graphs, may still read their device tables
    for (auto &kv : g_c2c) g_retired.push_back(kv.second);
    for (auto &kv : g_real) g_retired.push_back(kv.second);
    g_c2c.clear();
    g_real.clear();
  }
}

void plan_cache_stats(uint64_t *hits, uint64_t *writes) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  *hits = g_disk_hits;
  *writes = g_disk_writes;
}

// Host-only round trip of a plan record through the file format (no GPU):
// builds the record with an explicit LDS budget, writes it to `path`, reads
// it back and compares every field.  0 when equal.
int plan_record_roundtrip(uint64_t n, bool dbl, size_t budget, bool inter, bool real,
                          const std::string &path) {
  auto same_d = [](const DirectRec &x, const DirectRec &y) {
    return std::memcmp(&x.desc, &y.desc, sizeof x.desc) == 0 && x.tw == y.tw && x.gen == y.gen;
  };
  std::string f;
  if (real) {
    const RealRec r = build_real_rec(n, dbl, budget);
    write_file(path, encode(r));
    RealRec q;
    if (!read_file(path, &f) || !decode(f, n, dbl, budget, &q)) return 1;
    return (q.fused == r.fused && q.half == r.half && same_d(q.d, r.d) && q.rtw == r.rtw) ? 0 : 2;
  }
  const PlanRec r = build_c2c_rec(n, dbl, budget, inter);
  write_file(path, encode(r));
  PlanRec q;
  if (!read_file(path, &f) || !decode(f, n, dbl, budget, inter, &q)) return 1;
  const bool eq = q.kind == r.kind && q.n1 == r.n1 && q.n2 == r.n2 && q.bn2 == r.bn2 &&
                  q.shift == r.shift && same_d(q.main, r.main) && same_d(q.pa, r.pa) &&
                  same_d(q.pb, r.pb) && q.big_lo == r.big_lo && q.big_hi == r.big_hi &&
                  q.bk == r.bk && q.bkf == r.bkf && q.sub_n == r.sub_n &&
                  q.sub_inter == r.sub_inter;
  return eq ? 0 : 2;
}

}  // namespace rfh
