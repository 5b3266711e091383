// lineops.hip -- data-movement and elementwise operations beside the
// transforms: per-line scaling for the FFTLog (fast Hankel) transforms, and
// the cyclic shift behind roll / fftshift / ifftshift.
//
// The FFTLog transform of a real line a (rocket_fft/overloads.py:858-902,
// 1755-1809) is rfft(a * w_in) * u, back through irfft, reversed, times
// w_out.  Here the spectrum step and the reversal are one pass over the half
// spectrum: for y = irfft(X), y[n-1-j] = irfft(conj(X_k) e^{+2 pi i k / n})[j],
// so the front-end hands this kernel c_k = conj(u_k) e^{2 pi i k / n} (or
// e^{2 pi i k / n} / u_k for the inverse) and it forms conj(X_k) c_k in
// place; the bias weights w_in / w_out are the real mode.
#include <vector>

#include "device.hpp"

namespace rfh {
namespace {
using namespace dev;

// mode 0: real lines, x[l][j] *= w[j]; mode 1: complex lines,
// X[l][k] = conj(X[l][k]) * c[k].  stride: elements between line starts.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_line_scale(T *x, int64_t nlines, int64_t len,
                                                    int64_t stride, const T *w) {
  const int64_t total = nlines * len;
  for (int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; f < total;
       f += int64_t(gridDim.x) * blockDim.x) {
    const int64_t l = f / len, j = f - l * len;
    if constexpr (MODE == 0) {
      x[l * stride + j] *= w[j];
    } else {
      C2<T> *X = reinterpret_cast<C2<T> *>(x) + l * stride + j;
      const C2<T> c = reinterpret_cast<const C2<T> *>(w)[j];
      *X = cmul(conjc(*X), c);
    }
  }
}

template <typename T>
void line_scale(void *data, int64_t nlines, int64_t len, int64_t stride, const double *w,
                int mode, hipStream_t s) {
  const int64_t nw = mode == 0 ? len : 2 * len;
  std::vector<T> wt(static_cast<size_t>(nw));
  for (int64_t i = 0; i < nw; ++i) wt[size_t(i)] = T(w[i]);
  T *dw = nullptr;
  HIP_CHECK(hipMalloc(&dw, size_t(nw) * sizeof(T)));
  struct Free {
    T *p;
    ~Free() { (void)hipFree(p); }
  } fr{dw};
  HIP_CHECK(hipMemcpyAsync(dw, wt.data(), size_t(nw) * sizeof(T), hipMemcpyHostToDevice, s));
  const int64_t total = nlines * len;
  const unsigned grid = unsigned(std::min<int64_t>((total + 255) / 256, 8192));
  if (mode == 0)
    hipLaunchKernelGGL((k_line_scale<T, 0>), dim3(grid), dim3(256), 0, s, static_cast<T *>(data),
                       nlines, len, stride, dw);
  else
    hipLaunchKernelGGL((k_line_scale<T, 1>), dim3(grid), dim3(256), 0, s, static_cast<T *>(data),
                       nlines, len, stride, dw);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s));  // dw is freed on return
}

// roll: out (C-contiguous) [i_0, ..., i_{d-1}] = in[(i_k - sh_k) mod n_k] with
// arbitrary input strides (elements); one element per thread, consecutive
// threads on consecutive outputs.
constexpr int kRollDims = 32;
struct RollGeom {
  int ndim;
  int64_t shape[kRollDims];
  int64_t istr[kRollDims];
  int64_t shift[kRollDims];  // in [0, n_k)
};

template <typename E>
__global__ void __launch_bounds__(256) k_roll(const E *in, E *out, int64_t total,
                                              const RollGeom g) {
  for (int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; f < total;
       f += int64_t(gridDim.x) * blockDim.x) {
    int64_t rem = f, src = 0;
    for (int d = g.ndim - 1; d >= 0; --d) {
      const int64_t n = g.shape[d];
      const int64_t q = rem / n;
      int64_t i = rem - q * n - g.shift[d];
      rem = q;
      if (i < 0) i += n;
      src += i * g.istr[d];
    }
    out[f] = in[src];
  }
}

struct alignas(16) E16 {
  uint64_t a, b;
};

template <typename E>
void roll_launch(const void *in, void *out, int64_t total, const RollGeom &g, hipStream_t s) {
  const unsigned grid = unsigned(std::min<int64_t>((total + 255) / 256, 16384));
  hipLaunchKernelGGL((k_roll<E>), dim3(grid), dim3(256), 0, s, static_cast<const E *>(in),
                     static_cast<E *>(out), total, g);
  HIP_CHECK(hipGetLastError());
}

bool on_device(const void *p, int *dev) {
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess ||
      (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)) {
    (void)hipGetLastError();
    return false;
  }
  *dev = at.device;
  return true;
}

}  // namespace
}  // namespace rfh

extern "C" int rfh_roll(const void *in, void *out, int ndim, const int64_t *shape,
                        const int64_t *in_strides, const int64_t *shifts, int64_t elem_bytes,
                        void *stream) {
  using namespace rfh;
  try {
    fork_guard();
    if (ndim < 0 || ndim > kRollDims) RFH_THROW(1, "rfh_roll: ndim out of range");
    RollGeom g{};
    g.ndim = ndim;
    int64_t total = 1;
    for (int d = 0; d < ndim; ++d) {
      if (shape[d] < 0) RFH_THROW(1, "rfh_roll: negative extent");
      g.shape[d] = shape[d];
      g.istr[d] = in_strides[d];
      total *= shape[d];
    }
    if (total == 0) return 0;
    for (int d = 0; d < ndim; ++d) {
      const int64_t n = shape[d];
      int64_t sh = shifts[d] % n;
      if (sh < 0) sh += n;
      g.shift[d] = sh;
    }
    int dev_in = 0, dev_out = 0;
    if (!in || !out || !on_device(in, &dev_in) || !on_device(out, &dev_out) || dev_in != dev_out)
      RFH_THROW(1, "rfh_roll: in and out must be device memory of one device");
    int prev = 0;
    HIP_CHECK(hipGetDevice(&prev));
    if (prev != dev_out) HIP_CHECK(hipSetDevice(dev_out));
    struct Restore {
      int d, cur;
      ~Restore() {
        if (d != cur) (void)hipSetDevice(d);
      }
    } rs{prev, dev_out};
    const hipStream_t s = static_cast<hipStream_t>(stream);
    switch (elem_bytes) {
      case 1: roll_launch<uint8_t>(in, out, total, g, s); break;
      case 2: roll_launch<uint16_t>(in, out, total, g, s); break;
      case 4: roll_launch<uint32_t>(in, out, total, g, s); break;
      case 8: roll_launch<uint64_t>(in, out, total, g, s); break;
      case 16: roll_launch<E16>(in, out, total, g, s); break;
      default: RFH_THROW(1, "rfh_roll: element size must be 1, 2, 4, 8 or 16 bytes");
    }
    return 0;
  } catch (const Error &e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::exception &e) {
    set_error(1, e.what());
    return 1;
  }
}

extern "C" int rfh_line_scale(void *data, int64_t nlines, int64_t len, int64_t stride,
                              const double *w, int mode, bool double_precision, void *stream) {
  using namespace rfh;
  try {
    fork_guard();
    if (mode != 0 && mode != 1) RFH_THROW(1, "rfh_line_scale: mode must be 0 or 1");
    if (nlines < 0 || len < 0 || (nlines > 1 && stride < len))
      RFH_THROW(1, "rfh_line_scale: invalid geometry");
    if (nlines == 0 || len == 0) return 0;
    if (!data || !w) RFH_THROW(1, "rfh_line_scale: null pointer");
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, data) != hipSuccess ||
        (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)) {
      (void)hipGetLastError();
      RFH_THROW(1, "rfh_line_scale: data must be device memory");
    }
    int prev = 0;
    HIP_CHECK(hipGetDevice(&prev));
    if (prev != at.device) HIP_CHECK(hipSetDevice(at.device));
    struct Restore {
      int d, cur;
      ~Restore() {
        if (d != cur) (void)hipSetDevice(d);
      }
    } rs{prev, at.device};
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (double_precision)
      line_scale<double>(data, nlines, len, stride, w, mode, s);
    else
      line_scale<float>(data, nlines, len, stride, w, mode, s);
    return 0;
  } catch (const Error &e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::exception &e) {
    set_error(1, e.what());
    return 1;
  }
}
