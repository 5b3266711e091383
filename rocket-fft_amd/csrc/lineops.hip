// lineops.hip -- elementwise per-line scaling used by the FFTLog (fast
// Hankel) transforms of the front-end.
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

}  // namespace
}  // namespace rfh

extern "C" int rfh_line_scale(void *data, int64_t nlines, int64_t len, int64_t stride,
                              const double *w, int mode, bool double_precision, void *stream) {
  using namespace rfh;
  try {
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
