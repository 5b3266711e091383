// device.hpp -- device-side building blocks shared by the line kernels:
// complex arithmetic, radix butterflies, Stockham stages, the kernel
// argument block and the load/store helpers.
#pragma once

#include "lines.hpp"
#include "rfh.hpp"

namespace rfh {
namespace dev {

// ------------------------------------------------------------ complex math

template <typename T>
struct alignas(2 * sizeof(T)) C2 {
  T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> operator+(C2<T> a, C2<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ C2<T> operator-(C2<T> a, C2<T> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ __forceinline__ C2<T> operator*(T s, C2<T> a) {
  return {s * a.x, s * a.y};
}
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ C2<T> conjc(C2<T> a) {
  return {a.x, -a.y};
}
// a * (dir * i)
template <typename T>
__device__ __forceinline__ C2<T> rot(C2<T> a, T dir) {
  return {-dir * a.y, dir * a.x};
}
// w, conjugated when dir > 0 (tables hold forward-sign roots)
template <typename T>
__device__ __forceinline__ C2<T> dirw(C2<T> w, T dir) {
  return {w.x, dir < T(0) ? w.y : -w.y};
}

// Tile index of this workgroup with XCD-contiguous placement.  The
// dispatcher deals workgroups round-robin to the 8 XCDs (workgroup i runs on
// XCD i mod 8); renumbering them so that each XCD owns one contiguous run of
// tiles keeps the cache lines that neighbouring tiles share inside one L2.
__device__ __forceinline__ unsigned xcd_tile() {
  constexpr unsigned kXcd = 8;
  const unsigned g = gridDim.x, i = blockIdx.x;
  if (g < 2 * kXcd) return i;
  const unsigned x = i % kXcd, k = i / kXcd, q = g / kXcd, r = g % kXcd;
  return x * q + (x < r ? x : r) + k;
}

// q = u / d for 0 <= u < 2^22 via a float reciprocal and one correction.
__device__ __forceinline__ int fdiv(int u, int d, float inv) {
  int q = int(float(u) * inv);
  const int r = u - q * d;
  q += (r >= d) - (r < 0);
  return q;
}

// -------------------------------------------------------------- butterflies
// dir = -1: X_k = sum_m v_m e^{-2 pi i m k / R}; dir = +1 the conjugate kernel.

template <typename T>
__device__ __forceinline__ void dft2(C2<T> *v) {
  const C2<T> a = v[0];
  v[0] = a + v[1];
  v[1] = a - v[1];
}

template <typename T>
__device__ __forceinline__ void dft3(C2<T> *v, T dir) {
  const T s = T(0.866025403784438646763723170752936183);
  const C2<T> t = v[1] + v[2], d = v[1] - v[2];
  const C2<T> a = v[0] + T(-0.5) * t;
  v[0] = v[0] + t;
  const C2<T> b = rot(s * d, dir);
  v[1] = a + b;
  v[2] = a - b;
}

// Radix 9 as 3 x 3: DFT3 over x[n2], x[n2+3], x[n2+6] for each n2, twiddle
// W9^(n2 k1), DFT3 over n2; X[k1 + 3 k2].  W9^m = cos(2 pi m/9) + dir i sin(2 pi m/9).
template <typename T>
__device__ __forceinline__ void dft9(C2<T> *v, T dir) {
  const T c1 = T(0.766044443118978035202392650555416674),
          s1 = T(0.642787609686539326322643409907263433),
          c2 = T(0.173648177666930348851716626769314796),
          s2 = T(0.984807753012208059366743024589523014),
          c4 = T(-0.93969262078590838405410927732473147),
          s4 = T(0.342020143325668733044099614682259581);
  C2<T> a[3] = {v[0], v[3], v[6]}, b[3] = {v[1], v[4], v[7]}, c[3] = {v[2], v[5], v[8]};
  dft3(a, dir);
  dft3(b, dir);
  dft3(c, dir);
  b[1] = cmul(b[1], C2<T>{c1, dir * s1});
  b[2] = cmul(b[2], C2<T>{c2, dir * s2});
  c[1] = cmul(c[1], C2<T>{c2, dir * s2});
  c[2] = cmul(c[2], C2<T>{c4, dir * s4});
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    C2<T> t[3] = {a[k1], b[k1], c[k1]};
    dft3(t, dir);
    v[k1] = t[0];
    v[k1 + 3] = t[1];
    v[k1 + 6] = t[2];
  }
}

template <typename T>
__device__ __forceinline__ void dft4(C2<T> *v, T dir) {
  const C2<T> t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3];
  const C2<T> t3 = rot(v[1] - v[3], dir);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + t3;
  v[3] = t1 - t3;
}

template <typename T>
__device__ __forceinline__ void dft5(C2<T> *v, T dir) {
  const T c1 = T(0.309016994374947424102293417182819059);
  const T c2 = T(-0.809016994374947424102293417182819059);
  const T s1 = T(0.951056516295153572116439333379382143);
  const T s2 = T(0.587785252292473129168705954639072769);
  const C2<T> t1 = v[1] + v[4], t2 = v[2] + v[3], d1 = v[1] - v[4], d2 = v[2] - v[3];
  const C2<T> a1 = v[0] + c1 * t1 + c2 * t2, a2 = v[0] + c2 * t1 + c1 * t2;
  const C2<T> b1 = rot(s1 * d1 + s2 * d2, dir), b2 = rot(s2 * d1 - s1 * d2, dir);
  v[0] = v[0] + t1 + t2;
  v[1] = a1 + b1;
  v[4] = a1 - b1;
  v[2] = a2 + b2;
  v[3] = a2 - b2;
}

template <typename T>
__device__ __forceinline__ void dft8(C2<T> *v, T dir) {
  C2<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e, dir);
  dft4(o, dir);
  const T s = T(0.707106781186547524400844362104849039);
  const C2<T> o1 = {s * (o[1].x - dir * o[1].y), s * (o[1].y + dir * o[1].x)};
  const C2<T> o2 = rot(o[2], dir);
  const C2<T> o3 = {s * (-o[3].x - dir * o[3].y), s * (-o[3].y + dir * o[3].x)};
  v[0] = e[0] + o[0];
  v[4] = e[0] - o[0];
  v[1] = e[1] + o1;
  v[5] = e[1] - o1;
  v[2] = e[2] + o2;
  v[6] = e[2] - o2;
  v[3] = e[3] + o3;
  v[7] = e[3] - o3;
}

template <typename T>
__device__ __forceinline__ void dft16(C2<T> *v, T dir) {
  C2<T> e[8], o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  dft8(e, dir);
  dft8(o, dir);
  const T c1 = T(0.923879532511286756128183189396788933);
  const T s1 = T(0.382683432365089771728459984030398866);
  const T r = T(0.707106781186547524400844362104849039);
  const T cs[8] = {T(1), c1, r, s1, T(0), -s1, -r, -c1};
  const T sn[8] = {T(0), s1, r, c1, T(1), c1, r, s1};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const C2<T> w = {cs[k], dir * sn[k]};
    const C2<T> t = cmul(o[k], w);
    v[k] = e[k] + t;
    v[k + 8] = e[k] - t;
  }
}

// Odd prime P through symmetric pairs; g[q] = (cos 2 pi q/P, sin 2 pi q/P).
template <int P, typename T>
__device__ __forceinline__ void dftp(C2<T> *v, T dir, const C2<T> *__restrict__ g) {
  constexpr int H = (P - 1) / 2;
  C2<T> t[H], d[H];
  const C2<T> x0 = v[0];
  C2<T> sum = v[0];
#pragma unroll
  for (int m = 1; m <= H; ++m) {
    t[m - 1] = v[m] + v[P - m];
    d[m - 1] = v[m] - v[P - m];
    sum = sum + t[m - 1];
  }
  v[0] = sum;  // the inputs live on in x0, t and d
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    C2<T> a = x0, b = {T(0), T(0)};
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      const C2<T> w = g[(m * k) % P];
      a = a + w.x * t[m - 1];
      b = b + w.y * d[m - 1];
    }
    const C2<T> rb = rot(b, dir);
    v[k] = a + rb;
    v[P - k] = a - rb;
  }
}

template <int R, typename T>
__device__ __forceinline__ void dft(C2<T> *v, T dir, const C2<T> *g) {
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 3) dft3(v, dir);
  else if constexpr (R == 4) dft4(v, dir);
  else if constexpr (R == 5) dft5(v, dir);
  else if constexpr (R == 8) dft8(v, dir);
  else if constexpr (R == 9) dft9(v, dir);
  else if constexpr (R == 16) dft16(v, dir);
  else dftp<R>(v, dir, g);
}

// ------------------------------------------------------------------ stages
// Stockham stage (radix R, ns = product of earlier radices): butterfly j of
// N/R reads src[j + r N/R], multiplies by W_{ns R}^{(j mod ns) r}, and writes
// dst[(j - k) R + k + r ns] with k = j mod ns.  Output is in natural order
// after the last stage.

template <int R, typename T>
__device__ __forceinline__ void stage_fixed(const C2<T> *__restrict__ src,
                                            C2<T> *__restrict__ dst, int L, int ld, int N,
                                            int ns, const C2<T> *__restrict__ tw,
                                            const C2<T> *__restrict__ g, T dir) {
  const int NR = N / R;
  const int total = L * NR;
  const float invNR = 1.0f / float(NR), invns = 1.0f / float(ns);
  for (int u = threadIdx.x; u < total; u += blockDim.x) {
    const int l = fdiv(u, NR, invNR);
    const int j = u - l * NR;
    const int q = fdiv(j, ns, invns);
    const int k = j - q * ns;
    const C2<T> *s = src + l * ld + j;
    C2<T> *d = dst + l * ld + q * ns * R + k;
    C2<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[r * NR];
    if (ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], dirw(tw[(r - 1) * ns + k], dir));
    }
    dft<R>(v, dir, g);
#pragma unroll
    for (int r = 0; r < R; ++r) d[r * ns] = v[r];
  }
}

// Any odd radix at run time: inputs are twiddled in place in src, then each
// output pair (k, R - k) of each butterfly is accumulated straight from LDS by
// its own thread, from the symmetric sums x_m + x_{R-m} and x_m - x_{R-m}.
// O(R) work per element, spread over (R + 1) / 2 threads per butterfly so
// that a line with few butterflies still fills the workgroup; used for
// primes above 13.
template <typename T>
__device__ void stage_generic(C2<T> *__restrict__ src, C2<T> *__restrict__ dst, int L,
                              int ld, int N, int R, int ns, const C2<T> *__restrict__ tw,
                              const C2<T> *__restrict__ g, T dir) {
  const int NR = N / R;
  const int H = (R - 1) / 2;
  const float invNR = 1.0f / float(NR), invns = 1.0f / float(ns);
  if (ns > 1) {  // stage twiddles, in place, one (butterfly, input) pair per item
    const int tot = L * NR * (R - 1);
    const float inv = 1.0f / float(L * NR);
    for (int u = threadIdx.x; u < tot; u += blockDim.x) {
      const int r1 = fdiv(u, L * NR, inv);  // r - 1
      const int lj = u - r1 * (L * NR);
      const int l = fdiv(lj, NR, invNR);
      const int j = lj - l * NR;
      const int k = j - fdiv(j, ns, invns) * ns;
      C2<T> *p = src + l * ld + j + (r1 + 1) * NR;
      *p = cmul(*p, dirw(tw[r1 * ns + k], dir));
    }
    __syncthreads();
  }
  // one (butterfly, output pair kk / R - kk) per item, butterflies fastest so
  // that a wave shares kk and reads the same root g[kk m mod R]
  const int total = L * NR * (H + 1);
  const float invH = 1.0f / float(H + 1);
  for (int u = threadIdx.x; u < total; u += blockDim.x) {
    const int lk = fdiv(u, NR, invNR);
    const int j = u - lk * NR;
    const int l = fdiv(lk, H + 1, invH);
    const int kk = lk - l * (H + 1);
    const int q = fdiv(j, ns, invns);
    const int k = j - q * ns;
    const C2<T> *s = src + l * ld + j;
    C2<T> *d = dst + l * ld + q * ns * R + k;
    const C2<T> x0 = s[0];
    if (kk == 0) {
      C2<T> sum = x0;
      for (int r = 1; r < R; ++r) sum = sum + s[r * NR];
      d[0] = sum;
      continue;
    }
    C2<T> a = x0, b = {T(0), T(0)};
    int idx = 0;
    for (int m = 1; m <= H; ++m) {
      idx += kk;
      if (idx >= R) idx -= R;
      const C2<T> vm = s[m * NR], vp = s[(R - m) * NR];
      const C2<T> w = g[idx];
      a = a + w.x * (vm + vp);
      b = b + w.y * (vm - vp);
    }
    const C2<T> rb = rot(b, dir);
    d[kk * ns] = a + rb;
    d[(R - kk) * ns] = a - rb;
  }
}

// ----------------------------------------------------------- line kernel

enum LineOp {
  OP_C2C = 0,     // plain complex lines
  OP_PASSA = 1,   // four-step pass A (columns), inter-pass twiddle on store
  OP_PASSB = 2,   // four-step pass B (rows)
  OP_BLUE = 3,    // Bluestein entirely in LDS
  OP_R2C_EVEN = 4,
  OP_R2C_ODD = 5,
  OP_C2R_EVEN = 6,
  OP_C2R_ODD = 7,
  OP_R2R = 8,     // real line -> pre_elem -> FFT -> post_elem (DCT/DST, Hartley, FFTPACK)
  // elementwise passes of the multi-pass Bluestein form (no FFT; count = other)
  OP_BLUE_IN = 11,   // work[m] = x[m] * conj(chirp[m]) for m < n, zero up to bn2
  OP_BLUE_MUL = 12,  // work[m] *= kernel[m]
  OP_BLUE_OUT = 13,  // y[m] = fct * work[m] * conj(chirp[m]) for m < n
  // Bluestein rows of the three-pass form: forward FFT of the row, times the
  // kernel t2 at m = aux + k * other, inverse FFT, times the inverse four-step
  // twiddle (tables t0 / t1); written back in place
  OP_BLUEROW = 14,
};
// OP_R2R packs its modes into LineArgs::sub
constexpr int R2R_ORTHO = 1 << 16, R2R_CONJ = 1 << 17, R2R_FLIP = 1 << 18;
enum SubFlags {
  SUB_CHIRP = 1,   // pass A: multiply loads by the chirp, zero beyond n
  SUB_KERNEL = 2,  // pass B: multiply by the Bluestein kernel (natural order)
  SUB_FINAL = 4,   // pass A or B: multiply by the chirp and fct, keep m = aux + k other < n
  SUB_CONJ = 8,    // Bluestein backward / r2c conj output / c2r conj input
};

template <typename T>
struct LineArgs {
  int64_t nlines;
  int nb;
  int64_t bdim[kMaxBatchDims], bsi[kMaxBatchDims], bso[kMaxBatchDims];
  int64_t esi, eso;
  const T *in;
  T *out;
  int L, log2L, ld, nin, nout, lfast_in, lfast_out;
  int auxdim;    // batch dim whose index is a line's aux (four-step row / column)
  int twlds_lo, twlds_hi;  // pass A of the specialised kernel: four-step twiddle
                           // tables (t0 / t1 entries) copied to LDS; 0 = read globally
  FftDesc f;
  const C2<T> *tw, *gen;
  T dir, fct;
  int sub;
  int64_t bign;  // n of the transform (Bluestein / real), N of a four-step
  int64_t other; // four-step: N2 in pass A, N1 in pass B
  int shift;
  const C2<T> *t0, *t1, *t2;  // op tables
  int rot_in, rot_out;  // element-index rotations of the line set (0: none)
};

// (j + r) mod n for 0 <= j, r < n
__device__ __forceinline__ int rot_index(int j, int r, int n) {
  const int q = j + r;
  return q >= n ? q - n : q;
}

// Bluestein column passes (chirped loads of pass A, final stores of pass C):
// element j of column aux is element m = j other + aux of the transform
// line; with the line rotated by rot it is read from / written to element
// (m + rot) mod n, which lies (that - aux) line elements (es / other scalars
// each) from the column's start.  Offset in scalars from the column start.
__device__ __forceinline__ int64_t blue_col_offset(int j, int64_t m, int rot, int64_t n,
                                                   int aux, int64_t es, int64_t other) {
  if (!rot) return int64_t(j) * es;
  int64_t q = m + rot;
  if (q >= n) q -= n;
  return (q - aux) * (es / other);
}

template <bool V, typename T>
__device__ __forceinline__ C2<T> ldc(const T *p) {
  if constexpr (V) return *reinterpret_cast<const C2<T> *>(p);
  else return C2<T>{p[0], p[1]};
}
template <bool V, typename T>
__device__ __forceinline__ void stc(T *p, C2<T> v) {
  if constexpr (V) *reinterpret_cast<C2<T> *>(p) = v;
  else {
    p[0] = v.x;
    p[1] = v.y;
  }
}

// visit (l, j) for l < nl, j < cnt; l-fastest when lines interleave in memory
template <typename F>
__device__ __forceinline__ void for_tile(int nl, int cnt, int lfast, int log2L, F &&f) {
  if (lfast) {
    const int Lm = (1 << log2L) - 1;
    const int total = cnt << log2L;
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
      const int l = u & Lm, j = u >> log2L;
      if (l < nl) f(l, j);
    }
  } else {
    const int total = nl * cnt;
    const float inv = 1.0f / float(cnt);
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
      const int l = fdiv(u, cnt, inv);
      f(l, u - l * cnt);
    }
  }
}

template <typename T>
__device__ __forceinline__ C2<T> big_twiddle(const LineArgs<T> &a, int64_t e) {
  const int64_t mask = (int64_t(1) << a.shift) - 1;
  return cmul(a.t0[e & mask], a.t1[e >> a.shift]);
}

// e^{i pi num/den}: integer argument reduction, then sincospi in the
// working precision (FP64 for complex128, FP32 for complex64).
template <typename T>
__device__ __forceinline__ C2<T> expipi(int64_t num, int64_t den) {
  const int64_t per = 2 * den;
  int64_t r = num % per;
  if (r < 0) r += per;
  if constexpr (sizeof(T) == 8) {
    double s, c;
    sincospi(double(r) / double(den), &s, &c);
    return C2<T>{T(c), T(s)};
  } else {
    float s, c;
    sincospif(float(double(r) / double(den)), &s, &c);
    return C2<T>{c, s};
  }
}

// Element j of the complex work line built from input line x (element
// stride es) for the real-transform modes of PreMode; m = work length.
template <typename T>
__device__ __forceinline__ C2<T> pre_elem(int mode, const T *x, int64_t es, int64_t n,
                                          int64_t m, int64_t j, bool ortho, bool conj,
                                          bool flip) {
  const T sqrt2 = T(1.41421356237309504880168872420969808);
  C2<T> v = {T(0), T(0)};
  switch (mode) {
    case kPreRealAsComplex:
      if (j < n) {
        T xv = x[j * es];
        if (flip && j >= 2 && (j % 2) == 0) xv = -xv;
        v.x = xv;
      }
      break;
    case kPreHermitian: {  // half complex input X[0..n/2]
      const int64_t h = n / 2;
      if (j == 0) {
        v = C2<T>{x[0], T(0)};
      } else if (j < n - h || (2 * j < n)) {
        v = C2<T>{x[j * es], x[j * es + 1]};
        if (conj) v.y = -v.y;
      } else if (2 * j == n) {
        v = C2<T>{x[j * es], T(0)};
      } else {
        const int64_t k = n - j;
        v = C2<T>{x[k * es], x[k * es + 1]};
        if (!conj) v.y = -v.y;
      }
      break;
    }
    case kPreDct1: {
      const int64_t src = j < n ? j : m - j;
      T xv = x[src * es];
      if (ortho && (src == 0 || src == n - 1)) xv *= sqrt2;
      v.x = xv;
      break;
    }
    case kPreDst1:
      if (j >= 1 && j <= n) v.x = x[(j - 1) * es];
      else if (j >= n + 2) v.x = -x[(m - 1 - j) * es];
      break;
    case kPreType2:
      if (j < n) v.x = x[j * es];
      break;
    case kPreDct3:
      if (j < n) {
        T xv = x[j * es];
        if (j == 0 && ortho) xv *= sqrt2;
        const T w = j == 0 ? T(1) : T(2);
        v = (w * xv) * expipi<T>(j, 2 * n);
      }
      break;
    case kPreDst3:
      if (j >= 1 && j <= n) {
        T xv = x[(j - 1) * es];
        if (j == 1 && ortho) xv *= sqrt2;
        const T w = j == n ? T(1) : T(2);
        v = (w * xv) * expipi<T>(j, 2 * n);
      }
      break;
    case kPreType4:
      if (j < n) v = x[j * es] * expipi<T>(-j, 2 * n);
      break;
    case kPreHalfcomplex: {
      if (j == 0) {
        v.x = x[0];
      } else if (2 * j < n) {
        v = C2<T>{x[(2 * j - 1) * es], x[(2 * j) * es]};
        if (conj) v.y = -v.y;
      } else if (2 * j == n) {
        v.x = x[(n - 1) * es];
      } else {
        const int64_t k = n - j;
        v = C2<T>{x[(2 * k - 1) * es], -x[(2 * k) * es]};
        if (conj) v.y = -v.y;
      }
      break;
    }
    case kPreRealPairs:
      if (2 * j < n) v = C2<T>{x[2 * j * es], x[(2 * j + 1) * es]};
      break;
    case kPreDct2N:
    case kPreDst2N: {
      // v_j = x_2j for 2j < n, v_{n-1-j} = x_2j+1: V = FFT_n(v) gives
      // DCT-II_k = 2 Re(e^{-i pi k/2n} V_k); DST-II is DCT-II of (-1)^j x_j
      // read backwards
      const int64_t i = 2 * j < n ? 2 * j : 2 * (n - 1 - j) + 1;
      T xv = x[i * es];
      if (mode == kPreDst2N && (i & 1)) xv = -xv;
      v.x = xv;
      break;
    }
    case kPreDct3N:
    case kPreDst3N: {
      // a = x (DCT-III) or x reversed (DST-III), a_n = 0, ortho scales x_0 by
      // sqrt 2; V_j = (a_j - i a_{n-j}) e^{i pi j/2n} and the inverse FFT of
      // length n holds the even outputs up and the odd outputs down
      auto at = [&](int64_t q) -> T {
        if (q >= n) return T(0);
        const int64_t src = mode == kPreDct3N ? q : n - 1 - q;
        T xv = x[src * es];
        if (ortho && src == 0) xv *= sqrt2;
        return xv;
      };
      v = cmul(C2<T>{at(j), -at(n - j)}, expipi<T>(j, 2 * n));
      break;
    }
    case kPreHalfSpectrum: {
      // z_j = Σ_k [A_k + i e^{2πik/n} D_k] e^{2πikj/(n/2)} with A = X_k + conj X_{h-k},
      // D = X_k - conj X_{h-k}: the inverse FFT of length h = n/2 gives
      // y_2j + i y_2j+1.  Imaginary parts of bins 0 and h are ignored.
      const int64_t h = n / 2;
      C2<T> xj = {x[j * es], x[j * es + 1]}, xm = {x[(h - j) * es], x[(h - j) * es + 1]};
      if (conj) {
        xj.y = -xj.y;
        xm.y = -xm.y;
      }
      if (j == 0) {
        xj.y = T(0);
        xm.y = T(0);
      }
      const C2<T> A = xj + conjc(xm);
      const C2<T> B = cmul(xj - conjc(xm), expipi<T>(2 * j, n));
      v = C2<T>{A.x - B.y, A.y + B.x};
      break;
    }
  }
  return v;
}

// r2c bin k (0..n/2) of a real line of even length n from the length-n/2
// spectrum Z of its pairs z_j = x_2j + i x_2j+1 (get(q) = Z_q):
// X_k = (Z_k + conj Z_{h-k}) / 2 + e^{-2πik/n} (Z_k - conj Z_{h-k}) / 2i.
template <typename T, typename G>
__device__ __forceinline__ C2<T> half_spectrum_bin(G &&get, int64_t n, int64_t k) {
  const int64_t h = n / 2;
  if (k == 0 || k == h) {
    const C2<T> z = get(0);
    return C2<T>{k == 0 ? z.x + z.y : z.x - z.y, T(0)};
  }
  const C2<T> zk = get(k), zm = conjc(get(h - k));
  const C2<T> fe = T(0.5) * (zk + zm);
  const C2<T> dd = zk - zm;
  const C2<T> fo = C2<T>{T(0.5) * dd.y, T(-0.5) * dd.x};
  return fe + cmul(expipi<T>(-2 * k, n), fo);
}

// Real output element k from the transformed work line (get(q) = bin q)
// for the PostMode modes other than kPostComplexHalf; fct not applied.
template <typename T, typename G>
__device__ __forceinline__ T post_elem(int mode, G &&get, int64_t n, int64_t k, bool ortho,
                                       bool flip) {
  const T isqrt2 = T(0.707106781186547524400844362104849039);
  T r = T(0);
  switch (mode) {
    case kPostReal:
      r = get(k).x;
      if (flip && k >= 2 && (k % 2) == 0) r = -r;
      break;
    case kPostDct1:
      r = get(k).x;
      if (ortho && (k == 0 || k == n - 1)) r *= isqrt2;
      break;
    case kPostDst1:
      r = -get(k + 1).y;
      break;
    case kPostDct2: {
      const C2<T> z = cmul(get(k), expipi<T>(-k, 2 * n));
      r = T(2) * z.x;
      if (ortho && k == 0) r *= isqrt2;
      break;
    }
    case kPostDst2: {
      const C2<T> z = cmul(get(k + 1), expipi<T>(-(k + 1), 2 * n));
      r = -T(2) * z.y;
      if (ortho && k == 0) r *= isqrt2;
      break;
    }
    case kPostDct3:
      r = get(k).x;
      break;
    case kPostDst3:
      r = get(k).y;
      break;
    case kPostDct4:
    case kPostDst4: {
      const C2<T> z = cmul(get(k), expipi<T>(-(2 * k + 1), 4 * n));
      r = mode == kPostDct4 ? T(2) * z.x : -T(2) * z.y;
      break;
    }
    case kPostHartley:
      r = get(k).x + get(k).y;
      break;
    case kPostPackHalfcomplex:  // (the forward sign rule flips inputs only)
      if (k == 0) r = get(0).x;
      else if (k % 2) r = get((k + 1) / 2).x;
      else r = get(k / 2).y;
      break;
    case kPostRealPairs:
      r = (k % 2) ? get(k / 2).y : get(k / 2).x;
      break;
    case kPostDct2N:
    case kPostDst2N: {
      const int64_t q = mode == kPostDct2N ? k : n - 1 - k;
      const C2<T> z = cmul(get(q), expipi<T>(-q, 2 * n));
      r = T(2) * z.x;
      if (ortho && k == 0) r *= isqrt2;
      break;
    }
    case kPostDct3N:
    case kPostDst3N:
      r = get((k % 2 == 0) ? k / 2 : n - 1 - (k - 1) / 2).x;
      if (mode == kPostDst3N && (k & 1)) r = -r;
      break;
  }
  return r;
}

// Stage 0 of a line group fed straight from global memory: butterfly u of
// the group takes inputs j + r NR of line l through `load` (the op's
// pre-processing), and writes its radix-R outputs to dst (ns = 1: no
// twiddles, output position j R + r).  Missing lines (l >= nl) get zeros.
template <int R, typename T, typename Ld>
__device__ __forceinline__ void stage0_in(Ld &&load, C2<T> *__restrict__ dst, int L, int nl,
                                          int ld, int N, int lfast, int log2L,
                                          const C2<T> *__restrict__ g, T dir) {
  const int NR = N / R;
  const int total = L * NR;
  const float invNR = 1.0f / float(NR);
  for (int u = threadIdx.x; u < total; u += blockDim.x) {
    int l, j;
    if (lfast) {
      l = u & ((1 << log2L) - 1);
      j = u >> log2L;
    } else {
      l = fdiv(u, NR, invNR);
      j = u - l * NR;
    }
    C2<T> v[R];
    if (l < nl) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = load(l, j + r * NR);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = C2<T>{T(0), T(0)};
    }
    dft<R>(v, dir, g);
    C2<T> *d = dst + l * ld + j * R;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = v[r];
  }
}

// Last stage of a line group written straight to global memory: butterfly
// j (< ns, NR = ns) of line l reads src[j + r ns], twiddles by W^{j r},
// and hands natural-order bin j + r ns to `store` (the op's post-processing).
template <int R, typename T, typename St>
__device__ __forceinline__ void stage_last_out(St &&store, const C2<T> *__restrict__ src,
                                               int L, int nl, int ld, int ns, int lfast,
                                               int log2L, const C2<T> *__restrict__ tw,
                                               const C2<T> *__restrict__ g, T dir) {
  const int total = L * ns;
  const float invns = 1.0f / float(ns);
  for (int u = threadIdx.x; u < total; u += blockDim.x) {
    int l, j;
    if (lfast) {
      l = u & ((1 << log2L) - 1);
      j = u >> log2L;
    } else {
      l = fdiv(u, ns, invns);
      j = u - l * ns;
    }
    if (l >= nl) continue;
    const C2<T> *s = src + l * ld + j;
    C2<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[r * ns];
    if (ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], dirw(tw[(r - 1) * ns + j], dir));
    }
    dft<R>(v, dir, g);
#pragma unroll
    for (int r = 0; r < R; ++r) store(l, j + r * ns, v[r]);
  }
}

// Radices whose first / last stage may be fused with global memory: the
// hand-written butterflies (the symmetric-pair ones would raise the register
// count of every line kernel)
__device__ __forceinline__ bool fusable_radix(int r) {
  return r == 2 || r == 3 || r == 4 || r == 5 || r == 8 || r == 9 || r == 16;
}

template <typename T, typename Ld>
__device__ __forceinline__ void stage0_in_any(int R, Ld &&load, C2<T> *dst, int L, int nl,
                                              int ld, int N, int lfast, int log2L,
                                              const C2<T> *g, T dir) {
  switch (R) {
    case 2: stage0_in<2>(load, dst, L, nl, ld, N, lfast, log2L, g, dir); break;
    case 3: stage0_in<3>(load, dst, L, nl, ld, N, lfast, log2L, g, dir); break;
    case 4: stage0_in<4>(load, dst, L, nl, ld, N, lfast, log2L, g, dir); break;
    case 5: stage0_in<5>(load, dst, L, nl, ld, N, lfast, log2L, g, dir); break;
    case 8: stage0_in<8>(load, dst, L, nl, ld, N, lfast, log2L, g, dir); break;
    case 9: stage0_in<9>(load, dst, L, nl, ld, N, lfast, log2L, g, dir); break;
    default: stage0_in<16>(load, dst, L, nl, ld, N, lfast, log2L, g, dir); break;
  }
}

template <typename T, typename St>
__device__ __forceinline__ void stage_last_out_any(int R, St &&store, const C2<T> *src, int L,
                                                   int nl, int ld, int ns, int lfast,
                                                   int log2L, const C2<T> *tw,
                                                   const C2<T> *g, T dir) {
  switch (R) {
    case 2: stage_last_out<2>(store, src, L, nl, ld, ns, lfast, log2L, tw, g, dir); break;
    case 3: stage_last_out<3>(store, src, L, nl, ld, ns, lfast, log2L, tw, g, dir); break;
    case 4: stage_last_out<4>(store, src, L, nl, ld, ns, lfast, log2L, tw, g, dir); break;
    case 5: stage_last_out<5>(store, src, L, nl, ld, ns, lfast, log2L, tw, g, dir); break;
    case 8: stage_last_out<8>(store, src, L, nl, ld, ns, lfast, log2L, tw, g, dir); break;
    case 9: stage_last_out<9>(store, src, L, nl, ld, ns, lfast, log2L, tw, g, dir); break;
    default: stage_last_out<16>(store, src, L, nl, ld, ns, lfast, log2L, tw, g, dir); break;
  }
}

// Stages s0 .. s1-1 of f out of LDS (ping-pong a -> b); returns the buffer
// holding the result.
template <typename T>
__device__ C2<T> *run_stages(const FftDesc &f, int s0, int s1, C2<T> *a, C2<T> *b, int L,
                             int ld, const C2<T> *__restrict__ tw,
                             const C2<T> *__restrict__ gen, T dir) {
  for (int s = s0; s < s1; ++s) {
    const Stage st = f.st[s];
    const C2<T> *t = tw + st.twoff;
    const C2<T> *g = gen + st.genoff;
    switch (st.radix) {
      case 2: stage_fixed<2>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 3: stage_fixed<3>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 4: stage_fixed<4>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 5: stage_fixed<5>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 7: stage_fixed<7>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 8: stage_fixed<8>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 9: stage_fixed<9>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 11: stage_fixed<11>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 13: stage_fixed<13>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 17: stage_fixed<17>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      case 16: stage_fixed<16>(a, b, L, ld, f.n, st.ns, t, g, dir); break;
      default: stage_generic(a, b, L, ld, f.n, st.radix, st.ns, t, g, dir); break;
    }
    __syncthreads();
    C2<T> *x = a;
    a = b;
    b = x;
  }
  return a;
}

// One tile (a.L lines starting at line tile * a.L) of a line set: load with
// the op's pre-processing, all Stockham stages out of LDS (two buffers in
// smem_raw), store with the op's post-processing.  When the first (last)
// stage has a fixed radix it reads (writes) global memory directly instead
// of going through LDS; ops whose store needs two bins at once (r2c of even
// length, the real-to-real modes) and the Bluestein ops keep the LDS store.
// s_oi / s_oo / s_aux are a.L-entry LDS scratch for the per-line offsets.
template <typename T, int OP, bool VIN, bool VOUT>
__device__ __forceinline__ void line_body(const LineArgs<T> &a, int64_t tile,
                                          unsigned char *smem_raw, int64_t *s_oi,
                                          int64_t *s_oo, int *s_aux) {
  C2<T> *bufA = reinterpret_cast<C2<T> *>(smem_raw);
  C2<T> *bufB = bufA + a.L * a.ld;
  const int64_t b0 = tile * a.L;
  const int nl = int(min(int64_t(a.L), a.nlines - b0));
  // offsets of the group's lines; the slots of missing lines (the last
  // group of a line set) hold zeros and are never used to address memory
  for (int l = threadIdx.x; l < a.L; l += blockDim.x) {
    const LineOffsets o = group_slot(b0, l, a.L, a.nlines, a.nb, a.bdim, a.bsi, a.bso, a.auxdim);
    s_oi[l] = o.oi;
    s_oo[l] = o.oo;
    s_aux[l] = o.aux;
  }
  __syncthreads();

  const int ld = a.ld;
  const bool conjf = (a.sub & SUB_CONJ) != 0;
  // ---------------------------------------------------- per-element load
  auto load = [&](int l, int j) -> C2<T> {
    const T *base = a.in + s_oi[l];
    C2<T> v;
    if constexpr (OP == OP_C2C) {
      v = ldc<VIN>(base + rot_index(j, a.rot_in, a.nin) * a.esi);
    } else if constexpr (OP == OP_PASSB || OP == OP_BLUEROW) {
      v = ldc<VIN>(base + j * a.esi);
    } else if constexpr (OP == OP_PASSA) {
      if (a.sub & SUB_CHIRP) {
        const int64_t m = int64_t(j) * a.other + s_aux[l];
        if (m < a.bign) {
          const C2<T> c = a.t2[m];
          v = cmul(ldc<VIN>(base + blue_col_offset(j, m, a.rot_in, a.bign, s_aux[l], a.esi,
                                                   a.other)),
                   conjf ? c : conjc(c));
        } else {
          v = C2<T>{T(0), T(0)};
        }
      } else {
        v = ldc<VIN>(base + rot_index(j, a.rot_in, a.nin) * a.esi);
      }
    } else if constexpr (OP == OP_BLUE) {
      if (j < a.bign) {  // element (j + rot_in) mod n, times the chirp at j
        const C2<T> c = a.t0[j];
        v = cmul(ldc<VIN>(base + rot_index(j, a.rot_in, int(a.bign)) * a.esi),
                 conjf ? c : conjc(c));
      } else {
        v = C2<T>{T(0), T(0)};
      }
    } else if constexpr (OP == OP_R2C_EVEN) {
      if constexpr (VIN)  // contiguous real line: one 2-wide load
        v = ldc<true>(base + 2 * j);
      else
        v = C2<T>{base[2 * j * a.esi], base[(2 * j + 1) * a.esi]};
    } else if constexpr (OP == OP_R2C_ODD) {
      v = C2<T>{base[j * a.esi], T(0)};
    } else if constexpr (OP == OP_C2R_EVEN) {
      const int h = a.f.n;
      C2<T> xj = ldc<VIN>(base + j * a.esi), xm = ldc<VIN>(base + (h - j) * a.esi);
      if (conjf) {
        xj.y = -xj.y;
        xm.y = -xm.y;
      }
      if (j == 0) {
        xj.y = T(0);
        xm.y = T(0);
      }
      const C2<T> A = xj + conjc(xm);
      const C2<T> B = cmul(xj - conjc(xm), conjc(a.t0[j]));
      v = C2<T>{A.x - B.y, A.y + B.x};
    } else if constexpr (OP == OP_R2R) {
      v = pre_elem<T>(a.sub & 0xff, base, a.esi, a.bign, a.f.n, j, (a.sub & R2R_ORTHO) != 0,
                      (a.sub & R2R_CONJ) != 0, (a.sub & R2R_FLIP) != 0);
    } else {  // OP_C2R_ODD
      const int n = a.f.n;
      if (j == 0) {
        v = C2<T>{base[0], T(0)};
      } else if (2 * j < n) {
        v = ldc<VIN>(base + j * a.esi);
        if (conjf) v.y = -v.y;
      } else {
        v = ldc<VIN>(base + (n - j) * a.esi);
        if (!conjf) v.y = -v.y;
      }
    }
    return v;
  };
  // -------------------------------------- per-element store (bin k < nout)
  constexpr bool ELEM_STORE = OP == OP_C2C || OP == OP_PASSA || OP == OP_PASSB ||
                              OP == OP_R2C_ODD || OP == OP_C2R_EVEN || OP == OP_C2R_ODD;
  auto store = [&](int l, int k, C2<T> v) {
    if (k >= a.nout) return;
    T *base = a.out + s_oo[l];
    if constexpr (OP == OP_C2C) {
      stc<VOUT>(base + rot_index(k, a.rot_out, a.nout) * a.eso, a.fct * v);
    } else if constexpr (OP == OP_PASSA) {
      if (a.sub & SUB_FINAL) {  // Bluestein columns, last pass: chirp and fct, m < n
        const int64_t m = s_aux[l] + int64_t(k) * a.other;
        if (m < a.bign) {
          const C2<T> c = a.t0[m];
          stc<VOUT>(base + blue_col_offset(k, m, a.rot_out, a.bign, s_aux[l], a.eso, a.other),
                    a.fct * cmul(v, conjf ? c : conjc(c)));
        }
      } else {
        const C2<T> w = dirw(big_twiddle(a, int64_t(k) * s_aux[l]), a.dir);
        stc<VOUT>(base + k * a.eso, cmul(v, w));
      }
    } else if constexpr (OP == OP_BLUEROW) {
      const C2<T> w = dirw(big_twiddle(a, int64_t(k) * s_aux[l]), T(1));
      stc<VOUT>(base + k * a.eso, cmul(v, w));
    } else if constexpr (OP == OP_PASSB) {
      const int64_t m = s_aux[l] + int64_t(k) * a.other;
      if (a.sub & SUB_KERNEL) {
        const C2<T> w = a.t1[m];
        stc<VOUT>(base + k * a.eso, cmul(v, conjf ? conjc(w) : w));
      } else if (a.sub & SUB_FINAL) {
        if (m < a.bign) {
          const C2<T> c = a.t0[m];
          stc<VOUT>(base + k * a.eso, a.fct * cmul(v, conjf ? c : conjc(c)));
        }
      } else {
        stc<VOUT>(base + rot_index(k, a.rot_out, a.nout) * a.eso, a.fct * v);
      }
    } else if constexpr (OP == OP_BLUE) {  // bin k to (k + rot_out) mod n
      const C2<T> c = a.t0[k];
      stc<VOUT>(base + rot_index(k, a.rot_out, a.nout) * a.eso,
                a.fct * cmul(v, conjf ? c : conjc(c)));
    } else if constexpr (OP == OP_R2C_ODD) {
      if (conjf) v.y = -v.y;
      stc<VOUT>(base + k * a.eso, a.fct * v);
    } else if constexpr (OP == OP_C2R_EVEN) {
      const C2<T> z = a.fct * v;
      if constexpr (VOUT) {  // contiguous real line: one 2-wide store
        stc<true>(base + 2 * k, z);
      } else {
        base[2 * k * a.eso] = z.x;
        base[(2 * k + 1) * a.eso] = z.y;
      }
    } else if constexpr (OP == OP_C2R_ODD) {
      base[k * a.eso] = a.fct * v.x;
    }
  };

  // ---------------------------------------------------------- load + FFT
  const FftDesc &f = a.f;
  const bool two_ffts = OP == OP_BLUE || OP == OP_BLUEROW;
  // the real-to-real modes keep their (heavier) loads out of the fused stage
  const bool fuse_in = OP != OP_R2R && f.nst >= 1 && fusable_radix(f.st[0].radix);
  // the last stage writes global memory itself when it is not also the
  // fused first stage and the op stores bin by bin
  const bool fuse_out = ELEM_STORE && !two_ffts && f.nst >= 2 &&
                        fusable_radix(f.st[f.nst - 1].radix);
  C2<T> *res;
  if (fuse_in) {
    stage0_in_any(f.st[0].radix, load, bufA, a.L, nl, ld, f.n, a.lfast_in, a.log2L,
                  a.gen + f.st[0].genoff, a.dir);
    __syncthreads();
    res = run_stages<T>(f, 1, fuse_out ? f.nst - 1 : f.nst, bufA, bufB, a.L, ld, a.tw, a.gen,
                        a.dir);
  } else {
    for_tile(nl, a.nin, a.lfast_in, a.log2L,
             [&](int l, int j) { bufA[l * ld + j] = load(l, j); });
    __syncthreads();
    res = run_stages<T>(f, 0, fuse_out ? f.nst - 1 : f.nst, bufA, bufB, a.L, ld, a.tw, a.gen,
                        a.dir);
  }
  if (fuse_out) {
    const Stage st = f.st[f.nst - 1];
    stage_last_out_any(st.radix, store, res, a.L, nl, ld, st.ns, a.lfast_out, a.log2L,
                       a.tw + st.twoff, a.gen + st.genoff, a.dir);
    return;
  }
  if constexpr (OP == OP_BLUE || OP == OP_BLUEROW) {
    const int n2 = a.f.n;
    for_tile(nl, n2, 0, a.log2L, [&](int l, int k) {
      const C2<T> w = OP == OP_BLUE ? a.t1[k] : a.t2[s_aux[l] + int64_t(k) * a.other];
      res[l * ld + k] = cmul(res[l * ld + k], conjf ? conjc(w) : w);
    });
    __syncthreads();
    res = run_stages<T>(f, 0, f.nst, res, res == bufA ? bufB : bufA, a.L, ld, a.tw, a.gen,
                        -a.dir);
  }

  // --------------------------------------------------------------- store
  for_tile(nl, a.nout, a.lfast_out, a.log2L, [&](int l, int k) {
    const C2<T> *r = res + l * ld;
    if constexpr (OP == OP_R2C_EVEN) {
      T *base = a.out + s_oo[l];
      const int h = a.f.n;
      C2<T> X;
      if (k == 0 || k == h) {
        const C2<T> z = r[0];
        X = C2<T>{k == 0 ? z.x + z.y : z.x - z.y, T(0)};
      } else {
        const C2<T> zk = r[k], zm = conjc(r[h - k]);
        const C2<T> fe = T(0.5) * (zk + zm);
        const C2<T> dd = zk - zm;
        const C2<T> fo = C2<T>{T(0.5) * dd.y, T(-0.5) * dd.x};  // (zk - zm) / 2i
        X = fe + cmul(a.t0[k], fo);
      }
      if (conjf) X.y = -X.y;
      stc<VOUT>(base + k * a.eso, a.fct * X);
    } else if constexpr (OP == OP_R2R) {
      T *base = a.out + s_oo[l];
      base[k * a.eso] = a.fct * post_elem<T>((a.sub >> 8) & 0xff, [&](int64_t q) { return r[q]; },
                                             a.bign, k, (a.sub & R2R_ORTHO) != 0,
                                             (a.sub & R2R_FLIP) != 0);
    } else {
      store(l, k, r[k]);
    }
  });
}

}  // namespace dev

// Compile-time specialised kernels (fast.hip); false when none matches.
template <typename T>
bool launch_fast(int op, const dev::LineArgs<T> &a, bool lf_in, bool lf_out, bool vin,
                 bool vout, hipStream_t s);
extern bool g_fast_enabled;

// Persistent software-pipelined FP32 passes of the complex64 1024 x 1024
// four-step plans (pass1024.hip); false when the pass does not match.
bool launch_pass1024(int op, const dev::LineArgs<float> &a, bool lf_in, bool lf_out, bool vin,
                     bool vout, hipStream_t s);

// Persistent software-pipelined FP32 transform of contiguous complex64 lines
// of 4096 (pipe4096.hip); false when the pass does not match.
bool launch_pipe4096(int op, const dev::LineArgs<float> &a, bool lf_in, bool lf_out, bool vin,
                     bool vout, hipStream_t s);

// Split-precision MFMA kernels (mfma.hip): the kernel family launched, or
// nullptr when the mode is off or the pass has no matrix-core form
// (contiguous complex64 lines of 4096, the 1024-point four-step passes).
const char *launch_mfma(int op, const dev::LineArgs<float> &a, bool lf_in, bool lf_out,
                        bool vin, bool vout, hipStream_t s);
void mfma_probe(int bf, const float *a, const float *b, float *d, hipStream_t s);
extern int g_precision_mode;
// launches of the matrix-core kernels in precision mode `mode` (0: all modes)
uint64_t mfma_launch_count(int mode);
// while on (per thread), a missing operand table is an error instead of an
// allocation
void mfma_no_alloc(bool on);
// the FP64-derived DFT-16 operand table of a mode and direction (host only)
int mfma_operand_table(int mode, int dir, float *out, int cap);

}  // namespace rfh
