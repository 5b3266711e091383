// mfma.hip -- the split-precision form: radix-16 DFT stages as dense 16x16
// complex matrix products on the gfx950 matrix cores.
//
// A radix-16 butterfly is X = F v with F[k][r] = e^{dir 2 pi i k r / 16}.
// Sixteen butterflies of a wave are one matrix product D (16x16) = A (16x32)
// x B (32x16) with the complex product embedded in real arithmetic:
//   B = [Re v; Im v] (one column per butterfly),
//   A_re = [Re F, -Im F]  ->  D_re = Re X,     A_im = [Im F, Re F]  ->  D_im = Im X,
// on v_mfma_f32_16x16x32_{f16,bf16} with FP32 accumulation.  The four lanes
// of a butterfly each load four of its sixteen inputs (both parts) once;
// lanes l and l + 32 then swap the halves the operand layout wants
// (v_permlane32_swap), so no input is loaded or twiddled twice.
// FP16 form: every operand is split x = hi + lo (two FP16 values, ~22 bits),
// and D accumulates hi*hi + hi*lo + lo*hi; the operands of stage s are first
// scaled by 2^(10 - e - 4s), e the exponent of the line's largest input (a
// power of two: exact, undone after the product), keeping FP16 in range.
// BF16 form: bf16 has FP32's range but 8 mantissa bits, so operands are
// split three ways (hi + mid + lo), unscaled, and the six products down to
// 2^-24 are accumulated; a two-way BF16 split (three products, ~16 bits) is
// kept as a precision-study mode.
//
// Operand layout of v_mfma_f32_16x16x32_{f16,bf16} (64 lanes; lane l, g =
// l / 16): A row l % 16, columns 8g..8g+7; B rows 8g..8g+7, column l % 16;
// D rows 4g..4g+3, column l % 16.  k_mfma_probe lets a GPU test check this
// layout against a host matrix product.
//
// The kernel covers complex64 lines of 4096 (three radix-16 stages, 256
// threads per line, one line per workgroup); the global loads of stage 0,
// LDS exchanges with the padded layout, twiddle tables and stores of the
// last stage follow the specialised kernel of fast.hip.
#include <atomic>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <tuple>
#include <vector>

#include "bufio.hpp"
#include "passes.hpp"

namespace rfh {

int g_precision_mode = 0;  // 0 plain FP32, 1 FP16 split, 2 BF16 3-split, 3 BF16 2-split

namespace {
using namespace dev;

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf168 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

enum { MODE_F16 = 1, MODE_BF16_3 = 2, MODE_BF16_2 = 3 };

__global__ void k_mfma_probe(int bf, const float *a, const float *b, float *d) {
  // a, b: 64 lanes x 8 values in the operand layout above; d: 64 x 4
  const int l = threadIdx.x;
  f4 c = {0.f, 0.f, 0.f, 0.f};
  f4 r;
  if (bf) {
    bf168 x, y;
    for (int i = 0; i < 8; ++i) {
      x[i] = (__bf16)a[l * 8 + i];
      y[i] = (__bf16)b[l * 8 + i];
    }
    r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, c, 0, 0, 0);
  } else {
    half8 x, y;
    for (int i = 0; i < 8; ++i) {
      x[i] = (_Float16)a[l * 8 + i];
      y[i] = (_Float16)b[l * 8 + i];
    }
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(x, y, c, 0, 0, 0);
  }
  for (int i = 0; i < 4; ++i) d[l * 4 + i] = r[i];
}

template <int MODE>
struct Ops;

template <>
struct Ops<MODE_F16> {
  typedef half8 V;
  static constexpr int terms = 2;
  __device__ static V split(const float *x, int t) {
    V v;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const _Float16 h = (_Float16)x[i];
      v[i] = t == 0 ? h : (_Float16)(x[i] - (float)h);
    }
    return v;
  }
  __device__ static f4 mma(V a, V b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
  // eight values that are exact in FP16
  __device__ static V load(const float *p) {
    V v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (_Float16)p[i];
    return v;
  }
};

template <int TERMS>
struct OpsBf {
  typedef bf168 V;
  static constexpr int terms = TERMS;
  __device__ static V split(const float *x, int t) {
    V v;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      float r = x[i];
      __bf16 p = (__bf16)r;
      for (int s = 0; s < t; ++s) {
        r -= (float)p;
        p = (__bf16)r;
      }
      v[i] = p;
    }
    return v;
  }
  __device__ static f4 mma(V a, V b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  // eight values that are exact in BF16
  __device__ static V load(const float *p) {
    V v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (__bf16)p[i];
    return v;
  }
};
template <>
struct Ops<MODE_BF16_3> : OpsBf<3> {};
template <>
struct Ops<MODE_BF16_2> : OpsBf<2> {};

__device__ __forceinline__ int padf(int i) { return i + (i >> 4); }

// Sixteen radix-16 butterflies of a wave as one complex matrix product: lane
// l (column l % 16 = butterfly, kg = l / 16) holds inputs r0 .. r0 + 3 of
// its butterfly in v (twiddled, r0 = (kg & 1) * 8 + (kg >= 2 ? 4 : 0)),
// scaled by sc before the split; on return (zr, zi) hold its outputs
// 4 kg .. 4 kg + 3 (unscaled: multiply by 1 / sc).
//   B rows 8 kg .. 8 kg + 7: the lower lane (l < 32) takes the real parts of
//   the eight shared inputs (its own four, then those of lane l + 32), the
//   upper lane the imaginary parts (those of lane l - 32, then its own); one
//   v_permlane32_swap of (Re, Im) yields exactly that.
template <int MODE>
__device__ __forceinline__ void dft16_mma(const C2<float> (&v)[4], float sc,
                                          const typename Ops<MODE>::V *Are,
                                          const typename Ops<MODE>::V *Aim, f4 &zr, f4 &zi) {
  using O = Ops<MODE>;
  typedef typename O::V V;
  float x[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i].x * sc),
                                                    __float_as_uint(v[i].y * sc), false, false);
    x[i] = __uint_as_float(r[0]);
    x[4 + i] = __uint_as_float(r[1]);
  }
  V B[O::terms];
#pragma unroll
  for (int t = 0; t < O::terms; ++t) B[t] = O::split(x, t);
  zr = f4{0.f, 0.f, 0.f, 0.f};
  zi = f4{0.f, 0.f, 0.f, 0.f};
  // products down to the split's precision: sum of term indices < terms
#pragma unroll
  for (int ta = 0; ta < O::terms; ++ta)
#pragma unroll
    for (int tb = 0; tb < O::terms; ++tb)
      if (ta + tb < O::terms) {
        zr = O::mma(Are[ta], B[tb], zr);
        zi = O::mma(Aim[ta], B[tb], zi);
      }
}

template <int MODE>
__global__ void __launch_bounds__(256, 4) k_fft4096_mfma(const LineArgs<float> a,
                                                         const float *__restrict__ aop) {
  using O = Ops<MODE>;
  typedef typename O::V V;
  constexpr int N = 4096, NR = 256;
  __shared__ C2<float> line[N + (N >> 4) + 1];
  __shared__ float s_wmax[NR / 64];
  int line_ex = 0;  // FP16: exponent of the line's largest input
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, kg = lane >> 4;
  const bool upper = kg >= 2;  // lanes 32-63
  // this lane's four inputs (of sixteen) per butterfly: lanes kg and kg ^ 2
  // share inputs (kg & 1) * 8 .. + 7, the lower lane loading the first four
  const int r0 = (kg & 1) * 8 + (upper ? 4 : 0);
  const float dir = a.dir;
  const int64_t b = xcd_tile();
  const bool valid = b < a.nlines;
  const float *gin = a.in + b * a.bsi[a.nb > 0 ? a.nb - 1 : 0];
  float *gout = a.out + b * a.bso[a.nb > 0 ? a.nb - 1 : 0];

  // A operands (row = output k = lane % 16, columns 8 kg .. 8 kg + 7 of
  // [Re | Im]: A_re = [Re F, -Im F], A_im = [Im F, Re F]) from the FP64 host
  // table of their split parts (each part exact in the operand format)
  V Are[O::terms], Aim[O::terms];
#pragma unroll
  for (int t = 0; t < O::terms; ++t) {
    Are[t] = O::load(aop + ((2 * t) * 64 + lane) * 8);
    Aim[t] = O::load(aop + ((2 * t + 1) * 64 + lane) * 8);
  }

#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int NS = s == 0 ? 1 : (s == 1 ? 16 : 256);
    const C2<float> *tw = a.tw + a.f.st[s].twoff;
    f4 dre[4], dim[4];
    float inv[4];
    // all of this lane's inputs and twiddles of the stage first (loads in
    // flight together)
    C2<float> in[4][4], w[4][4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int j = wave * 64 + g * 16 + col;
      const int k = j % NS;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = r0 + i;
        if (s == 0)
          in[g][i] = valid ? ldc<true>(gin + 2 * (j + r * NR)) : C2<float>{0.f, 0.f};
        else
          in[g][i] = line[padf(j + r * NR)];
        w[g][i] = (s > 0 && r > 0) ? tw[(r - 1) * NS + k] : C2<float>{1.f, 0.f};
      }
    }
    if constexpr (MODE == MODE_F16) {
      if (s == 0) {  // one scale per line: the largest input, reduced over the block
        float m = 0.f;
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
          for (int i = 0; i < 4; ++i) m = fmaxf(m, fmaxf(fabsf(in[g][i].x), fabsf(in[g][i].y)));
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (lane == 0) s_wmax[wave] = m;
        __syncthreads();
        m = 0.f;
#pragma unroll
        for (int q = 0; q < NR / 64; ++q) m = fmaxf(m, s_wmax[q]);
        (void)frexpf(m, &line_ex);
        if (m == 0.f) line_ex = 10;
      }
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      C2<float> v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v[i] = in[g][i];
        if (s > 0) v[i] = cmul(v[i], dirw(w[g][i], dir));
      }
      // FP16 needs the operands scaled into its range: stage s inputs are at
      // most 16^s times the line's largest input, so 2^(10 - e - 4s) keeps
      // them below 2^10 (exact, undone after the product); BF16 has FP32's
      // exponent range and is split unscaled
      float sc = 1.f;
      inv[g] = 1.f;
      if constexpr (MODE == MODE_F16) {
        sc = ldexpf(1.f, 10 - line_ex - 4 * s);
        inv[g] = ldexpf(1.f, line_ex - 10 + 4 * s);
      }
      dft16_mma<MODE>(v, sc, Are, Aim, dre[g], dim[g]);
    }
    if (s == 2) {
      if (valid) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int j = wave * 64 + g * 16 + col;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int r = 4 * kg + q;
            stc<true>(gout + 2 * (j + r * NR),
                      C2<float>{a.fct * dre[g][q] * inv[g], a.fct * dim[g][q] * inv[g]});
          }
        }
      }
    } else {
      __syncthreads();  // every read of this stage is done
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int j = wave * 64 + g * 16 + col;
        const int k = j % NS;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 4 * kg + q;
          line[padf((j - k) * 16 + k + r * NS)] = C2<float>{dre[g][q] * inv[g], dim[g][q] * inv[g]};
        }
      }
      __syncthreads();
    }
  }
}

// Four-step passes over complex64 lines of 1024 = 16 x 16 x 4 (the passes of
// the 2^20 = 1024 x 1024 transforms of the precision configuration): both
// radix-16 stages as matrix products on the matrix cores, the radix-4 stage
// and the stores on the FP32 VALU.  A tile is 16 lines held in LDS by 1024
// threads (16 waves), with the padded layout of the specialised kernels.
//   Pass A (columns; lines interleaved on both sides): wave w takes
//   butterflies 4w .. 4w+3 of all 16 lines, the matrix column being the line,
//   so a wave's loads cover 128-byte row segments; the stores apply the
//   four-step twiddle W_N^(k1 c).
//   Pass B (rows in, interleaved out): wave w takes line w, the matrix column
//   being the butterfly (16 consecutive elements per load).
// The radix-4 stage and the stores use the interleaved mapping in both passes
// (thread t: line t % 16), so each store instruction writes 128-byte
// segments of 16 adjacent lines.
// Persistent and software-pipelined like the FP32 form (pass1024.hip): one
// 1024-thread workgroup per CU walks its 16-line tiles
// and issues the next tile's loads (buffer instructions, scalar per-element
// offsets) right after the current tile's first stage; no global load runs
// between a prefetch and its use (the memory counter completes in order), so
// stage 1's twiddles are applied in stage 0 from an LDS table, stage 2's come
// from LDS, and pass A's four-step tables sit in LDS.  Pass B's stage 0 -> 1
// exchange stays inside a wave (its line) and takes no workgroup barrier.
constexpr int kPipeL = 16, kPipeP = 1024 + 64 + 1, kPipeSt1 = 256, kPipeSt2 = 256;

template <int MODE, bool PASS_A>
__global__ void __launch_bounds__(1024, 1)
    k_pass1024_mfma_pipe(const LineArgs<float> a, const float *__restrict__ aop, int64_t ntiles,
                         uint32_t in_bytes, uint32_t out_bytes) {
  using O = Ops<MODE>;
  typedef typename O::V V;
  constexpr int L = kPipeL, P = kPipeP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<float> *lds = reinterpret_cast<C2<float> *>(smem_raw);
  __shared__ int64_t s_oi[3][L], s_oo[3][L];
  __shared__ int s_aux[3][L];
  __shared__ float s_max[16][L];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, col = lane & 15, kg = lane >> 4;
  const int r0 = (kg & 1) * 8 + (kg >= 2 ? 4 : 0);
  const int ml = PASS_A ? col : wave;  // line of this lane in the radix-16 stages
  auto bj = [&](int g) { return PASS_A ? wave * 4 + g : g * 16 + col; };
  const int l2 = t & 15, i2 = t >> 4;  // stage 2 and the stores
  const int64_t G = gridDim.x, b = blockIdx.x;
  const int64_t slot0 = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  if (slot0 >= ntiles) return;
  auto tile_of = [&](int64_t k) { return k * G + slot0; };
  auto fill = [&](int64_t tile, int s) {
    if (t < L) {
      const LineOffsets o =
          group_slot(tile * L, t, L, a.nlines, a.nb, a.bdim, a.bsi, a.bso, a.auxdim);
      s_oi[s][t] = o.oi;
      s_oo[s][t] = o.oo;
      s_aux[s][t] = o.aux;
    }
  };
  const float dir = a.dir;
  // behind the lines: stage-1 twiddles ([r 16 + k] = W_256^(r k), row 0
  // ones), stage 2's W_1024^j (j < 256), pass A's four-step lo / hi tables
  C2<float> *st1l = lds + L * P;
  C2<float> *st2l = st1l + kPipeSt1;
  C2<float> *twl = st2l + kPipeSt2;
  const __amdgpu_buffer_rsrc_t rin = bufio::make_rsrc(a.in, in_bytes);
  const __amdgpu_buffer_rsrc_t rout = bufio::make_rsrc(a.out, out_bytes);
  const uint32_t es_in = uint32_t(a.esi) * 4u, es_out = uint32_t(a.eso) * 4u;
  fill(tile_of(0), 0);
  if (tile_of(1) < ntiles) fill(tile_of(1), 1);
  __syncthreads();

  C2<float> in[4][4];
  // inputs bj(g) + 64 (r0 + i) of line ml; a missing line reads zeros
  auto load_tile = [&](int s, int64_t tile) {
    const bool ok = tile * L + ml < a.nlines;
    const uint32_t vo = ok ? uint32_t(s_oi[s][ml]) * 4u + uint32_t(bj(0) + 64 * r0) * es_in
                           : in_bytes;
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        in[g][i] = bufio::bload(rin, vo, uint32_t(bj(g) - bj(0) + 64 * i) * es_in);
  };
  load_tile(0, tile_of(0));
  // the tables' and operands' loads queue behind the first tile's, so both
  // are in flight together (the first stage waits for all of them anyway)
  {
    const C2<float> *tw1 = a.tw + a.f.st[1].twoff;
    const C2<float> *tw2 = a.tw + a.f.st[2].twoff;
    for (int q = t; q < kPipeSt1 + kPipeSt2; q += 1024)
      st1l[q] = q < 16 ? C2<float>{1.f, 0.f}
                       : dirw(q < kPipeSt1 ? tw1[q - 16] : tw2[q - kPipeSt1], dir);
  }
  if (PASS_A) {
    const int nt = a.twlds_lo + a.twlds_hi;
    for (int q = t; q < nt; q += 1024) twl[q] = q < a.twlds_lo ? a.t0[q] : a.t1[q - a.twlds_lo];
  }
  V Are[O::terms], Aim[O::terms];
#pragma unroll
  for (int k = 0; k < O::terms; ++k) {
    Are[k] = O::load(aop + ((2 * k) * 64 + lane) * 8);
    Aim[k] = O::load(aop + ((2 * k + 1) * 64 + lane) * 8);
  }
  __syncthreads();

  for (int64_t k = 0;; ++k) {
    int z;  // opaque zero: keeps the LDS twiddle reads in the loop
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    const int s = int(k % 3);
    const int64_t tile = tile_of(k);
    int ex = 0;  // FP16: exponent of the line's largest input
    if constexpr (MODE == MODE_F16) {
      float m = 0.f;
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int i = 0; i < 4; ++i) m = fmaxf(m, fmaxf(fabsf(in[g][i].x), fabsf(in[g][i].y)));
      m = fmaxf(m, __shfl_xor(m, 16));
      m = fmaxf(m, __shfl_xor(m, 32));
      if constexpr (PASS_A) {  // line = column: over the 16 waves
        if (kg == 0) s_max[wave][col] = m;
        __syncthreads();
        m = 0.f;
#pragma unroll
        for (int w = 0; w < 16; ++w) m = fmaxf(m, s_max[w][col]);
      } else {
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
      }
      (void)frexpf(m, &ex);
      if (m == 0.f) ex = 10;
    }
    // ---- stage 0 (radix 16) on the matrix cores; output r of butterfly j
    // lands at 16 j + r, stage 1's input j / 4, twiddled here by
    // W_256^((j / 4) r)
    {
      const float sc0 = MODE == MODE_F16 ? ldexpf(1.f, 10 - ex) : 1.f;
      const float inv0 = MODE == MODE_F16 ? ldexpf(1.f, ex - 10) : 1.f;
      C2<float> *line = lds + ml * P;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        f4 dre, dim;
        dft16_mma<MODE>(in[g], sc0, Are, Aim, dre, dim);
        const int j = bj(g);
        const C2<float> *tw = st1l + (j >> 2) * 16 + z;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 4 * kg + q;
          line[padf(j * 16 + r)] = cmul(C2<float>{dre[q] * inv0, dim[q] * inv0}, tw[r]);
        }
      }
    }
    if (tile_of(k + 2) < ntiles) fill(tile_of(k + 2), int((k + 2) % 3));
    const bool more = tile_of(k + 1) < ntiles;
    if (more) load_tile(int((k + 1) % 3), tile_of(k + 1));
    if constexpr (PASS_A) __syncthreads();
    else {
      __builtin_amdgcn_wave_barrier();
      asm volatile("" ::: "memory");
    }
    // ---- stage 1 (radix 16, NS 16), inputs already twiddled.  Output k of
    // butterfly j goes back to the slot of its input k (j + 64 k): the
    // positions a wave writes are the ones it read, so no barrier separates
    // the reads from the writes; stage 2 reads the Stockham order through
    // that map.
    {
      const float sc1 = MODE == MODE_F16 ? ldexpf(1.f, 6 - ex) : 1.f;
      const float inv1 = MODE == MODE_F16 ? ldexpf(1.f, ex - 6) : 1.f;
      C2<float> *ln = lds + ml * P;
#pragma unroll 1
      for (int g = 0; g < 4; ++g) {
        const int j = bj(g);
        C2<float> x[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = ln[padf(j + 64 * (r0 + i))];
        f4 dre, dim;
        dft16_mma<MODE>(x, sc1, Are, Aim, dre, dim);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          ln[padf(j + 64 * (4 * kg + q))] = C2<float>{dre[q] * inv1, dim[q] * inv1};
      }
    }
    __syncthreads();
    // ---- stage 2 (radix 4, NS 256) on the VALU, natural-order bins out
    if (tile * L + l2 < a.nlines) {
      const C2<float> *ln = lds + l2 * P;
      const uint32_t vo = uint32_t(s_oo[s][l2]) * 4u + uint32_t(i2) * es_out;
      const int aux = s_aux[s][l2];
#pragma unroll 1
      for (int q = 0; q < 4; ++q) {
        const int j = i2 + 64 * q;
        C2<float> u[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) u[r] = ln[padf(16 * r + (j & 15) + 64 * (j >> 4))];
        const C2<float> w1 = st2l[j + z], w2 = cmul(w1, w1), w3 = cmul(w2, w1);
        u[1] = cmul(u[1], w1);
        u[2] = cmul(u[2], w2);
        u[3] = cmul(u[3], w3);
        dft4(u, dir);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int kb = j + 256 * r;
          C2<float> xo;
          if constexpr (PASS_A) {
            const unsigned e = unsigned(kb) * unsigned(aux);
            const unsigned mask = (1u << a.shift) - 1u;
            const C2<float> w = cmul(twl[e & mask], twl[a.twlds_lo + int(e >> a.shift)]);
            xo = cmul(u[r], dirw(w, dir));
          } else {
            xo = a.fct * u[r];
          }
          bufio::bstore(rout, vo, uint32_t(64 * q + 256 * r) * es_out, xo);
        }
      }
    }
    if (!more) break;
    __syncthreads();
  }
}

// Contiguous complex64 lines of 4096 on the matrix cores, persistent and
// software-pipelined like the FP32 form (pipe4096.hip): one 1024-thread
// workgroup per CU walks tiles of four lines (four waves per line); a wave
// runs 64 butterflies of its line per stage, in four groups of sixteen (the
// matrix column being the butterfly).  The next tile's loads go out after
// stage 0.  Each stage's twiddles are applied as the stage before writes:
// stage 1's W_256^(r k) from an LDS table, stage 2's W_4096^(j2 (j / 16)) as
// A_j W_256^(q (j / 16)) (A_j = W_4096^((j mod 16)(j / 16)) from a second
// LDS table).  Stage 1 runs in place (each wave writes back exactly the
// slots it read), and stage 2 stores its bins straight to memory.
constexpr int kP4L = 4, kP4P = 4096 + 256 + 1, kP4St1 = 256, kP4Lo = 256;

template <int MODE>
__global__ void __launch_bounds__(1024, 1)
    k_pipe4096_mfma(const LineArgs<float> a, const float *__restrict__ aop, int64_t ntiles,
                    uint32_t in_bytes, uint32_t out_bytes) {
  using O = Ops<MODE>;
  typedef typename O::V V;
  constexpr int L = kP4L, P = kP4P, NR = 256;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<float> *lds = reinterpret_cast<C2<float> *>(smem_raw);
  __shared__ int64_t s_oi[3][L], s_oo[3][L];
  __shared__ float s_max[16];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, col = lane & 15, kg = lane >> 4;
  const int r0 = (kg & 1) * 8 + (kg >= 2 ? 4 : 0);
  const int l = wave >> 2, wl = wave & 3;  // line of the tile, wave within the line
  auto bj = [&](int g) { return wl * 64 + g * 16 + col; };
  const int64_t G = gridDim.x, b = blockIdx.x;
  const int64_t slot0 = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  if (slot0 >= ntiles) return;
  auto tile_of = [&](int64_t k) { return k * G + slot0; };
  auto fill = [&](int64_t tile, int s) {
    if (t < L) {
      const LineOffsets o =
          group_slot(tile * L, t, L, a.nlines, a.nb, a.bdim, a.bsi, a.bso, a.auxdim);
      s_oi[s][t] = o.oi;
      s_oo[s][t] = o.oo;
    }
  };
  const float dir = a.dir;
  // behind the lines: st1[r 16 + k] = W_256^(r k) (row 0 ones), lo[e] =
  // W_4096^e (e < 256, the first row of stage 2's table)
  C2<float> *st1l = lds + L * P;
  C2<float> *lo = st1l + kP4St1;
  const __amdgpu_buffer_rsrc_t rin = bufio::make_rsrc(a.in, in_bytes);
  const __amdgpu_buffer_rsrc_t rout = bufio::make_rsrc(a.out, out_bytes);
  const uint32_t es_in = uint32_t(a.esi) * 4u, es_out = uint32_t(a.eso) * 4u;
  fill(tile_of(0), 0);
  if (tile_of(1) < ntiles) fill(tile_of(1), 1);
  __syncthreads();

  C2<float> in[4][4];
  // inputs bj(g) + 256 (r0 + i) of line l; a missing line reads zeros
  auto load_tile = [&](int s, int64_t tile) {
    const bool ok = tile * L + l < a.nlines;
    const uint32_t vo = ok ? uint32_t(s_oi[s][l]) * 4u + uint32_t(bj(0) + NR * r0) * es_in
                           : in_bytes;
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        in[g][i] = bufio::bload(rin, vo, uint32_t(16 * g + NR * i) * es_in);
  };
  load_tile(0, tile_of(0));
  // the tables' and operands' loads queue behind the first tile's, so both
  // are in flight together (the first stage waits for all of them anyway)
  {
    const C2<float> *tw1 = a.tw + a.f.st[1].twoff;
    const C2<float> *tw2 = a.tw + a.f.st[2].twoff;
    for (int q = t; q < kP4St1 + kP4Lo; q += 1024)
      st1l[q] = q < 16 ? C2<float>{1.f, 0.f}
                       : dirw(q < kP4St1 ? tw1[q - 16] : tw2[q - kP4St1], dir);
  }
  V Are[O::terms], Aim[O::terms];
#pragma unroll
  for (int k = 0; k < O::terms; ++k) {
    Are[k] = O::load(aop + ((2 * k) * 64 + lane) * 8);
    Aim[k] = O::load(aop + ((2 * k + 1) * 64 + lane) * 8);
  }
  __syncthreads();

  for (int64_t k = 0;; ++k) {
    int z;  // opaque zero: keeps the LDS twiddle reads in the loop
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    const int s = int(k % 3);
    const int64_t tile = tile_of(k);
    C2<float> *ln = lds + l * P;
    int ex = 0;  // FP16: exponent of the line's largest input
    if constexpr (MODE == MODE_F16) {
      float m = 0.f;
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int i = 0; i < 4; ++i) m = fmaxf(m, fmaxf(fabsf(in[g][i].x), fabsf(in[g][i].y)));
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
      if (lane == 0) s_max[wave] = m;
      __syncthreads();
      m = fmaxf(fmaxf(s_max[4 * l], s_max[4 * l + 1]), fmaxf(s_max[4 * l + 2], s_max[4 * l + 3]));
      (void)frexpf(m, &ex);
      if (m == 0.f) ex = 10;
    }
    // ---- stage 0: output r of butterfly j lands at 16 j + r, stage 1's
    // input j / 16 of its butterfly 16 (j mod 16) + r, twiddled here by
    // W_256^((j / 16) r)
    {
      const float sc0 = MODE == MODE_F16 ? ldexpf(1.f, 10 - ex) : 1.f;
      const float inv0 = MODE == MODE_F16 ? ldexpf(1.f, ex - 10) : 1.f;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        f4 dre, dim;
        dft16_mma<MODE>(in[g], sc0, Are, Aim, dre, dim);
        const int j = bj(g);
        const C2<float> *tw = st1l + (j >> 4) * 16 + z;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 4 * kg + q;
          ln[padf(j * 16 + r)] = cmul(C2<float>{dre[q] * inv0, dim[q] * inv0}, tw[r]);
        }
      }
    }
    if (tile_of(k + 2) < ntiles) fill(tile_of(k + 2), int((k + 2) % 3));
    const bool more = tile_of(k + 1) < ntiles;
    if (more) load_tile(int((k + 1) % 3), tile_of(k + 1));
    __syncthreads();
    // ---- stage 1 (NS 16), inputs already twiddled; output q of butterfly
    // j goes back to j + 256 q with stage 2's twiddle A_j W_256^(q (j / 16))
    {
      const float sc1 = MODE == MODE_F16 ? ldexpf(1.f, 6 - ex) : 1.f;
      const float inv1 = MODE == MODE_F16 ? ldexpf(1.f, ex - 6) : 1.f;
#pragma unroll 1
      for (int g = 0; g < 4; ++g) {
        const int j = bj(g);
        C2<float> x[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = ln[padf(j + NR * (r0 + i))];
        f4 dre, dim;
        dft16_mma<MODE>(x, sc1, Are, Aim, dre, dim);
        const C2<float> aj = lo[(j & 15) * (j >> 4) + z];
        const C2<float> *tw = st1l + (j >> 4) * 16;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int qq = 4 * kg + q;
          ln[padf(j + NR * qq)] =
              cmul(C2<float>{dre[q] * inv1, dim[q] * inv1}, cmul(aj, tw[qq]));
        }
      }
    }
    __syncthreads();
    // ---- stage 2 (NS 256): butterfly j2 takes Stockham inputs j2 + 256 r
    // (at 16 r + j2 mod 16 + 256 (j2 / 16)) and stores bins j2 + 256 r
    if (tile * L + l < a.nlines) {
      const float sc2 = MODE == MODE_F16 ? ldexpf(1.f, 2 - ex) : 1.f;
      const float inv2 = MODE == MODE_F16 ? ldexpf(1.f, ex - 2) : 1.f;
      const uint32_t vo = uint32_t(s_oo[s][l]) * 4u + uint32_t(bj(0) + NR * 4 * kg) * es_out;
#pragma unroll 1
      for (int g = 0; g < 4; ++g) {
        const int j2 = bj(g);
        C2<float> x[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = ln[padf(16 * (r0 + i) + (j2 & 15) + NR * (j2 >> 4))];
        f4 dre, dim;
        dft16_mma<MODE>(x, sc2, Are, Aim, dre, dim);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          bufio::bstore(rout, vo, uint32_t(16 * g + NR * q) * es_out,
                        C2<float>{a.fct * dre[q] * inv2, a.fct * dim[q] * inv2});
      }
    }
    if (!more) break;
    __syncthreads();
  }
}

// ---------------------------------------------------------- host side

// x rounded to `bits` significant bits (ties to even), with a quantum of at
// least 2^qmin (FP16's subnormal step), in FP64.
double round_sig(double x, int bits, int qmin) {
  if (x == 0.0) return 0.0;
  int e = 0;
  (void)std::frexp(x, &e);  // |x| = m 2^e, 0.5 <= m < 1
  const int q = std::max(e - bits, qmin);
  return std::ldexp(std::nearbyint(std::ldexp(x, -q)), q);
}

// A operands of the radix-16 DFT product for one mode and direction: per
// lane l (col = l % 16, kg = l / 16) and i < 8, column kk = 8 kg + i of
// [Re F | -Im F] (A_re) and [Im F | Re F] (A_im), F[k][r] =
// e^{dir 2 pi i k r / 16} in FP64, split into `terms` parts each exact in
// the operand format (p_t = round(v - p_0 - ... - p_{t-1})), stored as
// floats at [(2 t + {0: re, 1: im}) * 64 + l] * 8 + i.
std::vector<float> dft16_operands(int mode, int dir) {
  const int terms = mode == MODE_F16 ? 2 : (mode == MODE_BF16_3 ? 3 : 2);
  const int bits = mode == MODE_F16 ? 11 : 8;
  const int qmin = mode == MODE_F16 ? -24 : -133;
  std::vector<float> t(size_t(terms) * 2 * 64 * 8);
  for (int l = 0; l < 64; ++l) {
    const int col = l & 15, kg = l >> 4;
    for (int i = 0; i < 8; ++i) {
      const int kk = 8 * kg + i, r = kk & 15, m = (col * r) & 15;
      double cs = std::cos(2.0 * M_PI * m / 16.0), sn = std::sin(2.0 * M_PI * m / 16.0);
      if (std::fabs(cs) < 1e-12) cs = 0.0;
      if (std::fabs(sn) < 1e-12) sn = 0.0;
      const double fr = cs, fi = dir * sn;
      const double v[2] = {kk < 16 ? fr : -fi, kk < 16 ? fi : fr};
      for (int c = 0; c < 2; ++c) {
        double rest = v[c];
        for (int k = 0; k < terms; ++k) {
          const double p = round_sig(rest, bits, qmin);
          t[(size_t(2 * k + c) * 64 + l) * 8 + i] = float(p);
          rest -= p;
        }
      }
    }
  }
  return t;
}

std::mutex g_aop_mu;
std::map<std::tuple<int, int, int>, std::unique_ptr<DevBuf>> g_aop;  // (device, mode, dir)
thread_local bool t_no_alloc = false;  // the calling entry point may not allocate

const float *operands(int mode, int dir) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_aop_mu);
  auto &b = g_aop[std::make_tuple(dev, mode, dir)];
  if (!b) {
    if (t_no_alloc)
      RFH_THROW(3, "the matrix-core tables of this precision mode are not on the device yet; "
                   "run the ordinary entry point once in this mode first");
    const std::vector<float> t = dft16_operands(mode, dir);
    std::unique_ptr<DevBuf> nb(new DevBuf());
    nb->upload(t.data(), t.size() * sizeof(float));
    b = std::move(nb);
  }
  return static_cast<const float *>(b->p);
}

std::atomic<uint64_t> g_mfma_launches[4];

}  // namespace

void mfma_no_alloc(bool on) { t_no_alloc = on; }

int mfma_operand_table(int mode, int dir, float *out, int cap) {
  if (mode < 1 || mode > 3 || (dir != 1 && dir != -1)) return -1;
  const std::vector<float> t = dft16_operands(mode, dir);
  for (int i = 0; i < int(t.size()) && i < cap; ++i) out[i] = t[i];
  return int(t.size());
}

uint64_t mfma_launch_count(int mode) {
  if (mode < 0 || mode > 3) return 0;
  if (mode == 0)
    return g_mfma_launches[1].load() + g_mfma_launches[2].load() + g_mfma_launches[3].load();
  return g_mfma_launches[mode].load();
}

namespace {
std::mutex g_pass_attr_mu;
std::set<std::pair<int, const void *>> g_pass_attr;  // (device, kernel)

void set_lds_attr(const void *fn) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_pass_attr_mu);
  if (!g_pass_attr.count({dev, fn})) {
    hipFuncAttributes fa;
    HIP_CHECK(hipFuncGetAttributes(&fa, fn));
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(lds_limit() - fa.sharedSizeBytes)));
    g_pass_attr.insert({dev, fn});
  }
}

// The persistent pipelined form: spans below 2^31 bytes with non-negative
// strides (buffer offsets), pass A's four-step tables in LDS beside the
// tile and the stage tables; otherwise the caller takes the one-tile form.
bool launch_pass1024_pipe(int mode, bool pass_a, const dev::LineArgs<float> &a0, hipStream_t s) {
  const int64_t si = bufio::line_span(a0, 1024, true), so = bufio::line_span(a0, 1024, false);
  if (si < 0 || so < 0 || si * 4 >= (int64_t(1) << 31) - (int64_t(8) << 20) ||
      so * 4 >= (int64_t(1) << 31))
    return false;
  dev::LineArgs<float> a = a0;
  const int64_t ntiles = (a.nlines + kPipeL - 1) / kPipeL;
  if (ntiles <= 0) return true;
  size_t lds = size_t(kPipeL * kPipeP + kPipeSt1 + kPipeSt2) * sizeof(C2<float>);
  a.twlds_lo = a.twlds_hi = 0;
  if (pass_a) {
    const int64_t big = int64_t(1024) * a.other;
    const int64_t lo_n = int64_t(1) << a.shift, hi_n = (big + lo_n - 1) / lo_n;
    const size_t extra = size_t(lo_n + hi_n) * sizeof(C2<float>);
    if (big >= (int64_t(1) << 31) || lds + extra + 2048 > lds_limit()) return false;
    a.twlds_lo = int(lo_n);
    a.twlds_hi = int(hi_n);
    lds += extra;
  }
  const void *fn = nullptr;
#define RFH_PIPE_FN(M)                                                                     \
  fn = pass_a ? reinterpret_cast<const void *>(&k_pass1024_mfma_pipe<M, true>)            \
              : reinterpret_cast<const void *>(&k_pass1024_mfma_pipe<M, false>)
  switch (mode) {
    case MODE_F16: RFH_PIPE_FN(MODE_F16); break;
    case MODE_BF16_3: RFH_PIPE_FN(MODE_BF16_3); break;
    default: RFH_PIPE_FN(MODE_BF16_2); break;
  }
#undef RFH_PIPE_FN
  set_lds_attr(fn);
  const float *aop = operands(mode, a.dir < 0 ? -1 : 1);
  int64_t grid = std::min<int64_t>(ntiles, bufio::device_cus());
  if (grid >= 8) grid -= grid % 8;
  uint32_t in_bytes = uint32_t(si * 4), out_bytes = uint32_t(so * 4);
  void *args[] = {&a, &aop, const_cast<int64_t *>(&ntiles), &in_bytes, &out_bytes};
  HIP_CHECK(hipLaunchKernel(fn, dim3(unsigned(grid)), dim3(1024), args, lds, s));
  return true;
}

// The four-step passes of complex64 lines of 1024 on the matrix cores; the
// kernel family launched, or nullptr.
const char *launch_pass1024(int mode, int op, const dev::LineArgs<float> &a, bool lf_in,
                            bool lf_out, hipStream_t s) {
  const bool pass_a = op == OP_PASSA && lf_in && lf_out && !(a.sub & (SUB_CHIRP | SUB_FINAL));
  const bool pass_b = op == OP_PASSB && !lf_in && lf_out &&
                      !(a.sub & (SUB_KERNEL | SUB_FINAL));
  if (!pass_a && !pass_b) return nullptr;
  if (a.f.nst != 3 || a.f.st[0].radix != 16 || a.f.st[1].radix != 16 || a.f.st[2].radix != 4)
    return nullptr;
  if (a.sub == 0 && launch_pass1024_pipe(mode, pass_a, a, s)) return "k_pass1024_mfma_pipe";
  return nullptr;  // spans or tables beyond the pipelined form: the FP32 kernels run it
}

// Contiguous complex64 lines of 4096 on the persistent pipelined kernel:
// sets of at least two tiles per CU, spans below 2^31 bytes with
// non-negative strides; otherwise the caller takes the one-line form.
bool launch_pipe4096_mfma(int mode, const dev::LineArgs<float> &a, hipStream_t s) {
  const int64_t ntiles = (a.nlines + kP4L - 1) / kP4L;
  const int cus = bufio::device_cus();
  if (ntiles < 2 * int64_t(cus)) return false;
  const int64_t si = bufio::line_span(a, 4096, true), so = bufio::line_span(a, 4096, false);
  if (si < 0 || so < 0 || si * 4 >= (int64_t(1) << 31) - (int64_t(8) << 20) ||
      so * 4 >= (int64_t(1) << 31))
    return false;
  const void *fn = mode == MODE_F16 ? reinterpret_cast<const void *>(&k_pipe4096_mfma<MODE_F16>)
                   : mode == MODE_BF16_3
                       ? reinterpret_cast<const void *>(&k_pipe4096_mfma<MODE_BF16_3>)
                       : reinterpret_cast<const void *>(&k_pipe4096_mfma<MODE_BF16_2>);
  const size_t lds = size_t(kP4L * kP4P + kP4St1 + kP4Lo) * sizeof(C2<float>);
  if (lds + 2048 > lds_limit()) return false;
  set_lds_attr(fn);
  const float *aop = operands(mode, a.dir < 0 ? -1 : 1);
  int64_t grid = std::min<int64_t>(ntiles, cus);
  if (grid >= 8) grid -= grid % 8;
  uint32_t in_bytes = uint32_t(si * 4), out_bytes = uint32_t(so * 4);
  dev::LineArgs<float> la = a;
  int64_t nt = ntiles;
  void *args[] = {&la, &aop, &nt, &in_bytes, &out_bytes};
  HIP_CHECK(hipLaunchKernel(fn, dim3(unsigned(grid)), dim3(1024), args, lds, s));
  return true;
}
}  // namespace

// Complex64 lines of 4096, contiguous, one line per group, and the four-step
// passes of complex64 lines of 1024: the MFMA form.
const char *launch_mfma(int op, const dev::LineArgs<float> &a, bool lf_in, bool lf_out,
                        bool vin, bool vout, hipStream_t s) {
  if (!g_precision_mode || !vin || !vout) return nullptr;
  if (a.rot_in || a.rot_out) return nullptr;  // rotated lines run on the FP32 kernels
  if (a.f.n == 1024 && (op == OP_PASSA || op == OP_PASSB)) {
    const char *fam = launch_pass1024(g_precision_mode, op, a, lf_in, lf_out, s);
    if (fam) g_mfma_launches[g_precision_mode].fetch_add(1);
    return fam;
  }
  if (op != OP_C2C || a.f.n != 4096 || lf_in || lf_out) return nullptr;
  if (a.f.nst != 3 || a.sub != 0) return nullptr;
  const int mode = g_precision_mode;
  if (launch_pipe4096_mfma(mode, a, s)) {
    g_mfma_launches[mode].fetch_add(1);
    return "k_pipe4096_mfma";
  }
  if (a.esi != 2 || a.eso != 2 || a.nb > 1) return nullptr;
  const float *aop = operands(mode, a.dir < 0 ? -1 : 1);
  const unsigned grid = unsigned(a.nlines);
  switch (mode) {
    case MODE_F16: hipLaunchKernelGGL(k_fft4096_mfma<MODE_F16>, dim3(grid), dim3(256), 0, s, a, aop); break;
    case MODE_BF16_3: hipLaunchKernelGGL(k_fft4096_mfma<MODE_BF16_3>, dim3(grid), dim3(256), 0, s, a, aop); break;
    default: hipLaunchKernelGGL(k_fft4096_mfma<MODE_BF16_2>, dim3(grid), dim3(256), 0, s, a, aop); break;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) RFH_THROW(2, std::string("mfma launch: ") + hipGetErrorString(e));
  g_mfma_launches[mode].fetch_add(1);
  return "k_fft4096_mfma";
}

void mfma_probe(int bf, const float *a, const float *b, float *d, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, bf, a, b, d);
  HIP_CHECK(hipGetLastError());
}

}  // namespace rfh
