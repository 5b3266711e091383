// fast.hip -- compile-time specialised line kernels for power-of-two lines.
//
// Same operations as k_lines (kernels.hip) but with the line length, radix
// schedule and thread mapping fixed at compile time:
//   * stage 0 reads its radix-R inputs straight from HBM into registers and
//     the last stage writes straight back, so a line crosses LDS only
//     between stages (S-1 exchanges instead of S+1);
//   * one LDS buffer per line, padded by one element every 16 (complex64) or
//     8 (complex128) elements and one per line, instead of two ping-pong
//     buffers -- twice the lines per CU;
//   * threads are mapped to (line, butterfly) so that global accesses are
//     coalesced: consecutive lanes take consecutive elements of one line
//     (contiguous lines), or the same element of adjacent lines
//     (interleaved lines, LFAST).  When input and output want different
//     mappings, or an operation needs two bins of a line at once (r2c), the
//     result goes through LDS and a store loop with the output mapping.
// Twiddles come from the same per-stage tables as k_lines (the plan's
// Direct for this length uses the same radix schedule).
#include <mutex>
#include <set>

#include "device.hpp"

namespace rfh {
namespace {
using namespace dev;

template <int... Rs>
struct Sched {
  static constexpr int count = sizeof...(Rs);
  static constexpr int at(int s) {
    constexpr int a[] = {Rs...};
    return a[s];
  }
  static constexpr int ns(int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= at(i);
    return p;
  }
};

template <typename T>
struct PadOf {
  static constexpr int shift = sizeof(T) == 4 ? 4 : 3;
};

template <typename T>
__device__ __forceinline__ int padi(int i) {
  return i + (i >> PadOf<T>::shift);
}

template <typename T, int N>
constexpr int pitch_of() {
  return N + (N >> PadOf<T>::shift) + 1;
}

// per-thread state of one fast-kernel invocation
template <typename T>
struct FCtx {
  const LineArgs<T> *a;
  const C2<T> *twl; // four-step twiddle tables in LDS (pass A), or null
  const C2<T> *rtw; // real-transform twiddles t0 (an LDS copy for short lines)
  C2<T> *line;     // this thread's line in LDS
  T *sline;        // ... in the split layout (one part at a time)
  T *soline;       // line of the output mapping in the split layout
  const int64_t *s_oo;  // per-line output offsets (store through LDS)
  const int *s_aux;
  int lo, io, nl;  // output mapping: line, index within it, lines in the tile
  const T *gin;    // this thread's line in global memory (input)
  T *gout;         // ... output
  int aux;
  bool valid;
  int i;           // index of the thread within its line
  T dir;           // sign of the stages being run (-1 forward)
};

// c2r: input p of the half-length complex transform from bins p and h - p
// of the half spectrum
template <typename T>
__device__ __forceinline__ C2<T> c2r_pre(const LineArgs<T> &a, const C2<T> *t0, C2<T> xj,
                                         C2<T> xm, int p) {
  if (a.sub & SUB_CONJ) {
    xj.y = -xj.y;
    xm.y = -xm.y;
  }
  if (p == 0) {
    xj.y = T(0);
    xm.y = T(0);
  }
  const C2<T> A = xj + conjc(xm);
  const C2<T> B = cmul(xj - conjc(xm), conjc(t0[p]));
  return C2<T>{A.x - B.y, A.y + B.x};
}

template <typename T, int OP, bool VIN>
__device__ __forceinline__ C2<T> fload(const FCtx<T> &c, int p) {
  const LineArgs<T> &a = *c.a;
  if constexpr (OP == OP_R2C_EVEN) {
    if constexpr (VIN)
      return ldc<true>(c.gin + 2 * p);
    else
      return C2<T>{c.gin[2 * p * a.esi], c.gin[(2 * p + 1) * a.esi]};
  } else if constexpr (OP == OP_C2R_EVEN) {
    const int h = a.f.n;
    return c2r_pre(a, c.rtw, ldc<VIN>(c.gin + p * a.esi), ldc<VIN>(c.gin + (h - p) * a.esi), p);
  } else if constexpr (OP == OP_PASSA) {
    if (a.sub & SUB_CHIRP) {  // Bluestein columns: x[m] times the chirp, zero from m = n on
      const int64_t m = int64_t(p) * a.other + c.aux;
      if (m >= a.bign) return C2<T>{T(0), T(0)};
      const C2<T> ch = a.t2[m];
      const int64_t off = blue_col_offset(p, m, a.rot_in, a.bign, c.aux, a.esi, a.other);
      return cmul(ldc<VIN>(c.gin + off), (a.sub & SUB_CONJ) ? ch : conjc(ch));
    }
    return ldc<VIN>(c.gin + ((p + a.rot_in) & (a.f.n - 1)) * a.esi);
  } else {  // OP_C2C (power-of-two lengths: the rotation is a mask)
    return ldc<VIN>(c.gin + ((p + a.rot_in) & (a.f.n - 1)) * a.esi);
  }
}

// four-step twiddle W_N^e from the LDS copy of the lo / hi tables, or from
// global memory
template <typename T>
__device__ __forceinline__ C2<T> ftwiddle(const LineArgs<T> &a, const C2<T> *twl, int64_t e) {
  if (twl) {
    const int64_t mask = (int64_t(1) << a.shift) - 1;
    return cmul(twl[e & mask], twl[a.twlds_lo + (e >> a.shift)]);
  }
  return big_twiddle(a, e);
}

// elementwise store of bin p of this thread's line (all ops except r2c)
template <typename T, int OP, bool VOUT>
__device__ __forceinline__ void fstore(const LineArgs<T> &a, const C2<T> *twl, T *gout, int aux,
                                       int p, C2<T> v) {
  if constexpr (OP == OP_C2C) {
    stc<VOUT>(gout + ((p + a.rot_out) & (a.f.n - 1)) * a.eso, a.fct * v);
  } else if constexpr (OP == OP_PASSA) {
    if (a.sub & SUB_FINAL) {  // Bluestein columns, last pass: chirp and fct, m < n
      const int64_t m = aux + int64_t(p) * a.other;
      if (m < a.bign) {
        const C2<T> ch = a.t0[m];
        stc<VOUT>(gout + blue_col_offset(p, m, a.rot_out, a.bign, aux, a.eso, a.other),
                  a.fct * cmul(v, (a.sub & SUB_CONJ) ? ch : conjc(ch)));
      }
    } else {
      const C2<T> w = dirw(ftwiddle(a, twl, int64_t(p) * aux), a.dir);
      stc<VOUT>(gout + p * a.eso, cmul(v, w));
    }
  } else if constexpr (OP == OP_BLUEROW) {  // inverse four-step twiddle, in place
    const C2<T> w = dirw(big_twiddle(a, int64_t(p) * aux), T(1));
    stc<VOUT>(gout + p * a.eso, cmul(v, w));
  } else if constexpr (OP == OP_PASSB) {
    const int64_t m = aux + int64_t(p) * a.other;
    const bool cj = (a.sub & SUB_CONJ) != 0;
    if (a.sub & SUB_KERNEL) {
      const C2<T> w = a.t1[m];
      stc<VOUT>(gout + p * a.eso, cmul(v, cj ? conjc(w) : w));
    } else if (a.sub & SUB_FINAL) {
      if (m < a.bign) {
        const C2<T> ch = a.t0[m];
        stc<VOUT>(gout + p * a.eso, a.fct * cmul(v, cj ? ch : conjc(ch)));
      }
    } else {
      stc<VOUT>(gout + ((p + a.rot_out) & (a.f.n - 1)) * a.eso, a.fct * v);
    }
  } else if constexpr (OP == OP_C2R_EVEN) {
    const C2<T> z = a.fct * v;
    if constexpr (VOUT) {
      stc<true>(gout + 2 * p, z);
    } else {
      gout[2 * p * a.eso] = z.x;
      gout[(2 * p + 1) * a.eso] = z.y;
    }
  }
}

// r2c post-processing of bin k (1..h-1) from bins k and h - k of the
// half-length spectrum z
template <typename T>
__device__ __forceinline__ C2<T> r2c_combine(const LineArgs<T> &a, const C2<T> *t0, C2<T> zk,
                                             C2<T> zhk, int k) {
  const C2<T> zm = conjc(zhk);
  const C2<T> fe = T(0.5) * (zk + zm);
  const C2<T> dd = zk - zm;
  const C2<T> fo = C2<T>{T(0.5) * dd.y, T(-0.5) * dd.x};
  C2<T> X = fe + cmul(t0[k], fo);
  if (a.sub & SUB_CONJ) X.y = -X.y;
  return a.fct * X;
}

// bins 0 and h, both from z_0
template <typename T>
__device__ __forceinline__ C2<T> r2c_edge(const LineArgs<T> &a, C2<T> z0, bool top) {
  return a.fct * C2<T>{top ? z0.x - z0.y : z0.x + z0.y, T(0)};
}

// r2c post-processing of bin k (0..h) from the natural-order half-length
// spectrum in LDS
template <typename T>
__device__ __forceinline__ C2<T> r2c_bin(const LineArgs<T> &a, const C2<T> *t0, const C2<T> *line,
                                         int k) {
  const int h = a.f.n;
  if (k == 0 || k == h) return r2c_edge(a, line[0], k == h);
  return r2c_combine(a, t0, line[padi<T>(k)], line[padi<T>(h - k)], k);
}

// r2c from the registers of the last stage: thread i holds bins
// k = (i + q TPL) + r NS (NS = Q TPL); bin N - k is slot (Q-1-q, R-1-r) of
// lane TPL - i, or for i = 0 slot (Q-q, R-1-r) (q >= 1) / (0, R-r) (q = 0)
// of the lane itself.  Every bin is stored straight from registers.
template <typename T, int N, int TPL, int Q, int R, bool VOUT>
__device__ __forceinline__ void r2c_store_shuffled(const FCtx<T> &c, const C2<T> (&v)[Q][R]) {
  const LineArgs<T> &a = *c.a;
  constexpr int NS = Q * TPL;
  const int i = c.i;
  const int lane = threadIdx.x & 63;
  const int src = i == 0 ? lane : lane + TPL - 2 * i;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const C2<T> w = v[Q - 1 - q][R - 1 - r];
      C2<T> m{__shfl(w.x, src), __shfl(w.y, src)};
      if (i == 0) m = q == 0 ? v[0][(R - r) % R] : v[Q - q][R - 1 - r];
      const int k = i + q * TPL + r * NS;
      const C2<T> X = (i == 0 && q == 0 && r == 0) ? r2c_edge(a, v[0][0], false)
                                                   : r2c_combine(a, c.rtw, v[q][r], m, k);
      if (c.valid) stc<VOUT>(c.gout + k * a.eso, X);
    }
  }
  if (c.valid && i == 0) stc<VOUT>(c.gout + N * a.eso, r2c_edge(a, v[0][0], true));
}

// c2r stage 0 with the bins h - p taken from the lanes that loaded them:
// with one butterfly per thread (p = j + r TPL, TPL lanes of a line adjacent
// in the wave), bin N - p is bin (TPL - j) + (R - 1 - r) TPL of lane
// TPL - j, or, for j = 0, bin (R - r) TPL of the lane itself (bin N loaded
// separately); each input is read from memory once instead of twice.
template <typename T, int N, int TPL, int R, bool VIN>
__device__ __forceinline__ void c2r_load_shuffled(const FCtx<T> &c, C2<T> (&v)[R]) {
  const LineArgs<T> &a = *c.a;
  const int j = c.i;
  C2<T> x[R];
#pragma unroll
  for (int r = 0; r < R; ++r)
    x[r] = c.valid ? ldc<VIN>(c.gin + (j + r * TPL) * a.esi) : C2<T>{T(0), T(0)};
  const C2<T> top = (c.valid && j == 0) ? ldc<VIN>(c.gin + N * a.esi) : C2<T>{T(0), T(0)};
  const int lane = threadIdx.x & 63;
  const int src = j == 0 ? lane : lane + TPL - 2 * j;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const C2<T> w = x[R - 1 - r];
    C2<T> m{__shfl(w.x, src), __shfl(w.y, src)};
    if (j == 0) m = r == 0 ? top : x[(R - r) % R];
    v[r] = c2r_pre(a, c.rtw, x[r], m, j + r * TPL);
  }
}

template <typename T, int N, int TPL, int OP, bool VIN, bool VOUT, bool DIRECT_OUT,
          typename SC, int S, bool FLAT, bool SHUF = false>
__device__ __forceinline__ void fstage(const FCtx<T> &c) {
  constexpr int R = SC::at(S);
  constexpr int NS = SC::ns(S);
  constexpr int NR = N / R;
  constexpr int Q = (NR + TPL - 1) / TPL;
  constexpr bool FIRST = S == 0;
  constexpr bool LAST = S == SC::count - 1;
  const LineArgs<T> &a = *c.a;
  const T dir = c.dir;
  C2<T> v[Q][R];
  static_assert(!SHUF || (FIRST && Q == 1 && NR == TPL && OP == OP_C2R_EVEN), "shuffled c2r load");
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = c.i + q * TPL;
    if (Q * TPL == NR || j < NR) {
      if constexpr (SHUF) {
        c2r_load_shuffled<T, N, TPL, R, VIN>(c, v[q]);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (FIRST && !FLAT)
            v[q][r] = c.valid ? fload<T, OP, VIN>(c, j + r * NR) : C2<T>{T(0), T(0)};
          else if constexpr (FIRST && OP == OP_C2R_EVEN)
            v[q][r] = c2r_pre(a, c.rtw, c.line[padi<T>(j + r * NR)],
                              c.line[padi<T>(N - j - r * NR)], j + r * NR);
          else
            v[q][r] = c.line[padi<T>(j + r * NR)];
        }
      }
      if constexpr (NS > 1) {
        const int k = j % NS;
        const C2<T> *tw = a.tw + a.f.st[S].twoff;
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], dirw(tw[(r - 1) * NS + k], dir));
      }
      dft<R>(v[q], dir, a.gen);
    }
  }
  if constexpr (LAST && DIRECT_OUT && OP == OP_R2C_EVEN) {
    static_assert(Q * TPL == NR && NR == NS, "r2c from registers needs full butterflies");
    r2c_store_shuffled<T, N, TPL, Q, R, VOUT>(c, v);
    return;
  } else if constexpr (LAST && DIRECT_OUT) {
    if (c.valid) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int j = c.i + q * TPL;
        if (Q * TPL == NR || j < NR) {
#pragma unroll
          for (int r = 0; r < R; ++r)
            fstore<T, OP, VOUT>(a, c.twl, c.gout, c.aux, j + r * NS, v[q][r]);
        }
      }
    }
    return;
  } else {
    if constexpr (!FIRST || FLAT) __syncthreads();  // every read of this stage is done
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = c.i + q * TPL;
      if (Q * TPL == NR || j < NR) {
        const int k = j % NS;
#pragma unroll
        for (int r = 0; r < R; ++r) c.line[padi<T>((j - k) * R + k + r * NS)] = v[q][r];
      }
    }
    __syncthreads();
    if constexpr (!LAST)
      fstage<T, N, TPL, OP, VIN, VOUT, DIRECT_OUT, SC, S + 1, FLAT>(c);
  }
}

// Split exchange: the same stages, but a line crosses LDS as its real parts
// and then its imaginary parts through a buffer of N scalars per line, so
// that a workgroup needs half the LDS (more workgroups per CU for complex128
// lines of 1024-4096).  The next stage's inputs are read into registers
// between the two halves; four barriers per exchange instead of two.
template <typename SC, int N, int TPL, int S>
struct StageShape {
  static constexpr int R = SC::at(S);
  static constexpr int NS = SC::ns(S);
  static constexpr int NR = N / R;
  static constexpr int Q = (NR + TPL - 1) / TPL;
};

template <typename T, typename SS, int TPL, int PART>
__device__ __forceinline__ void split_write(const FCtx<T> &c, const C2<T> (&v)[SS::Q][SS::R]) {
#pragma unroll
  for (int q = 0; q < SS::Q; ++q) {
    const int j = c.i + q * TPL;
    if (SS::Q * TPL == SS::NR || j < SS::NR) {
      const int k = j % SS::NS;
#pragma unroll
      for (int r = 0; r < SS::R; ++r)
        c.sline[padi<T>((j - k) * SS::R + k + r * SS::NS)] = PART ? v[q][r].y : v[q][r].x;
    }
  }
}

template <typename T, typename SS, int TPL, int PART>
__device__ __forceinline__ void split_read(const FCtx<T> &c, C2<T> (&w)[SS::Q][SS::R]) {
#pragma unroll
  for (int q = 0; q < SS::Q; ++q) {
    const int j = c.i + q * TPL;
    if (SS::Q * TPL == SS::NR || j < SS::NR) {
#pragma unroll
      for (int r = 0; r < SS::R; ++r) {
        const T x = c.sline[padi<T>(j + r * SS::NR)];
        if (PART)
          w[q][r].y = x;
        else
          w[q][r].x = x;
      }
    }
  }
}

template <typename T, int N, int TPL, int OP, bool VIN, bool VOUT, bool DIRECT_OUT,
          typename SC, int S>
__device__ __forceinline__ void fstage_split(
    const FCtx<T> &c, C2<T> (&v)[StageShape<SC, N, TPL, S>::Q][StageShape<SC, N, TPL, S>::R]) {
  using SS = StageShape<SC, N, TPL, S>;
  constexpr int R = SS::R, NS = SS::NS, NR = SS::NR, Q = SS::Q;
  constexpr bool FIRST = S == 0;
  constexpr bool LAST = S == SC::count - 1;
  static_assert(OP != OP_R2C_EVEN, "r2c reads two bins per output: full exchange only");
  const LineArgs<T> &a = *c.a;
  const T dir = c.dir;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = c.i + q * TPL;
    if (Q * TPL == NR || j < NR) {
      if constexpr (FIRST) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          v[q][r] = c.valid ? fload<T, OP, VIN>(c, j + r * NR) : C2<T>{T(0), T(0)};
      }
      if constexpr (NS > 1) {
        const int k = j % NS;
        const C2<T> *tw = a.tw + a.f.st[S].twoff;
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], dirw(tw[(r - 1) * NS + k], dir));
      }
      dft<R>(v[q], dir, a.gen);
    }
  }
  if constexpr (LAST && DIRECT_OUT) {
    if (c.valid) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int j = c.i + q * TPL;
        if (Q * TPL == NR || j < NR) {
#pragma unroll
          for (int r = 0; r < R; ++r)
            fstore<T, OP, VOUT>(a, c.twl, c.gout, c.aux, j + r * NS, v[q][r]);
        }
      }
    }
  } else if constexpr (LAST) {
    // natural-order result, read with the output mapping and stored
    constexpr int NK = N / TPL;
    C2<T> z[NK];
    const bool on = c.lo < c.nl;
    if constexpr (!FIRST) __syncthreads();
    split_write<T, SS, TPL, 0>(c, v);
    __syncthreads();
    if (on) {
#pragma unroll
      for (int kk = 0; kk < NK; ++kk) z[kk].x = c.soline[padi<T>(c.io + kk * TPL)];
    }
    __syncthreads();
    split_write<T, SS, TPL, 1>(c, v);
    __syncthreads();
    if (on) {
#pragma unroll
      for (int kk = 0; kk < NK; ++kk) z[kk].y = c.soline[padi<T>(c.io + kk * TPL)];
      T *g = a.out + c.s_oo[c.lo];
      const int aux = c.s_aux[c.lo];
#pragma unroll
      for (int kk = 0; kk < NK; ++kk)
        fstore<T, OP, VOUT>(a, c.twl, g, aux, c.io + kk * TPL, z[kk]);
    }
  } else {
    using SN = StageShape<SC, N, TPL, S + 1>;
    C2<T> w[SN::Q][SN::R];
    if constexpr (!FIRST) __syncthreads();  // every read of the last exchange is done
    split_write<T, SS, TPL, 0>(c, v);
    __syncthreads();
    split_read<T, SN, TPL, 0>(c, w);
    __syncthreads();
    split_write<T, SS, TPL, 1>(c, v);
    __syncthreads();
    split_read<T, SN, TPL, 1>(c, w);
    fstage_split<T, N, TPL, OP, VIN, VOUT, DIRECT_OUT, SC, S + 1>(c, w);
  }
}

// r2c / c2r of half-length N <= 512 keep their N post/pre twiddles in LDS
// behind the lines (read once per bin; from global memory they cost one
// more vector memory instruction per element)
template <typename T, int N, int OP>
constexpr bool rtw_in_lds() {
  return (OP == OP_R2C_EVEN || OP == OP_C2R_EVEN) && N <= 512;
}

// LDS bytes of the lines of a workgroup (the four-step twiddle tables of
// pass A, when copied, or the real-transform twiddles follow at this offset)
template <typename T, int N, int LPB, bool SPLIT>
constexpr size_t lines_bytes() {
  return SPLIT ? (size_t(LPB) * pitch_of<T, N>() * sizeof(T) + 15) / 16 * 16
               : size_t(LPB) * pitch_of<T, N>() * sizeof(C2<T>);
}

template <typename T, int N, int TPL, int LPB, bool LF_IN, bool LF_OUT, int OP, bool VIN,
          bool VOUT, typename SC, bool SPLIT>
__global__ void __launch_bounds__(TPL *LPB, SPLIT ? 4 : 1) k_fast(const LineArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int P = pitch_of<T, N>();
  // flat mapping for lines of at most 16 (one thread per line): the tile is
  // loaded and stored by consecutive lanes over consecutive elements (one
  // contiguous run per wave instead of one element per lane on lines 128 B
  // apart) through LDS.  Measured 1.6x on complex64 lines of 16; for longer
  // lines (8+ lanes per line) the extra LDS pass and index division cost
  // more than they save.
  constexpr bool FLAT = !LF_IN && !LF_OUT && TPL == 1 && VIN && VOUT && !SPLIT &&
                        (OP == OP_C2C || OP == OP_R2C_EVEN || OP == OP_C2R_EVEN);
  // r2c stores from registers when a line's lanes are adjacent in the wave
  // (mirrored bins by lane shuffles; the last stage covers every slot)
  constexpr bool R2C_REG =
      OP == OP_R2C_EVEN && !LF_IN && !LF_OUT && !SPLIT && TPL >= 2 && TPL <= 64 &&
      StageShape<SC, N, TPL, SC::count - 1>::Q * TPL == StageShape<SC, N, TPL, SC::count - 1>::NR;
  constexpr bool DIRECT_OUT = (LF_IN == LF_OUT) && (OP != OP_R2C_EVEN || R2C_REG) && !FLAT;
  C2<T> *lds = reinterpret_cast<C2<T> *>(smem_raw);
  T *slds = reinterpret_cast<T *>(smem_raw);
  __shared__ int64_t s_oo[LPB];
  __shared__ int64_t s_oi[FLAT ? LPB : 1];
  __shared__ int s_aux[LPB];
  const int t = threadIdx.x;
  const int l = LF_IN ? t % LPB : t / TPL;
  const int i = LF_IN ? t / LPB : t % TPL;
  const int64_t tile = xcd_tile();
  const int64_t b = tile * LPB + l;
  FCtx<T> c;
  c.a = &a;
  c.dir = a.dir;
  c.twl = nullptr;
  if constexpr (OP == OP_PASSA) {
    if (a.twlds_lo > 0) {  // copy the four-step twiddle tables behind the lines
      C2<T> *twl = reinterpret_cast<C2<T> *>(smem_raw + lines_bytes<T, N, LPB, SPLIT>());
      const int nt = a.twlds_lo + a.twlds_hi;
      for (int q = t; q < nt; q += TPL * LPB)
        twl[q] = q < a.twlds_lo ? a.t0[q] : a.t1[q - a.twlds_lo];
      __syncthreads();
      c.twl = twl;
    }
  }
  c.rtw = a.t0;
  if constexpr (rtw_in_lds<T, N, OP>()) {
    C2<T> *rtw = reinterpret_cast<C2<T> *>(smem_raw + lines_bytes<T, N, LPB, SPLIT>());
    for (int q = t; q < N; q += TPL * LPB) rtw[q] = a.t0[q];
    __syncthreads();
    c.rtw = rtw;
  }
  c.line = lds + l * P;
  c.sline = slds + l * P;
  c.valid = b < a.nlines;
  c.i = i;
  const LineOffsets lo_ = group_slot(tile * LPB, l, LPB, a.nlines, a.nb, a.bdim, a.bsi, a.bso,
                                     a.auxdim);
  const int64_t oi = lo_.oi, oo = lo_.oo;
  const int aux = lo_.aux;
  c.gin = a.in + oi;
  c.gout = a.out + oo;
  c.aux = aux;
  if constexpr (!DIRECT_OUT) {
    if (i == 0) {
      s_oo[l] = oo;
      s_aux[l] = aux;
    }
  }
  if constexpr (OP == OP_BLUEROW) {
    // Bluestein rows: forward FFT (global -> natural order in LDS), times
    // the kernel t2 at m = aux + k other, inverse FFT out of LDS, stored in
    // place with the inverse four-step twiddle; every load of a line comes
    // before the barriers that precede its stores
    fstage<T, N, TPL, OP, VIN, VOUT, false, SC, 0, false>(c);
    if (c.valid) {
      const bool cj = (a.sub & SUB_CONJ) != 0;
      for (int k = i; k < N; k += TPL) {
        const C2<T> w = a.t2[aux + int64_t(k) * a.other];
        c.line[padi<T>(k)] = cmul(c.line[padi<T>(k)], cj ? conjc(w) : w);
      }
    }
    __syncthreads();
    c.dir = -a.dir;
    fstage<T, N, TPL, OP, VIN, VOUT, true, SC, 0, true>(c);
    return;
  }
  if constexpr (FLAT) {
    if (i == 0) s_oi[l] = oi;
    __syncthreads();
    const int nl = int(min(int64_t(LPB), a.nlines - tile * LPB));
    constexpr int NIN = OP == OP_C2R_EVEN ? N + 1 : N;  // complex values per line
    for (int f = t; f < nl * NIN; f += TPL * LPB) {
      const int l2 = f / NIN, p = f - l2 * NIN;
      const T *g = a.in + s_oi[l2];
      const int q = OP == OP_C2C ? ((p + a.rot_in) & (N - 1)) : p;
      lds[l2 * P + padi<T>(p)] = ldc<true>(g + (OP == OP_R2C_EVEN ? 2 * p : q * a.esi));
    }
    __syncthreads();
  }
  if constexpr (SPLIT) {
    c.nl = int(min(int64_t(LPB), a.nlines - tile * LPB));
    c.lo = LF_OUT ? t % LPB : t / TPL;
    c.io = LF_OUT ? t / LPB : t % TPL;
    c.soline = slds + c.lo * P;
    c.s_oo = s_oo;
    c.s_aux = s_aux;
    using S0 = StageShape<SC, N, TPL, 0>;
    C2<T> v0[S0::Q][S0::R];
    fstage_split<T, N, TPL, OP, VIN, VOUT, DIRECT_OUT, SC, 0>(c, v0);
  } else {
    // c2r on contiguous lines with one stage-0 butterfly per thread: mirrored
    // bins by lane shuffles
    constexpr bool SHUF = OP == OP_C2R_EVEN && !LF_IN && !FLAT && TPL >= 2 && TPL <= 64 &&
                          StageShape<SC, N, TPL, 0>::NR == TPL;
    fstage<T, N, TPL, OP, VIN, VOUT, DIRECT_OUT, SC, 0, FLAT, SHUF>(c);
  }
  if constexpr (FLAT) {
    // natural-order result of every line is in LDS (the last stage ended
    // with a barrier): consecutive lanes store consecutive outputs
    const int nl = int(min(int64_t(LPB), a.nlines - tile * LPB));
    constexpr int NOUT = OP == OP_R2C_EVEN ? N + 1 : N;
    for (int f = t; f < nl * NOUT; f += TPL * LPB) {
      const int l2 = f / NOUT, k = f - l2 * NOUT;
      T *g = a.out + s_oo[l2];
      if constexpr (OP == OP_R2C_EVEN)
        stc<true>(g + k * a.eso, r2c_bin(a, c.rtw, lds + l2 * P, k));
      else
        fstore<T, OP, true>(a, c.twl, g, s_aux[l2], k, lds[l2 * P + padi<T>(k)]);
    }
  } else if constexpr (!DIRECT_OUT && !SPLIT) {
    // natural-order result of every line is in LDS; store with the output
    // mapping (the last stage ended with a barrier)
    const int nl = int(min(int64_t(LPB), a.nlines - tile * LPB));
    const int lo = LF_OUT ? t % LPB : t / TPL;
    const int io = LF_OUT ? t / LPB : t % TPL;
    if (lo < nl) {
      const C2<T> *ln = lds + lo * P;
      T *g = a.out + s_oo[lo];
      if constexpr (OP == OP_R2C_EVEN) {
        for (int k = io; k <= N; k += TPL) stc<VOUT>(g + k * a.eso, r2c_bin(a, c.rtw, ln, k));
      } else {
        for (int k = io; k < N; k += TPL)
          fstore<T, OP, VOUT>(a, c.twl, g, s_aux[lo], k, ln[padi<T>(k)]);
      }
    }
  }
}

// ------------------------------------------------------------ dispatch

std::mutex g_fast_mu;
std::set<std::pair<int, const void *>> g_fast_attr;  // (device, kernel)

template <typename T, int N, int TPL, int LPB, bool LF_IN, bool LF_OUT, int OP, bool VIN,
          bool VOUT, typename SC, bool SPLIT>
void go(const LineArgs<T> &a0, hipStream_t s) {
  const void *fn = reinterpret_cast<const void *>(
      &k_fast<T, N, TPL, LPB, LF_IN, LF_OUT, OP, VIN, VOUT, SC, SPLIT>);
  size_t lds = lines_bytes<T, N, LPB, SPLIT>();
  if constexpr (rtw_in_lds<T, N, OP>()) lds += size_t(N) * sizeof(C2<T>);
  LineArgs<T> a = a0;
  a.twlds_lo = a.twlds_hi = 0;
  if constexpr (OP == OP_PASSA) {
    // pass A's four-step twiddle tables go to LDS when they fit beside the
    // lines without lowering the number of workgroups per CU (the column
    // stores gather them at scattered indices)
    if (!(a.sub & SUB_FINAL)) {
      const int64_t big = int64_t(N) * a.other;
      const int64_t lo_n = int64_t(1) << a.shift, hi_n = (big + lo_n - 1) / lo_n;
      const size_t extra = size_t(lo_n + hi_n) * sizeof(C2<T>);
      static const size_t static_lds = [fn] {
        hipFuncAttributes fa;
        HIP_CHECK(hipFuncGetAttributes(&fa, fn));
        return size_t(fa.sharedSizeBytes);
      }();
      // only where the copy costs no workgroup slot per CU
      const size_t cap = lds_limit();
      if (lds + extra + static_lds <= cap &&
          cap / (lds + extra + static_lds) == cap / (lds + static_lds)) {
        a.twlds_lo = int(lo_n);
        a.twlds_hi = int(hi_n);
        lds += extra;
      }
    }
  }
  if (lds > 65536) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_fast_mu);
    if (!g_fast_attr.count({dev, fn})) {
      hipFuncAttributes fa;
      HIP_CHECK(hipFuncGetAttributes(&fa, fn));
      HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds_limit() - fa.sharedSizeBytes)));
      g_fast_attr.insert({dev, fn});
    }
  }
  const int64_t groups = (a.nlines + LPB - 1) / LPB;
  hipLaunchKernelGGL((k_fast<T, N, TPL, LPB, LF_IN, LF_OUT, OP, VIN, VOUT, SC, SPLIT>),
                     dim3(unsigned(groups)), dim3(TPL * LPB), lds, s, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    RFH_THROW(2, std::string("fast line kernel launch: ") + hipGetErrorString(e));
}

template <typename T, int N, int TPL, int LPB, int OP, typename SC, bool SPLIT>
bool go_map(const LineArgs<T> &a, bool lf_in, bool lf_out, bool vin, bool vout,
            hipStream_t s) {
#define RFH_GO(LI, LO, VI, VO)                                    \
  if (lf_in == LI && lf_out == LO && vin == VI && vout == VO) {   \
    go<T, N, TPL, LPB, LI, LO, OP, VI, VO, SC, SPLIT>(a, s);      \
    return true;                                                  \
  }
  RFH_GO(false, false, true, true)
  RFH_GO(true, true, true, true)
  RFH_GO(false, true, true, true)
  RFH_GO(true, false, true, true)
#undef RFH_GO
  return false;  // unaligned layouts use the generic kernel
}

// Thread and line shape of the specialised kernel for length N.
template <typename T, int N>
struct FastShape {
  static constexpr int E = 16;  // elements per thread (largest radix)
  static constexpr int TPL = N >= E ? N / E : 1;
  // contiguous lines: 256-thread blocks (one line per block from N = 4096)
  static constexpr int LPB_C = TPL >= 256 ? 1 : 256 / TPL;
  // interleaved lines: 128-byte row segments when the block (threads and
  // LDS) allows it, else at least 8 adjacent lines
  static constexpr int SEG = 128 / int(sizeof(C2<T>));
  static constexpr int LPB_8 = LPB_C >= 8 ? LPB_C : 8;
  static constexpr int LPB_W = LPB_8 >= SEG ? LPB_8 : SEG;
  static constexpr bool fits(int lpb) {
    return TPL * lpb <= 1024 && size_t(lpb) * pitch_of<T, N>() * sizeof(C2<T>) <= 152 * 1024;
  }
  static constexpr int LPB_I = fits(LPB_W) ? LPB_W : LPB_8;
  static constexpr bool interleaved_ok = fits(LPB_I);
  // split exchange (half the LDS, registers held to four waves per SIMD)
  // where it measured faster: complex128 lines of 4096, and interleaved
  // complex128 tiles that would otherwise take a whole CU's LDS; shorter
  // contiguous lines lose more to the extra barriers than they gain
  static constexpr size_t full_lds(int lpb) {
    return size_t(lpb) * pitch_of<T, N>() * sizeof(C2<T>);
  }
  static constexpr bool SPLIT_C = sizeof(T) == 8 && N >= 4096 && full_lds(LPB_C) > 65536;
  static constexpr bool SPLIT_I = sizeof(T) == 8 && full_lds(LPB_I) > 131072;
};

template <typename T, int N, int OP, typename SC>
bool go_len(const LineArgs<T> &a, bool lf_in, bool lf_out, bool vin, bool vout,
            hipStream_t s) {
  using F = FastShape<T, N>;
  constexpr bool r2c = OP == OP_R2C_EVEN;
  if constexpr (OP == OP_BLUEROW) {  // in place: one layout, the full exchange
    if (lf_in != lf_out) return false;
    if (lf_in) {
      if constexpr (F::interleaved_ok)
        return go_map<T, N, F::TPL, F::LPB_I, OP, SC, false>(a, lf_in, lf_out, vin, vout, s);
      return false;
    }
    return go_map<T, N, F::TPL, F::LPB_C, OP, SC, false>(a, lf_in, lf_out, vin, vout, s);
  }
  if (lf_in || lf_out) {
    if constexpr (F::interleaved_ok)
      return go_map<T, N, F::TPL, F::LPB_I, OP, SC, F::SPLIT_I && !r2c>(a, lf_in, lf_out, vin,
                                                                        vout, s);
    return false;
  }
  return go_map<T, N, F::TPL, F::LPB_C, OP, SC, F::SPLIT_C && !r2c>(a, lf_in, lf_out, vin, vout,
                                                                    s);
}

template <typename T>
bool interleaved_ok_t(uint64_t n) {
  switch (n) {
    case 2: return FastShape<T, 2>::interleaved_ok;
    case 4: return FastShape<T, 4>::interleaved_ok;
    case 8: return FastShape<T, 8>::interleaved_ok;
    case 16: return FastShape<T, 16>::interleaved_ok;
    case 32: return FastShape<T, 32>::interleaved_ok;
    case 64: return FastShape<T, 64>::interleaved_ok;
    case 128: return FastShape<T, 128>::interleaved_ok;
    case 256: return FastShape<T, 256>::interleaved_ok;
    case 512: return FastShape<T, 512>::interleaved_ok;
    case 1024: return FastShape<T, 1024>::interleaved_ok;
    case 2048: return FastShape<T, 2048>::interleaved_ok;
    case 4096: return FastShape<T, 4096>::interleaved_ok;
    default: return false;
  }
}

template <typename T, int OP>
bool go_op(const LineArgs<T> &a, bool lf_in, bool lf_out, bool vin, bool vout,
           hipStream_t s) {
  switch (a.f.n) {
    case 2: return go_len<T, 2, OP, Sched<2>>(a, lf_in, lf_out, vin, vout, s);
    case 4: return go_len<T, 4, OP, Sched<4>>(a, lf_in, lf_out, vin, vout, s);
    case 8: return go_len<T, 8, OP, Sched<8>>(a, lf_in, lf_out, vin, vout, s);
    case 16: return go_len<T, 16, OP, Sched<16>>(a, lf_in, lf_out, vin, vout, s);
    case 32: return go_len<T, 32, OP, Sched<16, 2>>(a, lf_in, lf_out, vin, vout, s);
    case 64: return go_len<T, 64, OP, Sched<16, 4>>(a, lf_in, lf_out, vin, vout, s);
    case 128: return go_len<T, 128, OP, Sched<16, 8>>(a, lf_in, lf_out, vin, vout, s);
    case 256: return go_len<T, 256, OP, Sched<16, 16>>(a, lf_in, lf_out, vin, vout, s);
    case 512: return go_len<T, 512, OP, Sched<16, 16, 2>>(a, lf_in, lf_out, vin, vout, s);
    case 1024: return go_len<T, 1024, OP, Sched<16, 16, 4>>(a, lf_in, lf_out, vin, vout, s);
    case 2048: return go_len<T, 2048, OP, Sched<16, 16, 8>>(a, lf_in, lf_out, vin, vout, s);
    case 4096: return go_len<T, 4096, OP, Sched<16, 16, 16>>(a, lf_in, lf_out, vin, vout, s);
    default: return false;
  }
}

}  // namespace

bool g_fast_enabled = true;

bool fast_interleaved_ok(uint64_t n, bool dbl) {
  return dbl ? interleaved_ok_t<double>(n) : interleaved_ok_t<float>(n);
}

// Launches the specialised kernel for (op, length, layout) when one exists.
template <typename T>
bool launch_fast(int op, const LineArgs<T> &a, bool lf_in, bool lf_out, bool vin, bool vout,
                 hipStream_t s) {
  if (!g_fast_enabled) return false;
  switch (op) {
    case OP_C2C: return go_op<T, OP_C2C>(a, lf_in, lf_out, vin, vout, s);
    case OP_PASSA: return go_op<T, OP_PASSA>(a, lf_in, lf_out, vin, vout, s);
    case OP_PASSB: return go_op<T, OP_PASSB>(a, lf_in, lf_out, vin, vout, s);
    case OP_BLUEROW: return go_op<T, OP_BLUEROW>(a, lf_in, lf_out, vin, vout, s);
    case OP_R2C_EVEN: return go_op<T, OP_R2C_EVEN>(a, lf_in, lf_out, vin, vout, s);
    case OP_C2R_EVEN: return go_op<T, OP_C2R_EVEN>(a, lf_in, lf_out, vin, vout, s);
    default: return false;
  }
}

template bool launch_fast<float>(int, const dev::LineArgs<float> &, bool, bool, bool, bool,
                                 hipStream_t);
template bool launch_fast<double>(int, const dev::LineArgs<double> &, bool, bool, bool, bool,
                                  hipStream_t);

}  // namespace rfh
