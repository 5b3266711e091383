// pipe4096.hip -- persistent, software-pipelined FP32 transform of contiguous
// complex64 lines of 4096 = 16 x 16 x 16 (the direct plan's line kernel).
//
// The one-line-per-workgroup kernel (256 threads, 34 KB of LDS, four per CU)
// leaves each workgroup's load, LDS and store phases back to back.  Here one
// 1024-thread workgroup per CU holds a tile of four lines (a line per 256
// threads, 136 KB) and walks its tiles in a loop; the loads of the next tile
// go out right after the current tile's first stage, so they are in flight
// while the last two stages and the stores of the current tile run.  Global
// memory goes through buffer instructions (scalar per-element offsets, the
// range check bounding every access to the line set's span).
//
// Stages (Stockham, natural order out): 0: radix 16 (NS 1) from registers,
// 1: radix 16 (NS 16), 2: radix 16 (NS 256) straight to global memory.  The
// twiddles come from LDS or registers, so that no global load runs between
// a prefetch and its use (the memory counter completes in order), and each
// stage's twiddles are applied as the stage before writes its outputs:
// stage 1's W_256^(r k) from the plan's table in LDS; stage 2's
// W_4096^(j2 r2) on stage 1's output q of butterfly j (j2 = j mod 16 + 16 q,
// r2 = j / 16) as A_j W_256^(q (j / 16)), A_j = W_4096^((j mod 16)(j / 16))
// held in one register pair per thread (from the plan's table) and W_256
// from the same LDS table.
#include <algorithm>
#include <mutex>
#include <set>

#include "bufio.hpp"
#include "passes.hpp"

namespace rfh {
namespace {
using namespace dev;

constexpr int kN = 4096, kL = 4, kT = 256;   // line length, lines per tile, threads per line
constexpr int kP = kN + kN / 16 + 1;         // padded line pitch
constexpr int kSt1 = 256;                    // stage twiddles kept in LDS
using bufio::bload;
using bufio::bstore;
using bufio::make_rsrc;

__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }

// in_bytes / out_bytes: spans of the line set from a.in / a.out (all
// offsets non-negative, below 2^31); ntiles = ceil(nlines / 4).  ROT: the
// line set carries element rotations (loads by rot_in, stores by rot_out),
// whose per-element offsets then live in vector registers.
template <bool ROT>
__global__ void __launch_bounds__(1024, 1)
    k_pipe4096(const LineArgs<float> a, int64_t ntiles, uint32_t in_bytes, uint32_t out_bytes) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<float> *lds = reinterpret_cast<C2<float> *>(smem_raw);
  // line offsets of the tiles in flight: a ring of three (current, next,
  // the one after), so that a slot is rewritten only after every reader
  // of its previous tile passed a barrier
  __shared__ int64_t s_oi[3][kL], s_oo[3][kL];
  const int t = threadIdx.x, l = t >> 8, i = t & (kT - 1);
  const int64_t G = gridDim.x;
  const int64_t b = blockIdx.x;
  // tiles k * G + slot: the workgroups of one XCD (b mod 8) take adjacent
  // tiles
  const int64_t slot0 = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  if (slot0 >= ntiles) return;  // whole workgroup, before any barrier
  auto tile_of = [&](int64_t k) { return k * G + slot0; };
  auto fill = [&](int64_t tile, int s) {
    if (t < kL) {
      const LineOffsets o =
          group_slot(tile * kL, t, kL, a.nlines, a.nb, a.bdim, a.bsi, a.bso, a.auxdim);
      s_oi[s][t] = o.oi;
      s_oo[s][t] = o.oo;
    }
  };
  // behind the lines, with the direction applied: st1[r 16 + k] =
  // W_256^(r k) (r, k < 16; row 0 ones).  Stage 2's butterfly j2 = (j mod
  // 16) + 16 q takes stage 1's output q of butterfly j as its input j / 16,
  // so its twiddle W_4096^(j2 (j / 16)) = A_j W_256^(q (j / 16)) with
  // A_j = W_4096^((j mod 16)(j / 16)), the (e < 256) first row of stage 2's
  // table.
  const float dir = a.dir;
  C2<float> *st1l = lds + kL * kP;
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(a.in, in_bytes);
  const __amdgpu_buffer_rsrc_t rout = make_rsrc(a.out, out_bytes);
  const uint32_t es_in = uint32_t(a.esi) * 4u, es_out = uint32_t(a.eso) * 4u;
  fill(tile_of(0), 0);
  if (tile_of(1) < ntiles) fill(tile_of(1), 1);
  __syncthreads();

  C2<float> v[16];
  // inputs x[i + 256 r] of line l of the tile in slot s; a missing line
  // reads past the span (zeros from the range check)
  auto load_tile = [&](int s, int64_t tile, int z) {
    const bool ok = tile * kL + l < a.nlines;
    if constexpr (ROT) {  // element (i + 256 r + rot_in) mod 4096
      const uint32_t base = ok ? uint32_t(s_oi[s][l]) * 4u : in_bytes;
#pragma unroll
      for (int r = 0; r < 16; ++r)
        v[r] = bload(rin, ok ? base + uint32_t((i + kT * r + a.rot_in + z) & (kN - 1)) * es_in : base,
                     0u);
    } else {
      const uint32_t vo = ok ? uint32_t(s_oi[s][l]) * 4u + uint32_t(i) * es_in : in_bytes;
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = bload(rin, vo, uint32_t(r) * uint32_t(kT) * es_in);
    }
  };
  load_tile(0, tile_of(0), 0);
  // the tables' loads queue behind the first tile's, so both are in flight
  // together (the first stage waits for all of them anyway)
  {
    const C2<float> *tw1 = a.tw + a.f.st[1].twoff;
    for (int q = t; q < kSt1; q += 1024)
      st1l[q] = q < 16 ? C2<float>{1.f, 0.f} : dirw(tw1[q - 16], dir);
  }
  const C2<float> aj = dirw(a.tw[a.f.st[2].twoff + (i & 15) * (i >> 4)], dir);
  __syncthreads();

  for (int64_t k = 0;; ++k) {
    // an opaque zero in the twiddle and rotated-element indices keeps their
    // reads and offsets in the loop: hoisted out, they would hold registers
    // the pipeline needs
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    const int s = int(k % 3);
    const int64_t tile = tile_of(k);
    C2<float> *ln = lds + l * kP;
    // ---- stage 0 (radix 16, NS 1) from registers into LDS.  Output r of
    // butterfly i lands at 16 i + r = (16 (i mod 16) + r) + 256 (i / 16):
    // stage 1's input i / 16 of its butterfly 16 (i mod 16) + r, whose
    // twiddle W_256^((i / 16) r) is applied here
    dft<16>(v, dir, static_cast<const C2<float> *>(nullptr));
    {
      const C2<float> *tw = st1l + (i >> 4) * 16 + z;  // row 0: ones
      ln[pad(i * 16)] = v[0];
#pragma unroll
      for (int r = 1; r < 16; ++r) ln[pad(i * 16 + r)] = cmul(v[r], tw[r]);
    }
    // offsets of tile k + 2 (read after this iteration's barriers)
    if (tile_of(k + 2) < ntiles) fill(tile_of(k + 2), int((k + 2) % 3));
    // loads of tile k + 1 go out now and land during stages 1 and 2
    const bool more = tile_of(k + 1) < ntiles;
    if (more) load_tile(int((k + 1) % 3), tile_of(k + 1), z);
    __syncthreads();
    // ---- stage 1 (radix 16, NS 16).  Output q of butterfly i goes back to
    // the slot of its input q (i + 256 q), which only this thread reads;
    // stage 2 reads the Stockham order (i / 16) 256 + i % 16 + 16 q through
    // that map.  Stage 2's twiddle A_i W_256^(q (i / 16)) is applied here.
    {
      C2<float> w[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) w[r] = ln[pad(i + kT * r)];  // twiddled in stage 0
      dft<16>(w, dir, static_cast<const C2<float> *>(nullptr));
      const C2<float> *tw = st1l + (i >> 4) * 16 + z;  // row 0: ones
      ln[pad(i)] = cmul(w[0], aj);
#pragma unroll
      for (int r = 1; r < 16; ++r) ln[pad(i + kT * r)] = cmul(w[r], cmul(aj, tw[r]));
    }
    __syncthreads();
    // ---- stage 2 (radix 16, NS 256): butterfly i takes Stockham inputs
    // i + 256 r (twiddled in stage 1) and stores bins i + 256 r
    if (tile * kL + l < a.nlines) {
      C2<float> u[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) u[r] = ln[pad(16 * r + (i & 15) + kT * (i >> 4))];
      dft<16>(u, dir, static_cast<const C2<float> *>(nullptr));
      const uint32_t vo = uint32_t(s_oo[s][l]) * 4u;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const C2<float> x = a.fct * u[r];
        if constexpr (ROT)  // bin (i + 256 r + rot_out) mod 4096
          bstore(rout, vo + uint32_t((i + kT * r + a.rot_out + z) & (kN - 1)) * es_out, 0u, x);
        else
          bstore(rout, vo + uint32_t(i) * es_out, uint32_t(kT * r) * es_out, x);
      }
    }
    if (!more) break;
    __syncthreads();  // LDS lines and this slot are free for the next tile
  }
}

std::mutex g_mu;
std::set<std::pair<int, const void *>> g_attr;  // (device, kernel) with the LDS limit raised

}  // namespace

bool launch_pipe4096(int op, const LineArgs<float> &a0, bool lf_in, bool lf_out, bool vin,
                     bool vout, hipStream_t s) {
  if (!g_fast_enabled || op != OP_C2C || a0.f.n != kN || a0.f.nst != 3) return false;
  if (a0.f.st[0].radix != 16 || a0.f.st[1].radix != 16 || a0.f.st[2].radix != 16) return false;
  if (lf_in || lf_out || !vin || !vout || a0.sub != 0) return false;
  if (a0.rot_in < 0 || a0.rot_in >= kN || a0.rot_out < 0 || a0.rot_out >= kN) return false;
  const int64_t ntiles = (a0.nlines + kL - 1) / kL;
  const int cus = bufio::device_cus();
  // a few tiles per CU at least: smaller sets run faster one line per
  // workgroup
  if (ntiles < 2 * int64_t(cus)) return false;
  const int64_t si = bufio::line_span(a0, kN, true), so = bufio::line_span(a0, kN, false);
  if (si < 0 || so < 0 || si * 4 >= (int64_t(1) << 31) - (int64_t(8) << 20) ||
      so * 4 >= (int64_t(1) << 31))
    return false;
  const bool rot = a0.rot_in != 0 || a0.rot_out != 0;
  const void *fn = rot ? reinterpret_cast<const void *>(&k_pipe4096<true>)
                       : reinterpret_cast<const void *>(&k_pipe4096<false>);
  const size_t lds = size_t(kL * kP + kSt1) * sizeof(C2<float>);
  {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_attr.count({dev, fn})) {
      hipFuncAttributes fa;
      HIP_CHECK(hipFuncGetAttributes(&fa, fn));
      HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds_limit() - fa.sharedSizeBytes)));
      g_attr.insert({dev, fn});
    }
  }
  if (lds + 2048 > lds_limit()) return false;
  int64_t grid = std::min<int64_t>(ntiles, cus);
  if (grid >= 8) grid -= grid % 8;
  const uint32_t in_bytes = uint32_t(si * 4), out_bytes = uint32_t(so * 4);
  const dim3 gd = dim3(unsigned(grid)), bd = dim3(1024);
  if (rot)
    hipLaunchKernelGGL((k_pipe4096<true>), gd, bd, lds, s, a0, ntiles, in_bytes, out_bytes);
  else
    hipLaunchKernelGGL((k_pipe4096<false>), gd, bd, lds, s, a0, ntiles, in_bytes, out_bytes);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) RFH_THROW(2, std::string("pipe4096 launch: ") + hipGetErrorString(e));
  return true;
}

}  // namespace rfh
