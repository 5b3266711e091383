// pass1024.hip -- persistent, software-pipelined FP32 four-step passes of
// complex64 lines of 1024 = 16 x 16 x 4: the two passes of the 2^20 =
// 1024 x 1024 plans (column pass A with the four-step twiddle on store, row
// pass B).
//
// A tile is 16 lines in LDS (padded pitch, 139 KB), so one 1024-thread
// workgroup fills a CU; in the one-tile-per-workgroup kernel the load, LDS
// and store phases of a CU are separated by workgroup barriers and never
// overlap.  Here each CU runs one resident workgroup that walks its tiles in
// a loop and issues the global loads of the next tile as soon as the current
// tile's first stage has left its registers, so those loads are in flight
// while the remaining two stages and the stores of the current tile run.
// Global memory goes through buffer instructions: the per-element offsets
// r * 64 * stride are scalar operands (one address register per thread), and
// the hardware range check of the buffer resource bounds every access to the
// line set's span.
//
// Stages (Stockham, natural order out): 0: radix 16 (NS 1) from registers,
// 1: radix 16 (NS 16), 2: radix 4 (NS 256) to global memory; the stage
// twiddles are the plan's, the same formulas as the other line kernels.
#include <algorithm>
#include <mutex>
#include <set>

#include "bufio.hpp"
#include "passes.hpp"

namespace rfh {
namespace {
using namespace dev;

constexpr int kN = 1024, kL = 16, kP = kN + kN / 16 + 1;  // padded line pitch
constexpr int kSt1 = 16 * 16, kSt2 = 256;  // stage twiddles kept in LDS
using bufio::bload;
using bufio::bstore;
using bufio::make_rsrc;

__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }

// The exchanges of stages 0 and 1 stay inside a line.  In pass B a wave owns
// its line there, and a wave's LDS operations complete in issue order, so
// only the compiler must keep them in order; pass A's lines span waves.
template <bool PASS_A>
__device__ __forceinline__ void line_sync() {
  if constexpr (PASS_A) {
    __syncthreads();
  } else {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
  }
}

// in_bytes / out_bytes: spans of the line set from a.in / a.out (all
// offsets non-negative, below 2^31); ntiles = ceil(nlines / 16).  ROT: the
// line set carries element rotations (pass A loads, pass B stores), whose
// per-element offsets then live in vector registers.
template <bool PASS_A, bool ROT>
__global__ void __launch_bounds__(1024, 1)
    k_pass1024(const LineArgs<float> a, int64_t ntiles, uint32_t in_bytes, uint32_t out_bytes) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<float> *lds = reinterpret_cast<C2<float> *>(smem_raw);
  // line offsets of the tiles in flight: a ring of three (current, next,
  // the one after), so that a slot is rewritten only after every reader
  // of its previous tile passed a barrier
  __shared__ int64_t s_oi[3][kL], s_oo[3][kL];
  __shared__ int s_aux[3][kL];
  const int t = threadIdx.x;
  // stages 0 and 1: pass A lines interleave in memory (lane t % 16 takes
  // line t % 16: 128-byte row segments); pass B lines are contiguous (a
  // wave per line: 512-byte runs)
  const int l0 = PASS_A ? (t & 15) : (t >> 6);
  const int i0 = PASS_A ? (t >> 4) : (t & 63);
  // stage 2 and the stores: interleaved on the output side in both passes
  const int l2 = t & 15, i2 = t >> 4;
  const int64_t G = gridDim.x;
  const int64_t b = blockIdx.x;
  // tiles k * G + slot: the workgroups of one XCD (b mod 8) take adjacent
  // tiles, so shared cache lines stay in one L2
  const int64_t slot0 = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  if (slot0 >= ntiles) return;  // whole workgroup, before any barrier
  auto tile_of = [&](int64_t k) { return k * G + slot0; };
  auto fill = [&](int64_t tile, int s) {
    if (t < kL) {
      const LineOffsets o =
          group_slot(tile * kL, t, kL, a.nlines, a.nb, a.bdim, a.bsi, a.bso, a.auxdim);
      s_oi[s][t] = o.oi;
      s_oo[s][t] = o.oo;
      s_aux[s][t] = o.aux;
    }
  };
  // behind the lines, with the direction applied: stage 1's twiddles
  // W_256^(r k) ([r 16 + k], r, k < 16), stage 2's W_1024^j (j < 256, whose
  // squares and cubes give its r = 2, 3 factors), then (pass A) the
  // four-step twiddle's lo / hi tables.  No global load runs between a
  // prefetch and its use: the memory counter completes in order, so waiting
  // for any later load would also wait for the whole prefetch.
  const float dir = a.dir;
  C2<float> *st1l = lds + kL * kP;
  C2<float> *st2l = st1l + kSt1;
  C2<float> *twl = st2l + kSt2;
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(a.in, in_bytes);
  const __amdgpu_buffer_rsrc_t rout = make_rsrc(a.out, out_bytes);
  const uint32_t es_in = uint32_t(a.esi) * 4u, es_out = uint32_t(a.eso) * 4u;
  fill(tile_of(0), 0);
  if (tile_of(1) < ntiles) fill(tile_of(1), 1);
  __syncthreads();

  C2<float> v[16];
  // inputs x[i0 + 64 r] of line l0 of the tile in slot s; a missing line
  // reads past the span (zeros from the range check)
  auto load_tile = [&](int s, int64_t tile) {
    const bool ok = tile * kL + l0 < a.nlines;
    if constexpr (ROT) {  // element (i0 + 64 r + rot_in) mod 1024
      const uint32_t base = ok ? uint32_t(s_oi[s][l0]) * 4u : in_bytes;
#pragma unroll
      for (int r = 0; r < 16; ++r)
        v[r] = bload(rin, ok ? base + uint32_t((i0 + 64 * r + a.rot_in) & (kN - 1)) * es_in : base,
                     0u);
    } else {
      const uint32_t vo = ok ? uint32_t(s_oi[s][l0]) * 4u + uint32_t(i0) * es_in : in_bytes;
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = bload(rin, vo, uint32_t(r) * 64u * es_in);
    }
  };
  load_tile(0, tile_of(0));
  // the tables' loads queue behind the first tile's, so both are in flight
  // together (the first stage waits for all of them anyway)
  {
    const C2<float> *tw1 = a.tw + a.f.st[1].twoff;
    const C2<float> *tw2 = a.tw + a.f.st[2].twoff;
    for (int q = t; q < kSt1 + kSt2; q += 1024)
      st1l[q] = q < 16 ? C2<float>{1.f, 0.f} : dirw(q < kSt1 ? tw1[q - 16] : tw2[q - kSt1], dir);
  }
  if (PASS_A) {
    const int nt = a.twlds_lo + a.twlds_hi;
    for (int q = t; q < nt; q += 1024) twl[q] = q < a.twlds_lo ? a.t0[q] : a.t1[q - a.twlds_lo];
  }
  __syncthreads();

  for (int64_t k = 0;; ++k) {
    // an opaque zero in the stage-twiddle indices keeps their reads in the
    // loop: hoisted out, they would hold 30+ registers the pipeline needs
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    const int s = int(k % 3);
    const int64_t tile = tile_of(k);
    // ---- stage 0 (radix 16, NS 1) from registers into LDS.  Output r of
    // butterfly i0 lands at position 16 i0 + r = (16 (i0 mod 4) + r) +
    // 64 (i0 / 4): stage 1's input i0 / 4 of its butterfly 16 (i0 mod 4) + r,
    // whose twiddle W_256^((i0 / 4) r) is applied here, while no prefetch
    // holds registers
    dft<16>(v, dir, static_cast<const C2<float> *>(nullptr));
    {
      C2<float> *ln = lds + l0 * kP;
      const C2<float> *tw = st1l + (i0 >> 2) * 16 + z;  // row 0: ones
      ln[pad(i0 * 16)] = v[0];
#pragma unroll
      for (int r = 1; r < 16; ++r) ln[pad(i0 * 16 + r)] = cmul(v[r], tw[r]);
    }
    // offsets of tile k + 2 (read after this iteration's barriers)
    if (tile_of(k + 2) < ntiles) fill(tile_of(k + 2), int((k + 2) % 3));
    // loads of tile k + 1 go out now and land during stages 1 and 2
    const bool more = tile_of(k + 1) < ntiles;
    if (more) load_tile(int((k + 1) % 3), tile_of(k + 1));
    line_sync<PASS_A>();
    // ---- stage 1 (radix 16, NS 16).  Output k of butterfly j goes back to
    // the slot of its input k (j + 64 k), which only this thread reads, so
    // no barrier separates the reads from the writes; stage 2 reads the
    // Stockham order (j / 16) 256 + j % 16 + 16 k through that map.
    {
      C2<float> w[16];
      C2<float> *ln = lds + l0 * kP;
#pragma unroll
      for (int r = 0; r < 16; ++r) w[r] = ln[pad(i0 + 64 * r)];  // twiddled in stage 0
      dft<16>(w, dir, static_cast<const C2<float> *>(nullptr));
#pragma unroll
      for (int r = 0; r < 16; ++r) ln[pad(i0 + 64 * r)] = w[r];
    }
    __syncthreads();
    // ---- stage 2 (radix 4, NS 256), natural-order bins to global memory
    if (tile * kL + l2 < a.nlines) {
      const C2<float> *ln = lds + l2 * kP;
      const uint32_t vo = uint32_t(s_oo[s][l2]) * 4u + uint32_t(i2) * es_out;
      const int aux = s_aux[s][l2];
#pragma unroll 1  // unrolled, the four groups hold more registers than are left
      for (int q = 0; q < 4; ++q) {
        const int j = i2 + 64 * q;
        C2<float> u[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) u[r] = ln[pad(16 * r + (j & 15) + 64 * (j >> 4))];
        const C2<float> w1 = st2l[j + z], w2 = cmul(w1, w1), w3 = cmul(w2, w1);
        u[1] = cmul(u[1], w1);
        u[2] = cmul(u[2], w2);
        u[3] = cmul(u[3], w3);
        dft4(u, dir);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int kb = j + 256 * r;
          C2<float> x;
          if constexpr (PASS_A) {  // W_N^e from the LDS tables, e < 2^31
            const unsigned e = unsigned(kb) * unsigned(aux);
            const unsigned mask = (1u << a.shift) - 1u;
            const C2<float> w = cmul(twl[e & mask], twl[a.twlds_lo + int(e >> a.shift)]);
            x = cmul(u[r], dirw(w, dir));
          } else {
            x = a.fct * u[r];
          }
          if constexpr (ROT)  // bin (kb + rot_out) mod 1024
            bstore(rout, vo - uint32_t(i2) * es_out + uint32_t((kb + a.rot_out) & (kN - 1)) * es_out,
                   0u, x);
          else
            bstore(rout, vo, uint32_t(64 * q + 256 * r) * es_out, x);
        }
      }
    }
    if (!more) break;
    __syncthreads();  // LDS lines and this slot are free for the next tile
  }
}

std::mutex g_mu;
std::set<std::pair<int, const void *>> g_attr;  // (device, kernel) with the LDS limit raised

}  // namespace

bool launch_pass1024(int op, const LineArgs<float> &a0, bool lf_in, bool lf_out, bool vin,
                     bool vout, hipStream_t s) {
  if (!g_fast_enabled || a0.f.n != kN || a0.f.nst != 3) return false;
  if (a0.f.st[0].radix != 16 || a0.f.st[1].radix != 16 || a0.f.st[2].radix != 4) return false;
  if (!vin || !vout || a0.sub != 0) return false;
  const bool pass_a = op == OP_PASSA;
  // rotations: pass A rotates its loads, pass B its stores (the four-step
  // split of a line set's rotation), each by less than 1024
  if ((pass_a ? a0.rot_out : a0.rot_in) != 0 || a0.rot_in < 0 || a0.rot_in >= kN ||
      a0.rot_out < 0 || a0.rot_out >= kN)
    return false;
  const bool rot = (pass_a ? a0.rot_in : a0.rot_out) != 0;
  if (pass_a ? !(lf_in && lf_out) : !(op == OP_PASSB && !lf_in && lf_out)) return false;
  const int64_t si = bufio::line_span(a0, kN, true), so = bufio::line_span(a0, kN, false);
  if (si < 0 || so < 0 || si * 4 >= (int64_t(1) << 31) - (int64_t(8) << 20) ||
      so * 4 >= (int64_t(1) << 31))
    return false;
  LineArgs<float> a = a0;
  const int64_t ntiles = (a.nlines + kL - 1) / kL;
  if (ntiles <= 0) return true;
  size_t lds = size_t(kL * kP + kSt1 + kSt2) * sizeof(C2<float>);
  a.twlds_lo = a.twlds_hi = 0;
  if (pass_a) {  // the four-step twiddle tables go to LDS; no room: decline
    const int64_t big = int64_t(kN) * a.other;
    const int64_t lo_n = int64_t(1) << a.shift, hi_n = (big + lo_n - 1) / lo_n;
    const size_t extra = size_t(lo_n + hi_n) * sizeof(C2<float>);
    if (big >= (int64_t(1) << 31) || lds + extra + 2048 > lds_limit()) return false;
    a.twlds_lo = int(lo_n);
    a.twlds_hi = int(hi_n);
    lds += extra;
  }
  const void *fn =
      pass_a ? (rot ? reinterpret_cast<const void *>(&k_pass1024<true, true>)
                    : reinterpret_cast<const void *>(&k_pass1024<true, false>))
             : (rot ? reinterpret_cast<const void *>(&k_pass1024<false, true>)
                    : reinterpret_cast<const void *>(&k_pass1024<false, false>));
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
  int64_t grid = std::min<int64_t>(ntiles, bufio::device_cus());
  if (grid >= 8) grid -= grid % 8;
  const uint32_t in_bytes = uint32_t(si * 4), out_bytes = uint32_t(so * 4);
  const dim3 gd = dim3(unsigned(grid)), bd = dim3(1024);
  if (pass_a && rot)
    hipLaunchKernelGGL((k_pass1024<true, true>), gd, bd, lds, s, a, ntiles, in_bytes, out_bytes);
  else if (pass_a)
    hipLaunchKernelGGL((k_pass1024<true, false>), gd, bd, lds, s, a, ntiles, in_bytes, out_bytes);
  else if (rot)
    hipLaunchKernelGGL((k_pass1024<false, true>), gd, bd, lds, s, a, ntiles, in_bytes, out_bytes);
  else
    hipLaunchKernelGGL((k_pass1024<false, false>), gd, bd, lds, s, a, ntiles, in_bytes,
                       out_bytes);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) RFH_THROW(2, std::string("pass1024 launch: ") + hipGetErrorString(e));
  return true;
}

}  // namespace rfh
