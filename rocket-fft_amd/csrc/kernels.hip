// kernels.hip -- gfx950 kernels of the transform engine.
//
// The workhorse is k_lines: one workgroup owns L whole lines, loads them
// (coalesced: consecutive lanes take consecutive elements of a line, or the
// same element of adjacent lines when the lines are interleaved in memory),
// runs every Stockham stage of the line FFT out of LDS (two ping-pong
// buffers, one barrier per stage), and stores.  Per-operation load/store
// hooks fuse the reference's pre/post processing into the same pass:
//   * c2c with the output factor fct                (ExecC2C, H:3646-3659)
//   * real -> n/2+1 bins and back through a half-length complex FFT
//     (general_r2c / general_c2r, H:3723-3852)
//   * Bluestein chirp, kernel product and second FFT (fftblue, H:2723-2762)
//   * the two passes of a four-step transform for lines longer than LDS,
//     with the inter-pass twiddle and the Bluestein steps folded in.
// (H = rocket_fft/_pocketfft_hdronly.h.)  Remaining real transforms (DCT/DST,
// Hartley, FFTPACK halfcomplex and real lengths without a fused plan) use
// k_pre -> complex lines -> k_post around the same c2c machinery.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "bufio.hpp"
#include "passes.hpp"

namespace rfh {
namespace {
using namespace dev;

thread_local const char *t_last_family = "k_lines";  // kernel family of the last launch

template <typename T, int OP, bool VIN, bool VOUT, int NT = kThreads>
__global__ void __launch_bounds__(NT) k_lines(const LineArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int64_t s_oi[kMaxLinesPerGroup], s_oo[kMaxLinesPerGroup];
  __shared__ int s_aux[kMaxLinesPerGroup];
  line_body<T, OP, VIN, VOUT>(a, int64_t(xcd_tile()), smem_raw, s_oi, s_oo, s_aux);
}

// ----------------------------------------------------------- host launch

int pow2floor(int64_t v) {
  int p = 1;
  while (int64_t(p) * 2 <= v) p *= 2;
  return p;
}

int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// lines per group: the largest power of two whose two LDS buffers fit
// `budget` (and are not wasted on missing lines)
int choose_lines(int ld, size_t elem, int64_t nlines, bool interleaved) {
  // interleaved lines want as many adjacent lines as LDS holds (row
  // segments); contiguous lines leave room for several groups per CU
  const size_t budget = interleaved ? lds_limit() - kLineStaticLds
                                    : std::min<size_t>(48 * 1024, lds_limit() - kLineStaticLds);
  int L = 1;
  while (L * 2 <= kMaxLinesPerGroup && size_t(2) * (L * 2) * ld * elem <= budget &&
         int64_t(L) < nlines)
    L *= 2;
  // fewer lines per group when the groups would leave CUs idle (down to one
  // line for contiguous lines, two adjacent lines for interleaved ones)
  const int floor_l = interleaved ? 2 : 1;
  while (L > floor_l && nlines < int64_t(L) * 512) L /= 2;
  return L;
}

bool lfast_for(int64_t es, const LineSet &g, bool input) {
  if (g.nb == 0) return false;
  const int64_t bs = input ? g.bsi[g.nb - 1] : g.bso[g.nb - 1];
  return std::llabs(es) > std::llabs(bs);
}

// complex accesses may use 2-wide loads when base and every stride are even
bool pairs_aligned(const void *p, size_t scalar, int64_t es, const int64_t *bs, int nb) {
  if (reinterpret_cast<uintptr_t>(p) % (2 * scalar)) return false;
  if (es % 2) return false;
  for (int d = 0; d < nb; ++d)
    if (bs[d] % 2) return false;
  return true;
}

std::mutex g_attr_mu;
std::set<std::pair<int, const void *>> g_attr_done;

// Raises a kernel's dynamic-LDS ceiling to what the device allows beside the
// kernel's static LDS (once per kernel and device).
void allow_lds(const void *fn, size_t bytes) {
  if (bytes <= 65536) return;
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_attr_mu);
  if (g_attr_done.count({dev, fn})) return;
  hipFuncAttributes fa;
  HIP_CHECK(hipFuncGetAttributes(&fa, fn));
  const int dyn = int(lds_limit()) - int(fa.sharedSizeBytes);
  if (int64_t(bytes) > dyn)
    RFH_THROW(3, "line kernel needs " + std::to_string(bytes) + " B of LDS beside " +
                     std::to_string(fa.sharedSizeBytes) + " B static; device allows " +
                     std::to_string(lds_limit()));
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
  if (e != hipSuccess)
    RFH_THROW(2, std::string("hipFuncSetAttribute(MaxDynamicSharedMemorySize, ") +
                     std::to_string(dyn) + "): " + hipGetErrorString(e));
  g_attr_done.insert({dev, fn});
}

// Threads of the wide line workgroups: 1024 for complex64 (the float line
// kernels stay within the 128 registers that allows), 512 for complex128
// (within 256).
template <typename T>
constexpr int kWideThreads = sizeof(T) == 4 ? 1024 : 512;

template <typename T, int OP, bool VIN, bool VOUT>
void launch_one(const LineArgs<T> &a, size_t lds, hipStream_t s) {
  const int64_t groups = (a.nlines + a.L - 1) / a.L;
  if (groups > 0x7fffffff / kThreads) RFH_THROW(3, "batch too large for one launch");
  bool wide = false;
  if constexpr (OP == OP_C2C || OP == OP_BLUE) {
    // At most one group per CU, the launch is bound by each group's stage
    // latency (four waves per CU at the narrow width): a wider group runs a
    // stage's butterflies in fewer rounds, as long as every thread still has
    // two (complex64) or four (complex128) elements of the group; below that
    // the wider barriers cost more than they save (256 lines of complex128
    // 1536: 10.1 us narrow, 10.8 us wide).
    constexpr int64_t kMinElems = (sizeof(T) == 4 ? 2 : 4) * kWideThreads<T>;
    wide = groups <= bufio::device_cus() && int64_t(a.L) * a.ld >= kMinElems;
    if (wide) {
      allow_lds(reinterpret_cast<const void *>(&k_lines<T, OP, VIN, VOUT, kWideThreads<T>>), lds);
      hipLaunchKernelGGL((k_lines<T, OP, VIN, VOUT, kWideThreads<T>>), dim3(unsigned(groups)),
                         dim3(kWideThreads<T>), lds, s, a);
    }
  }
  if (!wide) {
    allow_lds(reinterpret_cast<const void *>(&k_lines<T, OP, VIN, VOUT>), lds);
    hipLaunchKernelGGL((k_lines<T, OP, VIN, VOUT>), dim3(unsigned(groups)), dim3(kThreads), lds,
                       s, a);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    RFH_THROW(2, std::string("line kernel launch (") + std::to_string(groups) + " groups, " +
                     std::to_string(lds) + " B LDS, op " + std::to_string(OP) +
                     "): " + hipGetErrorString(e));
}

template <typename T, int OP>
void launch_var(const LineArgs<T> &a, size_t lds, bool vin, bool vout, hipStream_t s) {
  if (vin && vout) launch_one<T, OP, true, true>(a, lds, s);
  else if (vin) launch_one<T, OP, true, false>(a, lds, s);
  else if (vout) launch_one<T, OP, false, true>(a, lds, s);
  else launch_one<T, OP, false, false>(a, lds, s);
}

}  // namespace

template <typename T>
LineArgs<T> make_line_args(const PassSpec &ps, const LineSet &g) {
  LineArgs<T> a{};
  a.nlines = g.nlines;
  a.nb = g.nb;
  for (int d = 0; d < g.nb; ++d) {
    a.bdim[d] = g.bdim[d];
    a.bsi[d] = g.bsi[d];
    a.bso[d] = g.bso[d];
  }
  a.esi = g.esi;
  a.eso = g.eso;
  a.auxdim = g.auxdim >= 0 ? g.auxdim : g.nb - 1;
  a.in = static_cast<const T *>(g.in);
  a.out = static_cast<T *>(g.out);
  a.f = ps.d->desc;
  a.ld = a.f.n;
  a.nin = ps.nin;
  a.nout = ps.nout;
  a.lfast_in = lfast_for(g.esi, g, true);
  a.lfast_out = lfast_for(g.eso, g, false);
  const size_t elem = 2 * sizeof(T);
  a.L = ps.lines ? ps.lines
                 : choose_lines(a.ld, elem, g.nlines, a.lfast_in || a.lfast_out);
  a.log2L = ilog2(a.L);
  a.tw = static_cast<const C2<T> *>(ps.d->tw.p);
  a.gen = static_cast<const C2<T> *>(ps.d->gen.p);
  a.dir = T(ps.dir);
  a.fct = T(ps.fct);
  a.sub = ps.sub;
  a.bign = ps.bign;
  a.other = ps.other;
  a.shift = ps.shift;
  a.t0 = static_cast<const C2<T> *>(ps.t0);
  a.t1 = static_cast<const C2<T> *>(ps.t1);
  a.t2 = static_cast<const C2<T> *>(ps.t2);
  a.rot_in = int(g.rot_in);
  a.rot_out = int(g.rot_out);
  return a;
}

template LineArgs<float> make_line_args<float>(const PassSpec &, const LineSet &);
template LineArgs<double> make_line_args<double>(const PassSpec &, const LineSet &);

namespace {

// The elementwise Bluestein passes (chirp in, kernel product, chirp out) of
// the multi-pass form, one grid row per line.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_blue_elem(const LineArgs<T> a, int op) {
  const bool cj = (a.sub & SUB_CONJ) != 0;
  for (int64_t b = blockIdx.y; b < a.nlines; b += gridDim.y) {
    int64_t rem = b, oi = 0, oo = 0;
    for (int d = a.nb - 1; d >= 0; --d) {
      const int64_t q = rem / a.bdim[d];
      const int64_t idx = rem - q * a.bdim[d];
      rem = q;
      oi += idx * a.bsi[d];
      oo += idx * a.bso[d];
    }
    const T *x = a.in + oi;
    T *y = a.out + oo;
    for (int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; m < a.other;
         m += int64_t(gridDim.x) * blockDim.x) {
      C2<T> v = {T(0), T(0)};
      int64_t mo = m;  // output element
      if (op == OP_BLUE_IN) {
        if (m < a.bign) {  // element (m + rot_in) mod n of the line, times the chirp at m
          int64_t q = m + a.rot_in;
          if (q >= a.bign) q -= a.bign;
          const C2<T> c = a.t0[m];
          v = cmul(ldc<false>(x + q * a.esi), cj ? c : conjc(c));
        }
      } else if (op == OP_BLUE_MUL) {
        const C2<T> w = a.t1[m];
        v = cmul(ldc<false>(x + m * a.esi), cj ? conjc(w) : w);
      } else {  // bin m, chirped, to (m + rot_out) mod n
        const C2<T> c = a.t0[m];
        v = a.fct * cmul(ldc<false>(x + m * a.esi), cj ? c : conjc(c));
        mo = m + a.rot_out;
        if (mo >= a.other) mo -= a.other;
      }
      stc<false>(y + mo * a.eso, v);
    }
  }
}

// ---------------------------------------------------- kernel timing (opt-in)
// With timing on (rfh_kernel_timing), every line-kernel launch is bracketed by
// HIP events on the stream it runs on and filed under a label naming the
// kernel family, operation, length, precision, layout and algorithmic bytes;
// rfh_kernel_timing_report waits for the events and sums them per label.
std::atomic<bool> g_timing{false};
struct Timed {
  std::string label;
  hipEvent_t e0, e1;
  double bytes;
};
std::mutex g_timed_mu;
std::vector<Timed> g_timed;

const char *op_name(int op) {
  switch (op) {
    case OP_C2C: return "c2c";
    case OP_PASSA: return "passA";
    case OP_PASSB: return "passB";
    case OP_BLUE: return "blue";
    case OP_BLUEROW: return "bluerow";
    case OP_R2C_EVEN: return "r2c_even";
    case OP_R2C_ODD: return "r2c_odd";
    case OP_C2R_EVEN: return "c2r_even";
    case OP_C2R_ODD: return "c2r_odd";
    case OP_R2R: return "r2r";
    default: return "elem";
  }
}

// runs `launch` (returns the kernel family's name) between two events
template <typename F>
void timed_launch(hipStream_t s, const std::string &what, double bytes, F &&launch) {
  if (!g_timing.load()) {
    launch();
    return;
  }
  Timed t;
  t.bytes = bytes;
  HIP_CHECK(hipEventCreate(&t.e0));
  HIP_CHECK(hipEventCreate(&t.e1));
  HIP_CHECK(hipEventRecord(t.e0, s));
  const char *fam = launch();
  HIP_CHECK(hipEventRecord(t.e1, s));
  t.label = std::string(fam) + " " + what;
  std::lock_guard<std::mutex> lk(g_timed_mu);
  g_timed.push_back(t);
}

template <typename T>
void run_pass(const PassSpec &ps, const LineSet &g, hipStream_t s);

template <typename T>
void run_pass_timed(const PassSpec &ps, const LineSet &g, hipStream_t s) {
  if (!g_timing.load() || g.nlines <= 0) {
    run_pass<T>(ps, g, s);
    return;
  }
  const bool elem = ps.op == OP_BLUE_IN || ps.op == OP_BLUE_MUL || ps.op == OP_BLUE_OUT;
  const int64_t cin = elem ? ps.other : ps.nin, cout = elem ? ps.other : ps.nout;
  const double sc = sizeof(T);
  const double bytes = double(g.nlines) * ((ps.real_in ? 1 : 2) * sc * double(cin) +
                                           (ps.real_out ? 1 : 2) * sc * double(cout));
  int64_t minstride = 0;
  for (int d = 0; d < g.nb; ++d)
    minstride = d == 0 ? int64_t(std::llabs(g.bsi[d])) : std::min<int64_t>(minstride, std::llabs(g.bsi[d]));
  const bool inter = g.nb > 0 && (minstride < std::llabs(g.esi));
  const std::string what = std::string(op_name(ps.op)) + " n=" + std::to_string(ps.d->desc.n) +
                           (sizeof(T) == 8 ? " c128" : " c64") + (inter ? " strided" : "");
  timed_launch(s, what, bytes, [&]() -> const char * {
    run_pass<T>(ps, g, s);
    return t_last_family;
  });
}

template <typename T>
void run_pass(const PassSpec &ps, const LineSet &g, hipStream_t s) {
  if (g.nlines <= 0) return;
  t_last_family = "k_lines";
  if (ps.op == OP_BLUE_IN || ps.op == OP_BLUE_MUL || ps.op == OP_BLUE_OUT) {
    LineArgs<T> a = make_line_args<T>(ps, g);
    dim3 grid;
    grid.x = unsigned(std::max<int64_t>(1, std::min<int64_t>((ps.other + kThreads - 1) / kThreads,
                                                             1024)));
    grid.y = unsigned(std::max<int64_t>(1, std::min<int64_t>(g.nlines, 65535)));
    hipLaunchKernelGGL(k_blue_elem<T>, grid, dim3(kThreads), 0, s, a, ps.op);
    HIP_CHECK(hipGetLastError());
    t_last_family = "k_blue_elem";
    return;
  }
  LineArgs<T> a = make_line_args<T>(ps, g);
  const size_t elem = 2 * sizeof(T);
  const size_t lds = size_t(2) * a.L * a.ld * elem;
  if (lds + kLineStaticLds > lds_limit()) RFH_THROW(3, "line does not fit in LDS");
  bool vin, vout;
  if (ps.real_in)  // real pairs: 2-wide loads when the line is contiguous
    vin = g.esi == 1 && pairs_aligned(g.in, sizeof(T), 0, g.bsi, g.nb);
  else
    vin = pairs_aligned(g.in, sizeof(T), g.esi, g.bsi, g.nb);
  if (ps.real_out)
    vout = g.eso == 1 && pairs_aligned(g.out, sizeof(T), 0, g.bso, g.nb);
  else
    vout = pairs_aligned(g.out, sizeof(T), g.eso, g.bso, g.nb);
  if constexpr (sizeof(T) == 4) {
    if (const char *fam = launch_mfma(ps.op, a, a.lfast_in, a.lfast_out, vin, vout, s)) {
      t_last_family = fam;
      return;
    }
    if (launch_pass1024(ps.op, a, a.lfast_in, a.lfast_out, vin, vout, s)) {
      t_last_family = "k_pass1024";
      return;
    }
    if (launch_pipe4096(ps.op, a, a.lfast_in, a.lfast_out, vin, vout, s)) {
      t_last_family = "k_pipe4096";
      return;
    }
  }
  if (launch_fast<T>(ps.op, a, a.lfast_in, a.lfast_out, vin, vout, s)) {
    t_last_family = "k_fast";
    return;
  }
  switch (ps.op) {
    case OP_C2C: launch_var<T, OP_C2C>(a, lds, vin, vout, s); break;
    case OP_PASSA: launch_var<T, OP_PASSA>(a, lds, vin, vout, s); break;
    case OP_PASSB: launch_var<T, OP_PASSB>(a, lds, vin, vout, s); break;
    case OP_BLUE: launch_var<T, OP_BLUE>(a, lds, vin, vout, s); break;
    case OP_BLUEROW: launch_var<T, OP_BLUEROW>(a, lds, vin, vout, s); break;
    case OP_R2C_EVEN: launch_var<T, OP_R2C_EVEN>(a, lds, vin, vout, s); break;
    case OP_R2C_ODD: launch_var<T, OP_R2C_ODD>(a, lds, false, vout, s); break;
    case OP_C2R_EVEN: launch_var<T, OP_C2R_EVEN>(a, lds, vin, vout, s); break;
    case OP_C2R_ODD: launch_var<T, OP_C2R_ODD>(a, lds, vin, false, s); break;
    case OP_R2R: launch_one<T, OP_R2R, false, false>(a, lds, s); break;
  }
}

void run_pass_any(bool dbl, const PassSpec &ps, const LineSet &g, hipStream_t s) {
  if (dbl)
    run_pass_timed<double>(ps, g, s);
  else
    run_pass_timed<float>(ps, g, s);
}

// contiguous complex line set with g's batch shape, lines of length len
void contiguous_like(const LineSet &g, uint64_t len, int64_t *bs) {
  int64_t st = int64_t(2 * len);
  for (int d = g.nb - 1; d >= 0; --d) {
    bs[d] = st;
    st *= g.bdim[d];
  }
}

// Orders the batch dims by decreasing |stride| on the input (by_out false)
// or output side, so that the lines of one group are the closest in memory
// on that side; auxdim follows its dim.
void order_dims(LineSet &g, bool by_out) {
  int idx[kMaxBatchDims];
  for (int d = 0; d < g.nb; ++d) idx[d] = d;
  const int64_t *st = by_out ? g.bso : g.bsi;
  std::stable_sort(idx, idx + g.nb,
                   [&](int a, int b) { return std::llabs(st[a]) > std::llabs(st[b]); });
  LineSet o = g;
  o.auxdim = -1;
  for (int d = 0; d < g.nb; ++d) {
    o.bdim[d] = g.bdim[idx[d]];
    o.bsi[d] = g.bsi[idx[d]];
    o.bso[d] = g.bso[idx[d]];
    if (idx[d] == (g.auxdim >= 0 ? g.auxdim : g.nb - 1)) o.auxdim = d;
  }
  g = o;
}

// Lines interleaved on the input side: a batch dim is closer in memory
// than the elements of one line.
bool input_interleaved(const LineSet &g) {
  for (int d = 0; d < g.nb; ++d)
    if (std::llabs(g.bsi[d]) < std::llabs(g.esi)) return true;
  return false;
}

// Layout of a four-step intermediate of g's lines, each an n1 x n2 matrix
// (rows k1, columns c): line-major with rows of n2 (contiguous lines), or
// line-fastest (interleaved lines), so that the passes over it keep the
// same adjacent-line order as the user's arrays.
struct Tmp4 {
  void *p = nullptr;
  int64_t sb[kMaxBatchDims] = {};
  int64_t srow = 0, scol = 0;
};

Tmp4 tmp4(const LineSet &g, uint64_t n1, uint64_t n2, void *p, bool inter) {
  Tmp4 t;
  t.p = p;
  if (inter) {
    contiguous_like(g, 1, t.sb);
    t.scol = 2 * g.nlines;
    t.srow = t.scol * int64_t(n2);
  } else {
    contiguous_like(g, n1 * n2, t.sb);
    t.scol = 2;
    t.srow = 2 * int64_t(n2);
  }
  return t;
}

// Natural-order work lines of length len in the same two layouts.
struct Work {
  int64_t sb[kMaxBatchDims] = {};
  int64_t e = 0;
};

Work work_layout(const LineSet &g, uint64_t len, bool inter) {
  Work w;
  if (inter) {
    contiguous_like(g, 1, w.sb);
    w.e = 2 * g.nlines;
  } else {
    contiguous_like(g, len, w.sb);
    w.e = 2;
  }
  return w;
}

// Pass A of a four-step transform N = n1 * n2 over the lines of g: line
// (b, c) is column c of line b (n1 elements, n2 apart), written to row-major
// position (k1, c) of the intermediate t.  aux = c.
LineSet pass_a(const LineSet &g, uint64_t n1, uint64_t n2, const void *in, int64_t esi_line,
               const int64_t *inbs, const Tmp4 &t) {
  LineSet a{};
  a.in = in;
  a.out = t.p;
  a.nlines = g.nlines * int64_t(n2);
  a.nb = g.nb + 1;
  if (a.nb > kMaxBatchDims) RFH_THROW(3, "too many batch dimensions");
  for (int d = 0; d < g.nb; ++d) {
    a.bdim[d] = g.bdim[d];
    a.bsi[d] = inbs[d];
    a.bso[d] = t.sb[d];
  }
  a.bdim[g.nb] = int64_t(n2);
  a.bsi[g.nb] = esi_line;
  a.bso[g.nb] = t.scol;
  a.esi = esi_line * int64_t(n2);
  a.eso = t.srow;
  a.auxdim = g.nb;
  order_dims(a, false);
  return a;
}

// Pass B: line (b, k1) is row k1 of the intermediate (n2 elements), its
// bin k2 stored to position k1 + n1 * k2 of the output line.  aux = k1.
LineSet pass_b(const LineSet &g, uint64_t n1, uint64_t n2, const Tmp4 &t, void *out,
               int64_t eso_line, const int64_t *outbs) {
  LineSet b{};
  b.in = t.p;
  b.out = out;
  b.nlines = g.nlines * int64_t(n1);
  b.nb = g.nb + 1;
  if (b.nb > kMaxBatchDims) RFH_THROW(3, "too many batch dimensions");
  for (int d = 0; d < g.nb; ++d) {
    b.bdim[d] = g.bdim[d];
    b.bsi[d] = t.sb[d];
    b.bso[d] = outbs[d];
  }
  b.bdim[g.nb] = int64_t(n1);
  b.bsi[g.nb] = t.srow;
  b.bso[g.nb] = eso_line;
  b.esi = t.scol;
  b.eso = eso_line * int64_t(n1);
  b.auxdim = g.nb;
  order_dims(b, true);
  return b;
}

// Rows of the intermediate t transformed in place (lines (b, k1), n2
// elements).  aux = k1.
LineSet rows_inplace(const LineSet &g, uint64_t n1, const Tmp4 &t) {
  LineSet r{};
  r.in = t.p;
  r.out = t.p;
  r.nlines = g.nlines * int64_t(n1);
  r.nb = g.nb + 1;
  if (r.nb > kMaxBatchDims) RFH_THROW(3, "too many batch dimensions");
  for (int d = 0; d < g.nb; ++d) {
    r.bdim[d] = g.bdim[d];
    r.bsi[d] = r.bso[d] = t.sb[d];
  }
  r.bdim[g.nb] = int64_t(n1);
  r.bsi[g.nb] = r.bso[g.nb] = t.srow;
  r.esi = r.eso = t.scol;
  r.auxdim = g.nb;
  order_dims(r, false);
  return r;
}

// Columns of the intermediate t (lines (b, c), n1 elements) written to
// position c + n2 * r of the output line.  aux = c.
LineSet pass_c(const LineSet &g, uint64_t n2, const Tmp4 &t, void *out, int64_t eso_line,
               const int64_t *outbs) {
  LineSet c{};
  c.in = t.p;
  c.out = out;
  c.nlines = g.nlines * int64_t(n2);
  c.nb = g.nb + 1;
  if (c.nb > kMaxBatchDims) RFH_THROW(3, "too many batch dimensions");
  for (int d = 0; d < g.nb; ++d) {
    c.bdim[d] = g.bdim[d];
    c.bsi[d] = t.sb[d];
    c.bso[d] = outbs[d];
  }
  c.bdim[g.nb] = int64_t(n2);
  c.bsi[g.nb] = t.scol;
  c.bso[g.nb] = eso_line;
  c.esi = t.srow;
  c.eso = eso_line * int64_t(n2);
  c.auxdim = g.nb;
  order_dims(c, true);
  return c;
}

// Rows of a multi-pass intermediate t (lines (b, k1) of length M) as the
// input of the sub-plan, output to position k1 + n1 * k' of g's lines.
LineSet multi_rows(const LineSet &g, uint64_t n1, const Tmp4 &t) {
  LineSet r{};
  r.in = t.p;
  r.out = g.out;
  r.nlines = g.nlines * int64_t(n1);
  r.nb = g.nb + 1;
  if (r.nb > kMaxBatchDims) RFH_THROW(3, "too many batch dimensions");
  for (int d = 0; d < g.nb; ++d) {
    r.bdim[d] = g.bdim[d];
    r.bsi[d] = t.sb[d];
    r.bso[d] = g.bso[d];
  }
  r.bdim[g.nb] = int64_t(n1);
  r.bsi[g.nb] = t.srow;
  r.bso[g.nb] = g.eso;
  r.esi = t.scol;
  r.eso = g.eso * int64_t(n1);
  order_dims(r, false);
  return r;
}

}  // namespace

namespace {

// Two-pass transforms run in chunks of lines whose intermediate fits in
// kChunkBytes, so that pass B reads what pass A just wrote from the Infinity
// Cache (256 MB on MI355X) instead of HBM; chunking needs a single batch
// dimension to slice.  The complex64 1024 x 1024 plans, whose passes run on
// the persistent pipelined kernels, take kChunkBytesPipelined: each launch
// pays a pipeline fill and drain that outweighs the cache reuse (64 x 2^20:
// 0.443 ms in one 512 MiB chunk, 0.454 in 128 MiB chunks, 0.474 in 256).
constexpr size_t kChunkBytes = size_t(128) << 20;
constexpr size_t kChunkBytesPipelined = size_t(512) << 20;

int64_t chunk_lines(const C2CPlan &p, const LineSet &g, size_t chunk_bytes) {
  if (p.kind != kTwoPass && p.kind != kBlueTwoPass) return g.nlines;
  const size_t per = size_t(p.n1) * p.n2 * (p.dbl ? 16 : 8);
  const bool pipelined = p.kind == kTwoPass && !p.dbl && p.n1 == 1024 && p.n2 == 1024;
  const size_t cap = chunk_bytes ? chunk_bytes : (pipelined ? kChunkBytesPipelined : kChunkBytes);
  const int64_t c = std::max<int64_t>(1, int64_t(cap / per));
  return (g.nb == 1 && g.nlines > c) ? c : g.nlines;
}

LineSet slice_lines(const LineSet &g, int64_t c0, int64_t cn, size_t scalar) {
  LineSet s = g;
  s.nlines = cn;
  s.bdim[0] = cn;
  s.in = static_cast<const char *>(g.in) + c0 * g.bsi[0] * int64_t(scalar);
  s.out = static_cast<char *>(g.out) + c0 * g.bso[0] * int64_t(scalar);
  return s;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Lines per group of a four-step pass over lines of n: at least `base`
// adjacent lines (64-byte row segments), more for short lines until a group
// holds 2048 elements or its two LDS buffers reach 64 KB (or the plan's budget).
int pass_lines(const C2CPlan &p, uint64_t n, int base) {
  const size_t elem = p.dbl ? 16 : 8;
  const size_t cap = std::min<size_t>(p.budget ? p.budget : 65536, 65536);
  int L = base;
  while (L * 2 <= kMaxLinesPerGroup && uint64_t(L) * n < 2048 &&
         size_t(2) * size_t(L * 2) * n * elem <= cap)
    L *= 2;
  return L;
}

}  // namespace

size_t c2c_tmp_bytes(const C2CPlan &p, const LineSet &g, size_t chunk_bytes) {
  const size_t elem = p.dbl ? 16 : 8;
  const size_t lines = size_t(chunk_lines(p, g, chunk_bytes));
  switch (p.kind) {
    case kTwoPass: return lines * p.n1 * p.n2 * elem;
    case kBlueTwoPass: return lines * p.n1 * p.n2 * elem;
    case kMultiPass: {
      const size_t own = align256(size_t(g.nlines) * p.n1 * p.n2 * elem);
      const Tmp4 t = tmp4(g, p.n1, p.n2, nullptr, input_interleaved(g));
      return own + c2c_tmp_bytes(*p.sub, multi_rows(g, p.n1, t), chunk_bytes);
    }
    case kBlueMulti: {
      const size_t own = align256(size_t(g.nlines) * p.bn2 * elem);
      LineSet w{};
      w.nlines = g.nlines;
      w.nb = g.nb;
      for (int d = 0; d < g.nb; ++d) w.bdim[d] = g.bdim[d];
      const Work wl = work_layout(g, p.bn2, input_interleaved(g));
      for (int d = 0; d < g.nb; ++d) w.bsi[d] = w.bso[d] = wl.sb[d];
      w.esi = w.eso = wl.e;
      return own + c2c_tmp_bytes(*p.sub, w, chunk_bytes);
    }
    default: return 0;
  }
}

bool c2c_rotation_fusable(const C2CPlan &p, int64_t rot_in, int64_t rot_out) {
  if (rot_in < 0 || rot_out < 0 || uint64_t(rot_in) >= p.n || uint64_t(rot_out) >= p.n)
    return false;
  if (!rot_in && !rot_out) return true;
  if (p.kind == kDirect || p.kind == kBlueDirect || p.kind == kBlueTwoPass ||
      p.kind == kBlueMulti)
    return true;
  if (p.kind == kTwoPass)
    return uint64_t(rot_in) % p.n2 == 0 && uint64_t(rot_out) % p.n1 == 0;
  if (p.kind == kMultiPass && p.sub)
    return uint64_t(rot_in) % p.n2 == 0 && uint64_t(rot_out) % p.n1 == 0 &&
           c2c_rotation_fusable(*p.sub, 0, rot_out / int64_t(p.n1));
  return false;
}

std::vector<PassLaunch> describe_c2c(const C2CPlan &p, const LineSet &g, int dir, double fct,
                                     void *tmp, size_t chunk_bytes) {
  const int64_t cl = chunk_lines(p, g, chunk_bytes);
  if (cl < g.nlines) {  // one two-pass sequence per chunk, sharing tmp
    std::vector<PassLaunch> out;
    const size_t scalar = p.dbl ? 8 : 4;
    for (int64_t c0 = 0; c0 < g.nlines; c0 += cl) {
      const LineSet s = slice_lines(g, c0, std::min(cl, g.nlines - c0), scalar);
      for (const PassLaunch &pl : describe_c2c(p, s, dir, fct, tmp, chunk_bytes)) out.push_back(pl);
    }
    return out;
  }
  std::vector<PassLaunch> out;
  const bool dbl = p.dbl;
  const size_t elem = dbl ? 16 : 8;
  const int sub_conj = dir > 0 ? SUB_CONJ : 0;
  const int lines = dbl ? 4 : 8;
  const bool inter = input_interleaved(g);
  if ((g.rot_in || g.rot_out) && !c2c_rotation_fusable(p, g.rot_in, g.rot_out))
    RFH_THROW(3, "this plan cannot fuse the requested rotation into its passes");
  if (p.kind == kDirect) {
    PassSpec ps;
    ps.d = &p.main;
    ps.op = OP_C2C;
    ps.dir = dir;
    ps.fct = fct;
    ps.nin = ps.nout = int(p.n);
    out.push_back(PassLaunch{ps, g});
  } else if (p.kind == kBlueDirect) {
    PassSpec ps;
    ps.d = &p.main;
    ps.op = OP_BLUE;
    ps.dir = -1;  // forward FFT, kernel product, backward FFT (H:2732-2762)
    ps.sub = sub_conj;
    ps.fct = fct;
    ps.bign = int64_t(p.n);
    ps.t0 = p.bk.p;
    ps.t1 = p.bkf.p;
    ps.nin = int(p.bn2);
    ps.nout = int(p.n);
    out.push_back(PassLaunch{ps, g});
  } else if (p.kind == kTwoPass) {
    const uint64_t n1 = p.n1, n2 = p.n2;
    const Tmp4 t = tmp4(g, n1, n2, tmp, inter);
    PassSpec a;
    a.d = &p.pa;
    a.op = OP_PASSA;
    a.dir = dir;
    a.other = int64_t(n2);
    a.shift = p.shift;
    a.t0 = p.big_lo.p;
    a.t1 = p.big_hi.p;
    a.nin = a.nout = int(n1);
    a.lines = pass_lines(p, n1, lines);
    // rotations by whole rows / columns: input element j1 n2 + c of a line is
    // column c's element j1, so rot_in = t n2 rotates each column by t;
    // output bin k1 + n1 k2 is row k1's bin k2, so rot_out = u n1 rotates
    // each row by u
    LineSet la = pass_a(g, n1, n2, g.in, g.esi, g.bsi, t);
    la.rot_in = g.rot_in / int64_t(n2);
    out.push_back(PassLaunch{a, la});
    PassSpec b;
    b.d = &p.pb;
    b.op = OP_PASSB;
    b.dir = dir;
    b.fct = fct;
    b.other = int64_t(n1);
    b.nin = b.nout = int(n2);
    b.lines = pass_lines(p, n2, lines);
    LineSet lb = pass_b(g, n1, n2, t, g.out, g.eso, g.bso);
    lb.rot_out = g.rot_out / int64_t(n1);
    out.push_back(PassLaunch{b, lb});
  } else if (p.kind == kMultiPass) {
    // columns of n1 with the twiddle of N, then the sub-plan of M on the rows
    const uint64_t n1 = p.n1, M = p.n2;
    const Tmp4 t = tmp4(g, n1, M, tmp, inter);
    PassSpec a;
    a.d = &p.pa;
    a.op = OP_PASSA;
    a.dir = dir;
    a.other = int64_t(M);
    a.shift = p.shift;
    a.t0 = p.big_lo.p;
    a.t1 = p.big_hi.p;
    a.nin = a.nout = int(n1);
    a.lines = pass_lines(p, n1, lines);
    // rotations as in the four-step plan: rot_in = t M rotates each column
    // by t; rot_out = u n1 rotates the sub-plan's output of each row by u
    LineSet la = pass_a(g, n1, M, g.in, g.esi, g.bsi, t);
    la.rot_in = g.rot_in / int64_t(M);
    out.push_back(PassLaunch{a, la});
    void *rest = static_cast<char *>(tmp) + align256(size_t(g.nlines) * n1 * M * elem);
    LineSet rows = multi_rows(g, n1, t);
    rows.rot_out = g.rot_out / int64_t(n1);
    for (const PassLaunch &pl : describe_c2c(*p.sub, rows, dir, fct, rest, chunk_bytes))
      out.push_back(pl);
  } else if (p.kind == kBlueTwoPass) {
    // Bluestein in three passes over one intermediate t (n1 x n2 per line):
    // columns (x * chirp, forward, twiddle); rows in place (forward, kernel,
    // inverse, inverse twiddle); columns (inverse, chirp and fct, m < n).
    const uint64_t n1 = p.n1, n2 = p.n2;
    const Tmp4 t = tmp4(g, n1, n2, tmp, inter);
    PassSpec a;
    a.d = &p.pa;
    a.op = OP_PASSA;
    a.dir = -1;
    a.sub = SUB_CHIRP | sub_conj;
    a.bign = int64_t(p.n);
    a.other = int64_t(n2);
    a.shift = p.shift;
    a.t0 = p.big_lo.p;
    a.t1 = p.big_hi.p;
    a.t2 = p.bk.p;
    a.nin = a.nout = int(n1);
    a.lines = pass_lines(p, n1, lines);
    // rotations of the transform line: the chirped loads of pass A read
    // element (m + rot_in) mod n, the final stores of pass C write bin m to
    // (m + rot_out) mod n (whole-line element rotations, not column ones)
    LineSet ga = pass_a(g, n1, n2, g.in, g.esi, g.bsi, t);
    ga.rot_in = g.rot_in;
    ga.rot_out = 0;
    out.push_back(PassLaunch{a, ga});
    PassSpec r;
    r.d = &p.pb;
    r.op = OP_BLUEROW;
    r.dir = -1;
    r.sub = sub_conj;
    r.other = int64_t(n1);
    r.shift = p.shift;
    r.t0 = p.big_lo.p;
    r.t1 = p.big_hi.p;
    r.t2 = p.bkf.p;
    r.nin = r.nout = int(n2);
    r.lines = pass_lines(p, n2, lines);
    out.push_back(PassLaunch{r, rows_inplace(g, n1, t)});
    PassSpec c;
    c.d = &p.pa;
    c.op = OP_PASSA;
    c.dir = +1;
    c.sub = SUB_FINAL | sub_conj;
    c.fct = fct;
    c.bign = int64_t(p.n);
    c.other = int64_t(n2);
    c.t0 = p.bk.p;
    c.nin = c.nout = int(n1);
    c.lines = pass_lines(p, n1, lines);
    LineSet gc = pass_c(g, n2, t, g.out, g.eso, g.bso);
    gc.rot_in = 0;
    gc.rot_out = g.rot_out;
    out.push_back(PassLaunch{c, gc});
  } else {  // kBlueMulti: elementwise chirp and kernel passes around the sub-plan of bn2
    const uint64_t N = p.bn2;
    const Work wl = work_layout(g, N, inter);
    char *wp = static_cast<char *>(tmp);
    void *rest = wp + align256(size_t(g.nlines) * N * elem);
    LineSet w = g;  // the work lines, natural order, length bn2
    w.in = wp;
    w.out = wp;
    for (int d = 0; d < g.nb; ++d) w.bsi[d] = w.bso[d] = wl.sb[d];
    w.esi = w.eso = wl.e;
    w.auxdim = -1;
    w.rot_in = w.rot_out = 0;  // the work lines are never rotated
    LineSet gin = g;  // user input -> work lines (rotated loads)
    gin.out = wp;
    for (int d = 0; d < g.nb; ++d) gin.bso[d] = wl.sb[d];
    gin.eso = wl.e;
    gin.rot_out = 0;
    LineSet gout = g;  // work lines -> user output (rotated stores)
    gout.in = wp;
    for (int d = 0; d < g.nb; ++d) gout.bsi[d] = wl.sb[d];
    gout.esi = wl.e;
    gout.rot_in = 0;
    PassSpec e;
    e.d = &p.main;
    e.sub = sub_conj;
    e.bign = int64_t(p.n);
    e.t0 = p.bk.p;
    e.t1 = p.bkf.p;
    e.op = OP_BLUE_IN;
    e.other = int64_t(N);
    out.push_back(PassLaunch{e, gin});
    for (const PassLaunch &pl : describe_c2c(*p.sub, w, -1, 1.0, rest, chunk_bytes))
      out.push_back(pl);
    e.op = OP_BLUE_MUL;
    e.other = int64_t(N);
    out.push_back(PassLaunch{e, w});
    for (const PassLaunch &pl : describe_c2c(*p.sub, w, +1, 1.0, rest, chunk_bytes))
      out.push_back(pl);
    e.op = OP_BLUE_OUT;
    e.fct = fct;
    e.other = int64_t(p.n);
    out.push_back(PassLaunch{e, gout});
  }
  return out;
}

void check_tmp_spans(const std::vector<PassLaunch> &passes, const void *tmp, size_t bytes,
                     size_t scalar) {
  const char *lo = static_cast<const char *>(tmp), *hi = lo + bytes;
  auto check = [&](const void *ptr, const int64_t *bs, int64_t es, int64_t count,
                   const LineSet &g) {
    const char *q = static_cast<const char *>(ptr);
    if (!tmp || q < lo || q >= hi) return;  // not a workspace access
    int64_t span = (count - 1) * std::llabs(es);
    for (int d = 0; d < g.nb; ++d) span += (g.bdim[d] - 1) * std::llabs(bs[d]);
    if (q + size_t(span + 2) * scalar > hi)
      RFH_THROW(1, "internal error: a pass would reach past its workspace");
  };
  for (const PassLaunch &pl : passes) {
    const bool elem = pl.ps.op == OP_BLUE_IN || pl.ps.op == OP_BLUE_MUL ||
                      pl.ps.op == OP_BLUE_OUT;
    const int64_t cin = elem ? pl.ps.other : pl.ps.nin, cout = elem ? pl.ps.other : pl.ps.nout;
    check(pl.g.in, pl.g.bsi, pl.g.esi, cin, pl.g);
    check(pl.g.out, pl.g.bso, pl.g.eso, cout, pl.g);
  }
}

void launch_c2c(const C2CPlan &p, const LineSet &g, int dir, double fct, hipStream_t s,
                void *tmp, size_t chunk_bytes) {
  const auto passes = describe_c2c(p, g, dir, fct, tmp, chunk_bytes);
  check_tmp_spans(passes, tmp, c2c_tmp_bytes(p, g, chunk_bytes), p.dbl ? 8 : 4);
  for (const PassLaunch &pl : passes) run_pass_any(p.dbl, pl.ps, pl.g, s);
}

void launch_r2r_fused(const C2CPlan &p, const LineSet &g, PreMode pre, PostMode post,
                      const RealPrePost &rr, bool fft_forward, double fct, hipStream_t s) {
  PassSpec ps;
  ps.d = &p.main;
  ps.op = OP_R2R;
  ps.dir = fft_forward ? -1 : +1;
  ps.fct = fct;
  ps.bign = rr.n;
  ps.sub = int(pre) | (int(post) << 8) | (rr.ortho ? R2R_ORTHO : 0) |
           (rr.conj ? R2R_CONJ : 0) | (rr.flip ? R2R_FLIP : 0);
  ps.nin = int(rr.m);
  ps.nout = int(rr.n);
  run_pass_any(p.dbl, ps, g, s);
}

PassLaunch describe_real_fused(const RealPlan &rp, RealOp op, const LineSet &g, bool conj_flag,
                               double fct) {
  PassSpec ps;
  ps.d = &rp.d;
  ps.fct = fct;
  ps.sub = conj_flag ? SUB_CONJ : 0;
  ps.t0 = rp.rtw.p;
  const int64_t n = int64_t(rp.n);
  const bool even = (n % 2) == 0;
  if (op == kR2C) {
    ps.dir = -1;
    ps.op = even ? OP_R2C_EVEN : OP_R2C_ODD;
    ps.nin = int(even ? n / 2 : n);
    ps.nout = int(n / 2 + 1);
    ps.real_in = true;
  } else {
    ps.dir = +1;
    ps.op = even ? OP_C2R_EVEN : OP_C2R_ODD;
    ps.nin = int(even ? n / 2 : n);
    ps.nout = int(even ? n / 2 : n);
    ps.real_out = true;
  }
  return PassLaunch{ps, g};
}

void launch_real_fused(const RealPlan &rp, RealOp op, const LineSet &g, bool conj_flag,
                       double fct, hipStream_t s) {
  const PassLaunch pl = describe_real_fused(rp, op, g, conj_flag, fct);
  run_pass_any(rp.dbl, pl.ps, pl.g, s);
}

// ------------------------------------------------------- pre / post kernels

namespace {

template <typename T>
struct PrePostArgs {
  int64_t nlines;
  int nb;
  int64_t bdim[kMaxBatchDims], bs[kMaxBatchDims];
  int64_t es;
  const T *rin;   // pre: real / half-complex input line set
  T *rout;        // post: real / complex output line set
  C2<T> *tmp;     // contiguous complex work lines of length m
  int64_t n, m, count;
  int mode, ortho, conj, flip;
  T fct;
};

template <typename T>
__device__ __forceinline__ int64_t line_offset(const PrePostArgs<T> &a, int64_t b) {
  int64_t rem = b, o = 0;
  for (int d = a.nb - 1; d >= 0; --d) {
    const int64_t q = rem / a.bdim[d];
    o += (rem - q * a.bdim[d]) * a.bs[d];
    rem = q;
  }
  return o;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_pre(const PrePostArgs<T> a) {
  for (int64_t b = blockIdx.y; b < a.nlines; b += gridDim.y) {
    const T *x = a.rin + line_offset(a, b);
    C2<T> *t = a.tmp + b * a.m;
    const int64_t n = a.n;
    for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < a.m;
         j += int64_t(gridDim.x) * blockDim.x) {
      t[j] = pre_elem<T>(a.mode, x, a.es, n, a.m, j, a.ortho, a.conj, a.flip);
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_post(const PrePostArgs<T> a) {
  for (int64_t b = blockIdx.y; b < a.nlines; b += gridDim.y) {
    T *y = a.rout + line_offset(a, b);
    const C2<T> *t = a.tmp + b * a.m;
    const int64_t n = a.n;
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < a.count;
         k += int64_t(gridDim.x) * blockDim.x) {
      if (a.mode == kPostComplexHalf || a.mode == kPostHalfSpectrum) {
        C2<T> v = a.mode == kPostComplexHalf
                      ? t[k]
                      : half_spectrum_bin<T>([&](int64_t q) { return t[q]; }, n, k);
        if (a.conj) v.y = -v.y;
        y[k * a.es] = a.fct * v.x;
        y[k * a.es + 1] = a.fct * v.y;
        continue;
      }
      const T r = post_elem<T>(a.mode, [&](int64_t q) { return t[q]; }, n, k, a.ortho, a.flip);
      y[k * a.es] = a.fct * r;
    }
  }
}

template <typename T>
PrePostArgs<T> prepost_args(const RealPrePost &rp, const LineSet &g, bool pre) {
  PrePostArgs<T> a{};
  a.nlines = g.nlines;
  a.nb = g.nb;
  for (int d = 0; d < g.nb; ++d) {
    a.bdim[d] = g.bdim[d];
    a.bs[d] = pre ? g.bsi[d] : g.bso[d];
  }
  a.es = pre ? g.esi : g.eso;
  a.n = rp.n;
  a.m = rp.m;
  a.ortho = rp.ortho;
  a.conj = rp.conj;
  a.flip = rp.flip;
  return a;
}

void grid_for(int64_t count, int64_t nlines, dim3 &grid) {
  grid.x = unsigned(std::max<int64_t>(1, std::min<int64_t>((count + kThreads - 1) / kThreads, 64)));
  grid.y = unsigned(std::max<int64_t>(1, std::min<int64_t>(nlines, 65535)));
  grid.z = 1;
}

}  // namespace

void launch_pre(PreMode mode, const RealPrePost &rp, const LineSet &g, bool dbl,
                void *tmp_complex, hipStream_t s) {
  auto go = [&](auto zero) {
    using T = decltype(zero);
    PrePostArgs<T> a = prepost_args<T>(rp, g, true);
    a.rin = static_cast<const T *>(g.in);
    a.tmp = static_cast<C2<T> *>(tmp_complex);
    a.mode = mode;
    dim3 grid;
    grid_for(rp.m, g.nlines, grid);
    hipLaunchKernelGGL(k_pre<T>, grid, dim3(kThreads), 0, s, a);
    HIP_CHECK(hipGetLastError());
  };
  if (dbl)
    go(double(0));
  else
    go(float(0));
}

void launch_post(PostMode mode, const RealPrePost &rp, const LineSet &g, bool dbl,
                 const void *tmp_complex, double fct, hipStream_t s) {
  auto go = [&](auto zero) {
    using T = decltype(zero);
    PrePostArgs<T> a = prepost_args<T>(rp, g, false);
    a.rout = static_cast<T *>(g.out);
    a.tmp = static_cast<C2<T> *>(const_cast<void *>(tmp_complex));
    a.mode = mode;
    a.fct = T(fct);
    a.count = (mode == kPostComplexHalf || mode == kPostHalfSpectrum) ? rp.n / 2 + 1 : rp.n;
    dim3 grid;
    grid_for(a.count, g.nlines, grid);
    hipLaunchKernelGGL(k_post<T>, grid, dim3(kThreads), 0, s, a);
    HIP_CHECK(hipGetLastError());
  };
  if (dbl)
    go(double(0));
  else
    go(float(0));
}

// -------------------------------------------------------------- mirrors

namespace {

struct MirrorPos {
  int64_t self, mirror;  // offsets (scalar units) of idx and rev(idx)
  int cls;               // 0 untouched half, 1 upper half, 2 plane larger,
                         // 3 plane smaller, 4 plane fixed point
};

__device__ __forceinline__ MirrorPos mirror_pos(const MirrorGeom &g, int64_t t,
                                                const int64_t *str) {
  int64_t c[kMaxDims];
  int64_t rem = t;
  for (int d = g.ndim - 1; d >= 0; --d) {
    const int64_t q = rem / g.shape[d];
    c[d] = rem - q * g.shape[d];
    rem = q;
  }
  MirrorPos p{0, 0, 0};
  int cmp = 0;  // C-order comparison of idx against rev(idx)
  for (int d = 0; d < g.ndim; ++d) {
    const int64_t r = g.rev[d] ? (c[d] == 0 ? 0 : g.shape[d] - c[d]) : c[d];
    p.self += c[d] * str[d];
    p.mirror += r * str[d];
    if (!cmp && d != g.last && c[d] != r) cmp = c[d] > r ? 1 : -1;
  }
  const int64_t n = g.shape[g.last], h = n / 2 + 1, kl = c[g.last];
  if (kl >= h) p.cls = 1;
  else if (kl == 0 || (n % 2 == 0 && 2 * kl == n)) p.cls = cmp > 0 ? 2 : (cmp < 0 ? 3 : 4);
  else p.cls = 0;
  return p;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_sym_fill(const MirrorGeom g, int64_t total,
                                                      T *out) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const MirrorPos p = mirror_pos(g, t, g.so);
    // the reference's C-order loop out[rev(i)] = conj(out[i]) over the half
    // region leaves: upper half = conj(mirror); in the self-mirrored planes
    // the later (larger) position = conj(earlier), fixed points conjugated.
    if (p.cls == 1 || p.cls == 2) {
      out[p.self] = out[p.mirror];
      out[p.self + 1] = -out[p.mirror + 1];
    } else if (p.cls == 4) {
      out[p.self + 1] = -out[p.self + 1];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_hartley(const MirrorGeom g, int64_t total,
                                                     const T *tmp, T *out) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const MirrorPos po = mirror_pos(g, t, g.so);
    const MirrorPos pt = mirror_pos(g, t, g.st);
    T r;
    switch (po.cls) {
      case 1:  // upper: re - im of the mirrored bin
      case 3:  // plane, earlier position: overwritten by the later one
        r = tmp[pt.mirror] - tmp[pt.mirror + 1];
        break;
      case 4:
        r = tmp[pt.self] - tmp[pt.self + 1];
        break;
      default:  // untouched half and later plane position: re + im
        r = tmp[pt.self] + tmp[pt.self + 1];
        break;
    }
    out[po.self] = r;
  }
}

}  // namespace

void kernel_timing(bool on) { g_timing.store(on); }

std::string kernel_timing_report() {
  std::vector<Timed> done;
  {
    std::lock_guard<std::mutex> lk(g_timed_mu);
    done.swap(g_timed);
  }
  struct Acc {
    int n = 0;
    double ms = 0, bytes = 0;
  };
  std::map<std::string, Acc> acc;
  for (Timed &t : done) {
    float ms = 0;
    if (hipEventSynchronize(t.e1) == hipSuccess && hipEventElapsedTime(&ms, t.e0, t.e1) == hipSuccess) {
      Acc &a = acc[t.label];
      ++a.n;
      a.ms += ms;
      a.bytes += t.bytes;
    }
    (void)hipEventDestroy(t.e0);
    (void)hipEventDestroy(t.e1);
  }
  std::string out;
  char line[512];
  for (auto &kv : acc) {
    std::snprintf(line, sizeof line, "%s\t%d\t%.6f\t%.0f\n", kv.first.c_str(), kv.second.n,
                  kv.second.ms, kv.second.bytes);
    out += line;
  }
  return out;
}

void launch_sym_fill(const MirrorGeom &mg, bool dbl, void *out, hipStream_t s) {
  int64_t total = 1;
  for (int d = 0; d < mg.ndim; ++d) total *= mg.shape[d];
  if (total <= 0) return;
  const unsigned grid = unsigned(std::min<int64_t>((total + kThreads - 1) / kThreads, 65536));
  if (dbl)
    hipLaunchKernelGGL(k_sym_fill<double>, dim3(grid), dim3(kThreads), 0, s, mg, total,
                       static_cast<double *>(out));
  else
    hipLaunchKernelGGL(k_sym_fill<float>, dim3(grid), dim3(kThreads), 0, s, mg, total,
                       static_cast<float *>(out));
  HIP_CHECK(hipGetLastError());
}

void launch_hartley_combine(const MirrorGeom &mg, bool dbl, const void *tmp, void *out,
                            hipStream_t s) {
  int64_t total = 1;
  for (int d = 0; d < mg.ndim; ++d) total *= mg.shape[d];
  if (total <= 0) return;
  const unsigned grid = unsigned(std::min<int64_t>((total + kThreads - 1) / kThreads, 65536));
  if (dbl)
    hipLaunchKernelGGL(k_hartley<double>, dim3(grid), dim3(kThreads), 0, s, mg, total,
                       static_cast<const double *>(tmp), static_cast<double *>(out));
  else
    hipLaunchKernelGGL(k_hartley<float>, dim3(grid), dim3(kThreads), 0, s, mg, total,
                       static_cast<const float *>(tmp), static_cast<float *>(out));
  HIP_CHECK(hipGetLastError());
}

}  // namespace rfh
