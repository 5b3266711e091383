// ragged.hip -- many independent transforms of uneven lengths and kinds
// (complex c2c, real r2c / c2r) in persistent launches.
//
// A batch is planned once (rfh_batch_create): every job gets the passes the
// per-call engine would run for it (direct LDS line, Bluestein in LDS, the
// four-step passes, or pre -> c2c -> post for real lengths without a fused
// plan), all sized to one LDS budget, with its temporaries carved out of one
// arena.  The passes are cut into work units (a run of tiles of one pass)
// and grouped by level: level k holds the k-th pass of every job that has
// one, so a unit depends only on units of earlier levels.  Execution is one
// fill kernel that zeroes the levels' unit counters, then one persistent
// launch per level (a job has at most a handful of passes): resident
// workgroups pull units from that level's counter, largest job first, and
// run their tiles.  Kernel boundaries order the levels, so no workgroup ever
// waits on another and a large job's tiles spread over many CUs.  The LDS
// budget and the grid size are picked by a small autotuner that times the
// candidates on scratch copies of the batch's buffers.
//
// Execution allocates nothing and calls no host synchronisation, so it can
// be captured in a HIP graph.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <tuple>

#include "../../include/rocket_fft_hip.h"
#include "passes.hpp"

namespace rfh {
namespace {
using namespace dev;

constexpr int OP_PRE = 9;   // one line: real / half-complex input -> complex work line
constexpr int OP_POST = 10; // one line: complex work line -> real / half-complex output
// elements of work a unit gathers from consecutive tiles of one pass
constexpr int64_t kUnitElems = 4096;
// elements of a pre / post line per tile: the elementwise real-transform
// steps are cut into tiles like the line passes, so that a long line's
// pre / post step spreads over many workgroups instead of one
constexpr int64_t kPrePostChunk = 2048;

struct PassDev {
  int op;
  int tiles;
};
struct UnitDev {  // tiles tile0 .. tile0 + ntiles - 1 of pass `pass`
  int pass;
  int tile0;
  int ntiles;
};

// elements [tile C, (tile + 1) C) of a pre / post line (C = kPrePostChunk)
template <typename T>
__device__ void pre_line(const LineArgs<T> &a, int tile) {
  C2<T> *t = reinterpret_cast<C2<T> *>(a.out);
  const int mode = a.sub & 0xff;
  const int64_t j0 = int64_t(tile) * kPrePostChunk;
  const int64_t j1 = min(j0 + kPrePostChunk, int64_t(a.f.n));
  for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x)
    t[j] = pre_elem<T>(mode, a.in, a.esi, a.bign, a.f.n, j, false, (a.sub & R2R_CONJ) != 0,
                       false);
}

template <typename T>
__device__ void post_line(const LineArgs<T> &a, int tile) {
  const C2<T> *t = reinterpret_cast<const C2<T> *>(a.in);
  const int mode = (a.sub >> 8) & 0xff;
  const int64_t n = a.bign;
  const int64_t k0 = int64_t(tile) * kPrePostChunk;
  if (mode == kPostComplexHalf || mode == kPostHalfSpectrum) {
    const int64_t k1 = min(k0 + kPrePostChunk, n / 2 + 1);
    for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
      C2<T> v = mode == kPostComplexHalf
                    ? t[k]
                    : half_spectrum_bin<T>([&](int64_t q) { return t[q]; }, n, k);
      if (a.sub & R2R_CONJ) v.y = -v.y;
      a.out[k * a.eso] = a.fct * v.x;
      a.out[k * a.eso + 1] = a.fct * v.y;
    }
  } else {
    const int64_t k1 = min(k0 + kPrePostChunk, n);
    for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x)
      a.out[k * a.eso] =
          a.fct * post_elem<T>(mode, [&](int64_t q) { return t[q]; }, n, k, false, false);
  }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_ragged(const UnitDev *__restrict__ units,
                                                    int nunits,
                                                    const PassDev *__restrict__ passes,
                                                    const LineArgs<T> *__restrict__ args,
                                                    unsigned *queue) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int64_t s_oi[kMaxLinesPerGroup], s_oo[kMaxLinesPerGroup];
  __shared__ int s_aux[kMaxLinesPerGroup];
  __shared__ int s_unit;
  for (;;) {
    if (threadIdx.x == 0) s_unit = int(atomicAdd(queue, 1u));
    __syncthreads();
    const int q = __builtin_amdgcn_readfirstlane(s_unit);
    __syncthreads();  // every lane has read s_unit before the next grab
    if (q >= nunits) break;
    const UnitDev u = units[q];
    const PassDev pd = passes[u.pass];
    const LineArgs<T> &a = args[u.pass];
    for (int t = u.tile0; t < u.tile0 + u.ntiles; ++t) {
      switch (pd.op) {
        case OP_C2C: line_body<T, OP_C2C, true, true>(a, t, smem_raw, s_oi, s_oo, s_aux); break;
        case OP_PASSA: line_body<T, OP_PASSA, true, true>(a, t, smem_raw, s_oi, s_oo, s_aux); break;
        case OP_PASSB: line_body<T, OP_PASSB, true, true>(a, t, smem_raw, s_oi, s_oo, s_aux); break;
        case OP_BLUE: line_body<T, OP_BLUE, true, true>(a, t, smem_raw, s_oi, s_oo, s_aux); break;
        case OP_BLUEROW: line_body<T, OP_BLUEROW, true, true>(a, t, smem_raw, s_oi, s_oo, s_aux); break;
        case OP_R2C_EVEN: line_body<T, OP_R2C_EVEN, true, true>(a, t, smem_raw, s_oi, s_oo, s_aux); break;
        case OP_R2C_ODD: line_body<T, OP_R2C_ODD, false, true>(a, t, smem_raw, s_oi, s_oo, s_aux); break;
        case OP_C2R_EVEN: line_body<T, OP_C2R_EVEN, true, true>(a, t, smem_raw, s_oi, s_oo, s_aux); break;
        case OP_C2R_ODD: line_body<T, OP_C2R_ODD, true, false>(a, t, smem_raw, s_oi, s_oo, s_aux); break;
        case OP_PRE: pre_line<T>(a, t); break;
        default: post_line<T>(a, t); break;
      }
      __syncthreads();  // tile boundary: LDS and per-line slots are reused
    }
  }
}

__global__ void k_zero(unsigned *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// ------------------------------------------------------------------ host

struct Built {
  size_t budget = 0;          // dynamic LDS per workgroup
  DevBuf units, passes, args, queue, arena;
  std::vector<int> level_off;  // units of level k: [level_off[k], level_off[k + 1])
  std::vector<std::shared_ptr<C2CPlan>> c2c_keep;
  std::vector<std::shared_ptr<RealPlan>> real_keep;
  int njobs = 0;
  int lds_need = 0;
};

struct Batch {
  bool dbl = false;
  std::vector<rfh_job_t> jobs;
  std::unique_ptr<Built> b;
  int grid = 0;
  int autotuned = 0;
  double tuned_ms = 0;
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

LineSet one_line(const void *in, void *out, int64_t esi, int64_t eso) {
  LineSet g{};
  g.in = in;
  g.out = out;
  g.nlines = 1;
  g.nb = 0;
  g.esi = esi;
  g.eso = eso;
  return g;
}

// cost estimate for largest-first ordering: bytes moved by all passes
double job_cost(const std::vector<PassLaunch> &pl) {
  double c = 0;
  for (const auto &p : pl) c += double(p.g.nlines) * double(p.ps.nin + p.ps.nout);
  return c;
}

template <typename T>
std::unique_ptr<Built> build(const std::vector<rfh_job_t> &jobs, bool dbl, size_t budget) {
  auto B = std::make_unique<Built>();
  B->budget = budget;
  B->njobs = int(jobs.size());
  const size_t ce = 2 * sizeof(T);
  // pass 1: plans and temporaries
  std::vector<std::vector<PassLaunch>> per(jobs.size());
  std::vector<size_t> tmp_off(jobs.size()), work_off(jobs.size());
  size_t arena = 0;
  std::vector<std::shared_ptr<C2CPlan>> cplan(jobs.size());
  std::vector<std::shared_ptr<RealPlan>> rplan(jobs.size());
  for (size_t i = 0; i < jobs.size(); ++i) {
    const rfh_job_t &j = jobs[i];
    if (j.n == 0) continue;
    if (j.kind == RFH_JOB_C2C) {
      cplan[i] = get_c2c_plan(j.n, dbl, budget);
      tmp_off[i] = arena;
      arena = align_up(arena + c2c_tmp_bytes(*cplan[i], one_line(nullptr, nullptr, 2, 2)), 256);
    } else {
      rplan[i] = get_real_plan(j.n, dbl, budget);
      if (!rplan[i]->fused) {  // pre -> c2c(m) -> post through a work line of m
        const uint64_t m = rplan[i]->half ? j.n / 2 : j.n;
        cplan[i] = get_c2c_plan(m, dbl, budget);
        work_off[i] = arena;
        arena = align_up(arena + size_t(m) * ce, 256);
        tmp_off[i] = arena;
        arena = align_up(arena + c2c_tmp_bytes(*cplan[i], one_line(nullptr, nullptr, 2, 2)), 256);
      }
    }
  }
  B->arena.bytes = 0;
  if (arena) {
    HIP_CHECK(hipMalloc(&B->arena.p, arena));
    B->arena.bytes = arena;
  }
  char *ar = static_cast<char *>(B->arena.p);
  // pass 2: the passes of every job
  for (size_t i = 0; i < jobs.size(); ++i) {
    const rfh_job_t &j = jobs[i];
    if (j.n == 0) continue;
    const int dir = j.forward ? -1 : +1;
    if (j.kind == RFH_JOB_C2C) {
      per[i] = describe_c2c(*cplan[i], one_line(j.in, j.out, 2, 2), dir, j.fct, ar + tmp_off[i]);
      check_tmp_spans(per[i], ar + tmp_off[i],
                      c2c_tmp_bytes(*cplan[i], one_line(nullptr, nullptr, 2, 2)), dbl ? 8 : 4);
      continue;
    }
    const bool r2c = j.kind == RFH_JOB_R2C;
    if (rplan[i]->fused) {
      per[i].push_back(describe_real_fused(*rplan[i], r2c ? kR2C : kC2R,
                                           one_line(j.in, j.out, r2c ? 1 : 2, r2c ? 2 : 1),
                                           r2c ? !j.forward : j.forward, j.fct));
      continue;
    }
    void *work = ar + work_off[i];
    PassSpec pre{};
    pre.d = &cplan[i]->main;  // only desc.n (= m) is read
    pre.op = OP_PRE;
    pre.bign = int64_t(j.n);
    const bool half = rplan[i]->half;
    const int pre_mode = r2c ? (half ? kPreRealPairs : kPreRealAsComplex)
                             : (half ? kPreHalfSpectrum : kPreHermitian);
    const int post_mode = r2c ? (half ? kPostHalfSpectrum : kPostComplexHalf)
                              : (half ? kPostRealPairs : kPostReal);
    pre.sub = pre_mode | (!r2c && j.forward ? R2R_CONJ : 0);
    per[i].push_back(PassLaunch{pre, one_line(j.in, work, r2c ? 1 : 2, 2)});
    const auto cp = describe_c2c(*cplan[i], one_line(work, work, 2, 2), r2c ? -1 : +1, 1.0,
                                 ar + tmp_off[i]);
    check_tmp_spans(cp, ar + tmp_off[i],
                    c2c_tmp_bytes(*cplan[i], one_line(nullptr, nullptr, 2, 2)), dbl ? 8 : 4);
    for (const auto &pl : cp) per[i].push_back(pl);
    PassSpec post{};
    post.d = &cplan[i]->main;
    post.op = OP_POST;
    post.bign = int64_t(j.n);
    post.fct = j.fct;
    post.sub = (post_mode << 8) | (r2c && !j.forward ? R2R_CONJ : 0);
    per[i].push_back(PassLaunch{post, one_line(work, j.out, 2, r2c ? 2 : 1)});
  }
  // pass 3: device tables; units by level, the jobs of a level largest first
  std::vector<PassDev> pd;
  std::vector<LineArgs<T>> la;
  std::vector<std::pair<double, int>> cost;
  std::vector<int> first(jobs.size());
  int lds_need = 0, levels = 0;
  for (size_t i = 0; i < jobs.size(); ++i) {
    first[i] = int(pd.size());
    levels = std::max(levels, int(per[i].size()));
    for (const PassLaunch &pl : per[i]) {
      if (pl.ps.op == OP_BLUE_IN || pl.ps.op == OP_BLUE_MUL || pl.ps.op == OP_BLUE_OUT)
        RFH_THROW(3, "ragged batch: length " + std::to_string(jobs[i].n) +
                         " needs the multi-pass Bluestein form, which the persistent kernel "
                         "does not run");
      LineArgs<T> a = make_line_args<T>(pl.ps, pl.g);
      int tiles = int((pl.g.nlines + a.L - 1) / a.L);
      if (pl.ps.op == OP_PRE || pl.ps.op == OP_POST) {
        a.f.n = int(cplan[i]->n);  // work-line length m
        a.L = 1;
        // outputs of the step: the work line, or the half spectrum / real line
        const int post_mode = (a.sub >> 8) & 0xff;
        const int64_t count =
            pl.ps.op == OP_PRE ? int64_t(a.f.n)
            : (post_mode == kPostComplexHalf || post_mode == kPostHalfSpectrum) ? a.bign / 2 + 1
                                                                                 : a.bign;
        tiles = int((count + kPrePostChunk - 1) / kPrePostChunk);
      } else {
        lds_need = std::max<int>(lds_need, int(2 * a.L * a.ld * ce));
      }
      pd.push_back(PassDev{pl.ps.op, tiles});
      la.push_back(a);
    }
    cost.push_back({job_cost(per[i]), int(i)});
  }
  if (size_t(lds_need) > budget) RFH_THROW(3, "ragged batch: a pass exceeds the LDS budget");
  B->lds_need = lds_need;
  std::stable_sort(cost.begin(), cost.end(),
                   [](const std::pair<double, int> &x, const std::pair<double, int> &y) {
                     return x.first > y.first;
                   });
  std::vector<UnitDev> units;
  B->level_off.assign(1, 0);
  for (int lv = 0; lv < levels; ++lv) {
    for (const auto &c : cost) {
      const size_t i = size_t(c.second);
      if (int(per[i].size()) <= lv) continue;
      const int pid = first[i] + lv;
      const LineArgs<T> &a = la[size_t(pid)];
      // a unit holds enough tiles for about kUnitElems elements of work
      const bool pp = pd[size_t(pid)].op == OP_PRE || pd[size_t(pid)].op == OP_POST;
      const int64_t per_tile =
          pp ? kPrePostChunk : int64_t(std::max(1, a.L)) * std::max(1, a.f.n);
      const int tpu = int(std::max<int64_t>(1, kUnitElems / per_tile));
      for (int t0 = 0; t0 < pd[size_t(pid)].tiles; t0 += tpu)
        units.push_back(UnitDev{pid, t0, std::min(tpu, pd[size_t(pid)].tiles - t0)});
    }
    B->level_off.push_back(int(units.size()));
  }
  B->units.upload(units.data(), units.size() * sizeof(UnitDev));
  B->passes.upload(pd.data(), pd.size() * sizeof(PassDev));
  B->args.upload(la.data(), la.size() * sizeof(LineArgs<T>));
  std::vector<unsigned> zeros(size_t(std::max(levels, 1)), 0u);
  B->queue.upload(zeros.data(), zeros.size() * sizeof(unsigned));
  for (auto &p : cplan)
    if (p) B->c2c_keep.push_back(p);
  for (auto &p : rplan)
    if (p) B->real_keep.push_back(p);
  return B;
}

std::mutex g_attr_mu;
std::set<std::pair<int, int>> g_attr_set;  // (device, double precision)

template <typename T>
void launch_batch(const Built &B, int grid, hipStream_t s) {
  if (B.njobs == 0) return;
  const void *fn = reinterpret_cast<const void *>(&k_ragged<T>);
  {
    std::lock_guard<std::mutex> lk(g_attr_mu);
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (!g_attr_set.count({dev, int(sizeof(T) == 8)})) {
      hipFuncAttributes fa;
      HIP_CHECK(hipFuncGetAttributes(&fa, fn));
      HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds_limit() - fa.sharedSizeBytes)));
      g_attr_set.insert({dev, int(sizeof(T) == 8)});
    }
  }
  const int levels = int(B.level_off.size()) - 1;
  unsigned *queue = static_cast<unsigned *>(B.queue.p);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, queue, levels);
  for (int lv = 0; lv < levels; ++lv) {
    const int n = B.level_off[size_t(lv) + 1] - B.level_off[size_t(lv)];
    if (n == 0) continue;
    hipLaunchKernelGGL(k_ragged<T>, dim3(unsigned(std::min(grid, n))), dim3(kThreads),
                       size_t(B.lds_need), s,
                       static_cast<const UnitDev *>(B.units.p) + B.level_off[size_t(lv)], n,
                       static_cast<const PassDev *>(B.passes.p),
                       static_cast<const LineArgs<T> *>(B.args.p), queue + lv);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) RFH_THROW(2, std::string("ragged launch: ") + hipGetErrorString(e));
  }
}

int cu_count() {
  static int n = [] {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      v = 256;
    return v > 0 ? v : 256;
  }();
  return n;
}

void run(Batch &bt, int grid, hipStream_t s) {
  if (bt.dbl)
    launch_batch<double>(*bt.b, grid, s);
  else
    launch_batch<float>(*bt.b, grid, s);
}

double time_ms(Batch &bt, int grid, hipStream_t s, int reps) {
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  run(bt, grid, s);  // warm
  HIP_CHECK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) run(bt, grid, s);
  HIP_CHECK(hipEventRecord(e1, s));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return double(ms) / reps;
}

std::unique_ptr<Built> build_any(const Batch &bt, size_t budget) {
  return bt.dbl ? build<double>(bt.jobs, true, budget) : build<float>(bt.jobs, false, budget);
}

// Autotuner: LDS budget (which sets the four-step splits and how many
// workgroups share a CU) x grid size; each candidate is timed on the batch.
// Results are kept per batch signature (precision, then kind and length of
// every job) for the process and, when a cache directory is set, in a text
// file per GPU architecture and CU count that later processes read:
//   <dir>/ragged_tune_<arch>_<cus>cu.txt, one line per signature:
//   "rfh1 <fnv-1a hash> <second hash> <jobs> <lds budget> <grid>"
// Lines are appended with one write each; unreadable lines are ignored.
std::mutex g_tune_mu;
std::map<std::vector<uint64_t>, std::pair<size_t, int>> g_tuned;
std::string g_cache_dir;  // empty: no disk cache
struct DiskKey {
  uint64_t h1, h2, n;
  bool operator<(const DiskKey &o) const {
    return std::tie(h1, h2, n) < std::tie(o.h1, o.h2, o.n);
  }
};
std::map<DiskKey, std::pair<size_t, int>> g_disk;
std::string g_disk_loaded;  // file whose lines are in g_disk

std::vector<uint64_t> signature(const Batch &bt) {
  std::vector<uint64_t> sig{uint64_t(bt.dbl)};
  for (const auto &j : bt.jobs) sig.push_back((uint64_t(j.kind) << 60) ^ j.n);
  return sig;
}

DiskKey disk_key(const std::vector<uint64_t> &sig) {
  uint64_t h1 = 1469598103934665603ull, h2 = 0x9e3779b97f4a7c15ull;
  for (uint64_t v : sig) {
    for (int b = 0; b < 8; ++b) {
      h1 = (h1 ^ ((v >> (8 * b)) & 0xff)) * 1099511628211ull;
    }
    h2 ^= v + 0x9e3779b97f4a7c15ull + (h2 << 6) + (h2 >> 2);
    h2 *= 0xff51afd7ed558ccdull;
  }
  return DiskKey{h1, h2, uint64_t(sig.size())};
}

// LDS budgets the autotuner tries (the device limit less the kernel's static
// LDS caps the larger two)
std::vector<size_t> tune_budgets() {
  const size_t cap = lds_limit() - 8192;
  return {size_t(32768), size_t(65536), std::min<size_t>(98304, cap),
          std::min<size_t>(147456, cap)};
}

bool budget_ok(size_t b) {
  for (size_t c : tune_budgets())
    if (b == c) return true;
  return false;
}

std::string cache_file() {
  if (g_cache_dir.empty()) return {};
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t pr;
  std::string arch = "gpu";
  if (hipGetDeviceProperties(&pr, dev) == hipSuccess) {
    arch.clear();
    for (const char *c = pr.gcnArchName; *c && *c != ':'; ++c)
      arch += std::isalnum(static_cast<unsigned char>(*c)) ? *c : '_';
  }
  return g_cache_dir + "/ragged_tune_" + arch + "_" + std::to_string(cu_count()) + "cu.txt";
}

// with g_tune_mu held
void load_disk(const std::string &path) {
  if (path == g_disk_loaded) return;
  g_disk.clear();
  g_disk_loaded = path;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string tag;
    DiskKey k{};
    size_t budget = 0;
    int grid = 0;
    if (!(is >> tag >> std::hex >> k.h1 >> k.h2 >> std::dec >> k.n >> budget >> grid)) continue;
    if (tag != "rfh1" || !budget_ok(budget) || grid <= 0 || grid > 64 * cu_count()) continue;
    g_disk[k] = {budget, grid};
  }
}

void append_disk(const std::string &path, const DiskKey &k, size_t budget, int grid) {
  char buf[160];
  const int len = std::snprintf(buf, sizeof buf, "rfh1 %016llx %016llx %llu %zu %d\n",
                                (unsigned long long)k.h1, (unsigned long long)k.h2,
                                (unsigned long long)k.n, budget, grid);
  const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_APPEND, 0644);
  if (fd < 0) return;  // the cache is an optimisation: an unwritable directory is not an error
  const ssize_t w = ::write(fd, buf, size_t(len));
  (void)w;  // a short write leaves a line that later reads ignore
  ::close(fd);
}

void autotune(Batch &bt, hipStream_t s) {
  const auto sig = signature(bt);
  const DiskKey dk = disk_key(sig);
  std::string path;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(sig);
    if (it != g_tuned.end()) {
      bt.b = build_any(bt, it->second.first);
      bt.grid = it->second.second;
      bt.autotuned = 2;
      return;
    }
    path = cache_file();
    if (!path.empty()) {
      load_disk(path);
      auto d = g_disk.find(dk);
      if (d != g_disk.end()) {
        std::unique_ptr<Built> b;
        try {
          b = build_any(bt, d->second.first);
        } catch (const Error &) {
          // a cached budget that no longer plans (another build): tune again
        }
        if (b) {
          bt.b = std::move(b);
          bt.grid = d->second.second;
          bt.autotuned = 3;
          g_tuned[sig] = d->second;
          return;
        }
      }
    }
  }
  // The candidates run on scratch copies of the job buffers, never on the
  // caller's: a job may be in place or read another job's output, and
  // creating a batch must not transform anything.
  Batch tb;
  tb.dbl = bt.dbl;
  tb.jobs = bt.jobs;
  DevBuf scratch;
  {
    const size_t rs = bt.dbl ? 8 : 4, cs = 2 * rs;
    size_t total = 0;
    std::vector<std::pair<size_t, size_t>> off(tb.jobs.size());
    for (size_t i = 0; i < tb.jobs.size(); ++i) {
      const rfh_job_t &j = tb.jobs[i];
      const size_t h = size_t(j.n / 2 + 1);
      const size_t in_b = j.kind == RFH_JOB_C2C ? j.n * cs : (j.kind == RFH_JOB_R2C ? j.n * rs : h * cs);
      const size_t out_b = j.kind == RFH_JOB_C2C ? j.n * cs : (j.kind == RFH_JOB_R2C ? h * cs : j.n * rs);
      off[i].first = total;
      total = align_up(total + in_b, 256);
      off[i].second = total;
      total = align_up(total + out_b, 256);
    }
    if (total) {
      HIP_CHECK(hipMalloc(&scratch.p, total));
      scratch.bytes = total;
      HIP_CHECK(hipMemsetAsync(scratch.p, 0, total, s));
    }
    char *base = static_cast<char *>(scratch.p);
    for (size_t i = 0; i < tb.jobs.size(); ++i) {
      tb.jobs[i].in = base + off[i].first;
      tb.jobs[i].out = base + off[i].second;
    }
  }
  double best = 1e300;
  size_t best_budget = 0;
  int best_grid = 0;
  for (size_t budget : tune_budgets()) {
    try {
      tb.b = build_any(tb, budget);
    } catch (const Error &) {
      continue;
    }
    const int per_cu = std::max(1, int((lds_limit() - 8192) / std::max<size_t>(tb.b->lds_need, 1)));
    for (int wpc : {1, 2, 4}) {
      if (wpc > per_cu) continue;
      const int grid = cu_count() * wpc;
      const double ms = time_ms(tb, grid, s, 3);
      if (ms < best) {
        best = ms;
        best_budget = budget;
        best_grid = grid;
      }
    }
  }
  HIP_CHECK(hipStreamSynchronize(s));  // the timing runs are done with the scratch copies
  tb.b.reset();
  if (!best_budget) RFH_THROW(3, "ragged batch: no LDS budget fits every job");
  bt.b = build_any(bt, best_budget);
  bt.grid = best_grid;
  bt.autotuned = 1;
  bt.tuned_ms = best;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[sig] = {best_budget, best_grid};
  if (!path.empty() && path == cache_file()) {
    append_disk(path, dk, best_budget, best_grid);
    g_disk[dk] = {best_budget, best_grid};
  }
}

// device or managed memory (the persistent kernel dereferences job pointers
// directly; host memory would fault)
bool device_memory(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t at;
  std::memset(&at, 0, sizeof(at));
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

template <typename F>
int guarded_int(F &&f) {
  try {
    fork_guard();
    f();
    return 0;
  } catch (const Error &e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::exception &e) {
    set_error(1, e.what());
    return 1;
  }
}

}  // namespace
}  // namespace rfh

using namespace rfh;

extern "C" {

int rfh_batch_create(const rfh_job_t *jobs, int64_t njobs, bool double_precision, bool tune,
                     void *stream, void **handle) {
  return guarded_int([&] {
    auto *bt = new Batch();
    bt->dbl = double_precision;
    bt->jobs.assign(jobs, jobs + njobs);
    const size_t al = double_precision ? 16 : 8;
    for (const auto &j : bt->jobs) {
      if (j.kind < RFH_JOB_C2C || j.kind > RFH_JOB_C2R) {
        delete bt;
        RFH_THROW(3, "ragged batch: unknown job kind");
      }
      if ((reinterpret_cast<uintptr_t>(j.in) | reinterpret_cast<uintptr_t>(j.out)) % al) {
        delete bt;
        RFH_THROW(3, "ragged batch: job buffers must be aligned to the complex element size");
      }
      if (j.n > 0 && (!device_memory(j.in) || !device_memory(j.out))) {
        delete bt;
        RFH_THROW(3, "ragged batch: job buffers must be device (or managed) memory");
      }
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    try {
      if (tune) {
        autotune(*bt, s);
      } else {
        bt->b = build_any(*bt, 65536);
        bt->grid = cu_count() * 2;
      }
    } catch (...) {
      delete bt;
      throw;
    }
    *handle = bt;
  });
}

int rfh_batch_execute(void *handle, void *stream) {
  return guarded_int([&] {
    Batch *bt = static_cast<Batch *>(handle);
    run(*bt, bt->grid, static_cast<hipStream_t>(stream));
  });
}

int rfh_batch_execute_serial(void *handle, void *stream) {
  // one drop-in call per transform, as a caller of numba_c2c / numba_r2c /
  // numba_c2r would make them, enqueued on the same stream
  hipStream_t prev_s;
  bool prev_b;
  thread_stream(&prev_s, &prev_b);
  rfh_set_stream(stream, false);
  const int rc = guarded_int([&] {
    Batch *bt = static_cast<Batch *>(handle);
    const int64_t rs = bt->dbl ? 8 : 4, cs = 2 * rs;
    struct Arr1 {
      arystruct_t s;
      int64_t ss[2];
    };
    uint64_t axis0 = 0;
    Arr1 ax{};
    ax.s.nitems = 1;
    ax.s.itemsize = 8;
    ax.s.data = &axis0;
    ax.s.shape_and_strides[0] = 1;
    ax.s.shape_and_strides[1] = 8;
    for (const rfh_job_t &j : bt->jobs) {
      if (j.n == 0) continue;
      Arr1 ai{}, ao{};
      const bool c2c = j.kind == RFH_JOB_C2C, r2c = j.kind == RFH_JOB_R2C;
      const int64_t h = int64_t(j.n / 2 + 1);
      ai.s.itemsize = (c2c || !r2c) ? cs : rs;
      ai.s.shape_and_strides[0] = (c2c || r2c) ? int64_t(j.n) : h;
      ai.s.shape_and_strides[1] = ai.s.itemsize;
      ai.s.data = const_cast<void *>(j.in);
      ao.s.itemsize = (c2c || r2c) ? cs : rs;
      ao.s.shape_and_strides[0] = (c2c || !r2c) ? int64_t(j.n) : h;
      ao.s.shape_and_strides[1] = ao.s.itemsize;
      ao.s.data = j.out;
      if (c2c)
        numba_c2c(1, &ai.s, &ao.s, &ax.s, j.forward != 0, j.fct, 1);
      else if (r2c)
        numba_r2c(1, &ai.s, &ao.s, &ax.s, j.forward != 0, j.fct, 1);
      else
        numba_c2r(1, &ai.s, &ao.s, &ax.s, j.forward != 0, j.fct, 1);
      char msg[256];
      if (rfh_last_error(msg, sizeof msg)) RFH_THROW(3, std::string("per-transform call: ") + msg);
    }
  });
  rfh_set_stream(prev_s, prev_b);
  return rc;
}

int rfh_tune_cache(const char *dir, bool clear_memory) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (clear_memory) g_tuned.clear();
  g_cache_dir = dir ? std::string(dir) : std::string();
  while (g_cache_dir.size() > 1 && g_cache_dir.back() == '/') g_cache_dir.pop_back();
  g_disk.clear();
  g_disk_loaded.clear();
  return 0;
}

int rfh_batch_info(void *handle, int *grid, int *lds_bytes, int *autotuned, double *tuned_ms) {
  Batch *bt = static_cast<Batch *>(handle);
  *grid = bt->grid;
  *lds_bytes = bt->b ? bt->b->lds_need : 0;
  *autotuned = bt->autotuned;
  *tuned_ms = bt->tuned_ms;
  return 0;
}

int rfh_batch_destroy(void *handle) {
  return guarded_int([&] { delete static_cast<Batch *>(handle); });
}

}  // extern "C"
