// rfh.hpp -- internal declarations of the gfx950 transform engine.
//
// Vocabulary: a "line" is one 1-D transform of the N-D problem (the reference
// engine's multi_iter lines, _pocketfft_hdronly.h:3272-3331).  A line set is
// a batch of lines described by up to kMaxBatchDims batch dimensions with
// independent input/output strides, plus an element stride inside a line.
// All strides and offsets are in SCALAR units (float or double), so a
// contiguous complex line has element stride 2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace rfh {

constexpr int kMaxStages = 24;
constexpr int kMaxBatchDims = 10;
constexpr int kMaxDims = 32;
constexpr int kMaxLinesPerGroup = 256;
constexpr int kThreads = 256;
// LDS kept back from the line kernels' dynamic buffers for their static
// arrays: the per-line offset slots, 8 + 8 + 4 bytes for each of up to
// kMaxLinesPerGroup lines (5120 B), rounded up.  The launch checks the
// kernel's actual static size against the device limit.
constexpr size_t kLineStaticLds = 6144;
static_assert(kLineStaticLds >= kMaxLinesPerGroup * (8 + 8 + 4), "line kernel static LDS");

// ------------------------------------------------------------------ errors

void set_error(int code, const std::string &msg);
// the calling thread's stream settings (rfh_set_stream)
void thread_stream(hipStream_t *s, bool *blocking);
struct Error {
  int code;
  std::string msg;
};
#define RFH_THROW(code, msg) throw ::rfh::Error{(code), (msg)}
// Every entry point that may touch HIP calls this first.  The first call
// records the process; in a child forked after that, HIP state inherited
// from the parent is unusable, so the call fails with code 4 before any HIP
// call is made (a child of a parent that never used the library works).
void fork_guard();
#define HIP_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      RFH_THROW(2, std::string(#expr) + ": " + hipGetErrorString(e_));               \
  } while (0)

// ------------------------------------------------------------------ plans

// One Stockham stage of an LDS-resident FFT: radix, ns = product of the
// radices before it, offsets of its twiddles W_{ns*radix}^{k*r} (r = 1..radix-1,
// k < ns, laid out [(r-1)*ns + k]) and, for odd radices above 5, of its
// table of roots W_radix^q (q < radix) stored as (cos, sin).
struct Stage {
  int radix;
  int ns;
  int twoff;
  int genoff;
};

struct FftDesc {
  int n;
  int nst;
  Stage st[kMaxStages];
};

// Device memory owned by a plan.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf();
  void upload(const void *host, size_t nbytes);
};

// An FFT small enough to run entirely in LDS (two ping-pong buffers).
struct Direct {
  FftDesc desc{};
  DevBuf tw;   // stage twiddles (complex, plan precision)
  DevBuf gen;  // roots for radices 7, 11, 13 and generic odd primes
};

enum PlanKind {
  kDirect = 0,      // one LDS-resident FFT per line
  kBlueDirect = 1,  // Bluestein, its length-bn2 FFTs in LDS
  kTwoPass = 2,     // four-step N = n1 * n2, both passes in LDS
  kBlueTwoPass = 3, // Bluestein in four passes (two four-step FFTs of bn2)
  kMultiPass = 4,   // N = n1 * M: one column pass of n1, then the plan `sub` of M on the rows
  kBlueMulti = 5,   // Bluestein whose length bn2 needs `sub` (multi-pass): elementwise
                    // chirp / kernel passes around two transforms of bn2
};

// Complex transform plan for one length and precision on one device.
struct C2CPlan {
  uint64_t n = 0;
  bool dbl = false;
  int kind = kDirect;
  Direct main;           // kDirect: length n; kBlueDirect: length bn2
  uint64_t n1 = 0, n2 = 0;  // two-pass split of the executed length
  Direct pa, pb;         // pass A (length n1, strided columns), pass B (n2)
  DevBuf big_lo, big_hi; // W_N^e = lo[e & mask] * hi[e >> shift]  (forward sign)
  int shift = 0;
  uint64_t bn2 = 0;      // Bluestein length (0 = none)
  DevBuf bk, bkf;        // chirp e^{+i pi m^2/n} (n) and DFT(chirp)/bn2 (bn2)
  std::shared_ptr<const C2CPlan> sub;  // kMultiPass: rows of M = n2; kBlueMulti: length bn2
  size_t budget = 0;     // dynamic LDS the plan's passes may use per workgroup
};

// Real-input plan used by the fused r2c / c2r kernels: even n runs a
// complex FFT of n/2 (rtw = e^{-2 pi i k/n}, k <= n/2), odd n a complex FFT
// of n with zero imaginary part.  fused == false when that FFT does not fit
// in LDS; the engine then falls back to the generic real path.
struct RealPlan {
  uint64_t n = 0;
  bool dbl = false;
  bool fused = false;  // one LDS pass (d is the complex FFT of n/2 or n)
  bool half = false;   // even n otherwise: pairs -> c2c plan of n/2 -> twiddle
  Direct d;
  DevBuf rtw;
};

// process-wide switch of the specialised power-of-two kernels (fast.hip)
extern bool g_fast_enabled;
// whether the specialised kernel has a form for n (power of two) with
// interleaved lines
bool fast_interleaved_ok(uint64_t n, bool dbl);

// Plans whose LDS passes fit `lds_budget` bytes of dynamic LDS (0 = the
// device limit minus the kernels' static LDS).  `interleaved`: the lines
// lie side by side in memory (a strided axis), so a line group must hold
// enough adjacent lines for wide row segments; lengths whose direct form
// cannot then run as four-step passes instead.
std::shared_ptr<C2CPlan> get_c2c_plan(uint64_t n, bool dbl, size_t lds_budget = 0,
                                      bool interleaved = false);
std::shared_ptr<RealPlan> get_real_plan(uint64_t n, bool dbl, size_t lds_budget = 0);
// The cached plan of get_c2c_plan's arguments on the current device, or
// null when none was built yet (never allocates).
std::shared_ptr<C2CPlan> find_c2c_plan(uint64_t n, bool dbl, size_t lds_budget = 0,
                                       bool interleaved = false);
// On-disk cache of plan records under `dir` (empty: off, the default);
// `clear_memory` also drops the plans built so far in this process.
void plan_cache_dir(const std::string &dir, bool clear_memory);
// Plans read from / written to the disk cache so far.
void plan_cache_stats(uint64_t *hits, uint64_t *writes);
// Host-only: a plan record written to `path` and read back; 0 when every
// field survives, 1 when the file does not read back, 2 when it differs.
int plan_record_roundtrip(uint64_t n, bool dbl, size_t budget, bool inter, bool real,
                          const std::string &path);

// Radix schedule (execution order) for an LDS-resident FFT of length n.
std::vector<int> gpu_radices(uint64_t n);
// Largest line length that fits the LDS budget (0 = device) with `lines`
// lines per group.
uint64_t direct_max(bool dbl, int lines, size_t budget = 0);
size_t lds_limit();

// Reference plan facts restated (bit-exact).
uint64_t good_size_cmplx(uint64_t n);
uint64_t good_size_real(uint64_t n);
uint64_t largest_prime_factor(uint64_t n);
double cost_guess(uint64_t n);
std::vector<uint64_t> cfftp_factors(uint64_t n);
std::vector<uint64_t> rfftp_factors(uint64_t n);
bool plan_bluestein(uint64_t n, bool real, uint64_t *n2);

// ------------------------------------------------------------------ lines

struct LineSet {
  const void *in = nullptr;  // scalar base pointers
  void *out = nullptr;
  int64_t nlines = 0;
  int nb = 0;                // batch dims, innermost last
  int64_t bdim[kMaxBatchDims] = {};
  int64_t bsi[kMaxBatchDims] = {};
  int64_t bso[kMaxBatchDims] = {};
  int64_t esi = 0, eso = 0;  // element strides (scalar units)
  int auxdim = -1;           // batch dim whose index the four-step ops use (-1: nb - 1)
  // circular rotations of the element index (0 <= rot < line length): the
  // transform's input element j is read from position (j + rot_in) mod n of
  // its line and its bin k is written to position (k + rot_out) mod n, which
  // fuses numpy.roll / fftshift / ifftshift along the axis into the
  // transform's loads and stores
  int64_t rot_in = 0, rot_out = 0;
};

// Kernel launchers (kernels.hip).  All enqueue on `s`.
// dir: -1 forward, +1 backward.
// chunk_bytes: largest four-step intermediate per chunk of lines (0: the
// default, sized to stay in the Infinity Cache); tmp must hold
// c2c_tmp_bytes(p, g, chunk_bytes).
void launch_c2c(const C2CPlan &p, const LineSet &g, int dir, double fct, hipStream_t s,
                void *tmp, size_t chunk_bytes = 0);
size_t c2c_tmp_bytes(const C2CPlan &p, const LineSet &g, size_t chunk_bytes = 0);
// Whether plan p runs a line set with element rotations rot_in / rot_out
// (0 <= rot < n) inside its own passes: one-pass plans (direct and
// Bluestein) and every Bluestein plan always; four-step plans when rot_in is
// a whole number of columns (n2) and rot_out of rows (n1); multi-pass plans
// likewise when their row sub-plan takes the output rotation.
bool c2c_rotation_fusable(const C2CPlan &p, int64_t rot_in, int64_t rot_out);

// Opt-in per-launch timing of the line kernels (HIP events on the launching
// stream); the report waits for the recorded launches and lists, one line per
// kernel label, "label\tlaunches\ttotal ms\talgorithmic bytes".
void kernel_timing(bool on);
std::string kernel_timing_report();

enum RealOp { kR2C = 0, kC2R = 1 };
// Fused real transforms: r2c (real lines -> n/2+1 complex, conj when
// conj_flag) or c2r (n/2+1 complex -> real lines, input conjugated when
// conj_flag).  Requires rp.fused.
void launch_real_fused(const RealPlan &rp, RealOp op, const LineSet &g, bool conj_flag,
                       double fct, hipStream_t s);

// Pre/post kernels of the generic real path: real (or half complex) lines
// <-> contiguous complex lines of length m.
enum PreMode {
  kPreRealAsComplex = 0,  // a_j = x_j (j < n), zero beyond
  kPreHermitian,          // full Hermitian spectrum from n/2+1 bins (c2r)
  kPreDct1,
  kPreDst1,
  kPreType2,              // zero padded to 2n
  kPreDct3,
  kPreDst3,
  kPreType4,
  kPreHalfcomplex,        // FFTPACK halfcomplex -> Hermitian spectrum
  kPreRealPairs,          // r2c, even n: a_j = x_2j + i x_2j+1 (j < n/2)
  kPreHalfSpectrum,       // c2r, even n: work line of the length-n/2 inverse FFT
  kPreDct2N,              // DCT-II through an n-point FFT: even inputs up, odd inputs down
  kPreDst2N,              // DST-II: the same with the odd inputs negated
  kPreDct3N,              // DCT-III through an n-point inverse FFT: (x_j - i x_{n-j}) e^{i pi j/2n}
  kPreDst3N,              // DST-III: the same over the reversed input
};
enum PostMode {
  kPostComplexHalf = 0,   // bins 0..n/2 (r2c)
  kPostReal,              // Re
  kPostDct1,
  kPostDst1,
  kPostDct2,
  kPostDst2,
  kPostDct3,
  kPostDst3,
  kPostDct4,
  kPostDst4,
  kPostHartley,           // Re + Im
  kPostPackHalfcomplex,   // bins -> FFTPACK halfcomplex
  kPostHalfSpectrum,      // r2c, even n: bins 0..n/2 from the length-n/2 spectrum
  kPostRealPairs,         // c2r, even n: y_2j = Re z_j, y_2j+1 = Im z_j
  kPostDct2N,             // DCT-II: 2 Re(e^{-i pi k/2n} V_k)
  kPostDst2N,             // DST-II: the same at n-1-k
  kPostDct3N,             // DCT-III: y_2m = Re v_m, y_2m+1 = Re v_{n-1-m}
  kPostDst3N,             // DST-III: the same with odd outputs negated
};
struct RealPrePost {
  int64_t n = 0;      // logical line length (real)
  int64_t m = 0;      // complex work length
  bool ortho = false;
  bool conj = false;  // pre: conjugate input spectrum; post: conjugate output
  bool flip = false;  // negate entries 2, 4, 6, ... (FFTPACK sign rule)
};
// pre -> FFT of length rr.m -> post in one LDS pass; needs p.kind == kDirect
// with p.n == rr.m.
void launch_r2r_fused(const C2CPlan &p, const LineSet &g, PreMode pre, PostMode post,
                      const RealPrePost &rr, bool fft_forward, double fct, hipStream_t s);
void launch_pre(PreMode mode, const RealPrePost &rp, const LineSet &g, bool dbl,
                void *tmp_complex, hipStream_t s);
void launch_post(PostMode mode, const RealPrePost &rp, const LineSet &g, bool dbl,
                 const void *tmp_complex, double fct, hipStream_t s);

// Hermitian completion of c2c_sym and the combine step of the genuine
// Hartley transform (exact overwrite order of the reference's C-order loop).
struct MirrorGeom {
  int ndim = 0;
  int64_t shape[kMaxDims] = {};
  int64_t so[kMaxDims] = {};   // output strides (scalar units)
  int64_t st[kMaxDims] = {};   // temp strides (scalar units; genuine Hartley)
  unsigned char rev[kMaxDims] = {};
  int last = 0;                // axes.back()
};
void launch_sym_fill(const MirrorGeom &mg, bool dbl, void *out, hipStream_t s);
void launch_hartley_combine(const MirrorGeom &mg, bool dbl, const void *tmp, void *out,
                            hipStream_t s);

}  // namespace rfh
