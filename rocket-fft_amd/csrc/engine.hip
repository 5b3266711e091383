// engine.hip -- N-D scheduling and the C ABI of the gfx950 transform engine.
//
// Mirrors the reference's drop-in boundary rocket_fft/_pocketfft_numba.cpp
// (dispatch on itemsize, shape source per entry point) and the N-D
// semantics of rocket_fft/_pocketfft_hdronly.h ("H"):
//   * c2c / dct / dst / r2r: axes one after another, the first reading ain
//     and the rest transforming aout in place, fct on the first axis only
//     (general_nd, H:3568-3607);
//   * r2c: real transform of axes.back(), then c2c of the other axes in
//     place on aout (H:3955-3975);
//   * c2r: c2c of the leading axes into a temporary, then the real
//     back-transform of axes.back() (H:3995-4023);
//   * c2c_sym: r2c, then the Hermitian completion of
//     _pocketfft_numba.cpp:111-143;
//   * genuine Hartley: r2c into a temporary, then the combine loop of
//     H:4061-4091.
// Each axis is one batched launch over all lines of that axis ("line set"),
// never a loop over lines on the host.
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

#include "../../include/rocket_fft_hip.h"
#include "device.hpp"
#include "workspace.hpp"

namespace rfh {

// ------------------------------------------------------- per-thread state

namespace {
thread_local int t_err = 0;
thread_local std::string t_msg;
thread_local hipStream_t t_stream = nullptr;
thread_local bool t_blocking = true;
}  // namespace

void thread_stream(hipStream_t *s, bool *blocking) {
  *s = t_stream;
  *blocking = t_blocking;
}

namespace {
std::atomic<pid_t> g_owner_pid{0};  // process that first used the library
}  // namespace

void fork_guard() {
  const pid_t me = getpid();
  pid_t owner = g_owner_pid.load();
  if (owner == 0 && g_owner_pid.compare_exchange_strong(owner, me)) return;
  if (owner != me)
    RFH_THROW(4, "this process was forked from process " + std::to_string(owner) +
                     " after that process had used the GPU through this library; HIP state "
                     "does not survive fork(): use the library only in processes started "
                     "fresh (spawn) or forked before the parent's first call");
}

void set_error(int code, const std::string &msg) {
  t_err = code;
  t_msg = msg;
  // code 5 (a shift that cannot be fused) is an answer for the caller, who
  // then runs the unfused form, not a failure
  if (code != 5) std::fprintf(stderr, "rocket_fft_hip: %s\n", msg.c_str());
}

// -------------------------------------------------------------- workspace

namespace {

struct HipMem {
  hipStream_t stream = nullptr;
  void *alloc(size_t n) {
    void *p = nullptr;
    HIP_CHECK(hipMalloc(&p, n));
    return p;
  }
  void free(void *p) { (void)hipFree(p); }
  void sync() { HIP_CHECK(hipStreamSynchronize(stream)); }
};

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return cs != hipStreamCaptureStatusNone;
}

// Device scratch owned by one (device, stream): calls on the same stream are
// serialised by `mu` so their temporaries cannot interleave.  Buffers follow
// the graph-safe growth policy of WsSlot (workspace.hpp).
struct Workspace {
  std::mutex mu;
  std::unique_ptr<WsSlot<HipMem>> slot[6];
  hipStream_t stream = nullptr;
  bool capturing = false;  // set per call, with mu held
  void *get(int i, size_t bytes) {
    if (!slot[i]) slot[i].reset(new WsSlot<HipMem>(HipMem{stream}));
    WsSlot<HipMem>::Status st;
    void *p = slot[i]->get(bytes, capturing, &st);
    if (st == WsSlot<HipMem>::kGrowDuringCapture)
      RFH_THROW(3, "the transform needs a larger device workspace (" + std::to_string(bytes) +
                       " B) while the stream is capturing a graph; run the same call once "
                       "before capture");
    return p;
  }
};

std::mutex g_ws_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<Workspace>> g_ws;

Workspace &workspace(hipStream_t s) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto &w = g_ws[std::make_pair(dev, s)];
  if (!w) {
    w.reset(new Workspace());
    w->stream = s;
  }
  return *w;
}

enum Slot { kSlotTwoPass = 0, kSlotReal = 1, kSlotND = 2, kSlotHostIn = 3, kSlotHostOut = 4 };

// ------------------------------------------------------------- geometry

struct Arr {
  int ndim = 0;
  int64_t shape[kMaxDims] = {};
  int64_t str[kMaxDims] = {};  // scalar units
  char *data = nullptr;
};

int64_t prod(const int64_t *s, int nd) {
  int64_t p = 1;
  for (int d = 0; d < nd; ++d) p *= s[d];
  return p;
}

// Lines along `ax` of an array of shape `shape`: batch dims ordered by
// decreasing input stride (the innermost moves fastest through memory),
// size-1 dims dropped, neighbours merged where both layouts allow.
LineSet lines_for_axis(const int64_t *shape, int nd, int ax, const int64_t *si,
                       const int64_t *so, const void *in, void *out) {
  int dims[kMaxDims];
  int nb = 0;
  for (int d = 0; d < nd; ++d)
    if (d != ax && shape[d] > 1) dims[nb++] = d;
  std::stable_sort(dims, dims + nb, [&](int a, int b) {
    return std::llabs(si[a]) > std::llabs(si[b]);
  });
  LineSet g{};
  g.in = in;
  g.out = out;
  g.esi = si[ax];
  g.eso = so[ax];
  g.nlines = 1;
  int m = 0;
  int64_t bd[kMaxDims], bi[kMaxDims], bo[kMaxDims];
  for (int k = 0; k < nb; ++k) {
    const int d = dims[k];
    if (m > 0 && bi[m - 1] == si[d] * shape[d] && bo[m - 1] == so[d] * shape[d]) {
      bd[m - 1] *= shape[d];
      bi[m - 1] = si[d];
      bo[m - 1] = so[d];
    } else {
      bd[m] = shape[d];
      bi[m] = si[d];
      bo[m] = so[d];
      ++m;
    }
  }
  if (m > kMaxBatchDims - 1) RFH_THROW(3, "too many non-mergeable batch dimensions");
  g.nb = m;
  for (int k = 0; k < m; ++k) {
    g.bdim[k] = bd[k];
    g.bsi[k] = bi[k];
    g.bso[k] = bo[k];
    g.nlines *= bd[k];
  }
  return g;
}

LineSet contiguous_lines(void *buf, int64_t nlines, int64_t len) {
  LineSet g{};
  g.in = buf;
  g.out = buf;
  g.nlines = nlines;
  g.nb = 1;
  g.bdim[0] = nlines;
  g.bsi[0] = g.bso[0] = 2 * len;
  g.esi = g.eso = 2;
  return g;
}

void c_strides(const int64_t *shape, int nd, int64_t unit, int64_t *st) {
  int64_t s = unit;
  for (int d = nd - 1; d >= 0; --d) {
    st[d] = s;
    s *= shape[d];
  }
}

// --------------------------------------------------------------- context

struct Ctx {
  hipStream_t s;
  Workspace *ws;
  bool dbl;
  size_t scalar;
};

struct AxisPlan {
  LineSet g;
  std::shared_ptr<C2CPlan> plan;
};

AxisPlan c2c_axis_plan(Ctx &c, const int64_t *shape, int nd, int ax, const int64_t *si,
                       const int64_t *so, const void *in, void *out) {
  const uint64_t n = uint64_t(shape[ax]);
  AxisPlan r;
  r.g = lines_for_axis(shape, nd, ax, si, so, in, out);
  // lines side by side in memory (a batch dim closer than the elements of
  // a line, on either side) want the interleaved plan
  bool inter = false;
  for (int d = 0; d < r.g.nb; ++d)
    inter = inter || std::llabs(r.g.bsi[d]) < std::llabs(r.g.esi) ||
            std::llabs(r.g.bso[d]) < std::llabs(r.g.eso);
  r.plan = get_c2c_plan(n, c.dbl, 0, inter);
  return r;
}

void c2c_axis(Ctx &c, const int64_t *shape, int nd, int ax, const int64_t *si,
              const int64_t *so, const void *in, void *out, bool forward, double fct) {
  AxisPlan ap = c2c_axis_plan(c, shape, nd, ax, si, so, in, out);
  void *tmp = c.ws->get(kSlotTwoPass, c2c_tmp_bytes(*ap.plan, ap.g));
  launch_c2c(*ap.plan, ap.g, forward ? -1 : +1, fct, c.s, tmp);
}

void c2c_nd(Ctx &c, const int64_t *shape, int nd, const int64_t *si, const int64_t *so,
            const uint64_t *axes, int nax, const void *in, void *out, bool forward,
            double fct) {
  for (int i = 0; i < nax; ++i)
    c2c_axis(c, shape, nd, int(axes[i]), i == 0 ? si : so, so, i == 0 ? in : out, out,
             forward, i == 0 ? fct : 1.0);
}

// complex lines of length m in the real-path slot, transformed in place
void *complex_work(Ctx &c, int64_t nlines, int64_t m) {
  return c.ws->get(kSlotReal, size_t(nlines) * m * 2 * c.scalar);
}

void c2c_work(Ctx &c, void *work, int64_t nlines, int64_t m, bool forward) {
  auto plan = get_c2c_plan(uint64_t(m), c.dbl);
  LineSet g = contiguous_lines(work, nlines, m);
  void *tmp = c.ws->get(kSlotTwoPass, c2c_tmp_bytes(*plan, g));
  launch_c2c(*plan, g, forward ? -1 : +1, 1.0, c.s, tmp);
}

// real -> n/2+1 complex along ax (si real units, so complex units)
void r2c_axis(Ctx &c, const int64_t *shape, int nd, int ax, const int64_t *si,
              const int64_t *so, const void *in, void *out, bool forward, double fct) {
  const int64_t n = shape[ax];
  auto rp = get_real_plan(uint64_t(n), c.dbl);
  LineSet g = lines_for_axis(shape, nd, ax, si, so, in, out);
  if (rp->fused) {
    launch_real_fused(*rp, kR2C, g, !forward, fct, c.s);
    return;
  }
  // even n: FFT of the n/2 pairs, then the split twiddle; odd n: full length
  const int64_t m = rp->half ? n / 2 : n;
  void *work = complex_work(c, g.nlines, m);
  RealPrePost rr;
  rr.n = n;
  rr.m = m;
  launch_pre(rp->half ? kPreRealPairs : kPreRealAsComplex, rr, g, c.dbl, work, c.s);
  c2c_work(c, work, g.nlines, m, true);
  rr.conj = !forward;
  launch_post(rp->half ? kPostHalfSpectrum : kPostComplexHalf, rr, g, c.dbl, work, fct, c.s);
}

// n/2+1 complex -> real along ax; shape is the real shape
void c2r_axis(Ctx &c, const int64_t *shape, int nd, int ax, const int64_t *si,
              const int64_t *so, const void *in, void *out, bool forward, double fct) {
  const int64_t n = shape[ax];
  auto rp = get_real_plan(uint64_t(n), c.dbl);
  LineSet g = lines_for_axis(shape, nd, ax, si, so, in, out);
  if (rp->fused) {
    launch_real_fused(*rp, kC2R, g, forward, fct, c.s);
    return;
  }
  const int64_t m = rp->half ? n / 2 : n;
  void *work = complex_work(c, g.nlines, m);
  RealPrePost rr;
  rr.n = n;
  rr.m = m;
  rr.conj = forward;
  launch_pre(rp->half ? kPreHalfSpectrum : kPreHermitian, rr, g, c.dbl, work, c.s);
  c2c_work(c, work, g.nlines, m, false);
  rr.conj = false;
  launch_post(rp->half ? kPostRealPairs : kPostReal, rr, g, c.dbl, work, fct, c.s);
}

void r2c_nd(Ctx &c, const int64_t *shape, int nd, const int64_t *si, const int64_t *so,
            const uint64_t *axes, int nax, const void *in, void *out, bool forward,
            double fct) {
  const int last = int(axes[nax - 1]);
  r2c_axis(c, shape, nd, last, si, so, in, out, forward, fct);
  if (nax == 1) return;
  int64_t shp[kMaxDims];
  std::copy(shape, shape + nd, shp);
  shp[last] = shape[last] / 2 + 1;
  c2c_nd(c, shp, nd, so, so, axes, nax - 1, out, out, forward, 1.0);
}

// one real-to-real axis through pre -> c2c(m) -> post
void r2r_axis(Ctx &c, const int64_t *shape, int nd, int ax, const int64_t *si,
              const int64_t *so, const void *in, void *out, PreMode pre, PostMode post,
              int64_t m, bool fft_forward, RealPrePost rr, double fct) {
  LineSet g = lines_for_axis(shape, nd, ax, si, so, in, out);
  rr.n = shape[ax];
  rr.m = m;
  auto plan = get_c2c_plan(uint64_t(m), c.dbl);
  if (plan->kind == kDirect) {  // the whole work line fits in LDS: one launch
    launch_r2r_fused(*plan, g, pre, post, rr, fft_forward, fct, c.s);
    return;
  }
  void *work = complex_work(c, g.nlines, m);
  launch_pre(pre, rr, g, c.dbl, work, c.s);
  c2c_work(c, work, g.nlines, m, fft_forward);
  launch_post(post, rr, g, c.dbl, work, fct, c.s);
}

// ------------------------------------------------------------- host staging

bool is_device_ptr(const void *p, int *dev) {
  if (!p) return false;
  hipPointerAttribute_t at;
  std::memset(&at, 0, sizeof(at));
  const hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged) {
    *dev = at.device;
    return true;
  }
  return false;
}

void span_of(const Arr &a, size_t scalar, size_t elem_scalars, int64_t *lo, int64_t *hi) {
  int64_t l = 0, h = 0;
  for (int d = 0; d < a.ndim; ++d) {
    const int64_t ext = (a.shape[d] - 1) * a.str[d];
    if (ext < 0) l += ext;
    else h += ext;
  }
  *lo = l * int64_t(scalar);
  *hi = (h + int64_t(elem_scalars)) * int64_t(scalar);
}

// Runs `body` with device-visible data pointers; host arrays are staged
// through device buffers (input span copied in, output span copied in
// first so untouched gaps survive, then copied back).
template <typename F>
void with_device_arrays(Arr &in, size_t in_elem, Arr &out, size_t out_elem, size_t scalar,
                        F &&body) {
  int dev_in = -1, dev_out = -1;
  const bool din = is_device_ptr(in.data, &dev_in);
  const bool dout = is_device_ptr(out.data, &dev_out);
  int prev = 0;
  HIP_CHECK(hipGetDevice(&prev));
  const int dev = dout ? dev_out : (din ? dev_in : prev);
  if (din && dout && dev_in != dev_out) RFH_THROW(3, "input and output on different devices");
  if (dev != prev) HIP_CHECK(hipSetDevice(dev));
  struct Restore {
    int d, p;
    ~Restore() {
      if (d != p) (void)hipSetDevice(p);
    }
  } restore{dev, prev};
  hipStream_t s = t_stream;
  Workspace &ws = workspace(s);
  std::lock_guard<std::mutex> lk(ws.mu);
  ws.capturing = stream_capturing(s);
  const bool host = !din || !dout;
  char *in_host = in.data, *out_host = out.data;
  int64_t ilo = 0, ihi = 0, olo = 0, ohi = 0;
  char *out_dev_base = nullptr;
  if (!dout) {
    span_of(out, scalar, out_elem, &olo, &ohi);
    out_dev_base = static_cast<char *>(ws.get(kSlotHostOut, size_t(ohi - olo)));
    HIP_CHECK(hipMemcpyAsync(out_dev_base, out_host + olo, size_t(ohi - olo),
                             hipMemcpyHostToDevice, s));
    out.data = out_dev_base - olo;
  }
  if (!din) {
    span_of(in, scalar, in_elem, &ilo, &ihi);
    if (!dout && in_host + ilo >= out_host + olo && in_host + ihi <= out_host + ohi) {
      in.data = out.data + (in_host - out_host);  // inside the output span
    } else {
      char *b = static_cast<char *>(ws.get(kSlotHostIn, size_t(ihi - ilo)));
      HIP_CHECK(hipMemcpyAsync(b, in_host + ilo, size_t(ihi - ilo), hipMemcpyHostToDevice, s));
      in.data = b - ilo;
    }
  }
  Ctx c{s, &ws, false, scalar};
  body(c);
  if (!dout)
    HIP_CHECK(hipMemcpyAsync(out_host + olo, out_dev_base, size_t(ohi - olo),
                             hipMemcpyDeviceToHost, s));
  if (host || t_blocking) HIP_CHECK(hipStreamSynchronize(s));
  in.data = in_host;
  out.data = out_host;
}

// --------------------------------------------------------- arystruct views

Arr view(uint64_t ndim, const arystruct_t *a, const int64_t *shape_src, size_t scalar) {
  if (ndim > uint64_t(kMaxDims)) RFH_THROW(3, "too many dimensions");
  Arr r;
  r.ndim = int(ndim);
  r.data = static_cast<char *>(a->data);
  for (int d = 0; d < r.ndim; ++d) {
    r.shape[d] = shape_src[d];
    const int64_t b = a->shape_and_strides[ndim + d];
    if (b % int64_t(scalar)) RFH_THROW(3, "stride is not a multiple of the element size");
    r.str[d] = b / int64_t(scalar);
  }
  return r;
}

void check_axes(const arystruct_t *axes, uint64_t ndim, std::vector<uint64_t> &out) {
  out.assign(static_cast<const uint64_t *>(axes->data),
             static_cast<const uint64_t *>(axes->data) + axes->nitems);
  for (uint64_t a : out)
    if (a >= ndim) RFH_THROW(3, "bad axis number");
}

void check_fits(const Arr &a, const int64_t *need, const char *what) {
  for (int d = 0; d < a.ndim; ++d)
    if (a.shape[d] < need[d]) RFH_THROW(3, std::string(what) + " array is too small");
}

bool any_zero(const int64_t *s, int nd) {
  for (int d = 0; d < nd; ++d)
    if (s[d] == 0) return true;
  return false;
}

template <typename F>
void guarded(F &&f) {
  try {
    t_err = 0;
    fork_guard();
    f();
  } catch (const Error &e) {
    set_error(e.code, e.msg);
  } catch (const std::exception &e) {
    set_error(1, e.what());
  }
}

MirrorGeom mirror_geom(const int64_t *shape, const int64_t *so, int nd,
                       const std::vector<uint64_t> &axes) {
  MirrorGeom mg;
  mg.ndim = nd;
  for (int d = 0; d < nd; ++d) {
    mg.shape[d] = shape[d];
    mg.so[d] = so[d];
  }
  for (uint64_t a : axes) mg.rev[a] = 1;
  mg.last = int(axes.back());
  return mg;
}

}  // namespace

// --------------------------------------------------------------- operations

namespace ops {

void c2c(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
         bool forward, double fct) {
  const bool dbl = ain->itemsize == 16;
  const size_t sc = dbl ? 8 : 4;
  const int64_t *shape = ain->shape_and_strides;
  if (any_zero(shape, int(ndim))) return;
  std::vector<uint64_t> ax;
  check_axes(axes, ndim, ax);
  if (ax.empty()) return;
  Arr in = view(ndim, ain, shape, sc), out = view(ndim, aout, aout->shape_and_strides, sc);
  check_fits(out, shape, "output");
  for (int d = 0; d < out.ndim; ++d) out.shape[d] = shape[d];
  with_device_arrays(in, 2, out, 2, sc, [&](Ctx &c) {
    c.dbl = dbl;
    c2c_nd(c, in.shape, in.ndim, in.str, out.str, ax.data(), int(ax.size()), in.data,
           out.data, forward, fct);
  });
}

// c2c with numpy.roll shifts fused into the loads of each axis's first pass
// and the stores of its last pass: aout = roll(c2c(roll(ain, shift_in)),
// shift_out), the rolls along axes[i] by shift_in[i] / shift_out[i].
void c2c_shifted(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
                 bool forward, double fct, const int64_t *shift_in, const int64_t *shift_out) {
  const bool dbl = ain->itemsize == 16;
  const size_t sc = dbl ? 8 : 4;
  const int64_t *shape = ain->shape_and_strides;
  if (any_zero(shape, int(ndim))) return;
  std::vector<uint64_t> ax;
  check_axes(axes, ndim, ax);
  if (ax.empty()) return;
  for (size_t i = 0; i < ax.size(); ++i)
    for (size_t j = 0; j < i; ++j)
      if (ax[i] == ax[j]) RFH_THROW(3, "a shifted transform takes each axis once");
  std::vector<int64_t> rin(ax.size()), rout(ax.size());
  for (size_t i = 0; i < ax.size(); ++i) {
    const int64_t n = shape[ax[i]];
    const int64_t a = shift_in ? shift_in[i] % n : 0, b = shift_out ? shift_out[i] % n : 0;
    // element j of the rolled input is input element j - shift_in
    rin[i] = a <= 0 ? -a : n - a;
    rout[i] = b < 0 ? b + n : b;
  }
  Arr in = view(ndim, ain, shape, sc), out = view(ndim, aout, aout->shape_and_strides, sc);
  check_fits(out, shape, "output");
  for (int d = 0; d < out.ndim; ++d) out.shape[d] = shape[d];
  with_device_arrays(in, 2, out, 2, sc, [&](Ctx &c) {
    c.dbl = dbl;
    std::vector<AxisPlan> aps;
    for (size_t i = 0; i < ax.size(); ++i) {
      aps.push_back(c2c_axis_plan(c, in.shape, in.ndim, int(ax[i]), i == 0 ? in.str : out.str,
                                  out.str, i == 0 ? in.data : out.data, out.data));
      if (!c2c_rotation_fusable(*aps.back().plan, rin[i], rout[i]))
        RFH_THROW(5, "the shift along axis " + std::to_string(ax[i]) +
                         " cannot be fused into this length's passes");
    }
    for (size_t i = 0; i < aps.size(); ++i) {
      AxisPlan &ap = aps[i];
      ap.g.rot_in = rin[i];
      ap.g.rot_out = rout[i];
      void *tmp = c.ws->get(kSlotTwoPass, c2c_tmp_bytes(*ap.plan, ap.g));
      launch_c2c(*ap.plan, ap.g, forward ? -1 : +1, i == 0 ? fct : 1.0, c.s, tmp);
    }
  });
}

void r2c(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
         bool forward, double fct, bool sym) {
  const bool dbl = ain->itemsize == 8;
  const size_t sc = dbl ? 8 : 4;
  const int64_t *shape = ain->shape_and_strides;
  if (any_zero(shape, int(ndim))) return;
  std::vector<uint64_t> ax;
  check_axes(axes, ndim, ax);
  if (ax.empty()) RFH_THROW(3, "axes must not be empty");
  Arr in = view(ndim, ain, shape, sc), out = view(ndim, aout, aout->shape_and_strides, sc);
  int64_t need[kMaxDims];
  std::copy(shape, shape + ndim, need);
  if (!sym) need[ax.back()] = shape[ax.back()] / 2 + 1;
  check_fits(out, need, "output");  // out keeps its own extent for host staging
  with_device_arrays(in, 1, out, 2, sc, [&](Ctx &c) {
    c.dbl = dbl;
    r2c_nd(c, in.shape, in.ndim, in.str, out.str, ax.data(), int(ax.size()), in.data,
           out.data, forward, fct);
    if (sym) launch_sym_fill(mirror_geom(in.shape, out.str, int(ndim), ax), dbl, out.data, c.s);
  });
}

void c2r(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
         bool forward, double fct) {
  const bool dbl = ain->itemsize == 16;
  const size_t sc = dbl ? 8 : 4;
  const int64_t *shape = aout->shape_and_strides;  // real output shape
  if (any_zero(shape, int(ndim))) return;
  std::vector<uint64_t> ax;
  check_axes(axes, ndim, ax);
  if (ax.empty()) RFH_THROW(3, "axes must not be empty");
  const int last = int(ax.back());
  int64_t shp_in[kMaxDims];
  std::copy(shape, shape + ndim, shp_in);
  shp_in[last] = shape[last] / 2 + 1;
  Arr in = view(ndim, ain, ain->shape_and_strides, sc), out = view(ndim, aout, shape, sc);
  check_fits(in, shp_in, "input");
  for (int d = 0; d < in.ndim; ++d) in.shape[d] = shp_in[d];
  with_device_arrays(in, 2, out, 1, sc, [&](Ctx &c) {
    c.dbl = dbl;
    const void *src = in.data;
    const int64_t *ssi = in.str;
    int64_t st[kMaxDims];
    if (ax.size() > 1) {
      c_strides(shp_in, int(ndim), 2, st);
      void *tmp = c.ws->get(kSlotND, size_t(prod(shp_in, int(ndim))) * 2 * sc);
      c2c_nd(c, shp_in, int(ndim), in.str, st, ax.data(), int(ax.size()) - 1, in.data, tmp,
             forward, 1.0);
      src = tmp;
      ssi = st;
    }
    c2r_axis(c, shape, int(ndim), last, ssi, out.str, src, out.data, forward, fct);
  });
}

void dcst(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
          uint64_t type, double fct, bool ortho, bool cosine) {
  if (type < 1 || type > 4) RFH_THROW(3, cosine ? "invalid DCT type" : "invalid DST type");
  const bool dbl = ain->itemsize == 8;
  const size_t sc = dbl ? 8 : 4;
  const int64_t *shape = ain->shape_and_strides;
  if (any_zero(shape, int(ndim))) return;
  std::vector<uint64_t> ax;
  check_axes(axes, ndim, ax);
  if (ax.empty()) return;
  for (uint64_t a : ax)
    if (type == 1 && cosine && shape[a] < 2) RFH_THROW(3, "DCT-I needs length >= 2");
  Arr in = view(ndim, ain, shape, sc), out = view(ndim, aout, aout->shape_and_strides, sc);
  check_fits(out, shape, "output");
  for (int d = 0; d < out.ndim; ++d) out.shape[d] = shape[d];
  with_device_arrays(in, 1, out, 1, sc, [&](Ctx &c) {
    c.dbl = dbl;
    for (size_t i = 0; i < ax.size(); ++i) {
      const int a = int(ax[i]);
      const int64_t n = shape[a];
      PreMode pre;
      PostMode post;
      int64_t m = 2 * n;
      bool fwd = true;
      switch (type) {
        case 1:
          pre = cosine ? kPreDct1 : kPreDst1;
          post = cosine ? kPostDct1 : kPostDst1;
          m = cosine ? 2 * (n - 1) : 2 * (n + 1);
          break;
        case 2:  // an n-point FFT of the reordered input
          pre = cosine ? kPreDct2N : kPreDst2N;
          post = cosine ? kPostDct2N : kPostDst2N;
          m = n;
          break;
        case 3:  // an n-point inverse FFT, outputs reordered
          pre = cosine ? kPreDct3N : kPreDst3N;
          post = cosine ? kPostDct3N : kPostDst3N;
          m = n;
          fwd = false;
          break;
        default:
          pre = kPreType4;
          post = cosine ? kPostDct4 : kPostDst4;
          break;
      }
      RealPrePost rr;
      rr.ortho = ortho;
      r2r_axis(c, in.shape, in.ndim, a, i == 0 ? in.str : out.str, out.str,
               i == 0 ? in.data : out.data, out.data, pre, post, m, fwd, rr,
               i == 0 ? fct : 1.0);
    }
  });
}

void fftpack(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
             bool r2h, bool forward, double fct) {
  const bool dbl = ain->itemsize == 8;
  const size_t sc = dbl ? 8 : 4;
  const int64_t *shape = ain->shape_and_strides;
  if (any_zero(shape, int(ndim))) return;
  std::vector<uint64_t> ax;
  check_axes(axes, ndim, ax);
  if (ax.empty()) return;
  Arr in = view(ndim, ain, shape, sc), out = view(ndim, aout, aout->shape_and_strides, sc);
  check_fits(out, shape, "output");
  for (int d = 0; d < out.ndim; ++d) out.shape[d] = shape[d];
  with_device_arrays(in, 1, out, 1, sc, [&](Ctx &c) {
    c.dbl = dbl;
    for (size_t i = 0; i < ax.size(); ++i) {
      const int a = int(ax[i]);
      RealPrePost rr;
      // ExecR2R (H:3854-3873): FFTPACK's sign rule on entries 2, 4, ...
      rr.flip = forward ? !r2h : r2h;
      r2r_axis(c, in.shape, in.ndim, a, i == 0 ? in.str : out.str, out.str,
               i == 0 ? in.data : out.data, out.data,
               forward ? kPreRealAsComplex : kPreHalfcomplex,
               forward ? kPostPackHalfcomplex : kPostReal, shape[a], forward, rr,
               i == 0 ? fct : 1.0);
    }
  });
}

void hartley(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
             double fct, bool genuine) {
  const bool dbl = ain->itemsize == 8;
  const size_t sc = dbl ? 8 : 4;
  const int64_t *shape = ain->shape_and_strides;
  if (any_zero(shape, int(ndim))) return;
  std::vector<uint64_t> ax;
  check_axes(axes, ndim, ax);
  if (ax.empty()) {
    if (genuine) RFH_THROW(3, "axes must not be empty");
    return;
  }
  Arr in = view(ndim, ain, shape, sc), out = view(ndim, aout, aout->shape_and_strides, sc);
  check_fits(out, shape, "output");
  for (int d = 0; d < out.ndim; ++d) out.shape[d] = shape[d];
  with_device_arrays(in, 1, out, 1, sc, [&](Ctx &c) {
    c.dbl = dbl;
    if (!genuine || ax.size() == 1) {
      for (size_t i = 0; i < ax.size(); ++i) {
        const int a = int(ax[i]);
        RealPrePost rr;
        r2r_axis(c, in.shape, in.ndim, a, i == 0 ? in.str : out.str, out.str,
                 i == 0 ? in.data : out.data, out.data, kPreRealAsComplex, kPostHartley,
                 shape[a], true, rr, i == 0 ? fct : 1.0);
      }
      return;
    }
    int64_t tshp[kMaxDims], st[kMaxDims];
    std::copy(shape, shape + ndim, tshp);
    tshp[ax.back()] = shape[ax.back()] / 2 + 1;
    c_strides(tshp, int(ndim), 2, st);
    void *tmp = c.ws->get(kSlotND, size_t(prod(tshp, int(ndim))) * 2 * sc);
    r2c_nd(c, in.shape, in.ndim, in.str, st, ax.data(), int(ax.size()), in.data, tmp, true,
           fct);
    MirrorGeom mg = mirror_geom(in.shape, out.str, int(ndim), ax);
    for (int d = 0; d < int(ndim); ++d) mg.st[d] = st[d];
    launch_hartley_combine(mg, dbl, tmp, out.data, c.s);
  });
}

}  // namespace ops

// ------------------------------------------------- allocation-free variant

// Scratch of the allocation-free c2c entry point: the four-step
// intermediates live in this file-scope device array instead of a
// workspace allocated by the library.  Each GPU holds its own copy for the
// life of the process; its device address is looked up once per device.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (current).
void *scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) RFH_THROW(3, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      RFH_THROW(2, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

struct NoAllocScope {
  NoAllocScope() { mfma_no_alloc(true); }
  ~NoAllocScope() { mfma_no_alloc(false); }
};
}  // namespace

namespace ops {

void c2c_noalloc(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
                 bool forward, double fct, hipStream_t s) {
  const bool dbl = ain->itemsize == 16;
  const size_t sc = dbl ? 8 : 4;
  const int64_t *shape = ain->shape_and_strides;
  if (any_zero(shape, int(ndim))) return;
  std::vector<uint64_t> ax;
  check_axes(axes, ndim, ax);
  if (ax.empty()) return;
  Arr in = view(ndim, ain, shape, sc), out = view(ndim, aout, aout->shape_and_strides, sc);
  check_fits(out, shape, "output");
  for (int d = 0; d < out.ndim; ++d) out.shape[d] = shape[d];
  int dev_in = -1, dev_out = -1;
  if (!is_device_ptr(in.data, &dev_in) || !is_device_ptr(out.data, &dev_out) ||
      dev_in != dev_out)
    RFH_THROW(3, "the allocation-free c2c takes device (or managed) memory of one GPU only");
  int prev = 0;
  HIP_CHECK(hipGetDevice(&prev));
  if (prev != dev_out) HIP_CHECK(hipSetDevice(dev_out));
  struct Restore {
    int d, p;
    ~Restore() {
      if (d != p) (void)hipSetDevice(p);
    }
  } restore{dev_out, prev};
  void *tmp = scratch_on(dev_out);
  NoAllocScope no_alloc;
  for (size_t i = 0; i < ax.size(); ++i) {
    const int a = int(ax[i]);
    const int64_t *si = i == 0 ? in.str : out.str;
    LineSet g = lines_for_axis(in.shape, in.ndim, a, si, out.str, i == 0 ? in.data : out.data,
                               out.data);
    bool inter = false;
    for (int d = 0; d < g.nb; ++d)
      inter = inter || std::llabs(g.bsi[d]) < std::llabs(g.esi) ||
              std::llabs(g.bso[d]) < std::llabs(g.eso);
    const uint64_t n = uint64_t(in.shape[a]);
    auto plan = find_c2c_plan(n, dbl, 0, inter);
    if (!plan)
      RFH_THROW(3, "no plan on this GPU yet for length " + std::to_string(n) +
                       (dbl ? " complex128" : " complex64") +
                       (inter ? " (strided axis)" : "") +
                       ": run numba_c2c once with this shape first (plans are built, and "
                       "allocated, by the ordinary entry point)");
    const size_t need = c2c_tmp_bytes(*plan, g, kStaticScratchBytes);
    if (need > kStaticScratchBytes)
      RFH_THROW(3, "the transform needs " + std::to_string(need) +
                       " B of scratch; the allocation-free entry point has " +
                       std::to_string(kStaticScratchBytes) + " B");
    launch_c2c(*plan, g, forward ? -1 : +1, i == 0 ? fct : 1.0, s, tmp, kStaticScratchBytes);
  }
}

}  // namespace ops
}  // namespace rfh

// ================================================================== C ABI

using namespace rfh;

extern "C" {

uint64_t numba_good_size(uint64_t target, bool real) {
  return real ? good_size_real(target) : good_size_cmplx(target);
}

void numba_c2c(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               bool forward, double fct, uint64_t /*nthreads*/) {
  guarded([&] { ops::c2c(ndim, ain, aout, axes, forward, fct); });
}

void numba_dct(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               uint64_t type, double fct, bool ortho, uint64_t /*nthreads*/) {
  guarded([&] { ops::dcst(ndim, ain, aout, axes, type, fct, ortho, true); });
}

void numba_dst(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               uint64_t type, double fct, bool ortho, uint64_t /*nthreads*/) {
  guarded([&] { ops::dcst(ndim, ain, aout, axes, type, fct, ortho, false); });
}

void numba_r2c(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               bool forward, double fct, uint64_t /*nthreads*/) {
  guarded([&] { ops::r2c(ndim, ain, aout, axes, forward, fct, false); });
}

void numba_c2c_sym(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout,
                   arystruct_t *axes, bool forward, double fct, uint64_t /*nthreads*/) {
  guarded([&] { ops::r2c(ndim, ain, aout, axes, forward, fct, true); });
}

void numba_c2r(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
               bool forward, double fct, uint64_t /*nthreads*/) {
  guarded([&] { ops::c2r(ndim, ain, aout, axes, forward, fct); });
}

void numba_r2r_fftpack(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout,
                       arystruct_t *axes, bool real2hermitian, bool forward, double fct,
                       uint64_t /*nthreads*/) {
  guarded([&] { ops::fftpack(ndim, ain, aout, axes, real2hermitian, forward, fct); });
}

void numba_r2r_separable_hartley(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout,
                                 arystruct_t *axes, double fct, uint64_t /*nthreads*/) {
  guarded([&] { ops::hartley(ndim, ain, aout, axes, fct, false); });
}

void numba_r2r_genuine_hartley(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout,
                               arystruct_t *axes, double fct, uint64_t /*nthreads*/) {
  guarded([&] { ops::hartley(ndim, ain, aout, axes, fct, true); });
}

int rfh_c2c_noalloc(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
                    bool forward, double fct, void *stream) {
  guarded([&] {
    ops::c2c_noalloc(ndim, ain, aout, axes, forward, fct, static_cast<hipStream_t>(stream));
  });
  return t_err;
}

int rfh_last_error(char *buf, size_t len) {
  if (buf && len) {
    std::strncpy(buf, t_msg.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return t_err;
}

void rfh_clear_error(void) {
  t_err = 0;
  t_msg.clear();
}

int rfh_set_stream(void *stream, bool blocking) {
  t_stream = static_cast<hipStream_t>(stream);
  t_blocking = blocking;
  return 0;
}

int rfh_set_precision(int mode) {
  const int prev = g_precision_mode;
  if (mode >= 0 && mode <= 3) g_precision_mode = mode;
  return prev;
}

int rfh_mfma_probe(int bf16, const float *a, const float *b, float *d, void *stream) {
  int rc = 0;
  guarded([&] { mfma_probe(bf16, a, b, d, static_cast<hipStream_t>(stream)); });
  rc = t_err;
  return rc;
}

uint64_t rfh_mfma_launches(int mode) { return mfma_launch_count(mode); }

int rfh_kernel_timing(bool enabled) {
  kernel_timing(enabled);
  return 0;
}

int rfh_kernel_timing_report(char *buf, size_t len) {
  std::string r;
  guarded([&] { r = kernel_timing_report(); });
  if (buf && len) {
    std::strncpy(buf, r.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return int(r.size());
}

int rfh_mfma_operands(int mode, int direction, float *out, int cap) {
  return mfma_operand_table(mode, direction, out, cap);
}

int rfh_c2c_shifted(uint64_t ndim, const arystruct_t *ain, arystruct_t *aout, arystruct_t *axes,
                    bool forward, double fct, uint64_t nthreads, const int64_t *shift_in,
                    const int64_t *shift_out) {
  (void)nthreads;
  guarded([&] { ops::c2c_shifted(ndim, ain, aout, axes, forward, fct, shift_in, shift_out); });
  return t_err;
}

int rfh_plan_cache(const char *dir, bool clear_memory) {
  guarded([&] { plan_cache_dir(dir ? std::string(dir) : std::string(), clear_memory); });
  return t_err;
}

int rfh_plan_cache_stats(uint64_t *hits, uint64_t *writes) {
  plan_cache_stats(hits, writes);
  return 0;
}

int rfh_plan_record_roundtrip(uint64_t n, bool double_precision, uint64_t lds_budget,
                              bool interleaved, bool real, const char *path) {
  int rc = -1;
  if (!lds_budget || !path) return -1;
  guarded([&] {
    rc = plan_record_roundtrip(n, double_precision, size_t(lds_budget), interleaved, real, path);
  });
  return t_err ? -1 : rc;
}

int rfh_set_fast_path(bool enabled) {
  g_fast_enabled = enabled;
  return 0;
}

int rfh_cfftp_factors(uint64_t n, uint64_t *out, int cap) {
  auto f = cfftp_factors(n);
  for (int i = 0; i < int(f.size()) && i < cap; ++i) out[i] = f[i];
  return int(f.size());
}

int rfh_rfftp_factors(uint64_t n, uint64_t *out, int cap) {
  auto f = rfftp_factors(n);
  for (int i = 0; i < int(f.size()) && i < cap; ++i) out[i] = f[i];
  return int(f.size());
}

int rfh_plan_choice(uint64_t n, bool real, uint64_t *n2) {
  return plan_bluestein(n, real, n2) ? 1 : 0;
}

int rfh_gpu_schedule(uint64_t n, bool double_precision, int *kind, int *radices, int cap,
                     uint64_t *n1, uint64_t *n2) {
  int count = -1;
  guarded([&] {
    auto p = get_c2c_plan(n, double_precision);
    *kind = p->kind;
    const FftDesc &d = (p->kind == kDirect || p->kind == kBlueDirect) ? p->main.desc
                                                                      : p->pa.desc;
    for (int i = 0; i < d.nst && i < cap; ++i) radices[i] = d.st[i].radix;
    *n1 = p->n1;
    *n2 = p->n2;
    count = d.nst;
  });
  return count;
}

}  // extern "C"
