// passes.hpp -- host description of one kernel pass over a line set, shared
// by the per-call launchers (kernels.hip) and the ragged batch executor
// (ragged.hip).
#pragma once

#include <vector>

#include "device.hpp"

namespace rfh {

// One launch of a line kernel: which FFT (d), which load/store operation,
// direction, factor and the op's tables.  lines = lines per workgroup
// (0 = choose from the layout).
struct PassSpec {
  const Direct *d;
  int op;
  int sub = 0;
  int dir = -1;
  double fct = 1.0;
  int64_t bign = 0, other = 0;
  int shift = 0;
  const void *t0 = nullptr, *t1 = nullptr, *t2 = nullptr;
  int nin = 0, nout = 0;
  int lines = 0;        // 0 = choose
  bool real_in = false, real_out = false;
};


struct PassLaunch {
  PassSpec ps;
  LineSet g;
};

// The passes launch_c2c would run for plan p over line set g (tmp = its
// workspace), in order.
std::vector<PassLaunch> describe_c2c(const C2CPlan &p, const LineSet &g, int dir, double fct,
                                     void *tmp, size_t chunk_bytes = 0);
// Throws if a pass that reads or writes the workspace [tmp, tmp + bytes)
// would reach past its end.
void check_tmp_spans(const std::vector<PassLaunch> &passes, const void *tmp, size_t bytes,
                     size_t scalar);
// The single pass of a fused r2c / c2r transform.
PassLaunch describe_real_fused(const RealPlan &rp, RealOp op, const LineSet &g, bool conj_flag,
                               double fct);

// Kernel arguments of a pass (lines per group chosen unless ps.lines is set).
template <typename T>
dev::LineArgs<T> make_line_args(const PassSpec &ps, const LineSet &g);

}  // namespace rfh
