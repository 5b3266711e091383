// lines.hpp -- where line b of a line set starts, and the per-line slots of
// a group of L lines.  Plain C++ with host and device qualifiers under HIP,
// so that the same code is unit-tested on the host.
//
// A line set has nb batch dimensions (innermost last) of extents bdim and
// strides bsi (input) / bso (output), in scalar units; line b is the
// row-major multi-index of b over bdim.  `aux` is the index along batch dim
// auxdim (a four-step pass's row or column number).
//
// A group holds lines b0 .. b0 + L - 1 of which only nl = min(L, nlines - b0)
// exist (the last group of a set can be partial).  The slots of missing
// lines hold zeros: a kernel may run its loops over all L slots for
// arithmetic, but must address memory only through the first nl, and a
// stray read of a missing slot then still lands on line 0 of the set.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define RFH_HD __host__ __device__
#else
#define RFH_HD
#endif

namespace rfh {

struct LineOffsets {
  int64_t oi;  // input offset of the line (scalars)
  int64_t oo;  // output offset
  int aux;     // index along batch dim auxdim
};

RFH_HD inline LineOffsets line_offsets(int64_t b, int nb, const int64_t *bdim,
                                       const int64_t *bsi, const int64_t *bso, int auxdim) {
  LineOffsets r{0, 0, 0};
  int64_t rem = b;
  for (int d = nb - 1; d >= 0; --d) {
    const int64_t q = rem / bdim[d];
    const int64_t idx = rem - q * bdim[d];
    rem = q;
    r.oi += idx * bsi[d];
    r.oo += idx * bso[d];
    if (d == auxdim) r.aux = int(idx);
  }
  return r;
}

// Lines of the group starting at line b0 that exist.
RFH_HD inline int group_lines(int64_t b0, int L, int64_t nlines) {
  const int64_t left = nlines - b0;
  return left <= 0 ? 0 : (left < L ? int(left) : L);
}

// Slot l (< L) of the group starting at b0: the line's offsets, or zeros
// for a line past the end of the set.
RFH_HD inline LineOffsets group_slot(int64_t b0, int l, int L, int64_t nlines, int nb,
                                     const int64_t *bdim, const int64_t *bsi,
                                     const int64_t *bso, int auxdim) {
  if (l >= group_lines(b0, L, nlines)) return LineOffsets{0, 0, 0};
  return line_offsets(b0 + l, nb, bdim, bsi, bso, auxdim);
}

}  // namespace rfh
