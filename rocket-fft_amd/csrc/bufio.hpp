// bufio.hpp -- 64-bit complex64 loads and stores through buffer
// instructions: one address register per thread, per-element offsets as
// scalar operands, and the resource's range check bounding every access to
// [base, base + bytes) (out-of-range loads read zeros, stores are dropped).
// Both carry the non-temporal cache policy: the pipelined passes read each
// element once and write it once, and streaming them past the caches'
// normal replacement measured faster (64 x 2^20 c64: 0.444 -> 0.422 ms;
// pass A 234 -> 222 us, pass B 208 -> 200 us; lines of 4096 unchanged).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>

#include "device.hpp"

namespace rfh {
namespace bufio {

typedef __attribute__((ext_vector_type(2))) unsigned u2;

constexpr int kNonTemporal = 2;  // cache-policy operand of the buffer builtins: nt

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, int(bytes),
                                           0x00020000);
}

__device__ __forceinline__ dev::C2<float> bload(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                uint32_t soff) {
  const u2 v = __builtin_amdgcn_raw_buffer_load_b64(r, int(voff), int(soff), kNonTemporal);
  // components through scalar temporaries: a bit cast of a vector element
  // expression itself would read the vector's first element
  const unsigned lo = v.x, hi = v.y;
  return dev::C2<float>{__builtin_bit_cast(float, lo), __builtin_bit_cast(float, hi)};
}

__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                       dev::C2<float> x) {
  u2 v;
  v[0] = __builtin_bit_cast(unsigned, x.x);
  v[1] = __builtin_bit_cast(unsigned, x.y);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, int(voff), int(soff), kNonTemporal);
}

// Largest scalar offset (floats) reached from a.in (in) or a.out by a line
// set of n-element lines, or -1 when a stride is negative.
inline int64_t line_span(const dev::LineArgs<float> &a, int n, bool in) {
  const int64_t es = in ? a.esi : a.eso;
  if (es < 0) return -1;
  int64_t s = int64_t(n - 1) * es + 2;
  for (int d = 0; d < a.nb; ++d) {
    const int64_t st = in ? a.bsi[d] : a.bso[d];
    if (st < 0 || a.bdim[d] < 1) return -1;
    s += (a.bdim[d] - 1) * st;
  }
  return s;
}

// compute units of the current device (256 when the query fails), queried
// once per device
inline int device_cus() {
  constexpr int kDevs = 64;
  static std::atomic<int> cache[kDevs];
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev >= 0 && dev < kDevs && (v = cache[dev].load(std::memory_order_relaxed)) > 0) return v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      v <= 0)
    return 256;
  if (dev >= 0 && dev < kDevs) cache[dev].store(v, std::memory_order_relaxed);
  return v;
}

}  // namespace bufio
}  // namespace rfh
