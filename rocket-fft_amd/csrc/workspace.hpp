// workspace.hpp -- growth policy of one workspace slot (device temporaries of
// the per-call transforms: four-step intermediates, real-path work lines,
// N-D c2r temporaries, host staging).
//
// A buffer handed out while the stream was capturing a graph is referenced
// by that graph for as long as the graph lives, which the library cannot
// see.  So a slot
//   * refuses to grow during capture (allocating and synchronising are not
//     capturable): the call fails with a clear error and the caller warms
//     the transform up eagerly first;
//   * marks a buffer it hands out during capture as pinned;
//   * on growth outside capture frees the old buffer only when it is not
//     pinned (after the stream is synchronised); a pinned buffer is retired
//     instead and freed with the slot (at process end).
// Mem provides: void *alloc(size_t) (throws on failure), void free(void *),
// void sync() (waits for the slot's stream).  Header-only so that the
// policy is unit-tested on the host with a fake Mem.
#pragma once

#include <algorithm>
#include <cstddef>
#include <vector>

namespace rfh {

template <class Mem>
class WsSlot {
 public:
  enum Status { kOk = 0, kGrowDuringCapture = 1 };

  explicit WsSlot(Mem mem = Mem()) : mem_(mem) {}
  WsSlot(const WsSlot &) = delete;
  WsSlot &operator=(const WsSlot &) = delete;
  ~WsSlot() {
    for (void *p : retired_) mem_.free(p);
    if (p_) mem_.free(p_);
  }

  // A buffer of at least `bytes` (nullptr for 0), or nullptr with *st set
  // to kGrowDuringCapture.
  void *get(size_t bytes, bool capturing, Status *st) {
    *st = kOk;
    if (bytes == 0) return nullptr;
    if (bytes_ < bytes) {
      if (capturing) {
        *st = kGrowDuringCapture;
        return nullptr;
      }
      const size_t want = std::max(bytes, bytes_ + bytes_ / 2);
      void *np = mem_.alloc(want);
      if (p_) {
        if (pinned_) {
          retired_.push_back(p_);
        } else {
          mem_.sync();  // earlier calls on the stream may still use it
          mem_.free(p_);
        }
      }
      p_ = np;
      bytes_ = want;
      pinned_ = false;
    }
    if (capturing) pinned_ = true;
    return p_;
  }

  size_t bytes() const { return bytes_; }
  bool pinned() const { return pinned_; }
  size_t retired() const { return retired_.size(); }

 private:
  Mem mem_;
  void *p_ = nullptr;
  size_t bytes_ = 0;
  bool pinned_ = false;
  std::vector<void *> retired_;
};

}  // namespace rfh
