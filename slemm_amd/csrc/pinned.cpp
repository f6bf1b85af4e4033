// Pinned host scratch: page-locked buffers per host thread and slot, grown on demand and kept for
// the life of the process.  Host-to-device copies from them are direct DMA (no staging copy), and
// a buffer that is reused is not page-faulted in again on every call; pinning costs more than the
// small and medium streams that go through here are worth.  A driver runs on one host thread from
// start to end, so the thread's buffers are its own; slots keep independent uses apart.
#include <hip/hip_runtime.h>

#include <cstddef>

#include "driver.hpp"

namespace slemm {

void* pinned_scratch(int slot, size_t bytes) {
  constexpr int kSlots = 4;
  REQUIRE(slot >= 0 && slot < kSlots, "pinned scratch slot out of range");
  thread_local void* buf[kSlots] = {nullptr, nullptr, nullptr, nullptr};
  thread_local size_t cap[kSlots] = {0, 0, 0, 0};
  if (bytes > cap[slot]) {
    void* p = nullptr;
    HIPCHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    if (buf[slot]) (void)hipHostFree(buf[slot]);
    buf[slot] = p;
    cap[slot] = bytes;
  }
  return buf[slot];
}

}  // namespace slemm
