// The operator without device allocation, for callers that cannot call hipMalloc (code under HIP
// graph capture, tight serving loops, harnesses that forbid it).  No reference counterpart.
//
// slemm_gpu_op_noalloc runs exactly the kernels of slemm_gpu_op (same layouts, splits and
// launch shapes, so the same results bit for bit), but every workspace the operator pipeline
// needs -- packed operands, split-reduction partials, column sums, int8 digits, per-column kinds
// -- is carved from `scratch`, one static device array per GPU of kScratchBytes, instead of the
// handle's own hipMalloc'd buffers.  The handle's buffers are set aside for the call and put
// back afterwards, so the allocating entry points keep their own memory.
//
// Every kernel of the pipeline writes the whole region it later reads within the same call
// (padding rows and dead columns included), so nothing in the scratch has to start at zero and
// no call depends on what an earlier call left there.  All calls on one GPU share its array:
// calls that overlap on one GPU (other host threads, or other streams still running an earlier
// call) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "pipeline.hpp"

namespace slemm {

size_t op_workspace_bytes(const slemm_geno* h, int kind, const int* col_kind, int t, int precision);

namespace {

constexpr size_t kScratchBytes = 256ull << 20;
constexpr int kMaxDevices = 64;

// Device address of `scratch` on each GPU, looked up once per device with that device current.
void* scratch_address(int device);

// The handle's operator workspaces, swapped out for the lifetime of the scope: the pipeline's
// reserves then take scratch memory from the installed arena; the destructor puts the handle's
// own buffers (and its cached kinds) back.
class BorrowedWorkspaces {
 public:
  BorrowedWorkspaces(slemm_geno* h, void* base, size_t cap) : h_(h) {
    arena_.base = static_cast<char*>(base);
    arena_.cap = cap;
    DevBuf* b[kN] = {&h->x1, &h->x2, &h->part1, &h->part2, &h->colsum, &h->qtv,
                     &h->tmp, &h->kinds, &h->redpart, &h->partbuf, &h->i8buf};
    for (int i = 0; i < kN; ++i) {
      bufs_[i] = b[i];
      saved_[i] = {b[i]->p, b[i]->bytes, b[i]->borrowed};
      b[i]->p = nullptr;
      b[i]->bytes = 0;
      b[i]->borrowed = false;
    }
    kinds_host_.swap(h->kinds_host);  // the kinds are uploaded again, into the scratch
    thread_arena() = &arena_;
  }
  ~BorrowedWorkspaces() {
    thread_arena() = nullptr;
    for (int i = 0; i < kN; ++i) {
      bufs_[i]->p = saved_[i].p;
      bufs_[i]->bytes = saved_[i].bytes;
      bufs_[i]->borrowed = saved_[i].borrowed;
    }
    h_->kinds_host.swap(kinds_host_);
  }
  BorrowedWorkspaces(const BorrowedWorkspaces&) = delete;
  BorrowedWorkspaces& operator=(const BorrowedWorkspaces&) = delete;

 private:
  static constexpr int kN = 11;
  struct Saved {
    void* p;
    size_t bytes;
    bool borrowed;
  };
  slemm_geno* h_;
  ScratchArena arena_;
  DevBuf* bufs_[kN];
  Saved saved_[kN];
  std::vector<int> kinds_host_;
};

}  // namespace
}  // namespace slemm

static __device__ __align__(256) unsigned char scratch[slemm::kScratchBytes];

namespace slemm {
namespace {

void* scratch_address(int device) {
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  std::lock_guard<std::mutex> lock(mu);
  REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range for the no-allocation scratch");
  if (!addr[device]) {
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the calling thread has `device` current
    if (!p) fail(SLEMM_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    addr[device] = p;
  }
  return addr[device];
}

}  // namespace
}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_op_noalloc(slemm_geno* h, int kind, const int* col_kind, float tau0, const float* V,
                                    int64_t ldv, int t, float* out, int64_t ldo, int precision, void* stream) {
  return guard([&] {
    REQUIRE(h, "null handle");
    REQUIRE(t > 0, "number of columns must be positive");
    HIPCHECK(hipSetDevice(h->device));
    require_model(h);
    check_precision(precision);
    const size_t need = op_workspace_bytes(h, kind, col_kind, t, precision);
    if (need > kScratchBytes)
      fail(SLEMM_ERR_ARG, "slemm_gpu_op_noalloc: this call needs up to " + std::to_string(need >> 20) +
                              " MiB of workspace, more than its " + std::to_string(kScratchBytes >> 20) +
                              " MiB scratch; use slemm_gpu_op or fewer columns");
    void* base = scratch_address(h->device);
    BorrowedWorkspaces ws(h, base, kScratchBytes);
    run_op(h, kind, col_kind, tau0, V, ldv, t, out, ldo, precision, (hipStream_t)stream);
  });
}
