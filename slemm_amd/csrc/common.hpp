// Shared definitions of the gfx950 engine: status handling, handle layout, small helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/slemm_gpu.h"

namespace slemm {

// Reduction dimension (individuals for Z'V, SNPs for Z U) and row counts of both packed copies
// are padded to this many elements: one "super-chunk" of the block-product kernels, 128 bytes
// of 2-bit codes per row, and a whole number of 512-row workgroup tiles.
constexpr int64_t kPad = 512;

inline int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Status plumbing: every C entry point runs its body through guard(), which turns C++
// exceptions into a status code and a per-thread message.
struct Error {
  int code;
  std::string msg;
};
void set_error(const std::string& msg);
[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what, const char* file, int line);
#define HIPCHECK(x) ::slemm::hip_check((x), #x, __FILE__, __LINE__)
#define REQUIRE(cond, msg) \
  do { if (!(cond)) ::slemm::fail(SLEMM_ERR_ARG, (msg)); } while (0)

template <class F>
int guard(F&& f) {
  try {
    f();
    return SLEMM_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return SLEMM_ERR_ALLOC;
  } catch (const std::exception& e) {
    set_error(e.what());
    return SLEMM_ERR_ARG;
  }
}

// Bump allocator over device memory the library does not own (the no-allocation entry point's
// static scratch array).  While one is installed on the calling thread, DevBuf::reserve carves
// its memory from it instead of calling hipMalloc, and fails with SLEMM_ERR_ARG past `cap`.
struct ScratchArena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t nbytes);
};
ScratchArena*& thread_arena();

// Device buffer owned by a handle; grows on demand, never shrinks.  `borrowed` memory comes from
// a ScratchArena and is never freed here.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool borrowed = false;
  void reserve(size_t nbytes);
  void release();
  template <class T> T* as() const { return static_cast<T*>(p); }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace slemm

// Device-resident genotype set.  Two packed copies of the same 2-bit codes:
//  gs: SNP-major,        rows = SNPs (m_rows = round_up(m, kPad)),  n_pad/4 bytes per row;
//  gi: individual-major, rows = individuals (n_rows = n_pad),        m_pad/4 bytes per row.
// Individual i of a row sits in bits 2(i%4), 2(i%4)+1 of byte i/4 (PLINK order); padding
// codes are 00 (dosage 0), so padded rows/columns contribute nothing to any product.
struct slemm_geno {
  int device = 0;
  int64_t n = 0, m = 0;
  int64_t n_pad = 0, m_pad = 0;
  int64_t cap_n_pad = 0;  // individuals (padded) the allocation holds: n_pad <= cap_n_pad
  uint8_t* gs = nullptr;
  uint8_t* gi = nullptr;
  int64_t ld_gs = 0, ld_gi = 0;  // bytes per row

  // model terms (device): gm, gw padded to m_pad with zeros; rr padded to n_pad with zeros.
  int32_t* popcnt = nullptr;  // m
  int32_t* cnt2 = nullptr;    // m: genotypes with both bits set (dosage 2)
  float* gm = nullptr;
  float* gw = nullptr;
  float* rr = nullptr;
  float* Q = nullptr;  // n_pad x c, column-major, padding rows zero
  size_t q_bytes = 0;  // allocated bytes of Q (kept when c drops, so that resetting Q frees nothing)
  int c = 0;
  bool counts_ready = false;
  bool model_ready = false;
  bool rr_unit = true;  // no error-variance weights: r_rsqrt = 1

  // workspaces of the operator pipeline
  slemm::DevBuf x1, x2;        // packed block operands of pass 1 / pass 2
  slemm::DevBuf part1, part2;  // per-workgroup column partial sums
  slemm::DevBuf colsum;        // reduced column sums (double)
  slemm::DevBuf qtv;           // Q' V  (c x t, double)
  slemm::DevBuf tmp;           // n_pad x t float
  slemm::DevBuf kinds;         // per-column operator kinds
  slemm::DevBuf redpart;       // partials of column reductions (double)
  slemm::DevBuf partbuf;       // split-reduction partial products of the block products
  slemm::DevBuf i8buf;         // digit operand and column scales of the int8 narrow blocks
  std::vector<int> kinds_host;
  // block-product variant and grid target per (pass, slot: 0 = hi/lo-in-columns block, else column
  // tiles), set by the autotuner; the defaults are the kernels it picks on MI355X at 100k x 500k:
  // the 32-copy byte-table pipelined kernel (9 = BP_PIPER32) up to five tiles, the pipelined
  // register kernel (5 = BP_PIPE, which falls back to the register kernel) beyond, 2048 workgroups
  int bp_variant[2][9] = {{9, 9, 9, 9, 9, 9, 5, 5, 5}, {9, 9, 9, 9, 9, 9, 5, 5, 5}};
  int bp_target[2][9] = {{2048, 2048, 2048, 2048, 2048, 2048, 2048, 2048, 2048},
                         {2048, 2048, 2048, 2048, 2048, 2048, 2048, 2048, 2048}};
  // int8 block products, set by the autotuner (0: the built-in defaults): workgroup target of the
  // reduction split per (pass, kernel class of i8_class) and the narrow kernel's row tiles per pass
  int i8_target[2][3] = {{0, 0, 0}, {0, 0, 0}};
  int i8_rt[2] = {0, 0};
  std::vector<float> rr_host;    // r_rsqrt (n)
  std::vector<float> rvec_host;  // R (n)
};

namespace slemm {

// Operator pipeline pieces shared by op.hip and lanczos.hip.
void ensure_counts(slemm_geno* h, hipStream_t st);
// Replaces the covariate basis Q (host n x c column-major) of the model terms.
void set_Q(slemm_geno* h, const float* Q, int c);
// model terms of unit error variances, no covariates and the given per-SNP weights (device, m
// floats; zero for SNPs that drop out), as set_model leaves them for unweighted SNPs
void set_model_weights(slemm_geno* h, const float* gw_dev);
// a handle with room for cap_n individuals of src's SNPs, filled by geno_refill
slemm_geno* geno_alloc_like(const slemm_geno* src, int64_t cap_n);
// h <- the individuals keep_idx (strictly increasing) of hs, compacted on the device into h's
// allocation (the model terms must be set again)
void geno_refill(slemm_geno* h, const slemm_geno* hs, const int64_t* keep_idx, int64_t n_keep);
void require_model(const slemm_geno* h);
void run_op(slemm_geno* h, int kind, const int* col_kind, float tau0, const float* V, int64_t ldv,
            int t, float* out, int64_t ldo, int precision, hipStream_t st);

// Column dot products in FP64: res[k] = sum_i A[i,k]*B[i,k] for k < t over n rows.
void col_dots(slemm_geno* h, const float* A, int64_t lda, const float* B, int64_t ldb, int64_t n,
              int t, double* res_dev, hipStream_t st);

}  // namespace slemm
