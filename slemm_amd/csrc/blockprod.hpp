// Launch interface of the genotype block-product kernels (blockprod.hip).
#pragma once

#include "common.hpp"

namespace slemm {

enum { PREC_FP32 = SLEMM_PREC_FP32, PREC_BF16X2 = SLEMM_PREC_BF16X2 };
enum { EPI_PASS1 = 0, EPI_PASS2 = 1 };

struct BlockProdArgs {
  const uint8_t* G = nullptr;  // packed rows
  int64_t ldg = 0;             // bytes per row (= C/4)
  int64_t nsuper = 0;          // C / 512
  const void* X = nullptr;     // packed operand
  int64_t lo_offset16 = 0;     // BF16X2: offset of the lo half in 16-byte units
  int KP = 0;                  // padded column count (row stride of X)
  int t = 0;                   // live columns
  int64_t rows = 0;            // live rows
  const double* colsum = nullptr;  // pass 1: s_k = sum_i X_ik ; pass 2: u_k = sum_j gm_j X_jk
  // pass 1
  const double* xsum = nullptr;    // column sums of the un-centred operand (null: not centred)
  double xsum_n = 0.0;             // number of live reduction rows (n)
  const int32_t* popcnt = nullptr; // per-row bit counts (rows padded with zeros)
  const float* gm = nullptr;
  const float* gw = nullptr;   // null: unweighted
  void* out2 = nullptr;        // pass-2 operand (may be null)
  int out2_mode = 0;
  int64_t out2_lo_offset = 0;  // in bf16 elements
  double* upart = nullptr;     // per-workgroup sum_j gm_j T_jk (may be null)
  // both
  float* out = nullptr;        // column-major result (pass 1: optional, pass 2: required)
  int64_t ldo = 0;
  // pass 2
  const float* rr = nullptr;   // null: no scaling
  const float* v = nullptr;    // null: no tau0 v term
  int64_t ldv = 0;
  float tau0 = 0.f;
  const int* kinds = nullptr;  // per-column kinds (device); LDET columns get tau0 v here
};

// Per-column model terms of a batch of traits on individual subsets (all null: one model).
// Column k belongs to trait col_trait[k]; per-trait arrays are trait-major with the leading
// dimension given (n_pad for individual-indexed, m_pad for SNP-indexed arrays).
struct TraitCols {
  const int* col_trait = nullptr;  // K (device)
  const float* mask = nullptr;     // T x ld_n: 1 for members of the trait's subset, else 0
  const float* q = nullptr;        // T x ld_n: intercept basis mask / sqrt(n_t)
  const double* n_t = nullptr;     // T subset sizes
  const float* gm = nullptr;       // T x ld_m: 2 x A2 frequency over the subset
  const float* gw = nullptr;       // T x ld_m: SNP weights over the subset
  const int32_t* pc = nullptr;     // T x ld_m: dosage sums over the subset
  int64_t ld_n = 0, ld_m = 0;
  // per 16-column tile of the block product: one past the last member of any of its columns'
  // subsets, in 512-individual super-chunks (pass-1 reduction) and in rows (pass-2 output
  // rows); the MFMAs beyond are skipped because the operands there are zero
  const int32_t* tile_nsuper = nullptr;
  const int64_t* tile_rows = nullptr;
};

constexpr int64_t kPackRows = 64;  // rows per workgroup of the pack kernel (one partial sum each)

struct PackArgs {
  const float* V = nullptr;
  int64_t ldv = 0;
  int t = 0;
  int KP = 0;
  int64_t n = 0;
  int64_t rows_pad = 0;        // multiple of 256
  const float* rr = nullptr;
  const float* Q = nullptr;
  int64_t ldq = 0;
  int c = 0;
  const double* qtv = nullptr;  // c x KP (row q stride KP); null: no projection
  const int* kinds = nullptr;
  int mode = 0;
  void* X = nullptr;            // null: only the column sums of x are produced
  int64_t lo_offset16 = 0;
  // tail hi/lo (tail_r > 0): the lo half of vector k >= tail0 is column k + tail_r of the hi
  // block (columns [t, tail0 + 2 tail_r) hold those halves); pad columns beyond get their hi only
  int tail0 = 0, tail_r = 0;
  const double* xsum = nullptr; // if set, x_ik -= float(xsum_k / n) before packing
  double* part = nullptr;       // (rows_pad / kPackRows) x KP partial column sums
  const float* wsum = nullptr;  // optional row weights of the column sums (sum_i wsum_i x_ik)
  TraitCols tc;                 // batched traits: per-column subset, intercept and de-meaning
};

// ---- split-BF16 block product with LDS-staged vector block (blockprod_lds.hip)
constexpr int kLdsRT = 4;  // 16-row tiles per wave; a workgroup covers 256 rows

struct LdsArgs {
  const uint8_t* G = nullptr;
  int64_t ldg = 0;
  int64_t nsuper = 0;         // reduction length / 512
  int64_t ns_per_split = 0;   // super-chunks per gridDim.z slice
  const void* X = nullptr;    // BF16 hi block, lo block at lo_offset16 (16-byte units)
  int64_t lo_offset16 = 0;
  int KP = 0;
  int64_t rows_pad = 0;
  float* part = nullptr;      // nsplit x KP x rows_pad partial products (column-major per split)
  bool squared = false;       // unpack squared dosages {0,1,1,4} instead of {0,1,1,2}
  int ncols = 0;              // live (nonzero) columns of the block, 0: all; the pipelined variants
                              // skip B loads of the others
  bool hl = false;            // hi/lo in columns (one column tile): the block's columns are the hi
                              // halves of the t vectors followed by their lo halves, and each
                              // step runs one MFMA per row tile; X's lo block is not read
  int tail_tile = -1;         // pipelined variants: this tile holds hi and lo halves as columns
                              // (no lo MFMA), tiles after last_tile are all zero (no MFMA)
  int last_tile = -1;         // -1: all tiles live
  // optional per 16-column tile limits (tile = blockIdx.y * NT + j): reduction super-chunks that
  // can be nonzero, and rows from which the tile's products are known to be zero (skipped)
  const int32_t* tile_nsuper = nullptr;
  const int64_t* tile_rows = nullptr;
};

// rows of the pass-1 epilogue's workgroups, the granularity of its partial sums (upart)
constexpr int64_t kFin1Rows = 64;

struct Fin1Args {
  const float* part = nullptr;
  int nsplit = 1;
  int64_t rows_pad = 0, rows = 0;
  int KP = 0, t = 0;
  const double* colsum = nullptr;  // s_k
  const double* xsum = nullptr;    // un-centred sums (null: none)
  double xsum_n = 0.0;
  const int32_t* popcnt = nullptr;
  const float* gm = nullptr;
  const float* gw = nullptr;       // null: unweighted
  void* X2 = nullptr;              // pass-2 operand (null: none): split BF16, or FP32 with x2_fp32
  int64_t lo_offset16 = 0;
  bool x2_fp32 = false;            // X2 as FP32, element (row, k) at row KP + k (int8 kernels)
  double* upart = nullptr;         // (rows_pad/kFin1Rows) x KP partial sums of gm * represented T
  float* out = nullptr;            // optional column-major rows x t
  int64_t ldo = 0;
  bool raw = false;                // out = summed partial products, no centring or weighting
  int tail0 = 0, tail_r = 0;       // tail hi/lo: partials of vector k >= tail0's lo half are column
                                   // k + tail_r; the pass-2 operand is written in the same layout
  TraitCols tc;                    // batched traits: per-column gm, gw, popcounts, subset size
};

struct Fin2Args {
  const float* part = nullptr;
  int nsplit = 1;
  int64_t rows_pad = 0, rows = 0;
  int KP = 0, t = 0;
  const double* colsum = nullptr;  // u_k
  const float* rr = nullptr;
  const float* v = nullptr;
  int64_t ldv = 0;
  float tau0 = 0.f;
  const int* kinds = nullptr;
  float* out = nullptr;
  int64_t ldo = 0;
  int tail0 = 0, tail_r = 0;       // tail hi/lo: partials of vector k >= tail0's lo half are column k + tail_r
  TraitCols tc;                    // batched traits: rows outside the column's subset are zeroed
};

// Block-product variants (selected per shape by the autotuner, slemm_gpu_autotune):
//  BP_REG : B fragments read from global memory by every wave, no LDS (2 waves/SIMD)
//  BP_LDS4: B staged in LDS per 4-step group, double-buffered, codes prefetched a chunk ahead
//  BP_LDS2: B staged in LDS per 2-step group, double-buffered, no code prefetch
//  BP_REG8: as BP_REG with 8 row tiles per wave (half the B traffic per MFMA, 1 wave/SIMD)
//  BP_REGL: as BP_REG with the codes unpacked through a 16-entry LDS table (fewer VALU ops;
//           pays off when few MFMAs share each unpacked operand, i.e. small K)
//  BP_PIPE: as BP_REG, software-pipelined: B fragments requested one step ahead, codes half a
//           super-chunk ahead (NT <= 5; wider tiles run BP_REG)
//  BP_PERSIST: BP_PIPE as a persistent grid (two workgroups per CU walking the tiles in order,
//           so resident workgroups share the vector block in L2)
//  BP_PIPEL: BP_PIPE with the codes unpacked through a 256-entry LDS byte table (a third of the
//           VALU work; for narrow blocks, where the unpack and not the MFMA issue is the limit)
//  BP_PIPER16 / BP_PIPER32: BP_PIPEL with the table replicated 16 / 32 times (32 / 64 KB), lane l
//           reading copy l % R: the copies of an entry sit in consecutive banks, so the random
//           byte lookups of a wave collide less (16) or not at all (32) in LDS banks; BP_PIPER32
//           runs 512-thread workgroups on 512 rows so that the 64 KB table still allows four
//           waves per SIMD
// Under the hi/lo-in-columns layout (LdsArgs::hl) only the pipelined variants exist; the
// others run BP_PIPE.
enum { BP_REG = 0, BP_LDS4 = 1, BP_LDS2 = 2, BP_REG8 = 3, BP_REGL = 4, BP_PIPE = 5, BP_PERSIST = 6,
       BP_PIPEL = 7, BP_PIPER16 = 8, BP_PIPER32 = 9, BP_NVARIANTS = 10 };
// variants that implement the hi/lo-in-columns layout
constexpr bool bp_hl_variant(int v) {
  return v == BP_PIPE || v == BP_PERSIST || v == BP_PIPEL || v == BP_PIPER16 || v == BP_PIPER32;
}
static_assert(BP_PIPE == 5, "slemm_geno's default block-product variant is BP_PIPE");

int lds_nt(int t);
// number of reduction splits so that the grid has about target_wgs workgroups
int lds_splits(int64_t rows_pad, int64_t nsuper, int64_t target_wgs, int64_t& ns_per);
void launch_bp(const LdsArgs& a, int NT, int nsplit, int variant, hipStream_t st);

// Split-BF16 blocks on the int8 matrix cores: the operand X (hi + lo per element) is rendered
// per column as three int8 digits of a 23-bit fixed-point value in the scratch buffer, and the
// block product writes the split partials of column k (and zeros in the column of its lo half,
// where the operand keeps one) into part, as the split-BF16 kernels do.  red_pad: reduction
// length (rows of X, columns of G's packed rows).
constexpr int kI8RT = 8;   // narrow blocks: 16-row tiles per wave; a workgroup covers 512 rows
constexpr int kI8WRT = 4;  // wide blocks: 16-row tiles per wave; a workgroup covers 256 rows
constexpr int kI8NVG = 2;  // wide blocks: 16-vector tiles per column group (6 digit tiles + tail)
// digit columns of t vectors: nvt tiles of 16 vectors (three digit tiles each), then r8 <= 5
// vectors in one tail tile; ndc digit columns in all (nvt = 0, ndc = 3 t for t <= 5)
struct I8Digits {
  int nvt = 0, r8 = 0, ndc = 0;
};
I8Digits i8_digits(int t);
int i8_groups(int t);
// Weights of the digit column sums: row weight w[row] (null: 1), from row col_row[k] of a
// row-major array with leading dimension ldw when col_row is given (per-trait gm).
struct I8SumWeights {
  const float* w = nullptr;
  const int* col_row = nullptr;
  int64_t ldw = 0;
};

struct I8Launch {
  const uint8_t* G = nullptr;
  int64_t ldg = 0;
  int64_t red_pad = 0, rows_pad = 0;
  const void* X = nullptr;  // FP32 operand: element (row, k) at X[row KP + k], red_pad rows
  int KP = 0, t = 0;
  bool squared = false;
  int nsplit = 1;
  int64_t ns_per_split = 0;
  void* scratch = nullptr;  // i8_scratch_bytes(red_pad, t) bytes
  float* part = nullptr;    // nsplit x KP x rows_pad
  int tail0 = 0, tail_r = 0;  // must be 0: the FP32 operand has no lo columns
  // optional per 16-column tile limits of the operand layout (wide blocks only), as in LdsArgs
  const int32_t* tile_nsuper = nullptr;
  const int64_t* tile_rows = nullptr;
  // optional: sum_out[k] = sum over rows of sum_w(row) x (represented operand)_k, from the digits
  double* sum_out = nullptr;
  I8SumWeights sum_w;
  int rt = 0;  // narrow blocks: 16-row tiles per wave (2, 4 or 8; 0 = the default, kI8RT)
};
bool i8_narrow_enabled();
size_t i8_scratch_bytes(int64_t red_pad, int t);
// rt: narrow blocks' row tiles per wave as launched (0 = the default)
int i8_splits(int t, int64_t rows_pad, int64_t nsuper, int64_t target_wgs, bool squared, int rt, int64_t& ns_per);
// kernel class of a t-vector int8 block: 0 narrow (1-5 vectors), 1 one vector tile without a
// tail (9-16), 2 the general wide kernel -- the autotuner keeps one setting per class and pass
int i8_class(int t);
void launch_bp_i8(const I8Launch& l, hipStream_t st);
void launch_fin1(const Fin1Args& a, hipStream_t st);
void launch_fin2(const Fin2Args& a, hipStream_t st);

// Kernel timing hooks (HIP events on the launch stream), shared by all block-product launchers.
struct TimingToken {
  hipEvent_t beg = nullptr;
  bool on = false;
};
TimingToken timing_begin(hipStream_t st);
void timing_end(TimingToken& tok, hipStream_t st, int pass, int precision);

int blockprod_nt(int t);
int64_t blockprod_rows_per_wg(int NT);
void launch_blockprod(const BlockProdArgs& a, int precision, int epi, int NT, int64_t rows_pad, hipStream_t st);
void launch_pack(const PackArgs& a, hipStream_t st);
void launch_sum_parts(const double* part, int64_t nparts, int KP, double* out, hipStream_t st);
void tn_product(const float* A, int64_t lda, int c, const float* B, int64_t ldb, int t, int64_t n,
                double* out, int ldo, DevBuf& scratch, hipStream_t st);
void launch_quad_finish(float* out, int64_t ldo, const float* V, int64_t ldv, const float* Q, int64_t ldq,
                        int c, const double* b, int KP, const int* kinds, int t, int64_t n, float tau0,
                        hipStream_t st);

}  // namespace slemm
