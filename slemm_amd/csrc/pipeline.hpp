// Pieces of the operator pipeline shared by op.hip (single model) and traits.hip (batched traits
// on individual subsets): operand layout, pass-1 packing, the two block-product passes with their
// epilogues, and the Z' pass.
#pragma once

#include "blockprod.hpp"

namespace slemm {

struct Layout {
  int NT, KP;
  size_t x1_bytes, x2_bytes;
  int64_t lo1_16, lo2_16;  // lo-half offsets (16-byte units)
  // split-BF16 blocks of at most 8 vectors keep hi and lo halves as columns [0,t) and [t,2t) of
  // one 16-column block: one MFMA per step instead of a hi and a lo MFMA on mostly empty tiles
  bool hl;
  // tail hi/lo: when the last tile holds r <= 8 vectors (t = 16 q + r, q >= 1), their lo halves
  // are columns [t, t + r) of the hi block and that tile runs no lo MFMA (t = 65: 9 MFMA tiles per
  // step instead of 10); hl is the case q = 0.  tail_r = 0: none.
  int tail0, tail_r;
  int slot;                // autotuner table entry: NT, or 0 for hi/lo-in-columns blocks
  // int8 block products (blockprod_i8.hip): the pass operands are then FP32 (element (row, k) at
  // row KP + k; no hi/lo halves, no lo columns) and the int8 digits are rendered from them
  bool i8;
  int op_mode;             // PREC_FP32 when i8, else the precision
};

struct BpChoice {
  int variant;
  int64_t target;
};

// traits: blocks of the batched traits (their narrow blocks stay on the BF16 kernels);
// allow_i8 = false: the split-BF16 layout whatever the int8 switches say (the autotuner's).
Layout layout_for(const slemm_geno* h, int t, int precision, bool traits = false, bool allow_i8 = true);
// MFMA instructions a t-column block's product issues per 16 x 64 output-by-reduction tile
int op_column_tiles(const slemm_geno* h, int t, int precision);
void check_precision(int precision);
void check_dense(const float* p, int64_t ld, int64_t rows, int t, const char* what);
const int* upload_kinds(slemm_geno* h, int kind, const int* col_kind, int t, bool& any_quad, hipStream_t st);
bool use_lds(int precision);
BpChoice bp_choice(const slemm_geno* h, int pass, int slot);
int64_t pass_splits(const slemm_geno* h, int pass, int slot, int64_t& nsp);
void tail_args(const Layout& L, int t, LdsArgs& la);

// pass-1 operand from V (n x t): rows scaled by rr (optional), QUADFORM columns projected,
// column means removed (within each column's subset when tc is given).  colsum layout:
// [0,KP) sums of the represented centred operand, [KP,2KP) pass-2 centring sums, [2KP,3KP)
// sums of the un-centred operand.
void pack_pass1(slemm_geno* h, const Layout& L, const float* V, int64_t ldv, int t, const float* rr,
                const double* qtv, const int* kinds, int precision, hipStream_t st, bool center = true,
                const TraitCols* tc = nullptr);
// Z'X1 pass, epilogue T = W (Zc'x) -> pass-2 operand x2 / out / upart.
void pass1_lds(slemm_geno* h, const Layout& L, int t, bool weighted, void* x2, double* upart, float* out,
               int64_t ldo, hipStream_t st, bool raw = false, bool squared = false,
               const TraitCols* tc = nullptr);
// Z X2 pass, epilogue out = rr (P - u) (+ tau0 v for LDET columns), masked to each column's subset.
// On the int8 kernels u is replaced by the gm-weighted sums of the represented (digit) operand.
void pass2_lds(slemm_geno* h, const Layout& L, int t, double* u, const float* rr, const float* V,
               int64_t ldv, float tau0, const int* kinds, float* out, int64_t ldo, hipStream_t st,
               const TraitCols* tc = nullptr);
// (W) Zc' (R^-1/2) V -> out (m x t), one pass over the SNP-major copy.
void run_zt(slemm_geno* h, const float* V, int64_t ldv, int t, float* out, int64_t ldo, bool weighted,
            bool scale_rr, int precision, hipStream_t st);

}  // namespace slemm
