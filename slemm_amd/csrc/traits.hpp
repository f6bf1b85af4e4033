// Batched traits on individual subsets of one device-resident bed.
#pragma once

#include <vector>

#include "blockprod.hpp"

// T traits, each with its own subset of the handle's individuals, its own allele frequencies and
// SNP weights over that subset (src/reml.cpp:612-622 evaluated on the subset), and an intercept
// as its only covariate.  Vector columns of a trait are zero outside its subset.
struct slemm_traits {
  slemm_geno* h = nullptr;     // used by every batched call; destroy reads only `device`
  int device = 0;
  int T = 0;
  int plan = 0;                // SLEMM_TRAITS_AUTO, _BATCHED or _SEQUENTIAL (slemm_gpu_traits_set_plan)
  slemm::DevBuf mask, q;       // T x n_pad: membership (0/1) and intercept basis mask/sqrt(n_t)
  slemm::DevBuf n_t;           // T doubles
  slemm::DevBuf gm, gw, pc;    // T x m_pad: per-trait 2 x A2 frequency, weights, dosage sums
  slemm::DevBuf col_trait, qd; // per-column trait ids of the last batch; per-column Q'v
  slemm::DevBuf tile_nsuper, tile_rows;  // per 16-column tile of the last batch: extent of its subsets
  std::vector<int> col_trait_host;
  std::vector<int64_t> extent;           // one past the last member of each subset
  std::vector<int> active;               // live columns of the running seed (empty: all)
  std::vector<int64_t> n_host;           // subset sizes
  std::vector<int64_t> monomorphic;      // SNPs monomorphic within each subset (weight set to 0)
  std::vector<std::vector<int64_t>> members;  // member individuals of each subset, ascending

  slemm::TraitCols cols() const {
    slemm::TraitCols c;
    c.col_trait = col_trait.as<int>();
    c.mask = mask.as<float>();
    c.q = q.as<float>();
    c.n_t = n_t.as<double>();
    c.gm = gm.as<float>();
    c.gw = gw.as<float>();
    c.pc = pc.as<int32_t>();
    c.ld_n = h->n_pad;
    c.ld_m = h->m_pad;
    c.tile_nsuper = tile_nsuper.as<int32_t>();
    c.tile_rows = tile_rows.as<int64_t>();
    return c;
  }
};

namespace slemm {
// Batched operator: column k gets the lnr_op_select operator (src/reml.cpp:96-118) of trait
// col_trait[k] on that trait's subset, kind col_kind[k].
void run_op_traits(slemm_traits* tr, const int* col_trait, const int* col_kind, float tau0, const float* V,
                   int64_t ldv, int t, float* out, int64_t ldo, int precision, hipStream_t st);
// The batched SLDF seed block B (n x T (2 + n_V), leading dimension n, zero outside each subset):
// columns [y~_0 .. y~_{T-1}, q_0 .. q_{T-1}, the n_V probes of trait 0, of trait 1, ...], written on
// the device.  yq (host): per trait t its n_t projected phenotype values then its n_t intercept
// basis values, traits one after another; bits (host): the sign bits of the first max_t n_t n_V
// Rademacher draws (rademacher_sign_bits), trait t's probe j entry i being draw j n_t + i.
void traits_fill_seed_block(const slemm_traits* tr, int n_V, const std::vector<float>& yq,
                            const std::vector<uint8_t>& bits, float* B);
}  // namespace slemm
