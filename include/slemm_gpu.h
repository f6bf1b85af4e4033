/*
 * slemm_gpu.h -- C ABI of the MI355X (gfx950) engine for SLEMM's stochastic-Lanczos REML path.
 *
 * The reference (jiang18/slemm v0.90.1) has no plugin or FFI layer: its operator seam is the
 * set of C++ functions in src/reml.cpp that the Krylov engine and the REML drivers call
 * (lnr_op_select, rmatvec, weighted_rmatvec, matvec, L_Seed, L_Solve, SLQ_LDet) plus the
 * bed decoding and counting in src/input.cpp / src/weight.cpp and the GEBV loop in
 * src/pred.cpp.  Every entry point below names the reference function it replaces.
 *
 * Conventions
 *  - Every function returns an int status (SLEMM_OK == 0).  On failure a message is kept per
 *    host thread and returned by slemm_gpu_last_error().  No C++ exception crosses this ABI.
 *  - Dense matrices are column-major float32 with an explicit leading dimension, exactly like
 *    the reference's Eigen MatrixXf arguments.  Pointers documented as "device" live in HBM
 *    (hipMalloc'd or torch tensors); pointers documented as "host" are ordinary memory.
 *  - `stream` is a hipStream_t passed as void* (NULL = the default stream).  Calls that
 *    return host results synchronise that stream before returning.
 *  - A handle is used by one host thread at a time.
 *  - The REML drivers keep small page-locked host buffers per calling thread (per-step scalars,
 *    probe sign bits, Monte Carlo draws), grown on demand and held until the process ends.
 */
#ifndef SLEMM_GPU_H_
#define SLEMM_GPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SLEMM_OK = 0,
  SLEMM_ERR_ARG = 1,     /* invalid argument (sizes, null pointers, unknown kind) */
  SLEMM_ERR_HIP = 2,     /* a HIP runtime call failed */
  SLEMM_ERR_ALLOC = 3,   /* device allocation failed */
  SLEMM_ERR_STATE = 4,   /* handle not ready (e.g. model not set) */
  SLEMM_ERR_RANK = 5,    /* fixed-effects design not of full rank (src/reml.cpp:586,769) */
  SLEMM_ERR_DATA = 6,    /* malformed input data (e.g. truncated bed, src/input.cpp:366) */
  SLEMM_ERR_NUMERIC = 7  /* numerical breakdown (e.g. an information matrix not positive definite) */
};

/* Linear operator kinds of lnr_op_select (src/reml.cpp:96-118). */
enum {
  SLEMM_OP_LDET = 0,     /* H0_ldet_mv     (src/reml.cpp:52-94):  R^-1/2 Zc W Zc' R^-1/2 v + tau0 v */
  SLEMM_OP_QUADFORM = 1  /* H0_quadform_mv (src/reml.cpp:8-50):   (I-QQ')R^-1/2 Zc W Zc' R^-1/2 (I-QQ') v + tau0 v */
};

/* Arithmetic used for the genotype block products. */
enum {
  SLEMM_PREC_FP32 = 0,   /* FP32 operands on FP32 MFMA, FP32 accumulation (plain path) */
  SLEMM_PREC_BF16X2 = 1  /* vector block split into BF16 hi + lo, BF16 MFMA, FP32 accumulation */
};

typedef struct slemm_geno slemm_geno;   /* device-resident packed genotypes + model terms */
typedef struct slemm_lseed slemm_lseed; /* device-resident Lanczos seed system (struct lseed) */
typedef struct slemm_traits slemm_traits; /* batch of traits on individual subsets of one handle */

const char* slemm_gpu_last_error(void);
int slemm_gpu_version(void);
int slemm_gpu_device_count(int* count);

/* ---------------------------------------------------------------- genotypes ---------------- */

/* Replaces the bed part of read_plink_bin_geno_file_and_generate_pxvb_data
 * (src/input.cpp:343-386: individual compaction) and the marker reorder (:446-454).
 * bed_rows: host, m_raw rows of row_bytes = ceil(n_raw/4) bytes (the .bed body after the 3
 * magic bytes).  keep_idx: host, n_keep strictly increasing fam indices of retained
 * individuals (NULL = all).  snp_idx: host, the m bed rows to retain, in analysis order
 * (NULL = all rows in file order).  Codes are kept verbatim (00,01,10,11); dosage = low+high. */
int slemm_gpu_geno_create(const uint8_t* bed_rows, int64_t n_raw, int64_t m_raw, int64_t row_bytes,
                          const int64_t* keep_idx, int64_t n_keep,
                          const int64_t* snp_idx, int64_t m, slemm_geno** out);

/* New rows-subset handle built on the device from an existing one (MAF/HWE filtering and
 * group ordering of src/input.cpp:389-454 without re-reading the bed). */
int slemm_gpu_geno_subset(const slemm_geno* h, const int64_t* snp_idx, int64_t m, slemm_geno** out);

/* Handle from a PLINK .bed file (path; 3-byte SNP-major magic checked): the same result as
 * slemm_gpu_geno_create on the file's rows, but the body is streamed from disk through pinned
 * double buffers, overlapped with the upload and compaction, and only the SNPs in snp_idx (NULL:
 * all) are read -- the replacement of the byte-at-a-time ifstream of src/input.cpp:359-384 for
 * beds larger than host memory.  SLEMM_ERR_DATA for a missing, foreign or short file. */
int slemm_gpu_geno_load_bed(const char* bed_path, int64_t n_raw, int64_t m_raw, const int64_t* keep_idx,
                            int64_t n_keep, const int64_t* snp_idx, int64_t m, slemm_geno** out);

/* New handle with the individuals keep_idx (strictly increasing, host) of an existing one, all
 * SNPs: the individual compaction of src/input.cpp:360-377 applied on the device to resident
 * rows (analyses of individual subsets without re-reading the bed). */
int slemm_gpu_geno_keep(const slemm_geno* h, const int64_t* keep_idx, int64_t n_keep, slemm_geno** out);

/* Synthetic bed generated in HBM: per-SNP p ~ U(maf_lo, maf_hi), genotype ~ Binomial(2,p)
 * (A2 count), `missing` fraction written as code 01.  Counter-based hashing, so the content
 * depends only on (seed, n, m); no reference counterpart (test and benchmark input). */
int slemm_gpu_geno_synthesize(int64_t n, int64_t m, uint64_t seed, float maf_lo, float maf_hi,
                              float missing, slemm_geno** out);
/* Same, with the allele frequency of every SNP given (p: m floats in [0,1], host memory), e.g.
 * drawn from a Beta distribution on the host.  No reference counterpart. */
int slemm_gpu_geno_synthesize_p(int64_t n, int64_t m, uint64_t seed, const float* p, float missing,
                                slemm_geno** out);

int slemm_gpu_geno_destroy(slemm_geno* h);
int slemm_gpu_geno_dims(const slemm_geno* h, int64_t* n, int64_t* m);

/* Copy packed SNP-major rows [j0, j0+cnt) to host (ceil(n/4) bytes each, PLINK code layout,
 * padding codes 00).  Exposes the decoded kmat of src/input.cpp:374-377 for bit-exact checks. */
int slemm_gpu_geno_download(const slemm_geno* h, int64_t j0, int64_t cnt, uint8_t* rows);

/* Per-SNP bit counts.  popcnt = number of set bits (the nz_bits of src/reml.cpp:616-619 and
 * src/weight.cpp:8-11); n2 = both bits, n1 = exactly one bit, n0 = none (src/input.cpp:14-19).
 * gm (optional) = float(popcnt)/float(n), the genotype mean of src/reml.cpp:620.
 * All outputs are host arrays of length m; any may be NULL. */
int slemm_gpu_geno_counts(slemm_geno* h, int32_t* popcnt, int32_t* n0, int32_t* n1, int32_t* n2,
                          float* gm);

/* ---------------------------------------------------------------- model terms -------------- */

/* Sets the terms the operators need, all host arrays:
 *   persnp (m, NULL = all ones): per-SNP heritability weights; gw = (gm(2-gm)/2)^-1 / m *
 *     persnp/mean(persnp) as src/reml.cpp:622;
 *   rvec (n, NULL = all ones): error-variance weights R; r_rsqrt = rvec^-1/2 (src/reml.cpp:574);
 *   Q (n x c column-major, ldq): orthonormal basis of R^-1/2 X (c may be 0).
 * gm is recomputed on the device from the bit counts. */
int slemm_gpu_set_model(slemm_geno* h, const float* persnp, const float* rvec,
                        const float* Q, int64_t ldq, int c);

/* Host copies of the model terms (each may be NULL). */
int slemm_gpu_get_model(const slemm_geno* h, float* gm, float* gw, float* r_rsqrt);

/* Overrides the per-SNP weights W of the operators with gw (host, m floats, finite, >= 0) after
 * set_model, e.g. sigma_g^2-scaled or per-window weights under differentiation.  Replaces the
 * gw line of src/reml.cpp:622 for callers that form the weights themselves. */
int slemm_gpu_set_gw(slemm_geno* h, const float* gw);

/* ---------------------------------------------------------------- operators ---------------- */

/* lnr_op_select (src/reml.cpp:96-118) over all t columns in one batch.
 * V, out: device, n x t column-major (ldv, ldo >= n).  kind applies to every column unless
 * col_kind (host, t entries) is given, in which case each column uses its own kind, so
 * probes (LDET) and the phenotype (QUADFORM) share one pass over the genotypes. */
int slemm_gpu_op(slemm_geno* h, int kind, const int* col_kind, float tau0,
                 const float* V, int64_t ldv, int t, float* out, int64_t ldo,
                 int precision, void* stream);

/* slemm_gpu_op without device allocation.  No counterpart in the reference: it is for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Same arguments, kernels and results as slemm_gpu_op; every workspace of
 * the call (packed operands, split-reduction partials, column sums, int8 digits, kinds) comes from
 * one static 256 MiB device array per GPU, looked up on the first call on each device.  A call
 * whose workspace would exceed 256 MiB fails with SLEMM_ERR_ARG before any work is queued (for
 * example n = 100k x m = 500k fits one vector, about 180 MiB, but not the 65-column block).  All
 * calls on one GPU share its array and it keeps the last call's contents: calls that overlap on
 * one GPU (other threads, or an earlier call still running on another stream) race. */
int slemm_gpu_op_noalloc(slemm_geno* h, int kind, const int* col_kind, float tau0,
                         const float* V, int64_t ldv, int t, float* out, int64_t ldo,
                         int precision, void* stream);

/* rmatvec (src/reml.cpp:159-195): out (m x t, ldo) = Zc' V.
 * weighted_rmatvec (src/reml.cpp:120-157) when weighted != 0: out = W Zc' V.
 * V device n x t (ldv).  If scale_rr != 0 the rows of V are first multiplied by r_rsqrt. */
int slemm_gpu_zt(slemm_geno* h, const float* V, int64_t ldv, int t, float* out, int64_t ldo,
                 int weighted, int scale_rr, int precision, void* stream);

/* Adjoint product: out_j = sum_k coef_k (Zc'A)_jk (Zc'B)_jk, j < m (device doubles).
 * A, B device n x t (lda, ldb); with scale_rr != 0 their rows are first multiplied by r_rsqrt.
 * coef: t host doubles.  Both blocks go through one pass over the bed as a 2t-column Zc'
 * product (the rmatvec of src/reml.cpp:159-195), followed by an FP64 row reduction.  With
 * A = P V, B = V for probe columns V and A = B = Py this gives dl/dgw_j, the derivative of the
 * REML log-likelihood with respect to the SNP weights (no reference counterpart; the reference
 * has no derivatives). */
int slemm_gpu_adjoint(slemm_geno* h, const float* A, int64_t lda, const float* B, int64_t ldb, int t,
                      const double* coef, int scale_rr, double* out, int precision, void* stream);

/* matvec (src/reml.cpp:197-233): out (n x t, ldo) = Zc U with U device m x t (ldu). */
int slemm_gpu_z(slemm_geno* h, const float* U, int64_t ldu, int t, float* out, int64_t ldo,
                int precision, void* stream);

/* computeTotalGeneticValuePLINK inner loop (src/pred.cpp:112-130): gebv (device, n doubles)
 * = sum_j alpha_j * z_ij over raw dosages z (not centred), with z -> 2-z where flip_j != 0.
 * alpha device (m doubles), flip host or NULL (m bytes).  FP64 accumulation. */
int slemm_gpu_gebv(slemm_geno* h, const double* alpha, const uint8_t* flip, double* gebv,
                   void* stream);
/* The same GEBV at FP32 accuracy in one pass over the individual-major codes (no reference
 * counterpart: a caller that needs the scores to ~1e-6 relative rather than the reference's
 * summation order bit for bit): Z u with u_j = +-alpha_j on the block-product kernels (the int8
 * narrow kernel on the default path), plus 2 sum of alpha over the flipped SNPs in FP64.  Same
 * arguments as slemm_gpu_gebv; uses the handle's workspaces. */
int slemm_gpu_gebv_fast(slemm_geno* h, const double* alpha, const uint8_t* flip, double* gebv,
                        void* stream);

/* GRAMMAR-gamma score terms of bin/slemm_gamma.py:306-319 for every SNP j (device outputs,
 * m floats each): xy_j = x_c' Py, xx_j = || (I-QQ') R^-1/2 x_c ||^2, with x_c the mean-
 * centred dosage (mean = popcnt/n).  Py device (n floats).  skip_mac: SNPs with
 * popcnt <= skip_mac or 2n-popcnt <= skip_mac get NaN (the reference's `--mac` filter). */
int slemm_gpu_gwas_score(slemm_geno* h, const float* Py, float skip_mac, float* xy, float* xx,
                         void* stream);

/* ---------------------------------------------------------------- Krylov engine ------------ */

/* L_Seed (src/reml.cpp:246-339): t lock-step Lanczos/CG recurrences on the operator of
 * `kind` (or col_kind per column) at shift tau0, starting from B (device n x t, ldb).
 * Dot products, norms and all recurrence scalars are FP64.
 * col_group (host, t entries, or NULL = one group): several independent seed systems in one
 * batch -- each group stops on its own, exactly as a separate L_Seed call would, when
 * max over its columns of |rho|/||b|| <= tol, or after maxit-1 steps; the batch ends when
 * every group has stopped.  Once a group has stopped, the operator runs on the remaining
 * columns only when that needs fewer MFMA tiles (SLEMM_LSEED_COMPACT=0 keeps every column).
 * keep_cols: the leading keep_cols columns keep every basis vector U_j in HBM (needed by
 * slemm_gpu_lsolve); the other columns keep only their tridiagonal coefficients (SLQ probes;
 * the reference frees their U at :634-636). */
int slemm_gpu_lseed(slemm_geno* h, int kind, const int* col_kind, const int* col_group, float tau0,
                    const float* B, int64_t ldb, int t, float tol, int maxit, int keep_cols,
                    int precision, int verbose, void* stream, slemm_lseed** out);

/* t, the largest step count, and whether every group converged. */
int slemm_gpu_lseed_info(const slemm_lseed* s, int* t, int* steps, int* converged);
/* Per-column step count j_k (host, t ints). */
int slemm_gpu_lseed_col_steps(const slemm_lseed* s, int* steps);
/* beta (host, t x (steps+1), column-major) and delta (t x steps); entries past a column's own
 * step count are zero.  Same layout as struct lseed's beta/delta (src/reml.cpp:333-336). */
int slemm_gpu_lseed_coeffs(const slemm_lseed* s, double* beta, double* delta);
/* A seed keeps its own copy of everything it needs: it stays valid after its genotype handle has
 * been destroyed. */
int slemm_gpu_lseed_destroy(slemm_lseed* s);

/* L_Solve (src/reml.cpp:369-431): X (device n x t2, ldx) solving (A + sigma I) X = B by
 * replaying the shifted CG recurrence on the stored basis of columns [col0, col0+t2), which
 * must be kept columns.  B device n x t2 (ldb) -- the same vectors the seed started from. */
int slemm_gpu_lsolve(const slemm_lseed* s, int col0, int t2, const float* B, int64_t ldb,
                     float sigma, float tol, float* X, int64_t ldx, void* stream);

/* Tridiagonal spectral data for SLQ (src/reml.cpp:637-646) in FP64 on the host for column k:
 * D, W (host, steps_k each): eigenvalues and first eigenvector components. */
int slemm_gpu_lseed_spectrum(const slemm_lseed* s, int k, double* D, double* W);

/* SLQ_LDet (src/reml.cpp:440-451) in FP64 over the probe columns [col0, col0+nv):
 * n/nv * sum_k sum_i W_ki^2 log(D_ki + sigma). */
int slemm_gpu_slq_ldet(const slemm_lseed* s, int col0, int nv, double sigma, double* ldet);

/* ---------------------------------------------------------------- REML drivers ------------- */

typedef struct {
  float vg, ve, h2, lrt;        /* variance components, h2, LRT (0 for LMC) */
  int lanczos_steps[3];         /* Lanczos steps of the seed systems (y, Q, probes) */
  int criterion_evals;          /* REML_criterion / REML_first_order evaluations */
  double seconds_seed;          /* wall time of the seed systems */
  double seconds_total;         /* wall time from resident bed to final estimates */
} slemm_reml_stats;

/* SLDF_REML (src/reml.cpp:549-690).  Host inputs: X (n x c, ldx) covariates, y (n).  Probes:
 * Rademacher from std::mt19937(rng_seed+54321) as src/reml.cpp:597-605.  Host outputs:
 * Py (n), snp_blup (m), blue (c), var_blue (c x c).  persnp/rvec given by set_model. */
int slemm_gpu_sldf_reml(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                        int rng_seed, int n_V, float tol_L, float s2max, float s2min, float tol_VC,
                        int precision, int verbose,
                        float* Py, float* snp_blup, float* blue, float* var_blue,
                        slemm_reml_stats* stats);

/* LMC_REML without --lmm (src/reml.cpp:730-882).  MC samples from std::mt19937(rng_seed+54321)
 * with std::normal_distribution<float> (src/reml.cpp:795-809). */
int slemm_gpu_lmc_reml(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                       int rng_seed, int n_MC, float tol_L, float s2max, float s2min,
                       int precision, int verbose,
                       float* Py, float* snp_blup, float* blue, slemm_reml_stats* stats);
/* The same with the secant's stop chosen by the caller (src/reml.cpp:830 stops when the update of
 * log((1-h2)/h2) is below 0.002 and :824 after 9 evaluations, which slemm_gpu_lmc_reml keeps):
 * secant_tol > 0, secant_max_iter >= 2 evaluations in all.  A tighter stop takes h2 to the root
 * of REML_first_order instead of the iterate the reference reports. */
int slemm_gpu_lmc_reml_ex(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                          int rng_seed, int n_MC, float tol_L, float s2max, float s2min,
                          float secant_tol, int secant_max_iter, int precision, int verbose,
                          float* Py, float* snp_blup, float* blue, slemm_reml_stats* stats);

/* The Rademacher probes of SLDF_REML (src/reml.cpp:597-605), host only: n_V unit-norm columns of
 * n entries -1/sqrt(n) or +1/sqrt(n) (B, host, n x n_V column-major) from
 * std::mt19937(rng_seed + 54321) through uniform_real_distribution<float>(0, 1) < 0.5, value for
 * value as the reference draws and normalises them.  slemm_gpu_sldf_reml draws the same probes. */
int slemm_gpu_rademacher_probes(int rng_seed, int64_t n, int n_V, float* B);

/* The Monte Carlo draws of LMC_REML (src/reml.cpp:788-801), host only: u (m x n_MC, column-major)
 * then e (n x n_MC) from one std::mt19937(rng_seed + 54321) stream through
 * std::normal_distribution<float>(0, 1), value for value as libstdc++ draws them (the polar
 * method's logarithms taken in parallel host threads).  u, e: host. */
int slemm_gpu_lmc_draws(int rng_seed, int64_t m, int64_t n, int n_MC, float* u, float* e);

/* REML score and average-information matrix at (vg, ve) for the model of SLDF_REML (no
 * reference counterpart): score = (dl/dvg, dl/dve), ai = (AI_gg, AI_ge, AI_ee) with
 *   dl/dvg = -1/2 [tr(PA) - y'PAPy], dl/dve = -1/2 [tr(P) - y'PPy], AI_kl = 1/2 y'P V_k P V_l P y
 * in the R^-1/2-whitened model (A = R^-1/2 Zc W Zc' R^-1/2, V_g = A, V_e = I).  Traces are
 * n/n_V sum_v v'Mv over the probe columns: `probes` (device n x n_V, ldp; E[v v'] = I/n) or, when
 * NULL, unit-norm Rademacher probes from mt19937(rng_seed+54321) as SLDF_REML draws them. */
int slemm_gpu_reml_score(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y, double vg,
                         double ve, const float* probes, int64_t ldp, int n_V, int rng_seed, float tol_L,
                         int precision, double* score, double* ai);

/* AI-REML: Newton steps theta += AI^-1 score on theta = (vg, ve) from (vg0, ve0) (both <= 0:
 * h2 = 1/2 of the projected phenotypic variance) until the largest relative change is below
 * tol_VC or max_iter steps; h2 is kept at or below s2max, which also fixes the seed shift
 * (1-s2max)/s2max.  Outputs as SLDF_REML; history (may be NULL): max_iter rows of
 * (vg, ve, score_g, score_e, AI_gg, AI_ge, AI_ee) at each iterate.  No reference counterpart
 * (the reference maximises the same likelihood by golden-section search, src/reml.cpp:493-547). */
int slemm_gpu_ai_reml(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y, int rng_seed,
                      int n_V, float tol_L, float s2max, double vg0, double ve0, double tol_VC,
                      int max_iter, int precision, int verbose, float* Py, float* snp_blup, float* blue,
                      float* var_blue, double* history, slemm_reml_stats* stats);

/* ---------------------------------------------------------------- batched traits ----------- */
/* Several traits on individual subsets of one handle, run as one batch (no reference
 * counterpart: the reference analyses one trait per run on its own individual set,
 * src/input.cpp:389-454).  masks: T x n host bytes (row t = trait t, 1 = the individual belongs
 * to the trait's subset).  Each trait gets gm and gw of src/reml.cpp:612-622 over its subset (a
 * SNP without variation in a subset gets weight 0 and is counted in `monomorphic`) and an
 * intercept as its only covariate; R = I. */
int slemm_gpu_traits_create(slemm_geno* h, int T, const uint8_t* masks, slemm_traits** out);
/* A traits handle uses its genotype handle in every call: destroy it before that handle. */
int slemm_gpu_traits_destroy(slemm_traits* tr);
/* Host outputs (T each, may be NULL): subset sizes and per-subset monomorphic SNP counts. */
int slemm_gpu_traits_info(const slemm_traits* tr, int64_t* n_t, int64_t* monomorphic);
/* Host copies (m each, may be NULL) of one trait's gm, gw and dosage sums over its subset. */
int slemm_gpu_traits_model(const slemm_traits* tr, int trait, float* gm, float* gw, int32_t* popcnt);
/* Batched lnr_op_select: column k (device V, n x t) is a vector of trait col_trait[k] (zero
 * outside its subset) and gets that trait's operator of kind col_kind[k] (NULL: all LDET), the
 * QUADFORM projection being off the trait's intercept.  All columns share one split-BF16 block
 * product per pass (SLEMM_PREC_BF16X2 only).  col_trait, col_kind: host, t entries. */
int slemm_gpu_traits_op(slemm_traits* tr, const int* col_trait, const int* col_kind, float tau0,
                        const float* V, int64_t ldv, int t, float* out, int64_t ldo, int precision,
                        void* stream);
/* L_Seed over the batched operator; arguments as slemm_gpu_lseed plus col_trait. */
int slemm_gpu_traits_lseed(slemm_traits* tr, const int* col_trait, const int* col_kind, const int* col_group,
                           float tau0, const float* B, int64_t ldb, int t, float tol, int maxit, int keep_cols,
                           int precision, int verbose, void* stream, slemm_lseed** out);

/* Plans of slemm_gpu_traits_sldf_reml (no reference counterpart):
 *  - BATCHED: one batched seed over all traits' phenotype, intercept and probe columns, then the
 *    golden-section search per trait on replays of it;
 *  - SEQUENTIAL: one trait at a time on a handle compacted to its individuals;
 *  - AUTO (the default): BATCHED while the batched block products' work (each 16-column tile over
 *    the rows up to the last member of its traits) is at most 1.25 times the one-at-a-time work
 *    (each trait's n_V + 1 columns over its individuals), else SEQUENTIAL.  Nested subsets
 *    (leading individuals) batch; random subsets of very different sizes go one at a time. */
#define SLEMM_TRAITS_AUTO 0
#define SLEMM_TRAITS_BATCHED 1
#define SLEMM_TRAITS_SEQUENTIAL 2
int slemm_gpu_traits_set_plan(slemm_traits* tr, int plan);
/* The plan the next slemm_gpu_traits_sldf_reml with n_V probes runs, and the work ratio
 * (batched over one at a time) the automatic choice reads; work_ratio may be NULL. */
int slemm_gpu_traits_plan(const slemm_traits* tr, int n_V, int* plan, double* work_ratio);

/* SLDF_REML (src/reml.cpp:549-690) for every trait of the batch, on the handle's plan (above).
 * y: T x n host (row t = trait t's phenotype; entries outside the subset are ignored).  Probes
 * of trait t are drawn from mt19937(rng_seed+54321) over its subset, as a single-trait run on
 * those individuals draws them.  Host outputs: vg, ve, h2, lrt (T each; lrt may be NULL),
 * steps (T x 3: Lanczos steps of y, intercept, probes; may be NULL), seconds[2] (seed, total;
 * may be NULL; one at a time, the seed time is the sum over traits). */
int slemm_gpu_traits_sldf_reml(slemm_traits* tr, const float* y, int rng_seed, int n_V, float tol_L,
                               float s2max, float s2min, float tol_VC, int precision, int verbose,
                               float* vg, float* ve, float* h2, float* lrt, int* steps, double* seconds);

/* --pred (src/pred.cpp:4-151): reads the estimate CSV (SNP id column 0, allele column 3, effect
 * column 9), bfile_prefix.fam/.bim/.bed, flips SNPs whose bim allele 1 differs from the
 * estimate's allele, computes GEBV = sum_j z_ij alpha_j in FP64 on the GPU and writes
 * "IID,GEBV" lines to output_file with the reference's stream formatting.  Host outputs (may be
 * NULL): number of individuals, number of SNPs used. */
int slemm_gpu_pred_plink(const char* estimate_file, const char* bfile_prefix, const char* output_file,
                         int64_t* n_out, int64_t* m_used);

/* --iter_weighting (src/option.cpp:417-528): equal-weight fit (SLDF when lrt and n < 5000,
 * else LMC with n_MC samples); per-SNP variances 2 maf (1-maf) b^2 normalised to mean 1;
 * clumping -- the largest remaining variance above sig_chisq becomes a fixed effect (centred
 * dosages) if more than indep_r2 of its sum of squares survives regression on those already
 * taken; weights persnp * 0.001/(0.001 + 0.999 exp(-mean_window/2)) over window_size+1 SNPs,
 * clumped SNPs 0, normalised; refit (SLDF when lrt, else LMC) with the extra covariates, whose
 * estimates replace those SNPs' BLUPs.  When m < window_size+1 the weighting is disabled and
 * one fit with persnp/mean(persnp) runs, as in the reference.  Host inputs: X (n x c), y,
 * persnp (m, NULL = ones), rvec (n, NULL = ones), maf (m).  Host outputs: Py (n), snp_blup (m),
 * blue (c: the original covariates), persnp_out (m, the final weights), sig_idx (m capacity) and
 * n_sig (the clumped SNPs in the order taken), stats of the final fit. */
int slemm_gpu_iter_weighting_reml(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                                  const float* persnp, const float* rvec, const float* maf, int window_size,
                                  float sig_chisq, float indep_r2, int lrt, int rng_seed, int n_V, int n_MC,
                                  float tol_L, float s2max, float s2min, float tol_VC, int precision, int verbose,
                                  float* Py, float* snp_blup, float* blue, float* persnp_out, int64_t* sig_idx,
                                  int* n_sig, slemm_reml_stats* stats);

/* ---------------------------------------------------------------- LMM quadratic forms ------ */

/* Quadratic forms of the --lmm association approximation (src/reml.cpp:884-982, real-genotype
 * branch): for each SNP snps[i] (host, k entries; NULL = the reference's own sample of k SNPs with
 * MAC > 20, see slemm_gpu_lmm_sample), after a fit with heritability
 * h2 (tau1 = (1-h2)/h2) whose Py (host, n) it returns: xPy = x'Py, xPx = v'H^-1 v (Lanczos seed
 * at tau1 with tolerance tol_L, then L_Solve), xSx = v'v with v = (I-QQ') R^-1/2 (x - gm), and
 * xBx = (v'v - z'(Zw'Zw + tau1 diag(1/gw))^-1 z)/tau1 over the window_size SNPs around the SNP
 * (src/reml.cpp:944-968; FP64 window systems through rocBLAS/rocSOLVER).  X (n x c, ldx)
 * covariates as given to the fit.  Host outputs xPy, xPx, xBx, xSx (k each); the gstat file's
 * ChiSq is xPy^2 / (vg xPx) (src/output.cpp:113). */
int slemm_gpu_lmm_gstat(slemm_geno* h, const float* X, int64_t ldx, int c, const float* Py, float h2,
                        const int64_t* snps, int k, int window_size, float tol_L, int precision, int verbose,
                        float* xPy, float* xPx, float* xBx, float* xSx);

/* The --lmm SNP sample of src/reml.cpp:920-925, host only: the SNPs i with gm_i*n > 20 and
 * (2-gm_i)*n > 20 (float arithmetic) in the order std::random_shuffle leaves them when rand() is
 * in its initial state (the reference never seeds it), the first min(num_qf_markers, N) of them.
 * gm: host, m floats (the genotype means of slemm_gpu_geno_counts).  snps (host, capacity
 * num_qf_markers; NULL to ask only for the count) receives the indices, *k their number.
 * slemm_gpu_lmm_gstat with snps == NULL uses this sample of the handle's own gm. */
int slemm_gpu_lmm_sample(const float* gm, int64_t m, int64_t n, int num_qf_markers, int64_t* snps, int* k);

/* --fake_geno branch of the --lmm quadratic forms (src/reml.cpp:886-918), after an LMC_REML fit
 * with seed rng_seed and n_MC samples: k simulated markers continue the fit's random stream
 * (mt19937(rng_seed+54321) after its n_MC*m + n_MC*n normal draws), each with allele frequency
 * uniform_real_distribution<float>(0.01, 0.99) and n binomial(2, p) dosages, centred by their
 * float mean; xPy, xPx, xSx (host, k each) as slemm_gpu_lmm_gstat computes them (no window
 * term: the reference reports xBx as NA and names the markers fake-1, fake-2, ...). */
int slemm_gpu_lmm_gstat_fake(slemm_geno* h, const float* X, int64_t ldx, int c, const float* Py, float h2, int k,
                             int rng_seed, int n_MC, float tol_L, int precision, int verbose, float* xPy,
                             float* xPx, float* xSx);

/* Windowed score tests of slemm_gwa step 2 (bin/slemm_gwa.py:322-431): model SNPs model_snps
 * (host, K, increasing) form blocks of block_size (reduced to an even number <= K) overlapping by
 * half; block b tests the SNPs from the previous block's bound to its own (model SNP
 * block_size - block_size/4 of the block), the first from test_begin and the last to test_end
 * (inclusive, handle SNP indices).  Per block H = Zb'Zb + model_snp_ct tau diag((2-p)p/2) in
 * FP64 (batched rocBLAS DGEMMs, rocSOLVER Cholesky); per tested SNP x (centred by its mean dosage)
 * xy = x'y and, with x scaled by R^-1/2 and projected off Q(X), xx = (x'x - z'H^-1 z)/tau,
 * z = Zb'x.  SNPs whose allele count is <= mac either way are NaN.  Host inputs X (n x c), y (n:
 * the phenotype column of the fit's output); host outputs xy, xx (test_end - test_begin + 1). */
int slemm_gpu_gwa_window(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                         const int64_t* model_snps, int K, int block_size, int64_t test_begin, int64_t test_end,
                         float tau, int64_t model_snp_ct, float mac, float* xy, float* xx);

/* ---------------------------------------------------------------- tuning ------------------- */

/* Kernel family that a t-vector operator block at this precision runs on (no reference
 * counterpart), as selected from t and the SLEMM_I8_NARROW / SLEMM_I8_WIDE / SLEMM_HL_COLUMNS
 * switches at the time of the call. */
#define SLEMM_KERNEL_FP32 0      /* FP32 operand, FP32 kernels (precision SLEMM_PREC_FP32) */
#define SLEMM_KERNEL_BF16X2 1    /* split-BF16 operand on BF16 MFMAs, FP32 accumulation */
#define SLEMM_KERNEL_I8_NARROW 2 /* int8 digits, narrow kernel (1-5 vectors) */
#define SLEMM_KERNEL_I8_TILE 3   /* int8 digits, wide kernel, one 16-vector tile without a tail */
#define SLEMM_KERNEL_I8_WIDE 4   /* int8 digits, general wide kernel */
int slemm_gpu_block_kernel(const slemm_geno* h, int t, int precision, int* kernel);

/* Autotuner of the block products (no reference counterpart; the reference's only blocking knob
 * is --subset_size, src/option.cpp:75,297), for a t-column operator on this handle's shape; the
 * fastest setting per pass is kept on the handle for later calls.
 *  - Blocks that run on the int8 kernels (1-5 and 9 or more vectors unless switched off): the
 *    workgroup target of the reduction split (1024, 2048, 4096, 8192, 16384) and, for narrow
 *    blocks (1-5 vectors), the row tiles per wave (4 or 8), each timed over full passes (digit
 *    preparation + block product) on random FP32 operands; kept per pass and kernel class
 *    (narrow, one vector tile, general wide).  Outputs: variant[2] = the narrow row tiles (0 for
 *    wide blocks), target[2], ms[0..1] = best per pass, ms[2 + pass*30 + rt_index*5 + target_index]
 *    the timings (rt_index 0: 4 row tiles, 1: 8).
 *  - Split-BF16 blocks: every kernel variant (B operand from global memory with 4 or 8 row tiles
 *    per wave, with table-driven code unpacking, or software-pipelined as an ordinary or a
 *    persistent grid or with a byte-table unpack from one shared or a 16- or 32-fold replicated
 *    table; or LDS-staged in 4- or 2-step groups) at three grid sizes (512, 1024, 2048) for both
 *    passes.  Blocks of t <= 8 vectors (hi and lo halves in one 16-column tile) have only the
 *    pipelined variants.  Outputs: variant[2], target[2], ms[0..1] = best per pass, then
 *    ms[2 + pass*30 + variant*3 + grid-size index].
 * Entries not timed read -1.  Host outputs, any may be NULL. */
#define SLEMM_AUTOTUNE_MS 62
int slemm_gpu_autotune(slemm_geno* h, int t, int reps, int* variant, int* target, double* ms);

/* ---------------------------------------------------------------- instrumentation ---------- */

/* Kernel timing of the genotype block products (no reference counterpart; the reference only
 * prints wall-clock stamps, src/option.cpp:21-34).  When enabled, every block-product launch
 * is bracketed by HIP events on its own stream.  pass: 0 = Z'V pass, 1 = Z U pass. */
int slemm_gpu_timing(int enable);
/* Matrix-core issue ceiling (no reference counterpart): a register-only loop of `chains`
 * independent accumulators per wave (1, 2, 4 or 8), `waves_per_simd` waves on every SIMD, `iters`
 * MFMAs per chain, of instruction instr (0 v_mfma_i32_16x16x64_i8, 1 v_mfma_i32_32x32x32_i8,
 * 2 v_mfma_f32_16x16x32_bf16, 3 v_mfma_f32_32x32x16_bf16; instr + 4: the same instruction with
 * per-chain random operands -- int8 dosages 0/1/2 against full-range digits, BF16 values with
 * random mantissas -- whose bit switching draws the power of real data) on the current device.
 * Host outputs:
 * tops (tera-operations per second, 2 per multiply-add), clock_ratio (shader-clock ticks per
 * 100 MHz real-time tick over the loop of one wave: the clock in units of 100 MHz; may be NULL), ms
 * (kernel time; may be NULL). */
int slemm_gpu_mfma_ceiling(int instr, int chains, int waves_per_simd, int iters, double* tops,
                           double* clock_ratio, double* ms);
int slemm_gpu_timing_read(int pass, int precision, double* total_ms, int* launches);
int slemm_gpu_timing_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* SLEMM_GPU_H_ */
