// Host-side pieces shared by the REML drivers (reml.cpp, aireml.hip): device scratch vectors,
// the covariate basis, projections and the Py / SNP BLUP / BLUE finish of src/reml.cpp:655-677.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <vector>

#include "common.hpp"
#include "hostla.hpp"

namespace slemm {

inline void abi(int st) {
  if (st != SLEMM_OK) fail(st, slemm_gpu_last_error());
}

struct DevF {
  float* p = nullptr;
  explicit DevF(size_t count) { HIPCHECK(hipMalloc(&p, sizeof(float) * std::max<size_t>(count, 1))); }
  ~DevF() {
    if (p) (void)hipFree(p);
  }
  DevF(const DevF&) = delete;
  DevF& operator=(const DevF&) = delete;
};

inline void upload(float* d, const std::vector<float>& h) {
  HIPCHECK(hipMemcpy(d, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
}
inline void download(std::vector<float>& h, const float* d, size_t count) {
  h.resize(count);
  HIPCHECK(hipMemcpy(h.data(), d, sizeof(float) * count, hipMemcpyDeviceToHost));
}

struct SeedGuard {
  slemm_lseed* s = nullptr;
  ~SeedGuard() {
    if (s) slemm_gpu_lseed_destroy(s);
  }
};

// Q R = diag(rr) X (n x c), Q stored in the handle for the operators.  Returns R.
inline MatD covariate_basis(slemm_geno* h, const float* X, int64_t ldx, int c, MatD& Q) {
  const int64_t n = h->n;
  MatD A(n, c), R;
  for (int j = 0; j < c; ++j)
    for (int64_t i = 0; i < n; ++i) A(i, j) = (double)h->rr_host[i] * (double)X[i + (size_t)j * ldx];
  if (!thin_qr(A, Q, R, 1e-6)) fail(SLEMM_ERR_RANK, "Fixed-effects design matrix is not of full rank.");
  std::vector<float> qf((size_t)n * c);
  for (size_t e = 0; e < qf.size(); ++e) qf[e] = (float)Q.a[e];
  set_Q(h, qf.data(), c);
  return R;
}

// X - Q (Q' X) in place on an n x t float block (FP64 accumulation)
inline void project_out(const MatD& Q, std::vector<float>& V, int64_t n, int t) {
  for (int k = 0; k < t; ++k) {
    float* v = &V[(size_t)k * n];
    for (int q = 0; q < Q.c; ++q) {
      double s = 0.0;
      for (int64_t i = 0; i < n; ++i) s += Q(i, q) * v[i];
      for (int64_t i = 0; i < n; ++i) v[i] = (float)((double)v[i] - Q(i, q) * s);
    }
  }
}

inline double sq_norm(const float* v, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += (double)v[i] * v[i];
  return s;
}

inline double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Dosages (0, 1, 2 as float) of the SNPs snps, n x k column-major on the host (gstat.hip).
void extract_dosage_cols(slemm_geno* h, const std::vector<int64_t>& snps, std::vector<float>& out);

// Rademacher probes of unit norm from mt19937(seed+54321), written to columns of B (ld n),
// draw order and normalisation as src/reml.cpp:597-605 (probes.hip).
void rademacher_probes(int rng_seed, int64_t n, int n_V, float* B);
// The same probes written to the device block dB (columns of length n, leading dimension ldb).
void rademacher_probes_device(int rng_seed, int64_t n, int n_V, float* dB, int64_t ldb);
// The signs of the first `total` Rademacher draws of that stream, one bit per draw in draw order
// (bit e % 8 of byte e / 8 set where draw e is -1; ceil(total / 8) bytes): probe column j of an
// n-row block holds draws j n .. j n + n - 1, so blocks of different n share the prefix of one
// stream; every entry of such a column is -rademacher_scale(n) or +rademacher_scale(n) (signs.cpp).
std::vector<uint8_t> rademacher_sign_bits(int rng_seed, int64_t total);
// The same bits into bits (ceil(total / 8) bytes).
void rademacher_sign_bits_into(int rng_seed, int64_t total, uint8_t* bits);

// Page-locked host memory of at least `bytes`, per host thread and slot (0 Lanczos scalars, 1 LMC
// draws, 2 probe sign bits), kept for the life of the process (pinned.cpp).
void* pinned_scratch(int slot, size_t bytes);
float rademacher_scale(int64_t n);

// REML_criterion at h2 (src/reml.cpp:457-489): returns the 2 logLL value the golden-section
// search compares (see criterion_value), sets vg, ve and the LRT.
using Criterion = std::function<double(float h2, float& vg, float& ve, float& lrt)>;
// The criterion as the search compares it: in double by default, which lets the search resolve
// the flat top of the likelihood below float's resolution (a criterion of magnitude 1e3..1e5
// rounded to float moves the optimum by ~1e-4 relative); rounded to float as src/reml.cpp:486
// returns it with SLEMM_CRITERION_FP32=1 (read per call), to replay the reference's search path.
double criterion_value(double c);
// True when SLEMM_CRITERION_FP32=1: the criterion replays the reference's arithmetic (y'Py from
// the FP32 L_Solve solution, criterion rounded to float).
bool criterion_replay();

// y'(A + sigma I)^-1 y from the seed column of y itself: Gauss quadrature on its FP64 Lanczos
// tridiagonal T_J, ||y||^2 e1'(T_J + sigma I)^-1 e1 = ||y||^2 sum_q W_q^2 / (D_q + sigma).  In exact
// arithmetic this is y' x_J for the CG iterate x_J that L_Solve replays (x_J = ||y|| U_J
// (T_J + sigma)^-1 e1 and U_J' y = ||y|| e1); unlike y' x_J formed from the FP32-rounded
// solution it is a smooth function of sigma, so the golden-section search does not see rounding
// jitter of ~(n - c) 1e-9 in the criterion.
struct SeedQuadrature {
  std::vector<double> D, W;
  double norm2 = 0.0;
  SeedQuadrature(const slemm_lseed* s, int col, double y_norm2);
  double inverse(double sigma) const;
};

// REML_GSS (src/reml.cpp:493-547): golden-section search on (s2min, s2max) to tol_VC; vg, ve
// and lrt are those of the last evaluated point, as in the reference.
void golden_section(float s2min, float s2max, float tol_VC, int verbose, const Criterion& criterion, float& vg,
                    float& ve, float& lrt);

// SLDF_REML (src/reml.cpp:549-690) as slemm_gpu_sldf_reml runs it; with estimates = false the fit
// stops at the variance components (no Py, SNP BLUP or BLUE; those outputs may be null).
void sldf_reml_run(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y, int rng_seed, int n_V,
                   float tol_L, float s2max, float s2min, float tol_VC, int precision, int verbose, float* Py,
                   float* snp_blup, float* blue, float* var_blue, slemm_reml_stats* stats, bool estimates);

// Final estimates at tau1 = ve/vg from a seed whose column 0 is the projected phenotype and
// columns 1..c the covariate basis (src/reml.cpp:655-677): Py = R^-1/2 (I-QQ') H^-1 y~,
// snp_blup = W Zc' Py, BLUE and (optionally) its covariance scaled by vg.
void finish_estimates(slemm_geno* h, slemm_lseed* seed, const float* dB, float tau0, float tau1, const MatD& Q,
                      const MatD& qrR, const float* X, int64_t ldx, int c, const float* y, float vg,
                      int precision, float* Py, float* snp_blup, float* blue, float* var_blue);

}  // namespace slemm
