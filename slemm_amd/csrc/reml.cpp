// REML driver over the device engine: SLDF_REML (src/reml.cpp:457-690).  LMC_REML
// (src/reml.cpp:692-882) is in lmc.hip.
//
// The control flow -- covariate QR, probe draws from the same libstdc++ generators, the three
// seed systems, golden-section search and the final Py, SNP BLUP and BLUE -- follows the
// reference step by step.  Differences, all deliberate:
//  * the seed systems run as one batched device Lanczos (y, Q and probes share every pass over
//    the genotypes; each group still stops on its own tolerance);
//  * Lanczos/CG scalars, dot products, the tridiagonal eigensolver and the SLQ sum are FP64;
//  * the covariate QR is an unpivoted Householder QR (the reference's column pivoting only
//    permutes R, and Q R = R^-1/2 X either way).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "blockprod.hpp"
#include "driver.hpp"

using namespace slemm;

namespace slemm {

bool criterion_replay() {
  const char* e = std::getenv("SLEMM_CRITERION_FP32");
  return e && e[0] == '1';
}

double criterion_value(double c) { return criterion_replay() ? (double)(float)c : c; }

SeedQuadrature::SeedQuadrature(const slemm_lseed* s, int col, double y_norm2) : norm2(y_norm2) {
  std::vector<int> steps;
  int t = 0;
  abi(slemm_gpu_lseed_info(s, &t, nullptr, nullptr));
  steps.resize(t);
  abi(slemm_gpu_lseed_col_steps(s, steps.data()));
  D.resize(steps[col]);
  W.resize(steps[col]);
  abi(slemm_gpu_lseed_spectrum(s, col, D.data(), W.data()));
}

double SeedQuadrature::inverse(double sigma) const {
  double q = 0.0;
  for (size_t i = 0; i < D.size(); ++i) q += W[i] * W[i] / (D[i] + sigma);
  return norm2 * q;
}

void golden_section(float s2min, float s2max, float tol_VC, int verbose, const Criterion& criterion, float& vg,
                    float& ve, float& lrt) {
  const double invphi = 0.618033988749894848204586834365638117720309179805762862135;
  const double invphi2 = 0.381966011250105151795413165634361882279690820194237137864;
  float a = s2min, b = s2max;
  double fm = 0;
  if (verbose) printf("\nDoing golden section search in (%g, %g)\n", a, b);
  float hh = b - a;
  const int nitr = hh < tol_VC ? 1 : (int)std::ceil(std::log(tol_VC / hh) / std::log(invphi));
  float cc = (float)(a + invphi2 * hh), dd = (float)(a + invphi * hh);
  double fc = criterion(cc, vg, ve, lrt);
  double fd = criterion(dd, vg, ve, lrt);
  for (int i = 0; i < nitr; ++i) {
    if (fc > fd) {
      b = dd;
      dd = cc;
      fd = fc;
      hh = (float)(hh * invphi);
      cc = (float)(a + invphi2 * hh);
      fc = criterion(cc, vg, ve, lrt);
      fm = fc;
    } else {
      a = cc;
      cc = dd;
      fc = fd;
      hh = (float)(hh * invphi);
      dd = (float)(a + invphi * hh);
      fd = criterion(dd, vg, ve, lrt);
      fm = fd;
    }
    if (verbose) printf("Itr %d: h2 in (%g, %g) and 2logLL = %g\n", i + 1, a, b, fm);
  }
}

void finish_estimates(slemm_geno* h, slemm_lseed* seed, const float* dB, float tau0, float tau1, const MatD& Q,
                      const MatD& qrR, const float* X, int64_t ldx, int c, const float* y, float vg,
                      int precision, float* Py, float* snp_blup, float* blue, float* var_blue) {
  const int64_t n = h->n, m = h->m;
  DevF dX((size_t)n * std::max(c, 1)), dv(n);
  // Py, SNP BLUP (:658-662)
  abi(slemm_gpu_lsolve(seed, 0, 1, dB, n, tau1 - tau0, 1e-6f, dX.p, n, nullptr));
  std::vector<float> py;
  download(py, dX.p, n);
  project_out(Q, py, n, 1);
  for (int64_t i = 0; i < n; ++i) py[i] *= h->rr_host[i];
  std::copy(py.begin(), py.end(), Py);
  upload(dv.p, py);
  // BLUE of the fixed effects (:674-677): the solves first, then the SNP BLUP pass is queued and
  // runs while the host forms X'H^-1X
  std::vector<float> hx;
  abi(slemm_gpu_lsolve(seed, 1, c, dB + n, n, tau1 - tau0, 1e-6f, dX.p, n, nullptr));
  download(hx, dX.p, (size_t)n * c);
  DevF dbl(m);
  abi(slemm_gpu_zt(h, dv.p, n, 1, dbl.p, m, 1, 0, precision, nullptr));
  MatD HinvX(n, c);
  for (int j = 0; j < c; ++j)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int q = 0; q < c; ++q) s += (double)hx[i + (size_t)q * n] * qrR(q, j);
      HinvX(i, j) = h->rr_host[i] * s;
    }
  std::vector<double> XtHX((size_t)c * c), rhs(c);
  for (int a2 = 0; a2 < c; ++a2) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += HinvX(i, a2) * y[i];
    rhs[a2] = s;
    for (int b2 = 0; b2 < c; ++b2) {
      double t = 0.0;
      for (int64_t i = 0; i < n; ++i) t += (double)X[i + (size_t)a2 * ldx] * HinvX(i, b2);
      XtHX[a2 + (size_t)b2 * c] = t;
    }
  }
  std::vector<double> bl = rhs, inv;
  if (!chol_solve(XtHX, c, bl)) fail(SLEMM_ERR_RANK, "X' H^-1 X is not positive definite");
  if (!invert(XtHX, c, inv)) fail(SLEMM_ERR_RANK, "X' H^-1 X is singular");
  for (int q = 0; q < c; ++q) blue[q] = (float)bl[q];
  if (var_blue)
    for (int e = 0; e < c * c; ++e) var_blue[e] = (float)(inv[e] * vg);
  HIPCHECK(hipDeviceSynchronize());
  HIPCHECK(hipMemcpy(snp_blup, dbl.p, sizeof(float) * m, hipMemcpyDeviceToHost));
}

}  // namespace slemm

void slemm::sldf_reml_run(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y, int rng_seed, int n_V,
                          float tol_L, float s2max, float s2min, float tol_VC, int precision, int verbose, float* Py,
                          float* snp_blup, float* blue, float* var_blue, slemm_reml_stats* stats, bool estimates) {
  {
    REQUIRE(h && X && y, "null argument");
    REQUIRE(!estimates || (Py && snp_blup && blue), "null argument");
    REQUIRE(c >= 1 && ldx >= h->n, "covariates: need c >= 1 and ldx >= n");
    REQUIRE(n_V >= 1, "number of probes must be positive");
    REQUIRE(s2min > 0 && s2max < 1 && s2min < s2max, "heritability range must satisfy 0 < s2min < s2max < 1");
    HIPCHECK(hipSetDevice(h->device));
    require_model(h);
    const auto t_start = std::chrono::steady_clock::now();
    const int64_t n = h->n;
    if (verbose) printf("\nStarted stochastic Lanczos DF-REML.\n");

    MatD Q;
    const MatD qrR = covariate_basis(h, X, ldx, c, Q);  // src/reml.cpp:583-588

    // projected phenotype (:591-592)
    std::vector<float> yp(n);
    for (int64_t i = 0; i < n; ++i) yp[i] = h->rr_host[i] * y[i];
    project_out(Q, yp, n, 1);

    // [y~, Q, Rademacher probes from mt19937(seed+54321) (:597-605)], the probes written on the
    // device from their signs
    const int T = 1 + c + n_V;
    std::vector<float> B((size_t)n * (1 + c));
    std::copy(yp.begin(), yp.end(), B.begin());
    for (int q = 0; q < c; ++q)
      for (int64_t i = 0; i < n; ++i) B[(size_t)(1 + q) * n + i] = (float)Q(i, q);
    DevF dB((size_t)n * T), dX((size_t)n * std::max(c, 1));
    upload(dB.p, B);
    rademacher_probes_device(rng_seed, n, n_V, dB.p + (size_t)(1 + c) * n, n);
    if (verbose) printf("Generated %d Rademacher probing vectors.\n", n_V);

    // seed systems y (QUADFORM), Q (LDET), probes (LDET) in one batch (:625-631)
    const float tau0 = (1 - s2max) / s2max;
    std::vector<int> kinds(T, SLEMM_OP_LDET), groups(T, 2);
    kinds[0] = SLEMM_OP_QUADFORM;
    groups[0] = 0;
    for (int q = 0; q < c; ++q) groups[1 + q] = 1;
    SeedGuard sg;
    abi(slemm_gpu_lseed(h, SLEMM_OP_LDET, kinds.data(), groups.data(), tau0, dB.p, n, T, tol_L, 600, 1 + c,
                        precision, verbose, nullptr, &sg.s));
    std::vector<int> csteps(T);
    abi(slemm_gpu_lseed_col_steps(sg.s, csteps.data()));
    const double t_seed = seconds_since(t_start);

    // tridiagonal spectra of the probe systems (:637-646), FP64
    std::vector<std::vector<double>> Dk(n_V), Wk(n_V);
    for (int j = 0; j < n_V; ++j) {
      const int J = csteps[1 + c + j];
      Dk[j].resize(J);
      Wk[j].resize(J);
      abi(slemm_gpu_lseed_spectrum(sg.s, 1 + c + j, Dk[j].data(), Wk[j].data()));
    }
    const double ySy = sq_norm(yp.data(), n);
    int evals = 0;
    std::vector<float> hx;
    const bool replay = criterion_replay();
    const SeedQuadrature quad_y(sg.s, 0, ySy);

    DevBuf dqtx;
    dqtx.reserve(sizeof(double) * c * c);
    std::vector<double> qtx_t((size_t)c * c);
    // One covariate: Q'H^-1Q from the quadrature of Q's own seed column, as y'Py (smooth in sigma,
    // no device work per evaluation); several: the cross terms need the solutions.  The
    // reference's arithmetic (replay) forms both from L_Solve solutions.
    std::unique_ptr<SeedQuadrature> quad_q;
    if (c == 1 && !replay) quad_q = std::make_unique<SeedQuadrature>(sg.s, 1, sq_norm(B.data() + n, n));

    // REML_criterion (:457-489)
    auto criterion = [&](float ss, float& vg, float& ve, float& lrt) -> double {
      ++evals;
      const float tau = (1 - ss) / ss;
      const float sigma = tau - tau0;
      double yPy = 0.0;
      if (replay) {  // y'x from the L_Solve solution, as :477
        abi(slemm_gpu_lsolve(sg.s, 0, 1, dB.p, n, sigma, 1e-6f, dX.p, n, nullptr));
        download(hx, dX.p, n);
        for (int64_t i = 0; i < n; ++i) yPy += (double)yp[i] * hx[i];
      } else {
        yPy = quad_y.inverse(sigma);
      }
      vg = (float)(yPy / (double)(n - c));
      ve = vg * tau;
      double acc = 0.0;
      for (int j = 0; j < n_V; ++j)
        for (size_t q = 0; q < Dk[j].size(); ++q) acc += Wk[j][q] * Wk[j][q] * std::log(Dk[j][q] + sigma);
      const double ldetH = acc * (double)n / (double)n_V;
      lrt = (float)((double)n * std::log(ySy / yPy) - ldetH);
      std::vector<double> QtX((size_t)c * c);
      if (quad_q) {
        QtX[0] = quad_q->inverse(sigma);
      } else {
        abi(slemm_gpu_lsolve(sg.s, 1, c, dB.p + n, n, sigma, 1e-6f, dX.p, n, nullptr));
        // Q' H^-1 Q from the solution in FP64 on the device (the basis as the handle holds it),
        // c x c doubles back instead of the n x c solution
        tn_product(h->Q, h->n_pad, c, dX.p, n, c, n, dqtx.as<double>(), c, h->redpart, nullptr);
        HIPCHECK(hipMemcpy(qtx_t.data(), dqtx.p, sizeof(double) * c * c, hipMemcpyDeviceToHost));
        for (int a = 0; a < c; ++a)
          for (int b = 0; b < c; ++b) QtX[a + (size_t)b * c] = qtx_t[(size_t)a * c + b];  // Q_a' x_b
      }
      return criterion_value((double)(n - c) * std::log(ySy / yPy) - log_abs_det(QtX, c) - ldetH);
    };

    // REML_GSS (:493-547): golden-section search, returns the last evaluated point
    float vg = 0, ve = 0, lrt = 0;
    golden_section(s2min, s2max, tol_VC, verbose, criterion, vg, ve, lrt);

    // Py, SNP BLUP, BLUE (:655-677)
    const float tau1 = ve / vg;
    if (estimates)
      finish_estimates(h, sg.s, dB.p, tau0, tau1, Q, qrR, X, ldx, c, y, vg, precision, Py, snp_blup, blue,
                       var_blue);
    if (verbose) printf("\nVg,%g\nVe,%g\nh2,%g\nLRT,%g\n", vg, ve, vg / (vg + ve), lrt);
    if (verbose) printf("\nCompleted stochastic Lanczos DF-REML.\n");
    if (s2max - vg / (vg + ve) < tol_VC && verbose)
      printf("\nWarning: the h2 estimate is virtually the same as the max you specified.\n"
             "Rerun with a higher --max_heritability.\n");
    if (stats) {
      stats->vg = vg;
      stats->ve = ve;
      stats->h2 = vg / (vg + ve);
      stats->lrt = lrt;
      stats->lanczos_steps[0] = csteps[0];
      stats->lanczos_steps[1] = csteps[1];
      stats->lanczos_steps[2] = csteps[1 + c];
      stats->criterion_evals = evals;
      stats->seconds_seed = t_seed;
      stats->seconds_total = seconds_since(t_start);
    }
  }
}

extern "C" int slemm_gpu_sldf_reml(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                                   int rng_seed, int n_V, float tol_L, float s2max, float s2min, float tol_VC,
                                   int precision, int verbose, float* Py, float* snp_blup, float* blue,
                                   float* var_blue, slemm_reml_stats* stats) {
  return guard([&] {
    sldf_reml_run(h, X, ldx, c, y, rng_seed, n_V, tol_L, s2max, s2min, tol_VC, precision, verbose, Py, snp_blup,
                  blue, var_blue, stats, true);
  });
}

