// REML drivers over the device engine: SLDF_REML (src/reml.cpp:457-690) and LMC_REML
// (src/reml.cpp:692-882, without the --lmm association part).
//
// The control flow -- covariate QR, probe/MC draws from the same libstdc++ generators, the
// three seed systems, golden-section search / secant iterations and the final Py, SNP BLUP and
// BLUE -- follows the reference step by step.  Differences, all deliberate:
//  * the seed systems run as one batched device Lanczos (y, Q and probes share every pass over
//    the genotypes; each group still stops on its own tolerance);
//  * Lanczos/CG scalars, dot products, the tridiagonal eigensolver and the SLQ sum are FP64;
//  * the covariate QR is an unpivoted Householder QR (the reference's column pivoting only
//    permutes R, and Q R = R^-1/2 X either way).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "driver.hpp"

using namespace slemm;

namespace slemm {

void rademacher_probes(int rng_seed, int64_t n, int n_V, float* B) {
  std::mt19937 rng(rng_seed + 54321);
  std::uniform_real_distribution<float> runif(0.0, 1.0);
  for (int j = 0; j < n_V; ++j) {
    float* v = B + (size_t)j * n;
    for (int64_t i = 0; i < n; ++i) v[i] = runif(rng) < 0.5 ? -1.0f : 1.0f;
  }
  for (int j = 0; j < n_V; ++j) {
    float* v = B + (size_t)j * n;
    float ss = 0.f;
    for (int64_t i = 0; i < n; ++i) ss += v[i] * v[i];
    const float nr = std::sqrt(ss);
    for (int64_t i = 0; i < n; ++i) v[i] /= nr;
  }
}

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
  {
    DevF dbl(m);
    abi(slemm_gpu_zt(h, dv.p, n, 1, dbl.p, m, 1, 0, precision, nullptr));
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(snp_blup, dbl.p, sizeof(float) * m, hipMemcpyDeviceToHost));
  }
  // BLUE of the fixed effects (:674-677)
  std::vector<float> hx;
  abi(slemm_gpu_lsolve(seed, 1, c, dB + n, n, tau1 - tau0, 1e-6f, dX.p, n, nullptr));
  download(hx, dX.p, (size_t)n * c);
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
}

}  // namespace slemm

extern "C" int slemm_gpu_sldf_reml(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                                   int rng_seed, int n_V, float tol_L, float s2max, float s2min, float tol_VC,
                                   int precision, int verbose, float* Py, float* snp_blup, float* blue,
                                   float* var_blue, slemm_reml_stats* stats) {
  return guard([&] {
    REQUIRE(h && X && y && Py && snp_blup && blue, "null argument");
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

    // Rademacher probes from mt19937(seed+54321) (:597-605)
    const int T = 1 + c + n_V;
    std::vector<float> B((size_t)n * T);
    std::copy(yp.begin(), yp.end(), B.begin());
    for (int q = 0; q < c; ++q)
      for (int64_t i = 0; i < n; ++i) B[(size_t)(1 + q) * n + i] = (float)Q(i, q);
    rademacher_probes(rng_seed, n, n_V, &B[(size_t)(1 + c) * n]);
    if (verbose) printf("Generated %d Rademacher probing vectors.\n", n_V);
    DevF dB((size_t)n * T), dX((size_t)n * std::max(c, 1));
    upload(dB.p, B);

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
      abi(slemm_gpu_lsolve(sg.s, 1, c, dB.p + n, n, sigma, 1e-6f, dX.p, n, nullptr));
      download(hx, dX.p, (size_t)n * c);
      std::vector<double> QtX((size_t)c * c, 0.0);
      for (int a = 0; a < c; ++a)
        for (int b = 0; b < c; ++b) {
          double s = 0.0;
          for (int64_t i = 0; i < n; ++i) s += Q(i, a) * hx[i + (size_t)b * n];
          QtX[a + (size_t)b * c] = s;
        }
      return criterion_value((double)(n - c) * std::log(ySy / yPy) - log_abs_det(QtX, c) - ldetH);
    };

    // REML_GSS (:493-547): golden-section search, returns the last evaluated point
    float vg = 0, ve = 0, lrt = 0;
    golden_section(s2min, s2max, tol_VC, verbose, criterion, vg, ve, lrt);

    // Py, SNP BLUP, BLUE (:655-677)
    const float tau1 = ve / vg;
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
  });
}

extern "C" int slemm_gpu_lmc_reml_ex(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                                     int rng_seed, int n_MC, float tol_L, float s2max, float s2min, float secant_tol,
                                     int secant_max_iter, int precision, int verbose, float* Py, float* snp_blup,
                                     float* blue, slemm_reml_stats* stats) {
  return guard([&] {
    REQUIRE(h && X && y && Py && snp_blup && blue, "null argument");
    REQUIRE(c >= 1 && ldx >= h->n, "covariates: need c >= 1 and ldx >= n");
    REQUIRE(n_MC >= 1, "number of MC samples must be positive");
    REQUIRE(s2max > 0 && s2max < 1, "max heritability must be in (0, 1)");
    REQUIRE(secant_tol > 0 && secant_max_iter >= 2, "secant: need tol > 0 and at least 2 evaluations");
    (void)s2min;
    HIPCHECK(hipSetDevice(h->device));
    require_model(h);
    const auto t_start = std::chrono::steady_clock::now();
    const int64_t n = h->n, m = h->m;
    if (verbose) printf("\nStarted Lanczos MC-REML.\n");
    MatD Q;
    covariate_basis(h, X, ldx, c, Q);  // src/reml.cpp:765-769
    std::vector<float> gw(m);
    HIPCHECK(hipMemcpy(gw.data(), h->gw, sizeof(float) * m, hipMemcpyDeviceToHost));

    // V = [y, Zc u_MC, e_MC] (:788-813)
    const int T = 1 + 2 * n_MC;
    std::vector<float> V((size_t)n * T);
    std::copy(y, y + n, V.begin());
    std::mt19937 rng(rng_seed + 54321);
    std::normal_distribution<float> rnorm(0.0, 1.0);
    {
      std::vector<float> u((size_t)m * n_MC);
      for (int j = 0; j < n_MC; ++j)
        for (int64_t i = 0; i < m; ++i) u[(size_t)j * m + i] = rnorm(rng);
      for (int j = 0; j < n_MC; ++j)
        for (int64_t i = 0; i < m; ++i) u[(size_t)j * m + i] *= std::sqrt(gw[i]);
      DevF du(u.size()), dzu((size_t)n * n_MC);
      upload(du.p, u);
      abi(slemm_gpu_z(h, du.p, m, n_MC, dzu.p, n, precision, nullptr));
      HIPCHECK(hipDeviceSynchronize());
      HIPCHECK(hipMemcpy(&V[(size_t)n], dzu.p, sizeof(float) * n * n_MC, hipMemcpyDeviceToHost));
    }
    for (int j = 1 + n_MC; j < T; ++j)
      for (int64_t i = 0; i < n; ++i) V[(size_t)j * n + i] = rnorm(rng);
    for (int j = 0; j < 1 + n_MC; ++j)
      for (int64_t i = 0; i < n; ++i) V[(size_t)j * n + i] *= h->rr_host[i];
    project_out(Q, V, n, T);

    // one seed system at 10x the tolerance, as BOLT-LMM relaxes CG (:817-819)
    const float tau0 = (1 - s2max) / s2max;
    DevF dV((size_t)n * T), dPV((size_t)n * T), dU((size_t)m * (1 + n_MC));
    upload(dV.p, V);
    SeedGuard sg;
    abi(slemm_gpu_lseed(h, SLEMM_OP_QUADFORM, nullptr, nullptr, tau0, dV.p, n, T, 10 * tol_L, 600, T, precision,
                        verbose, nullptr, &sg.s));
    int steps = 0;
    abi(slemm_gpu_lseed_info(sg.s, nullptr, &steps, nullptr));
    const double t_seed = seconds_since(t_start);

    int evals = 0;
    std::vector<float> PyMC((size_t)n * (1 + n_MC)), pv, ub;
    // REML_first_order (:696-728)
    auto first_order = [&](float ss) -> float {
      ++evals;
      const float tau = (1 - ss) / ss;
      const float sigma = tau - tau0;
      abi(slemm_gpu_lsolve(sg.s, 0, T, dV.p, n, sigma, 1e-6f, dPV.p, n, nullptr));
      download(pv, dPV.p, (size_t)n * T);
      const float st = std::sqrt(tau);
      for (int64_t i = 0; i < n; ++i) PyMC[i] = pv[i];
      for (int k = 1; k <= n_MC; ++k)
        for (int64_t i = 0; i < n; ++i)
          PyMC[(size_t)k * n + i] = pv[(size_t)k * n + i] + st * pv[(size_t)(n_MC + k) * n + i];
      project_out(Q, PyMC, n, 1 + n_MC);
      double e0 = sq_norm(PyMC.data(), n), eMC = 0.0;
      for (int k = 1; k <= n_MC; ++k) eMC += sq_norm(&PyMC[(size_t)k * n], n);
      for (int k = 0; k <= n_MC; ++k)
        for (int64_t i = 0; i < n; ++i) PyMC[(size_t)k * n + i] *= h->rr_host[i];
      upload(dPV.p, PyMC);
      abi(slemm_gpu_zt(h, dPV.p, n, 1 + n_MC, dU.p, m, 0, 0, precision, nullptr));
      HIPCHECK(hipDeviceSynchronize());
      download(ub, dU.p, (size_t)m * (1 + n_MC));
      double u0 = 0.0, uMC = 0.0;
      for (int64_t i = 0; i < m; ++i) {
        const double g = gw[i];
        u0 += (double)ub[i] * ub[i] * g;
        for (int k = 1; k <= n_MC; ++k) uMC += (double)ub[(size_t)k * m + i] * ub[(size_t)k * m + i] * g;
      }
      return (float)std::log((u0 / e0) / (uMC / eMC));
    };

    // secant iterations on log((1-h2)/h2) (:821-853)
    float s2 = s2max / 2;
    float log_tau1 = std::log((1 - s2) / s2);
    float f1 = first_order(s2);
    if (verbose) printf("  Iter 1: h2=%.4f log((1-h2)/h2)=%.4f funct=%.4f\n", s2, log_tau1, f1);
    s2 = f1 < 0 ? s2max / 4 : s2max;
    float log_tau2 = std::log((1 - s2) / s2);
    float f2 = first_order(s2);
    if (verbose) printf("  Iter 2: h2=%.4f log((1-h2)/h2)=%.4f funct=%.4f\n", s2, log_tau2, f2);
    for (int si = 3; si <= secant_max_iter; ++si) {
      float lt = (log_tau1 * f2 - log_tau2 * f1) / (f2 - f1);
      if (lt > 9) lt = 9;
      else if (lt < -9) lt = -9;
      if (std::fabs(lt - log_tau2) < secant_tol) break;
      s2 = 1 / (std::exp(lt) + 1);
      f1 = f2;
      f2 = first_order(s2);
      log_tau1 = log_tau2;
      log_tau2 = lt;
      if (verbose) printf("  Iter %d: h2=%.4f log((1-h2)/h2)=%.4f funct=%.4f\n", si, s2, lt, f2);
    }
    if (std::fabs(f2) > 0.005f && verbose)
      printf("\nWarning: secant iterations not converged. h2 close to 0 or small sample size?\n");

    // Py, variance components, SNP BLUP (:855-866)
    std::copy(PyMC.begin(), PyMC.begin() + n, Py);
    double ypy = 0.0;
    for (int64_t i = 0; i < n; ++i) ypy += (double)y[i] * Py[i];
    const float vg = (float)(ypy / (double)(n - c));
    const float tau1 = (1 - s2) / s2;
    const float ve = tau1 * vg;
    std::vector<float> py(Py, Py + n);
    upload(dPV.p, py);
    abi(slemm_gpu_zt(h, dPV.p, n, 1, dU.p, m, 1, 0, precision, nullptr));
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(snp_blup, dU.p, sizeof(float) * m, hipMemcpyDeviceToHost));
    if (verbose) printf("\nVg,%g\nVe,%g\nh2,%g\n", vg, ve, s2);

    // BLUE: residual y - Zc snp_blup - tau1 R Py regressed on X (:871-875)
    abi(slemm_gpu_z(h, dU.p, m, 1, dPV.p, n, precision, nullptr));
    HIPCHECK(hipDeviceSynchronize());
    std::vector<float> zb;
    download(zb, dPV.p, n);
    std::vector<double> XtX((size_t)c * c), rhs(c);
    for (int a = 0; a < c; ++a) {
      double s = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        const double r = (double)y[i] - ((double)zb[i] + (double)tau1 * h->rvec_host[i] * Py[i]);
        s += (double)X[i + (size_t)a * ldx] * r;
      }
      rhs[a] = s;
      for (int b = 0; b < c; ++b) {
        double t = 0.0;
        for (int64_t i = 0; i < n; ++i) t += (double)X[i + (size_t)a * ldx] * X[i + (size_t)b * ldx];
        XtX[a + (size_t)b * c] = t;
      }
    }
    if (!chol_solve(XtX, c, rhs)) fail(SLEMM_ERR_RANK, "X'X is not positive definite");
    for (int q = 0; q < c; ++q) blue[q] = (float)rhs[q];
    if (verbose) printf("\nCompleted Lanczos MC-REML.\n");
    if (s2 > s2max && verbose)
      printf("\nWarning: the h2 estimate is larger than the max you specified.\n"
             "Can rerun with a higher --max_heritability to gain accuracy.\n");
    if (stats) {
      stats->vg = vg;
      stats->ve = ve;
      stats->h2 = s2;
      stats->lrt = 0;
      stats->lanczos_steps[0] = steps;
      stats->lanczos_steps[1] = 0;
      stats->lanczos_steps[2] = 0;
      stats->criterion_evals = evals;
      stats->seconds_seed = t_seed;
      stats->seconds_total = seconds_since(t_start);
    }
  });
}

extern "C" int slemm_gpu_lmc_reml(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y, int rng_seed,
                                  int n_MC, float tol_L, float s2max, float s2min, int precision, int verbose,
                                  float* Py, float* snp_blup, float* blue, slemm_reml_stats* stats) {
  // the reference's secant schedule: stop at |d log tau| < 0.002, at most 9 evaluations (:824-836)
  return slemm_gpu_lmc_reml_ex(h, X, ldx, c, y, rng_seed, n_MC, tol_L, s2max, s2min, 0.002f, 9, precision, verbose,
                               Py, snp_blup, blue, stats);
}
