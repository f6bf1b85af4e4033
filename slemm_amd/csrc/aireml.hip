// AI-REML: average-information Newton iterations on (sigma_g^2, sigma_e^2) over the device
// engine, for the model of SLDF_REML (src/reml.cpp:549-690):
//   y = X b + g + e,  var(g) = sigma_g^2 Zc W Zc',  var(e) = sigma_e^2 R.
// The reference has no AI-REML (it maximises the REML criterion by golden-section search,
// :493-547, or solves a first-order condition by secant, :821-853); covariates, probes, the
// seed systems and the outputs (Py, SNP BLUP, BLUE) are those of SLDF_REML.
//
// Whitened by R^-1/2 and projected off Q (Q R0 = R^-1/2 X), with A = R^-1/2 Zc W Zc' R^-1/2,
// tau = sigma_e^2/sigma_g^2 and Hq = (I-QQ') A (I-QQ') + tau I:
//   P u = (I-QQ') Hq^-1 (I-QQ') u / sigma_g^2
//   dl/dsigma_g^2 = -1/2 [tr(P A) - y'PAPy],  dl/dsigma_e^2 = -1/2 [tr(P) - y'PPy]
//   AI_kl = 1/2 y'P V_k P V_l P y   (V_g = A, V_e = I)
// Traces are Hutchinson estimates over unit-norm probes, tr(M) ~ n/nV sum_v v'Mv.  All solves
// at every tau replay one Lanczos seed (shift tau - tau0); the two AI solves per iteration run a
// fresh two-column seed at tau.  Dot products and the 2x2 update are FP64.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "blockprod.hpp"
#include "driver.hpp"

namespace slemm {
namespace {

struct AiScore {
  double score_g = 0, score_e = 0, ai_gg = 0, ai_ge = 0, ai_ee = 0, tr_pa = 0, tr_p = 0;
};

// The seed system [y~, Q, probes] (n x T) and the fixed products A v of the probes.
struct AiSystem {
  slemm_geno* h = nullptr;
  int64_t n = 0;
  int c = 0, nV = 0, T = 0, precision = 0;
  float tau0 = 0, tol_L = 0;
  MatD Q, qrR;
  DevBuf dB, dAV, dX, dW, dots;
  SeedGuard sg;
  std::vector<int> csteps;

  AiSystem(slemm_geno* hh, const float* X, int64_t ldx, int cc, const float* y, const float* probes, int64_t ldp,
           int n_V, int rng_seed, float tau_seed, float tol, int prec, int verbose)
      : h(hh), n(hh->n), c(cc), nV(n_V), T(1 + cc + n_V), precision(prec), tau0(tau_seed), tol_L(tol) {
    qrR = covariate_basis(h, X, ldx, c, Q);
    std::vector<float> B((size_t)n * T);
    for (int64_t i = 0; i < n; ++i) B[i] = h->rr_host[i] * y[i];
    project_out(Q, B, n, 1);
    for (int q = 0; q < c; ++q)
      for (int64_t i = 0; i < n; ++i) B[(size_t)(1 + q) * n + i] = (float)Q(i, q);
    if (!probes) rademacher_probes(rng_seed, n, nV, &B[(size_t)(1 + c) * n]);
    dB.reserve(sizeof(float) * n * T);
    HIPCHECK(hipMemcpy(dB.p, B.data(), sizeof(float) * n * (probes ? 1 + c : T), hipMemcpyHostToDevice));
    if (probes)
      HIPCHECK(hipMemcpy2D(dB.as<float>() + (size_t)(1 + c) * n, sizeof(float) * n, probes, sizeof(float) * ldp,
                           sizeof(float) * n, nV, hipMemcpyDeviceToDevice));
    // A v for the probes (unprojected operator, no shift)
    dAV.reserve(sizeof(float) * n * nV);
    run_op(h, SLEMM_OP_LDET, nullptr, 0.f, probe_cols(), n, nV, dAV.as<float>(), n, precision, nullptr);
    // y~ (QUADFORM), Q (LDET, for the BLUE), probes (QUADFORM: traces of P), one batched seed
    std::vector<int> kinds(T, SLEMM_OP_QUADFORM), groups(T, 2);
    groups[0] = 0;
    for (int q = 0; q < c; ++q) {
      kinds[1 + q] = SLEMM_OP_LDET;
      groups[1 + q] = 1;
    }
    abi(slemm_gpu_lseed(h, SLEMM_OP_QUADFORM, kinds.data(), groups.data(), tau0, dB.as<float>(), n, T, tol_L, 600,
                        T, precision, verbose, nullptr, &sg.s));
    csteps.resize(T);
    abi(slemm_gpu_lseed_col_steps(sg.s, csteps.data()));
    dX.reserve(sizeof(float) * n * T);
    dW.reserve(sizeof(float) * n * 4);
    dots.reserve(sizeof(double) * (2 * (size_t)nV + 8));
  }

  const float* probe_cols() const { return dB.as<float>() + (size_t)(1 + c) * n; }

  // X <- (I - QQ') X on an n x t device block
  void project(float* Xd, int t) {
    if (c == 0) return;
    h->qtv.reserve(sizeof(double) * c * t);
    tn_product(h->Q, h->n_pad, c, Xd, n, t, n, h->qtv.as<double>(), t, h->redpart, nullptr);
    launch_quad_finish(Xd, n, Xd, n, h->Q, h->n_pad, c, h->qtv.as<double>(), t, nullptr, t, n, 0.f, nullptr);
  }

  std::vector<double> dots_of(const float* A, const float* B, int t) {
    col_dots(h, A, n, B, n, n, t, dots.as<double>(), nullptr);
    std::vector<double> r(t);
    HIPCHECK(hipMemcpy(r.data(), dots.p, sizeof(double) * t, hipMemcpyDeviceToHost));
    return r;
  }

  double dot1(const float* a, const float* b) { return dots_of(a, b, 1)[0]; }

  AiScore evaluate(double vg, double ve, int verbose) {
    AiScore s;
    const double tau = ve / vg;
    const float sigma = (float)(tau - (double)tau0);
    float* X = dX.as<float>();
    const float* Bd = dB.as<float>();
    // Hq^-1 y~ and Hq^-1 v, projected
    abi(slemm_gpu_lsolve(sg.s, 0, 1, Bd, n, sigma, 1e-6f, X, n, nullptr));
    abi(slemm_gpu_lsolve(sg.s, 1 + c, nV, Bd + (size_t)(1 + c) * n, n, sigma, 1e-6f, X + (size_t)(1 + c) * n, n,
                         nullptr));
    project(X, 1);
    project(X + (size_t)(1 + c) * n, nV);
    const float* xv = X + (size_t)(1 + c) * n;
    const std::vector<double> ta = dots_of(xv, dAV.as<float>(), nV);
    const std::vector<double> ti = dots_of(xv, probe_cols(), nV);
    double sa = 0, si = 0;
    for (int k = 0; k < nV; ++k) {
      sa += ta[k];
      si += ti[k];
    }
    s.tr_pa = sa * (double)n / nV / vg;
    s.tr_p = si * (double)n / nV / vg;
    // A x_y and the two AI solves Hq^-1 [A x_y, x_y]
    float* W = dW.as<float>();  // [A x_y, x_y, solves (2)]
    run_op(h, SLEMM_OP_LDET, nullptr, 0.f, X, n, 1, W, n, precision, nullptr);
    HIPCHECK(hipMemcpy(W + n, X, sizeof(float) * n, hipMemcpyDeviceToDevice));
    const double xAx = dot1(X, W), xx = dot1(X, X);
    const double vg2 = vg * vg, vg3 = vg2 * vg;
    s.score_g = -0.5 * (s.tr_pa - xAx / vg2);
    s.score_e = -0.5 * (s.tr_p - xx / vg2);
    SeedGuard sa2;
    abi(slemm_gpu_lseed(h, SLEMM_OP_QUADFORM, nullptr, nullptr, (float)tau, W, n, 2, tol_L, 600, 2, precision, 0,
                        nullptr, &sa2.s));
    abi(slemm_gpu_lsolve(sa2.s, 0, 2, W, n, 0.f, 1e-6f, W + 2 * n, n, nullptr));
    project(W + 2 * n, 2);
    const float *Ax = W, *x = W + n, *z0 = W + 2 * n, *z1 = W + 3 * n;
    s.ai_gg = 0.5 * dot1(Ax, z0) / vg3;
    s.ai_ge = 0.5 * dot1(Ax, z1) / vg3;
    s.ai_ee = 0.5 * dot1(x, z1) / vg3;
    if (verbose)
      printf("  vg=%.6g ve=%.6g  score=(%.6g, %.6g)  AI=[%.6g %.6g; %.6g %.6g]\n", vg, ve, s.score_g, s.score_e,
             s.ai_gg, s.ai_ge, s.ai_ge, s.ai_ee);
    return s;
  }
};

void check_common(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y, int n_V) {
  REQUIRE(h && X && y, "null argument");
  REQUIRE(c >= 1 && ldx >= h->n, "covariates: need c >= 1 and ldx >= n");
  REQUIRE(n_V >= 1, "number of probes must be positive");
  HIPCHECK(hipSetDevice(h->device));
  require_model(h);
}

}  // namespace
}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_reml_score(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y, double vg,
                                    double ve, const float* probes, int64_t ldp, int n_V, int rng_seed, float tol_L,
                                    int precision, double* score, double* ai) {
  return guard([&] {
    check_common(h, X, ldx, c, y, n_V);
    REQUIRE(score, "score is null");
    REQUIRE(vg > 0 && ve > 0, "variance components must be positive");
    REQUIRE(!probes || ldp >= h->n, "ldp < n");
    AiSystem S(h, X, ldx, c, y, probes, ldp, n_V, rng_seed, (float)(ve / vg), tol_L, precision, 0);
    const AiScore s = S.evaluate(vg, ve, 0);
    score[0] = s.score_g;
    score[1] = s.score_e;
    if (ai) {
      ai[0] = s.ai_gg;
      ai[1] = s.ai_ge;
      ai[2] = s.ai_ee;
    }
  });
}

extern "C" int slemm_gpu_ai_reml(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y, int rng_seed,
                                 int n_V, float tol_L, float s2max, double vg0, double ve0, double tol_VC,
                                 int max_iter, int precision, int verbose, float* Py, float* snp_blup, float* blue,
                                 float* var_blue, double* history, slemm_reml_stats* stats) {
  return guard([&] {
    check_common(h, X, ldx, c, y, n_V);
    REQUIRE(Py && snp_blup && blue, "null argument");
    REQUIRE(s2max > 0 && s2max < 1, "max heritability must be in (0, 1)");
    REQUIRE(max_iter >= 1 && tol_VC > 0, "max_iter and tol_VC must be positive");
    const auto t_start = std::chrono::steady_clock::now();
    const int64_t n = h->n;
    if (verbose) printf("\nStarted AI-REML.\n");
    const float tau0 = (1 - s2max) / s2max;
    AiSystem S(h, X, ldx, c, y, nullptr, 0, n_V, rng_seed, tau0, tol_L, precision, verbose);
    const double t_seed = seconds_since(t_start);

    double vg = vg0, ve = ve0;
    if (!(vg > 0 && ve > 0)) {  // start at h2 = 1/2 of the projected phenotypic variance
      std::vector<float> yp;
      download(yp, S.dB.as<float>(), n);
      vg = ve = 0.5 * sq_norm(yp.data(), n) / (double)(n - c);
    }
    int it = 0;
    bool at_bound = false;
    for (; it < max_iter; ++it) {
      const AiScore s = S.evaluate(vg, ve, verbose);
      if (history) {
        double* hrow = history + (size_t)it * 7;
        hrow[0] = vg; hrow[1] = ve; hrow[2] = s.score_g; hrow[3] = s.score_e;
        hrow[4] = s.ai_gg; hrow[5] = s.ai_ge; hrow[6] = s.ai_ee;
      }
      const double det = s.ai_gg * s.ai_ee - s.ai_ge * s.ai_ge;
      if (!(det > 0)) fail(SLEMM_ERR_NUMERIC, "average-information matrix is not positive definite");
      double dg = (s.ai_ee * s.score_g - s.ai_ge * s.score_e) / det;
      double de = (s.ai_gg * s.score_e - s.ai_ge * s.score_g) / det;
      // step halving keeps both components positive
      double step = 1.0;
      while (step > 1e-6 && !(vg + step * dg > 0 && ve + step * de > 0)) step *= 0.5;
      double vg1 = vg + step * dg, ve1 = ve + step * de;
      at_bound = false;
      if (ve1 / vg1 < tau0) {  // h2 above s2max: stay on the boundary of the seed's range
        const double tot = vg1 + ve1;
        vg1 = s2max * tot;
        ve1 = tot - vg1;
        at_bound = true;
      }
      const double rel = std::max(std::fabs(vg1 - vg) / vg, std::fabs(ve1 - ve) / ve);
      vg = vg1;
      ve = ve1;
      if (verbose) printf("AI iteration %d: vg=%.6g ve=%.6g h2=%.6f rel change %.3g\n", it + 1, vg, ve,
                          vg / (vg + ve), rel);
      if (rel < tol_VC) {
        ++it;
        break;
      }
    }
    finish_estimates(h, S.sg.s, S.dB.as<float>(), tau0, (float)(ve / vg), S.Q, S.qrR, X, ldx, c, y, (float)vg,
                     precision, Py, snp_blup, blue, var_blue);
    if (verbose) {
      printf("\nVg,%g\nVe,%g\nh2,%g\n", vg, ve, vg / (vg + ve));
      if (at_bound) printf("\nWarning: the h2 estimate reached the max you specified.\n");
      printf("\nCompleted AI-REML.\n");
    }
    if (stats) {
      stats->vg = vg;
      stats->ve = ve;
      stats->h2 = vg / (vg + ve);
      stats->lrt = 0;
      stats->lanczos_steps[0] = S.csteps[0];
      stats->lanczos_steps[1] = S.csteps[1];
      stats->lanczos_steps[2] = S.csteps[1 + c];
      stats->criterion_evals = it;
      stats->seconds_seed = t_seed;
      stats->seconds_total = seconds_since(t_start);
    }
  });
}
