// --iter_weighting (src/option.cpp:417-528): an equal-weight fit, per-SNP genetic variances from
// its SNP effects, clumping of the largest ones into fixed effects, window-averaged logistic SNP
// weights, and the weighted refit.  Host driver over the REML drivers of this library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "driver.hpp"

using namespace slemm;

namespace {

// one REML fit with the handle's current weights: SLDF (--lrt) or LMC
void fit(slemm_geno* h, bool sldf, const float* X, int64_t ldx, int c, const float* y, int rng_seed, int n_V,
         int n_MC, float tol_L, float s2max, float s2min, float tol_VC, int precision, int verbose, float* Py,
         float* blup, float* blue, slemm_reml_stats* st) {
  if (sldf) {
    std::vector<float> vb((size_t)c * c);
    abi(slemm_gpu_sldf_reml(h, X, ldx, c, y, rng_seed, n_V, tol_L, s2max, s2min, tol_VC, precision, verbose, Py,
                            blup, blue, vb.data(), st));
  } else {
    abi(slemm_gpu_lmc_reml(h, X, ldx, c, y, rng_seed, n_MC, tol_L, s2max, s2min, precision, verbose, Py, blup,
                           blue, st));
  }
}

float mean_of(const std::vector<float>& v) {
  double s = 0.0;
  for (float x : v) s += x;
  return (float)(s / (double)v.size());
}

}  // namespace

extern "C" int slemm_gpu_iter_weighting_reml(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                                             const float* persnp, const float* rvec, const float* maf,
                                             int window_size, float sig_chisq, float indep_r2, int lrt, int rng_seed,
                                             int n_V, int n_MC, float tol_L, float s2max, float s2min, float tol_VC,
                                             int precision, int verbose, float* Py, float* snp_blup, float* blue,
                                             float* persnp_out, int64_t* sig_idx, int* n_sig,
                                             slemm_reml_stats* stats) {
  return guard([&] {
    REQUIRE(h && X && y && maf && Py && snp_blup && blue && persnp_out && sig_idx && n_sig && stats,
            "null argument");
    REQUIRE(c >= 1 && ldx >= h->n, "covariates: need c >= 1 and ldx >= n");
    REQUIRE(window_size >= 0, "--window_size is negative.");
    HIPCHECK(hipSetDevice(h->device));
    const int64_t n = h->n, m = h->m;
    std::vector<float> base(m, 1.f);
    if (persnp) std::copy(persnp, persnp + m, base.begin());
    *n_sig = 0;
    if (m < window_size + 1) {  // :417-421: weighting disabled
      if (verbose) printf("\nWarning: --window_size >= total # of SNPs in --iter_weighting.\n");
      const float mu = mean_of(base);
      for (auto& v : base) v /= mu;
      abi(slemm_gpu_set_model(h, base.data(), rvec, nullptr, n, 0));
      fit(h, lrt != 0, X, ldx, c, y, rng_seed, n_V, n_MC, tol_L, s2max, s2min, tol_VC, precision, verbose, Py,
          snp_blup, blue, stats);
      std::copy(base.begin(), base.end(), persnp_out);
      return;
    }

    // equal-weight fit (:440-452): SLDF only for --lrt below 5,000 individuals
    std::vector<float> ones(m, 1.f);
    abi(slemm_gpu_set_model(h, ones.data(), rvec, nullptr, n, 0));
    std::vector<float> blup0(m), blue0(c), py0(n);
    slemm_reml_stats st0;
    fit(h, lrt != 0 && n < 5000, X, ldx, c, y, rng_seed, n_V, n_MC, tol_L, s2max, s2min, tol_VC, precision, verbose,
        py0.data(), blup0.data(), blue0.data(), &st0);

    // per-SNP genetic variance 2 p (1 - p) b^2, normalised to mean 1 (:456-462)
    std::vector<float> snpvar(m);
    for (int64_t i = 0; i < m; ++i) snpvar[i] = 2.0f * maf[i] * (1.0f - maf[i]) * blup0[i] * blup0[i];
    {
      const float mu = mean_of(snpvar);
      for (auto& v : snpvar) v /= mu;
    }

    // clumping (:464-492): largest remaining variance above sig_chisq joins the fixed effects if
    // its centred dosages keep more than indep_r2 of their sum of squares after regression on
    // the SNPs already taken
    std::vector<float> tmp = snpvar;
    std::vector<int64_t> sig;
    std::vector<double> S;  // n x |sig| centred dosages, FP64
    for (;;) {
      const int64_t j = (int64_t)(std::max_element(tmp.begin(), tmp.end()) - tmp.begin());
      if (!(tmp[j] > sig_chisq)) break;
      tmp[j] = 0.f;
      std::vector<float> z;
      extract_dosage_cols(h, {j}, z);
      double mu = 0.0;
      for (int64_t i = 0; i < n; ++i) mu += z[i];
      const float muf = (float)(mu / (double)n);
      std::vector<double> yc(n);
      for (int64_t i = 0; i < n; ++i) yc[i] = (double)(z[i] - muf);
      bool take = sig.empty();
      if (!take) {
        const int k = (int)sig.size();
        MatD A(n, k);
        std::copy(S.begin(), S.end(), A.a.begin());
        MatD Qs, R;
        thin_qr(A, Qs, R, 0.0);
        std::vector<double> r = yc;
        for (int q = 0; q < Qs.c; ++q) {
          double s = 0.0;
          for (int64_t i = 0; i < n; ++i) s += Qs(i, q) * r[i];
          for (int64_t i = 0; i < n; ++i) r[i] -= Qs(i, q) * s;
        }
        double rr2 = 0.0, yy2 = 0.0;
        for (int64_t i = 0; i < n; ++i) {
          rr2 += r[i] * r[i];
          yy2 += yc[i] * yc[i];
        }
        take = rr2 / yy2 > (double)indep_r2;
      }
      if (take) {
        sig.push_back(j);
        S.insert(S.end(), yc.begin(), yc.end());
      }
    }

    // window means of the variances over window_size + 1 SNPs, clamped at the ends, then the
    // logistic weights (:494-502); clumped SNPs get weight 0 (:508-513)
    std::vector<double> pre(m + 1, 0.0);
    for (int64_t i = 0; i < m; ++i) pre[i + 1] = pre[i] + snpvar[i];
    const int64_t w = window_size, half = window_size / 2;
    auto win = [&](int64_t a) { return (float)((pre[a + w + 1] - pre[a]) / (double)(w + 1)); };
    for (int64_t i = 0; i < m; ++i) {
      float v;
      if (i < half) v = win(0);
      else if (i > m - half - 1) v = win(m - w - 1);
      else v = win(i - half);
      tmp[i] = 0.001f / (0.001f + 0.999f * std::exp(v * -0.5f));
    }
    for (int64_t j : sig) tmp[j] = 0.f;
    std::vector<float> pw(m);
    for (int64_t i = 0; i < m; ++i) pw[i] = base[i] * tmp[i];
    {
      const float mu = mean_of(pw);
      for (auto& v : pw) v /= mu;
    }

    // weighted refit with the clumped SNPs as extra covariates (:505-524)
    const int k = (int)sig.size();
    std::vector<float> Xn((size_t)n * (c + k));
    for (int a = 0; a < c; ++a)
      for (int64_t i = 0; i < n; ++i) Xn[(size_t)a * n + i] = X[i + (size_t)a * ldx];
    for (int a = 0; a < k; ++a)
      for (int64_t i = 0; i < n; ++i) Xn[(size_t)(c + a) * n + i] = (float)S[(size_t)a * n + i];
    abi(slemm_gpu_set_model(h, pw.data(), rvec, nullptr, n, 0));
    std::vector<float> bl(c + k);
    fit(h, lrt != 0, Xn.data(), n, c + k, y, rng_seed, n_V, n_MC, tol_L, s2max, s2min, tol_VC, precision, verbose,
        Py, snp_blup, bl.data(), stats);
    for (int a = 0; a < k; ++a) snp_blup[sig[a]] = bl[c + a];  // :523
    std::copy(bl.begin(), bl.begin() + c, blue);
    std::copy(pw.begin(), pw.end(), persnp_out);
    std::copy(sig.begin(), sig.end(), sig_idx);
    *n_sig = k;
  });
}
