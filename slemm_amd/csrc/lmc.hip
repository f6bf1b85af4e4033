// LMC_REML (src/reml.cpp:692-882, without the --lmm association part) over the device engine,
// with every n- and m-length vector of the secant iterations resident on the device.
//
// The reference's steps, in its order: covariate QR (:765-769); the Monte Carlo draws u (m x n_MC)
// and e (n x n_MC) from one mt19937(seed + 54321) stream through normal_distribution<float>
// (:788-801); V = (I - QQ') [R^-1/2 y, R^-1/2 Zc (sqrt(gw) u), e] (:802-813); one Lanczos seed of
// the projected operator at 10 tol_L (:817-819); per secant point REML_first_order (:696-728) by
// L_Solve, Py_k = x_k + sqrt(tau) x_{n_MC+k}, projection, e = Py, u = sqrt(gw) Zc' R^-1/2 Py and
// f = log[(|u_0|^2/|e_0|^2) / (sum |u_k|^2 / sum |e_k|^2)]; the secant on log tau (:821-853); vg,
// ve, the SNP BLUP and the BLUE (:855-877).  Here the per-point work stays on the device: the
// combination, projection and rr scaling are kernels, the squared norms are FP64 column
// reductions, and only 2 (1 + n_MC) doubles come back per point.
//
// The normal draws reproduce libstdc++'s normal_distribution<float> over std::mt19937 value for
// value (Marsaglia's polar method, the second value of each pair cached for the next call): the
// engine's words are generated in whole twisted blocks, and the attempts are then tested and
// transformed in parallel host threads (normal_draws below).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "blockprod.hpp"
#include "driver.hpp"
#include "rng.hpp"

namespace slemm {

// count values of std::normal_distribution<float>(0, 1) drawn from std::mt19937(seed), in draw
// order: each attempt takes engine words 2j (x) and 2j + 1 (y) and is accepted when
// r2 = x^2 + y^2 is in (0, 1]; accepted pair p gives out[2p] = y m, out[2p+1] = x m with
// m = sqrt(-2 log(r2) / r2), all in float.  The engine's words are generated in order (the
// only sequential part); the attempts are tested, counted and transformed in parallel chunks,
// the counts' prefix sums placing each chunk's pairs.
void normal_draws(uint32_t seed, int64_t count, float* out) {
  if (count <= 0) return;
  const int64_t npairs = (count + 1) / 2;
  const int threads = npairs < (1 << 14) ? 1 : host_threads();
  Mt19937Blocks mt(seed);
  std::vector<uint32_t> raw;
  constexpr int64_t kChunk = 1 << 14;  // attempts per chunk
  std::vector<int64_t> acc;
  // attempts: pi/4 of them are accepted; start a little above the expected need and extend
  int64_t want = npairs + npairs / 4 + npairs / 16 + 2048;
  for (;;) {
    const int64_t words = 2 * want;
    const int64_t have = (int64_t)raw.size();
    if (words > have) {
      const int64_t nblk = (words - have + Mt19937Blocks::kN - 1) / Mt19937Blocks::kN;
      raw.resize((size_t)(have + nblk * Mt19937Blocks::kN));
      mt_fill_blocks(mt, nblk, &raw[(size_t)have]);
    }
    const int64_t attempts = (int64_t)raw.size() / 2;
    const int64_t nchunk = (attempts + kChunk - 1) / kChunk;
    acc.assign((size_t)nchunk, 0);
    parallel_parts(nchunk, threads, [&](int64_t c) {
      int64_t a = 0;
      for (int64_t j = c * kChunk; j < std::min(attempts, (c + 1) * kChunk); ++j) {
        const float x = polar_coord(raw[2 * j]), y = polar_coord(raw[2 * j + 1]);
        const float xx = x * x;  // separate statements: no fused multiply-add, as the reference's
        const float yy = y * y;  // baseline x86-64 build evaluates them
        const float r2 = xx + yy;
        a += !(r2 > 1.0f || r2 == 0.0f);
      }
      acc[c] = a;
    });
    int64_t total = 0;
    for (auto& a : acc) {
      const int64_t t = a;
      a = total;  // first pair index of the chunk
      total += t;
    }
    if (total >= npairs) break;
    want = attempts + (npairs - total) * 2 + 4096;
  }
  const int64_t attempts = (int64_t)raw.size() / 2;
  parallel_parts((int64_t)acc.size(), threads, [&](int64_t c) {
    int64_t p = acc[c];
    for (int64_t j = c * kChunk; j < std::min(attempts, (c + 1) * kChunk) && p < npairs; ++j) {
      const float x = polar_coord(raw[2 * j]), y = polar_coord(raw[2 * j + 1]);
      const float xx = x * x;
      const float yy = y * y;
      const float r2 = xx + yy;
      if (r2 > 1.0f || r2 == 0.0f) continue;
      const float lg = -2.0f * std::log(r2);
      const float mult = std::sqrt(lg / r2);
      const float a = y * mult, b = x * mult;
      out[2 * p] = a * 1.0f + 0.0f;  // ret * stddev + mean
      if (2 * p + 1 < count) out[2 * p + 1] = b * 1.0f + 0.0f;
      ++p;
    }
  });
}

namespace {

// A[i, k] *= s_i (sqrt(s_i) with SQRT) for i < rows, k < t; IEEE round-to-nearest products.
template <bool SQRT>
__global__ __launch_bounds__(256) void k_scale_rows(float* __restrict__ A, int64_t lda, int64_t rows, int t,
                                                    const float* __restrict__ s) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  const float f = SQRT ? __fsqrt_rn(s[i]) : s[i];
  for (int k = 0; k < t; ++k) A[i + (int64_t)k * lda] = __fmul_rn(A[i + (int64_t)k * lda], f);
}

// P[:, 0] = X[:, 0];  P[:, k] = X[:, k] + st X[:, n_MC + k] (k = 1..n_MC), float product then sum
__global__ __launch_bounds__(256) void k_lmc_combine(const float* __restrict__ X, int64_t n, int n_mc, float st,
                                                     float* __restrict__ P) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  P[i] = X[i];
  for (int k = 1; k <= n_mc; ++k)
    P[i + k * n] = __fadd_rn(X[i + k * n], __fmul_rn(st, X[i + (int64_t)(n_mc + k) * n]));
}

// Weighted squared column norms, stage 1: block (k, b) sums w_i A[i,k]^2 over rows
// [b rpb, (b+1) rpb) in FP64; stage 2 sums the blocks of each column in fixed order.
__global__ __launch_bounds__(256) void k_wsq1(const float* __restrict__ A, int64_t lda, const float* __restrict__ w,
                                              int64_t rows, int64_t rpb, int t, double* __restrict__ part) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  const int64_t i0 = (int64_t)blockIdx.y * rpb, i1 = min(rows, i0 + rpb);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    const double a = A[i + (int64_t)k * lda];
    s += a * a * (double)w[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * t + k] = red[0];
}

__global__ void k_wsq2(const double* __restrict__ part, int64_t nblk, int t, double* __restrict__ res) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t) return;
  double s = 0.0;
  for (int64_t b = 0; b < nblk; ++b) s += part[b * t + k];
  res[k] = s;
}

unsigned blocks_of(int64_t rows) { return (unsigned)((rows + 255) / 256); }

// The device state of one LMC fit: V (n x T), the L_Solve solutions X and, for up to kMaxPts
// secant points at once, the combined Py blocks P (n x (1 + n_MC) each), Zc' P and the FP64
// reduction buffers.
constexpr int kMaxPts = 3;
struct LmcSystem {
  slemm_geno* h;
  int64_t n, m;
  int n_mc, T, c;
  DevBuf dV, dX, dP, dU, dots, part;

  LmcSystem(slemm_geno* hh, int nmc, int cc)
      : h(hh), n(hh->n), m(hh->m), n_mc(nmc), T(1 + 2 * nmc), c(cc) {
    dV.reserve(sizeof(float) * n * T);
    dX.reserve(sizeof(float) * n * T);
    dP.reserve(sizeof(float) * n * (1 + n_mc) * kMaxPts);
    dU.reserve(sizeof(float) * m * (1 + n_mc) * kMaxPts);
    dots.reserve(sizeof(double) * 2 * (1 + n_mc) * kMaxPts);
  }

  // A <- (I - QQ') A on an n x t device block (FP64 Q'A)
  void project(float* A, int t) {
    if (c == 0) return;
    h->qtv.reserve(sizeof(double) * c * t);
    tn_product(h->Q, h->n_pad, c, A, n, t, n, h->qtv.as<double>(), t, h->redpart, nullptr);
    launch_quad_finish(A, n, A, n, h->Q, h->n_pad, c, h->qtv.as<double>(), t, nullptr, t, n, 0.f, nullptr);
  }

  void scale_rr(float* A, int t) {
    if (h->rr_unit) return;
    k_scale_rows<false><<<blocks_of(n), 256>>>(A, n, n, t, h->rr);
    HIPCHECK(hipGetLastError());
  }

  // sum_i w_i A[i,k]^2, k < t, over `rows` rows, into res (device doubles)
  void wsq(const float* A, int64_t lda, const float* w, int64_t rows, int t, double* res) {
    const int64_t rpb = 16384, nblk = (rows + rpb - 1) / rpb;
    part.reserve(sizeof(double) * nblk * t);
    k_wsq1<<<dim3((unsigned)t, (unsigned)nblk), 256>>>(A, lda, w, rows, rpb, t, part.as<double>());
    HIPCHECK(hipGetLastError());
    k_wsq2<<<(t + 255) / 256, 256>>>(part.as<double>(), nblk, t, res);
    HIPCHECK(hipGetLastError());
  }
};

}  // namespace
}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_lmc_draws(int rng_seed, int64_t m, int64_t n, int n_MC, float* u, float* e) {
  return guard([&] {
    REQUIRE(m >= 0 && n >= 0 && n_MC >= 0, "sizes must be non-negative");
    REQUIRE((u || m * n_MC == 0) && (e || n * n_MC == 0), "null argument");
    const int64_t nu = m * n_MC, ne = n * n_MC;
    std::vector<float> all((size_t)(nu + ne));
    normal_draws((uint32_t)(rng_seed + 54321), nu + ne, all.data());
    std::copy(all.begin(), all.begin() + nu, u);
    std::copy(all.begin() + nu, all.end(), e);
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
    LmcSystem L(h, n_MC, c);
    const int T = L.T;
    float* V = L.dV.as<float>();

    // V = (I - QQ') [R^-1/2 y, R^-1/2 Zc (sqrt(gw) u), e] (:788-813); u drawn first, then e
    {
      const int64_t nu = m * n_MC, ne = n * n_MC;
      float* draws = static_cast<float*>(pinned_scratch(1, sizeof(float) * (size_t)(nu + ne)));
      normal_draws((uint32_t)(rng_seed + 54321), nu + ne, draws);
      DevF du((size_t)nu);
      HIPCHECK(hipMemcpy(du.p, draws, sizeof(float) * nu, hipMemcpyHostToDevice));
      k_scale_rows<true><<<blocks_of(m), 256>>>(du.p, m, m, n_MC, h->gw);
      HIPCHECK(hipGetLastError());
      abi(slemm_gpu_z(h, du.p, m, n_MC, V + n, n, precision, nullptr));
      HIPCHECK(hipMemcpy(V, y, sizeof(float) * n, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(V + (size_t)(1 + n_MC) * n, draws + nu, sizeof(float) * ne, hipMemcpyHostToDevice));
    }
    L.scale_rr(V, 1 + n_MC);
    L.project(V, T);

    // one seed system at 10x the tolerance, as BOLT-LMM relaxes CG (:817-819)
    const float tau0 = (1 - s2max) / s2max;
    SeedGuard sg;
    abi(slemm_gpu_lseed(h, SLEMM_OP_QUADFORM, nullptr, nullptr, tau0, V, n, T, 10 * tol_L, 600, T, precision,
                        verbose, nullptr, &sg.s));
    int steps = 0;
    abi(slemm_gpu_lseed_info(sg.s, nullptr, &steps, nullptr));
    const double t_seed = seconds_since(t_start);

    int evals = 0;
    float* Xs = L.dX.as<float>();
    float* P = L.dP.as<float>();
    float* U = L.dU.as<float>();
    double* dots = L.dots.as<double>();
    const int w = 1 + n_MC;
    std::vector<double> nr(2 * (size_t)w * kMaxPts);
    // REML_first_order (:696-728) at npt heritabilities in one batch: each point's solves and
    // combined block, then one Zc' pass over all the points' columns (each column's products are
    // independent of the others in the block)
    auto first_order_pts = [&](const float* ss, int npt, float* f) {
      for (int p = 0; p < npt; ++p) {
        const float tau = (1 - ss[p]) / ss[p];
        const float sigma = tau - tau0;
        abi(slemm_gpu_lsolve(sg.s, 0, T, V, n, sigma, 1e-6f, Xs, n, nullptr));
        k_lmc_combine<<<blocks_of(n), 256>>>(Xs, n, n_MC, std::sqrt(tau), P + (size_t)p * w * n);
        HIPCHECK(hipGetLastError());
      }
      const int t = npt * w;
      L.project(P, t);
      col_dots(h, P, n, P, n, n, t, dots, nullptr);  // |e_k|^2
      L.scale_rr(P, t);
      abi(slemm_gpu_zt(h, P, n, t, U, m, 0, 0, precision, nullptr));
      L.wsq(U, m, h->gw, m, t, dots + t);  // |sqrt(gw) Zc' R^-1/2 Py_k|^2
      HIPCHECK(hipMemcpy(nr.data(), dots, sizeof(double) * 2 * t, hipMemcpyDeviceToHost));
      for (int p = 0; p < npt; ++p) {
        const double* e = nr.data() + (size_t)p * w;
        const double* u = nr.data() + t + (size_t)p * w;
        double eMC = 0.0, uMC = 0.0;
        for (int k = 1; k <= n_MC; ++k) {
          eMC += e[k];
          uMC += u[k];
        }
        f[p] = (float)std::log((u[0] / e[0]) / (uMC / eMC));
      }
    };
    float* Plast = P;  // the last evaluated point's block (its column 0 is R^-1/2 P y)
    auto first_order = [&](float ss) -> float {
      ++evals;
      float f = 0.f;
      first_order_pts(&ss, 1, &f);
      Plast = P;
      return f;
    };

    // secant iterations on log((1-h2)/h2) (:821-853).  The second point depends on the sign at
    // the first (s2max / 4 or s2max), so the first batch evaluates the first point and both
    // candidates for the second: one Zc' pass instead of two
    float s2 = s2max / 2;
    float log_tau1 = std::log((1 - s2) / s2);
    const float first3[3] = {s2, s2max / 4, s2max};
    float fb[3];
    first_order_pts(first3, 3, fb);
    evals += 2;
    float f1 = fb[0];
    if (verbose) printf("  Iter 1: h2=%.4f log((1-h2)/h2)=%.4f funct=%.4f\n", s2, log_tau1, f1);
    s2 = f1 < 0 ? s2max / 4 : s2max;
    float log_tau2 = std::log((1 - s2) / s2);
    float f2 = f1 < 0 ? fb[1] : fb[2];
    Plast = P + (size_t)(f1 < 0 ? 1 : 2) * w * n;
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

    // Py (the last point's R^-1/2 P y, column 0 of its block), variance components, SNP BLUP
    // (:855-866)
    HIPCHECK(hipMemcpy(Py, Plast, sizeof(float) * n, hipMemcpyDeviceToHost));
    double ypy = 0.0;
    for (int64_t i = 0; i < n; ++i) ypy += (double)y[i] * Py[i];
    const float vg = (float)(ypy / (double)(n - c));
    const float tau1 = (1 - s2) / s2;
    const float ve = tau1 * vg;
    abi(slemm_gpu_zt(h, Plast, n, 1, U, m, 1, 0, precision, nullptr));
    HIPCHECK(hipMemcpy(snp_blup, U, sizeof(float) * m, hipMemcpyDeviceToHost));
    if (verbose) printf("\nVg,%g\nVe,%g\nh2,%g\n", vg, ve, s2);

    // BLUE: residual y - Zc snp_blup - tau1 R Py regressed on X (:871-875)
    abi(slemm_gpu_z(h, U, m, 1, Xs, n, precision, nullptr));
    std::vector<float> zb;
    download(zb, Xs, n);
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
