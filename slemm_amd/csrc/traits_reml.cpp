// SLDF_REML (src/reml.cpp:549-690) for a batch of traits on individual subsets, on one of two
// plans:
//  - batched: one Lanczos seed carries every trait's phenotype, intercept and probe columns
//    through the shared block products, and the golden-section search then runs per trait on
//    replays of that seed;
//  - one trait at a time: each trait on a handle compacted to its individuals (device-side), with
//    the trait's SNP weights, through the single-trait driver.
// Per trait the computation is the single-trait driver's on the trait's individuals either way:
// the same probes (drawn from mt19937(seed+54321) over the subset), the same criterion and search.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "driver.hpp"
#include "traits.hpp"

using namespace slemm;

namespace {

// A batched block product costs each 16-column tile its subsets' extent (rows up to the last
// member of any of its traits); one trait at a time costs each trait its n_V + 1 columns over its
// own individuals.  The batched products run at higher efficiency per column-row (one wide block
// instead of T narrower ones, one graph per Lanczos step, one seed block built on the device): on
// n = 200,000 individuals, m = 300,000 SNPs, 32 traits of 32 probes they took 1.61 s against 1.81 s
// at a work ratio of 1.065 (nested subsets) and 2.87 s against 1.80 s at 2.35 (random subsets).
// The line through the two batched times crosses the one-at-a-time time at a ratio of 1.27, so
// the automatic plan batches up to a ratio of 1.25.
constexpr double kBatchedWorkRatio = 1.25;

double batched_work_ratio(const slemm_traits* tr, int n_V) {
  const int T = tr->T;
  std::vector<int64_t> ext;  // per column of the batched block, in its order
  for (int t = 0; t < T; ++t) ext.push_back(tr->extent[t]);
  for (int t = 0; t < T; ++t) ext.push_back(tr->extent[t]);
  for (int t = 0; t < T; ++t)
    for (int j = 0; j < n_V; ++j) ext.push_back(tr->extent[t]);
  double wb = 0, ws = 0;
  for (size_t c0 = 0; c0 < ext.size(); c0 += 16)
    wb += 16.0 * (double)*std::max_element(ext.begin() + c0, ext.begin() + std::min(ext.size(), c0 + 16));
  for (int t = 0; t < T; ++t) ws += (double)tr->n_host[t] * (n_V + 1);
  return wb / ws;
}

int resolve_plan(const slemm_traits* tr, int n_V, double* ratio) {
  const double r = batched_work_ratio(tr, n_V);
  if (ratio) *ratio = r;
  if (tr->plan != SLEMM_TRAITS_AUTO) return tr->plan;
  return r <= kBatchedWorkRatio ? SLEMM_TRAITS_BATCHED : SLEMM_TRAITS_SEQUENTIAL;
}

struct GenoGuard {
  slemm_geno* g = nullptr;
  ~GenoGuard() {
    if (g) slemm_gpu_geno_destroy(g);
  }
};

// one trait at a time: the trait's individuals compacted on the device into one handle sized for
// the largest subset (allocated once: a fresh multi-GB allocation per trait costs more than the
// trait's fit), with the trait's weights (monomorphic SNPs of the subset at weight 0), then the
// single-trait driver with an intercept
void traits_one_at_a_time(slemm_traits* tr, const float* y, int rng_seed, int n_V, float tol_L, float s2max,
                          float s2min, float tol_VC, int precision, int verbose, float* vg_out, float* ve_out,
                          float* h2_out, float* lrt_out, int* steps_out, double* seconds) {
  slemm_geno* h = tr->h;
  const auto t_start = std::chrono::steady_clock::now();
  double t_seed = 0;
  GenoGuard gk;
  gk.g = geno_alloc_like(h, *std::max_element(tr->n_host.begin(), tr->n_host.end()));
  for (int t = 0; t < tr->T; ++t) {
    const std::vector<int64_t>& mem = tr->members[t];
    const int64_t nt = (int64_t)mem.size();
    geno_refill(gk.g, h, mem.data(), nt);
    set_model_weights(gk.g, tr->gw.as<float>() + (size_t)t * h->m_pad);
    std::vector<float> X(nt, 1.0f), yt(nt);
    for (int64_t i = 0; i < nt; ++i) yt[i] = y[(size_t)t * h->n + mem[i]];
    slemm_reml_stats st;
    std::memset(&st, 0, sizeof(st));
    // the variance components only, as the batched plan returns (no Py, SNP BLUP or BLUE)
    sldf_reml_run(gk.g, X.data(), nt, 1, yt.data(), rng_seed, n_V, tol_L, s2max, s2min, tol_VC, precision,
                  verbose, nullptr, nullptr, nullptr, nullptr, &st, false);
    vg_out[t] = st.vg;
    ve_out[t] = st.ve;
    h2_out[t] = st.h2;
    if (lrt_out) lrt_out[t] = st.lrt;
    if (steps_out)
      for (int q = 0; q < 3; ++q) steps_out[3 * t + q] = st.lanczos_steps[q];
    t_seed += st.seconds_seed;
  }
  if (seconds) {
    seconds[0] = t_seed;
    seconds[1] = seconds_since(t_start);
  }
}

}  // namespace

extern "C" int slemm_gpu_traits_set_plan(slemm_traits* tr, int plan) {
  return guard([&] {
    REQUIRE(tr, "null handle");
    REQUIRE(plan == SLEMM_TRAITS_AUTO || plan == SLEMM_TRAITS_BATCHED || plan == SLEMM_TRAITS_SEQUENTIAL,
            "unknown plan");
    tr->plan = plan;
  });
}

extern "C" int slemm_gpu_traits_plan(const slemm_traits* tr, int n_V, int* plan, double* work_ratio) {
  return guard([&] {
    REQUIRE(tr && plan, "null argument");
    REQUIRE(n_V >= 1, "number of probes must be positive");
    *plan = resolve_plan(tr, n_V, work_ratio);
  });
}

extern "C" int slemm_gpu_traits_sldf_reml(slemm_traits* tr, const float* y, int rng_seed, int n_V, float tol_L,
                                          float s2max, float s2min, float tol_VC, int precision, int verbose,
                                          float* vg_out, float* ve_out, float* h2_out, float* lrt_out,
                                          int* steps_out, double* seconds) {
  return guard([&] {
    REQUIRE(tr && y && vg_out && ve_out && h2_out, "null argument");
    REQUIRE(n_V >= 1, "number of probes must be positive");
    REQUIRE(s2min > 0 && s2max < 1 && s2min < s2max, "heritability range must satisfy 0 < s2min < s2max < 1");
    slemm_geno* h = tr->h;
    HIPCHECK(hipSetDevice(h->device));
    if (resolve_plan(tr, n_V, nullptr) == SLEMM_TRAITS_SEQUENTIAL) {
      traits_one_at_a_time(tr, y, rng_seed, n_V, tol_L, s2max, s2min, tol_VC, precision, verbose, vg_out, ve_out,
                           h2_out, lrt_out, steps_out, seconds);
      return;
    }
    const auto t_start = std::chrono::steady_clock::now();
    const int T = tr->T;
    const int64_t n = h->n;
    const int ncol = T * (2 + n_V);
    // columns: [y~_0 .. y~_{T-1}, q_0 .. q_{T-1}, probes of trait 0, probes of trait 1, ...],
    // written on the device from each trait's y~ and q values and one stream of probe signs
    // (trait t's probes, drawn from mt19937(seed+54321) over its n_t individuals, are the first
    // n_t n_V draws of that stream)
    std::vector<float> yq;
    std::vector<int> col_trait(ncol), kinds(ncol, SLEMM_OP_LDET), groups(ncol);
    std::vector<std::vector<float>> yp(T);
    std::vector<MatD> Qt(T);
    int64_t nmax = 0;
    for (int t = 0; t < T; ++t) {
      const std::vector<int64_t>& mem = tr->members[t];
      const int64_t nt = (int64_t)mem.size();
      // intercept basis and projected phenotype over the subset, as the single-trait driver
      // forms them on a handle holding only these individuals (src/reml.cpp:583-592)
      MatD ones(nt, 1), R;
      for (int64_t i = 0; i < nt; ++i) ones(i, 0) = 1.0;
      if (!thin_qr(ones, Qt[t], R, 1e-6)) fail(SLEMM_ERR_RANK, "Fixed-effects design matrix is not of full rank.");
      yp[t].resize(nt);
      for (int64_t i = 0; i < nt; ++i) yp[t][i] = y[(size_t)t * n + mem[i]];
      project_out(Qt[t], yp[t], nt, 1);
      yq.insert(yq.end(), yp[t].begin(), yp[t].end());
      for (int64_t i = 0; i < nt; ++i) yq.push_back((float)Qt[t](i, 0));
      nmax = std::max(nmax, nt);
      col_trait[t] = t;
      kinds[t] = SLEMM_OP_QUADFORM;
      groups[t] = t;
      col_trait[T + t] = t;
      groups[T + t] = T + t;
      for (int j = 0; j < n_V; ++j) {
        col_trait[2 * T + t * n_V + j] = t;
        groups[2 * T + t * n_V + j] = 2 * T + t;
      }
    }
    DevF dB((size_t)n * ncol), dX(n);
    traits_fill_seed_block(tr, n_V, yq, rademacher_sign_bits(rng_seed, nmax * n_V), dB.p);
    if (verbose) printf("Generated %d Rademacher probing vectors for each of %d traits.\n", n_V, T);
    const float tau0 = (1 - s2max) / s2max;
    SeedGuard sg;
    abi(slemm_gpu_traits_lseed(tr, col_trait.data(), kinds.data(), groups.data(), tau0, dB.p, n, ncol, tol_L, 600,
                               2 * T, precision, verbose, nullptr, &sg.s));
    std::vector<int> csteps(ncol);
    abi(slemm_gpu_lseed_col_steps(sg.s, csteps.data()));
    const double t_seed = seconds_since(t_start);

    DevBuf dots;
    dots.reserve(sizeof(double));
    auto dot_dev = [&](const float* a, const float* b) {
      col_dots(h, a, n, b, n, n, 1, dots.as<double>(), nullptr);
      double r = 0;
      HIPCHECK(hipMemcpy(&r, dots.p, sizeof(double), hipMemcpyDeviceToHost));
      return r;
    };
    for (int t = 0; t < T; ++t) {
      const int64_t nt = (int64_t)tr->members[t].size();
      std::vector<std::vector<double>> Dk(n_V), Wk(n_V);
      for (int j = 0; j < n_V; ++j) {
        const int col = 2 * T + t * n_V + j;
        Dk[j].resize(csteps[col]);
        Wk[j].resize(csteps[col]);
        abi(slemm_gpu_lseed_spectrum(sg.s, col, Dk[j].data(), Wk[j].data()));
      }
      const double ySy = sq_norm(yp[t].data(), nt);
      const float* ycol = dB.p + (size_t)t * n;
      const float* qcol = dB.p + (size_t)(T + t) * n;
      int evals = 0;
      const bool replay = criterion_replay();
      const SeedQuadrature quad_y(sg.s, t, ySy);
      // q'H^-1q of the trait's intercept from the quadrature of its own seed column, as y'Py (the
      // reference's arithmetic, replay: from the L_Solve solution)
      double qSq = 0.0;
      for (int64_t i = 0; i < nt; ++i) qSq += (double)(float)Qt[t](i, 0) * (double)(float)Qt[t](i, 0);
      const SeedQuadrature quad_q(sg.s, T + t, qSq);
      // REML_criterion (src/reml.cpp:457-489) on the trait's subset (n_t individuals, c = 1)
      auto criterion = [&](float ss, float& vg, float& ve, float& lrt) -> double {
        ++evals;
        const float tau = (1 - ss) / ss;
        const float sigma = tau - tau0;
        double yPy = 0.0;
        if (replay) {
          abi(slemm_gpu_lsolve(sg.s, t, 1, ycol, n, sigma, 1e-6f, dX.p, n, nullptr));
          yPy = dot_dev(ycol, dX.p);
        } else {
          yPy = quad_y.inverse(sigma);
        }
        vg = (float)(yPy / (double)(nt - 1));
        ve = vg * tau;
        double acc = 0.0;
        for (int j = 0; j < n_V; ++j)
          for (size_t q = 0; q < Dk[j].size(); ++q) acc += Wk[j][q] * Wk[j][q] * std::log(Dk[j][q] + sigma);
        const double ldetH = acc * (double)nt / (double)n_V;
        lrt = (float)((double)nt * std::log(ySy / yPy) - ldetH);
        double QtX = 0.0;
        if (replay) {
          abi(slemm_gpu_lsolve(sg.s, T + t, 1, qcol, n, sigma, 1e-6f, dX.p, n, nullptr));
          QtX = dot_dev(qcol, dX.p);
        } else {
          QtX = quad_q.inverse(sigma);
        }
        return criterion_value((double)(nt - 1) * std::log(ySy / yPy) - std::log(std::fabs(QtX)) - ldetH);
      };
      float vg = 0, ve = 0, lrt = 0;
      golden_section(s2min, s2max, tol_VC, verbose, criterion, vg, ve, lrt);
      vg_out[t] = vg;
      ve_out[t] = ve;
      h2_out[t] = vg / (vg + ve);
      if (lrt_out) lrt_out[t] = lrt;
      if (steps_out) {
        steps_out[3 * t] = csteps[t];
        steps_out[3 * t + 1] = csteps[T + t];
        steps_out[3 * t + 2] = csteps[2 * T + t * n_V];
      }
      if (verbose) printf("Trait %d (n=%ld): Vg %g Ve %g h2 %g after %d evaluations\n", t, (long)nt, vg, ve,
                          vg / (vg + ve), evals);
    }
    if (seconds) {
      seconds[0] = t_seed;
      seconds[1] = seconds_since(t_start);
    }
  });
}
