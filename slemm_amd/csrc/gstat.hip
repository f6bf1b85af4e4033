// Quadratic forms of the --lmm association approximation (src/reml.cpp:884-982, real-genotype
// branch :920-980): for sampled SNPs x_i, xPy = x_i' Py, xPx = v_i' H^-1 v_i and xSx = v_i' v_i
// with v_i = (I - QQ') R^-1/2 (x_i - gm_i) and H the quadform operator at tau1, and the local
// form xBx = (xSx - z' (Zw'Zw + tau1 diag(1/gw_w))^-1 z) / tau1 over the window Zw of
// window_size SNPs around x_i (z = Zw' v_i).  H^-1 v uses one Lanczos seed over all sampled
// SNPs and L_Solve.  The window systems are formed and solved in FP64, with v_i taken as its
// own column of the projected window (x_i lies in its window): the reference works in FP32,
// where xSx - z'G^-1 z cancels most of its digits.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "driver.hpp"

namespace slemm {
namespace {

// out[i + c ldo] = dosage(snps[c], i) - gm[snps[c]] in FP32, as src/reml.cpp:930-935
__global__ void k_extract_centered(const uint8_t* __restrict__ gs, int64_t ld_gs, const int64_t* __restrict__ snps,
                                   int64_t n, const float* __restrict__ gm, float* __restrict__ out, int64_t ldo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  const int64_t j = snps[c];
  const uint32_t code = (gs[j * ld_gs + (i >> 2)] >> (2 * (i & 3))) & 3u;
  out[i + (int64_t)c * ldo] = (float)((code & 1u) + (code >> 1)) - gm[j];
}

// Zw[i + j ld] = rr_i (dosage(wds + j, i) - gm) in FP64 (the FP32 centred value, then scaled)
__global__ void k_extract_window(const uint8_t* __restrict__ gs, int64_t ld_gs, int64_t wds, int64_t n,
                                 const float* __restrict__ gm, const float* __restrict__ rr, double* __restrict__ zw,
                                 int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= n) return;
  const uint32_t code = (gs[(wds + j) * ld_gs + (i >> 2)] >> (2 * (i & 3))) & 3u;
  const float x = (float)((code & 1u) + (code >> 1)) - gm[wds + j];
  zw[i + j * ld] = (double)rr[i] * (double)x;
}

// G_jj += tau1 / gw_(wds+j)  (gw infinite for monomorphic SNPs adds 0, as in the reference)
__global__ void k_add_ridge(double* __restrict__ G, int64_t ld, int w, const float* __restrict__ gw, int64_t wds,
                            double tau1) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < w) G[j + (int64_t)j * ld] += tau1 / (double)gw[wds + j];
}

void rb(rocblas_status s, const char* what) {
  if (s != rocblas_status_success) fail(SLEMM_ERR_HIP, std::string("rocBLAS/rocSOLVER failed in ") + what);
}

struct Blas {
  rocblas_handle h = nullptr;
  Blas() { rb(rocblas_create_handle(&h), "create_handle"); }
  ~Blas() {
    if (h) (void)rocblas_destroy_handle(h);
  }
};

template <typename T>
struct DevT {
  T* p = nullptr;
  explicit DevT(size_t count) { HIPCHECK(hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1))); }
  ~DevT() {
    if (p) (void)hipFree(p);
  }
  DevT(const DevT&) = delete;
  DevT& operator=(const DevT&) = delete;
};

// G[e] = sum over the S partial products W[s][e], in order
__global__ void k_sum_slices(const double* __restrict__ W, int64_t slice, int S, double* __restrict__ G) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= slice) return;
  double acc = 0.0;
  for (int q = 0; q < S; ++q) acc += W[(int64_t)q * slice + e];
  G[e] = acc;
}

// QA[q + j c] = sum_i Q[i + q n] A[i + j n] (c x w): one workgroup per column j of A, fixed-order
// reduction (rocBLAS runs this c-row product as a GEMM with a handful of output tiles)
__global__ __launch_bounds__(256) void k_qt_a(const double* __restrict__ Q, int c, const double* __restrict__ A,
                                              int64_t n, double* __restrict__ QA) {
  __shared__ double red[256];
  const int64_t j = blockIdx.x;
  const double* a = A + j * n;
  for (int q = 0; q < c; ++q) {
    const double* qq = Q + (int64_t)q * n;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += qq[i] * a[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) QA[q + j * c] = red[0];
    __syncthreads();
  }
}

// out (cols x rows, ld cols) = in' for in (rows x cols, ld rows), FP64, through 32 x 32 LDS tiles
__global__ void k_transpose_f64(const double* __restrict__ in, int64_t rows, int cols, double* __restrict__ out) {
  __shared__ double t[32][33];
  const int64_t r0 = (int64_t)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int64_t r = r0 + threadIdx.x;
    const int c = c0 + k;
    if (r < rows && c < cols) t[k][threadIdx.x] = in[r + (int64_t)c * rows];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int c = c0 + threadIdx.x;
    const int64_t r = r0 + k;
    if (r < rows && c < cols) out[c + r * cols] = t[threadIdx.x][k];
  }
}

void transpose_f64(const double* in, int64_t rows, int cols, double* out) {
  k_transpose_f64<<<dim3((unsigned)((rows + 31) / 32), (unsigned)((cols + 31) / 32)), dim3(32, 8)>>>(in, rows, cols,
                                                                                                   out);
  HIPCHECK(hipGetLastError());
}

// C = A'B (p x q, ld p) for tall FP64 A (n x p) and B (n x q), given as their transposes AT
// (p x n, ld p) and BT (q x n, ld q).  rocBLAS runs this shape well only as NT products on the
// transposed operands and with the long reduction split: S row chunks as one strided-batched
// GEMM (plus the leftover rows) into separate slices of work, summed in a fixed order.
struct TallProduct {
  int64_t n;
  int S;
  int64_t ch, rem;
  DevT<double> work;
  TallProduct(int64_t n_, int64_t max_pq)
      : n(n_), S((int)std::max<int64_t>(1, std::min<int64_t>(32, n_ / 2048))), ch(n_ / S), rem(n_ - ch * S),
        work((size_t)(S + 1) * max_pq) {}
  void operator()(rocblas_handle bh, const double* AT, int p, const double* BT, int q, double* C) {
    const double one = 1.0, zero = 0.0;
    const int64_t slice = (int64_t)p * q;
    rb(rocblas_dgemm_strided_batched(bh, rocblas_operation_none, rocblas_operation_transpose, p, q, (rocblas_int)ch,
                                     &one, AT, p, ch * p, BT, q, ch * q, &zero, work.p, p, slice, S),
       "dgemm_strided_batched A'B");
    int nsl = S;
    if (rem > 0) {
      rb(rocblas_dgemm(bh, rocblas_operation_none, rocblas_operation_transpose, p, q, (rocblas_int)rem, &one,
                       AT + ch * S * p, p, BT + ch * S * q, q, &zero, work.p + (int64_t)S * slice, p),
         "dgemm A'B rest");
      nsl = S + 1;
    }
    k_sum_slices<<<(unsigned)((slice + 255) / 256), 256>>>(work.p, slice, nsl, C);
    HIPCHECK(hipGetLastError());
  }
};

// xPy = V' Py (V: centred markers, n x k, before scaling; :936), then v = (I - QQ') R^-1/2 V
// (:938-939), xSx = ||v||^2 (:942) and xPx = v' H^-1 v with one Lanczos seed over all columns
// at tau1 and L_Solve at shift 0 (:940-941).  V is overwritten by v; dV holds v on exit.
void quad_forms(slemm_geno* h, const MatD& Q, std::vector<float>& V, int k, const float* Py, float tau1,
                float tol_L, int precision, int verbose, DevF& dV, float* xPy, float* xPx, float* xSx) {
  const int64_t n = h->n;
  for (int i = 0; i < k; ++i) {
    double s = 0.0;
    for (int64_t r = 0; r < n; ++r) s += (double)V[(size_t)i * n + r] * Py[r];
    xPy[i] = (float)s;
  }
  for (int i = 0; i < k; ++i)
    for (int64_t r = 0; r < n; ++r) V[(size_t)i * n + r] *= h->rr_host[r];
  project_out(Q, V, n, k);
  for (int i = 0; i < k; ++i) xSx[i] = (float)sq_norm(&V[(size_t)i * n], n);
  upload(dV.p, V);
  DevF dW((size_t)n * k);
  SeedGuard sg;
  abi(slemm_gpu_lseed(h, SLEMM_OP_QUADFORM, nullptr, nullptr, tau1, dV.p, n, k, tol_L, 600, k, precision, verbose,
                      nullptr, &sg.s));
  abi(slemm_gpu_lsolve(sg.s, 0, k, dV.p, n, 0.0f, 1e-6f, dW.p, n, nullptr));
  HIPCHECK(hipDeviceSynchronize());
  std::vector<float> W;
  download(W, dW.p, (size_t)n * k);
  for (int i = 0; i < k; ++i) {
    double s = 0.0;
    for (int64_t r = 0; r < n; ++r) s += (double)V[(size_t)i * n + r] * W[(size_t)i * n + r];
    xPx[i] = (float)s;
  }
}

// glibc's rand(): the TYPE_3 additive lagged-Fibonacci generator behind random(), in the state
// srand(1) (or no srand call at all) leaves it.  Published definition: r_0 = seed,
// r_i = 16807 r_{i-1} mod (2^31 - 1) for i = 1..30, then x_i = x_{i-31} + x_{i-3} (mod 2^32)
// with the 31-word state read by a front index 3 ahead of the rear one, the first 310 outputs
// discarded, and each output the 32-bit sum shifted right by one (31 bits).
class GlibcRand {
 public:
  explicit GlibcRand(uint32_t seed) {
    int64_t r = seed == 0 ? 1 : seed;
    state_[0] = (uint32_t)r;
    for (int i = 1; i < kDeg; ++i) {
      r = (16807 * r) % 2147483647;
      state_[i] = (uint32_t)r;
    }
    front_ = kSep;
    rear_ = 0;
    for (int i = 0; i < 10 * kDeg; ++i) (void)next();
  }
  int32_t next() {
    state_[front_] += state_[rear_];
    const uint32_t out = state_[front_] >> 1;
    front_ = front_ + 1 == kDeg ? 0 : front_ + 1;
    rear_ = rear_ + 1 == kDeg ? 0 : rear_ + 1;
    return (int32_t)out;
  }

 private:
  static constexpr int kDeg = 31, kSep = 3;
  uint32_t state_[kDeg];
  int front_, rear_;
};

}  // namespace

// The --lmm SNP sample of src/reml.cpp:920-925: the indices i with gm_i * n > 20 and
// (2 - gm_i) * n > 20 (float products, as the reference's float gm times int n), shuffled by
// std::random_shuffle -- libstdc++'s loop `for i in 1..N-1: j = rand() % (i + 1); swap(v[i],
// v[j]) if i != j` -- on rand() in its initial state (the reference seeds it nowhere), first
// min(k, N) kept.  A fresh generator per call gives every call the sample of a fresh run.
std::vector<int64_t> reference_lmm_sample(const std::vector<float>& gm, int64_t n, int k) {
  std::vector<int64_t> v;
  const float fn = (float)n;
  for (size_t i = 0; i < gm.size(); ++i)
    if (gm[i] * fn > 20.0f && (2.0f - gm[i]) * fn > 20.0f) v.push_back((int64_t)i);
  GlibcRand rng(1);
  for (size_t i = 1; i < v.size(); ++i) {
    const size_t j = (size_t)((int64_t)rng.next() % (int64_t)(i + 1));
    if (i != j) std::swap(v[i], v[j]);
  }
  if ((int64_t)v.size() > k) v.resize((size_t)k);
  return v;
}

}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_lmm_sample(const float* gm, int64_t m, int64_t n, int num_qf_markers, int64_t* snps,
                                    int* k) {
  return guard([&] {
    REQUIRE(gm && k, "null argument");
    REQUIRE(m >= 0 && n >= 1, "need m >= 0 and n >= 1");
    REQUIRE(num_qf_markers >= 1, "num_qf_markers must be positive");
    const std::vector<int64_t> v = reference_lmm_sample(std::vector<float>(gm, gm + m), n, num_qf_markers);
    *k = (int)v.size();
    if (snps) std::copy(v.begin(), v.end(), snps);
  });
}

extern "C" int slemm_gpu_lmm_gstat(slemm_geno* h, const float* X, int64_t ldx, int c, const float* Py, float h2,
                                   const int64_t* snps, int k, int window_size, float tol_L, int precision,
                                   int verbose, float* xPy, float* xPx, float* xBx, float* xSx) {
  return guard([&] {
    REQUIRE(h && X && Py && xPy && xPx && xBx && xSx, "null argument");
    REQUIRE(c >= 1 && ldx >= h->n, "covariates: need c >= 1 and ldx >= n");
    REQUIRE(k >= 1, "need at least one SNP");
    REQUIRE(h2 > 0 && h2 < 1, "h2 must be in (0, 1)");
    HIPCHECK(hipSetDevice(h->device));
    require_model(h);
    const int64_t n = h->n, m = h->m;
    std::vector<int64_t> sample;
    if (!snps) {  // the reference's own sample (src/reml.cpp:920-925)
      std::vector<float> gm(m);
      HIPCHECK(hipMemcpy(gm.data(), h->gm, sizeof(float) * m, hipMemcpyDeviceToHost));
      sample = reference_lmm_sample(gm, n, k);
      REQUIRE((int)sample.size() == k,
              "fewer SNPs with MAC > 20 than requested; slemm_gpu_lmm_sample gives the sample size");
      snps = sample.data();
    }
    REQUIRE(window_size >= 2 && window_size % 2 == 0, "window_size must be even and positive");
    REQUIRE(window_size < m, "--window_size >= # of SNPs kept for --lmm.");
    for (int i = 0; i < k; ++i) REQUIRE(snps[i] >= 0 && snps[i] < m, "SNP index out of range");
    const float tau1 = (1 - h2) / h2;  // src/reml.cpp:859
    MatD Q;
    covariate_basis(h, X, ldx, c, Q);

    // V = centred dosages of the sampled SNPs, then xPy, xSx, xPx
    DevT<int64_t> dsnps(k);
    HIPCHECK(hipMemcpy(dsnps.p, snps, sizeof(int64_t) * k, hipMemcpyHostToDevice));
    DevF dV((size_t)n * k);
    k_extract_centered<<<dim3((unsigned)((n + 255) / 256), (unsigned)k), 256>>>(h->gs, h->ld_gs, dsnps.p, n, h->gm,
                                                                                dV.p, n);
    HIPCHECK(hipGetLastError());
    std::vector<float> V;
    download(V, dV.p, (size_t)n * k);
    quad_forms(h, Q, V, k, Py, tau1, tol_L, precision, verbose, dV, xPy, xPx, xSx);

    // xBx over the window around each SNP (:944-968), FP64 on the device
    const int w = window_size;
    Blas blas;
    DevT<double> zw((size_t)n * w), G((size_t)w * w), qd((size_t)n * std::max(c, 1)), qz((size_t)c * w),
        z((size_t)w);
    DevT<rocblas_int> info(1);
    TallProduct tall(n, (int64_t)w * w);
    DevT<double> zwT((size_t)n * w);
    {
      std::vector<double> qh((size_t)n * c);
      for (size_t e = 0; e < qh.size(); ++e) qh[e] = (double)(float)Q.a[e];  // the handle's FP32 basis
      HIPCHECK(hipMemcpy(qd.p, qh.data(), sizeof(double) * qh.size(), hipMemcpyHostToDevice));
    }
    const double one = 1.0, minus_one = -1.0;
    for (int i = 0; i < k; ++i) {
      const int64_t j = snps[i];
      int64_t wds;
      if (j < w / 2) wds = 0;
      else if (j > m - w / 2) wds = m - w;
      else wds = j - w / 2;
      k_extract_window<<<dim3((unsigned)((n + 255) / 256), (unsigned)w), 256>>>(h->gs, h->ld_gs, wds, n, h->gm,
                                                                               h->rr, zw.p, n);
      HIPCHECK(hipGetLastError());
      // Zw <- Zw - Q (Q' Zw)
      k_qt_a<<<(unsigned)w, 256>>>(qd.p, c, zw.p, n, qz.p);
      HIPCHECK(hipGetLastError());
      rb(rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, w, c, &minus_one, qd.p,
                       (rocblas_int)n, qz.p, c, &one, zw.p, (rocblas_int)n), "dgemm Zw - Q(Q'Zw)");
      // G = Zw'Zw (from the transposed window) + tau1 diag(1/gw); z = Zw'v with v = Zw's column jj
      // (the SNP's own column) is G's column jj before the ridge
      const int64_t jj = j - wds;
      transpose_f64(zw.p, n, w, zwT.p);
      tall(blas.h, zwT.p, w, zwT.p, w, G.p);
      HIPCHECK(hipMemcpy(z.p, G.p + jj * w, sizeof(double) * w, hipMemcpyDeviceToDevice));
      k_add_ridge<<<(w + 255) / 256, 256>>>(G.p, w, w, h->gw, wds, (double)tau1);
      HIPCHECK(hipGetLastError());
      std::vector<double> zh(w);
      HIPCHECK(hipMemcpy(zh.data(), z.p, sizeof(double) * w, hipMemcpyDeviceToHost));
      // G^-1 z by Cholesky (Eigen's llt in the reference)
      rb(rocsolver_dpotrf(blas.h, rocblas_fill_lower, w, G.p, w, info.p), "dpotrf");
      rocblas_int inf = 0;
      HIPCHECK(hipMemcpy(&inf, info.p, sizeof(rocblas_int), hipMemcpyDeviceToHost));
      if (inf != 0) fail(SLEMM_ERR_NUMERIC, "window system of the LMM quadratic form is not positive definite");
      rb(rocsolver_dpotrs(blas.h, rocblas_fill_lower, w, 1, G.p, w, z.p, w), "dpotrs");
      std::vector<double> sol(w);
      HIPCHECK(hipMemcpy(sol.data(), z.p, sizeof(double) * w, hipMemcpyDeviceToHost));
      double q = 0.0;
      for (int e = 0; e < w; ++e) q += zh[e] * sol[e];
      xBx[i] = (float)((zh[jj] - q) / (double)tau1);  // zh[jj] = v'v
      if (verbose && (i + 1) % 10 == 0) printf("  gstat: %d of %d SNPs\n", i + 1, k);
    }
  });
}

// --fake_geno branch of --lmm (src/reml.cpp:886-918): k simulated markers drawn from the fit's
// random stream after the MC samples of LMC_REML (:795-809, replayed here: mt19937(seed+54321),
// n_MC*m then n_MC*n normal_distribution<float> draws), each with allele frequency
// uniform_real_distribution<float>(0.01, 0.99) and n binomial_distribution<int>(2, p) dosages;
// centred by their float mean, then the same xPy, xPx, xSx as for real SNPs (no window term).
extern "C" int slemm_gpu_lmm_gstat_fake(slemm_geno* h, const float* X, int64_t ldx, int c, const float* Py, float h2,
                                        int k, int rng_seed, int n_MC, float tol_L, int precision, int verbose,
                                        float* xPy, float* xPx, float* xSx) {
  return guard([&] {
    REQUIRE(h && X && Py && xPy && xPx && xSx, "null argument");
    REQUIRE(c >= 1 && ldx >= h->n, "covariates: need c >= 1 and ldx >= n");
    REQUIRE(k >= 1, "need at least one marker");
    REQUIRE(n_MC >= 1, "n_MC must be positive");
    REQUIRE(h2 > 0 && h2 < 1, "h2 must be in (0, 1)");
    HIPCHECK(hipSetDevice(h->device));
    require_model(h);
    const int64_t n = h->n, m = h->m;
    const float tau1 = (1 - h2) / h2;
    MatD Q;
    covariate_basis(h, X, ldx, c, Q);
    std::vector<float> V((size_t)n * k);
    {
      std::mt19937 rng(rng_seed + 54321);
      std::normal_distribution<float> rnorm(0.0, 1.0);
      for (int64_t e = 0; e < (int64_t)n_MC * m; ++e) (void)rnorm(rng);
      for (int64_t e = 0; e < (int64_t)n_MC * n; ++e) (void)rnorm(rng);
      std::uniform_real_distribution<float> runif(0.01, 0.99);
      for (int j = 0; j < k; ++j) {
        const float a1_freq = runif(rng);
        std::binomial_distribution<int> rbiom(2, a1_freq);
        for (int64_t i = 0; i < n; ++i) V[(size_t)j * n + i] = (float)rbiom(rng);
      }
    }
    // column means as one float divide of the dosage sum (Eigen's colwise mean); the sum is exact
    // in any order up to n = 2^23, beyond which the reference's float summation order decides
    // its last bit
    for (int j = 0; j < k; ++j) {
      float* v = &V[(size_t)j * n];
      int64_t s = 0;
      for (int64_t i = 0; i < n; ++i) s += (int64_t)v[i];
      const float mu = (float)s / (float)n;
      for (int64_t i = 0; i < n; ++i) v[i] -= mu;
    }
    DevF dV((size_t)n * k);
    quad_forms(h, Q, V, k, Py, tau1, tol_L, precision, verbose, dV, xPy, xPx, xSx);
  });
}

namespace slemm {

// Dosages (0, 1, 2 as float) of the SNPs snps, n x k column-major on the host.
void extract_dosage_cols(slemm_geno* h, const std::vector<int64_t>& snps, std::vector<float>& out) {
  const int64_t n = h->n, k = (int64_t)snps.size();
  out.assign((size_t)n * k, 0.f);
  if (k == 0) return;
  DevT<int64_t> ds(k);
  HIPCHECK(hipMemcpy(ds.p, snps.data(), sizeof(int64_t) * k, hipMemcpyHostToDevice));
  DevT<float> zero((size_t)h->m);
  HIPCHECK(hipMemset(zero.p, 0, sizeof(float) * h->m));
  DevF dout((size_t)n * k);
  k_extract_centered<<<dim3((unsigned)((n + 255) / 256), (unsigned)k), 256>>>(h->gs, h->ld_gs, ds.p, n, zero.p,
                                                                              dout.p, n);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(out.data(), dout.p, sizeof(float) * n * k, hipMemcpyDeviceToHost));
}

}  // namespace slemm

// ------------------------------------------------------------------------------------------
// Windowed score tests of slemm_gwa step 2 (bin/slemm_gwa.py:322-431): blocks of block_size
// model SNPs overlapping by half, each with H = Zb'Zb + model_snp_ct tau diag((2 - p) p / 2) (p the
// SNP's mean dosage), and every tested SNP of the block's range scored as
// xx = (x'x - z'H^-1 z) / tau with z = Zb'x; x centred, scaled by R^-1/2 and projected off Q.

namespace slemm {
namespace {

// out[i + c n] = dosage(snps[c], i) - gm[snps[c]] in FP64 (the FP32 mean, as the script's
// float32 vec_x -= sum / n)
__global__ void k_extract_f64(const uint8_t* __restrict__ gs, int64_t ld_gs, const int64_t* __restrict__ snps,
                              int64_t n, const float* __restrict__ gm, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  const int64_t j = snps[c];
  const uint32_t code = (gs[j * ld_gs + (i >> 2)] >> (2 * (i & 3))) & 3u;
  out[i + (int64_t)c * n] = (double)((float)((code & 1u) + (code >> 1)) - gm[j]);
}

__global__ void k_scale_rows(double* __restrict__ a, int64_t n, int t, const float* __restrict__ rr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i + (int64_t)blockIdx.y * n] *= (double)rr[i];
}

// xx_c = (sum_i x_ic^2 - sum_k z_kc s_kc) / tau for the t columns; one block per column
__global__ void k_window_xx(const double* __restrict__ X, int64_t n, const double* __restrict__ Z,
                            const double* __restrict__ S, int b, double tau, double* __restrict__ xx) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const double v = X[i + (int64_t)c * n];
    s += v * v;
  }
  for (int k = threadIdx.x; k < b; k += 256) s -= Z[k + (int64_t)c * b] * S[k + (int64_t)c * b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) xx[c] = red[0] / tau;
}

// H_kk += model_snp_ct tau rvar_k
__global__ void k_add_diag(double* __restrict__ H, int b, const double* __restrict__ d) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < b) H[k + (int64_t)k * b] += d[k];
}

}  // namespace
}  // namespace slemm

extern "C" int slemm_gpu_gwa_window(slemm_geno* h, const float* X, int64_t ldx, int c, const float* y,
                                    const int64_t* model_snps, int K, int block_size, int64_t test_begin,
                                    int64_t test_end, float tau, int64_t model_snp_ct, float mac, float* xy,
                                    float* xx) {
  return guard([&] {
    REQUIRE(h && X && y && model_snps && xy && xx, "null argument");
    REQUIRE(c >= 1 && ldx >= h->n, "covariates: need c >= 1 and ldx >= n");
    REQUIRE(K >= 2, "need at least two model SNPs");
    REQUIRE(tau > 0, "tau must be positive");
    HIPCHECK(hipSetDevice(h->device));
    const int64_t n = h->n, m = h->m;
    REQUIRE(test_begin >= 0 && test_end >= test_begin && test_end < m, "tested SNP range out of bounds");
    for (int i = 0; i < K; ++i) {
      REQUIRE(model_snps[i] >= 0 && model_snps[i] < m, "model SNP index out of range");
      REQUIRE(i == 0 || model_snps[i] > model_snps[i - 1], "model SNPs must be strictly increasing");
    }
    ensure_counts(h, nullptr);
    std::vector<int32_t> pc(m);
    std::vector<float> gmh(m);
    HIPCHECK(hipMemcpy(pc.data(), h->popcnt, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(gmh.data(), h->gm, sizeof(float) * m, hipMemcpyDeviceToHost));
    // the model's R^-1/2 and Q of the covariates
    if (h->rr_host.empty()) abi(slemm_gpu_set_model(h, nullptr, nullptr, nullptr, n, 0));
    MatD Q;
    covariate_basis(h, X, ldx, c, Q);

    // blocks of model SNPs and the tested range of each (bin/slemm_gwa.py:322-345)
    int b = block_size;
    if (b > K) b = K / 2 * 2;
    REQUIRE(b >= 2, "block size must be at least 2");
    const int step = b / 2, ext = step / 2;
    std::vector<std::vector<int64_t>> blocks;
    const int nb = (K - step) / step;
    for (int i = 0; i < nb; ++i) blocks.emplace_back(model_snps + (size_t)i * step, model_snps + (size_t)i * step + b);
    if ((K - step) % step != 0) blocks.emplace_back(model_snps + (K - b), model_snps + K);
    const int nblk = (int)blocks.size();
    std::vector<int64_t> lo(nblk), hi(nblk);
    lo[0] = test_begin;
    hi[0] = blocks[0][b - ext];
    for (int i = 1; i < nblk; ++i) {
      lo[i] = hi[i - 1];
      hi[i] = blocks[i][b - ext];
    }
    hi[nblk - 1] = test_end + 1;
    for (int i = 0; i < nblk; ++i) {  // outputs cover [test_begin, test_end] only
      lo[i] = std::max(lo[i], test_begin);
      hi[i] = std::min(hi[i], test_end + 1);
    }

    Blas blas;
    const double one = 1.0, zero = 0.0, minus_one = -1.0;
    const int cq = Q.c;
    DevT<double> qd((size_t)n * cq), Zb((size_t)n * b), H((size_t)b * b), qz((size_t)cq * std::max<int64_t>(b, 4096)),
        dy((size_t)n), diag((size_t)b);
    {
      std::vector<double> qh((size_t)n * cq), yh(n);
      for (size_t e = 0; e < qh.size(); ++e) qh[e] = (double)(float)Q.a[e];
      for (int64_t i = 0; i < n; ++i) yh[i] = (double)y[i];
      HIPCHECK(hipMemcpy(qd.p, qh.data(), sizeof(double) * qh.size(), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(dy.p, yh.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    const int64_t TCH = 1024;  // tested SNPs per batch (a block's range holds about b / 2)
    DevT<int64_t> didx((size_t)std::max<int64_t>(b, TCH));
    DevT<double> Xt((size_t)n * TCH), Zx((size_t)b * TCH), Sx((size_t)b * TCH), dxy(TCH), dxx(TCH);
    DevT<rocblas_int> info(1);
    TallProduct tall(n, (int64_t)b * std::max<int64_t>(b, TCH));
    DevT<double> ZbT((size_t)n * b), XtT((size_t)n * TCH);
    // centre, scale by R^-1/2 and project out Q (in place on an n x t FP64 block)
    auto prepare = [&](double* A, int t) {
      k_scale_rows<<<dim3((unsigned)((n + 255) / 256), (unsigned)t), 256>>>(A, n, t, h->rr);
      HIPCHECK(hipGetLastError());
      k_qt_a<<<(unsigned)t, 256>>>(qd.p, cq, A, n, qz.p);
      HIPCHECK(hipGetLastError());
      rb(rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, t, cq, &minus_one,
                       qd.p, (rocblas_int)n, qz.p, cq, &one, A, (rocblas_int)n), "dgemm A - Q(Q'A)");
    };
    for (int blk = 0; blk < nblk; ++blk) {
      if (lo[blk] >= hi[blk]) continue;  // no tested SNP in this block's range
      // Zb and H = Zb'Zb + model_snp_ct tau rvars (:363-383, :387)
      HIPCHECK(hipMemcpy(didx.p, blocks[blk].data(), sizeof(int64_t) * b, hipMemcpyHostToDevice));
      k_extract_f64<<<dim3((unsigned)((n + 255) / 256), (unsigned)b), 256>>>(h->gs, h->ld_gs, didx.p, n, h->gm, Zb.p);
      HIPCHECK(hipGetLastError());
      prepare(Zb.p, b);
      transpose_f64(Zb.p, n, b, ZbT.p);
      tall(blas.h, ZbT.p, b, ZbT.p, b, H.p);
      std::vector<double> dh(b);
      for (int k = 0; k < b; ++k) {
        const float mu = gmh[blocks[blk][k]];
        const float rv = (2.0f - mu) * mu / 2.0f;
        dh[k] = (double)model_snp_ct * (double)tau * (double)rv;
      }
      HIPCHECK(hipMemcpy(diag.p, dh.data(), sizeof(double) * b, hipMemcpyHostToDevice));
      k_add_diag<<<(b + 255) / 256, 256>>>(H.p, b, diag.p);
      rb(rocsolver_dpotrf(blas.h, rocblas_fill_lower, b, H.p, b, info.p), "dpotrf");
      rocblas_int inf = 0;
      HIPCHECK(hipMemcpy(&inf, info.p, sizeof(rocblas_int), hipMemcpyDeviceToHost));
      if (inf != 0) fail(SLEMM_ERR_NUMERIC, "block system of the windowed scan is not positive definite");
      // tested SNPs of the block's range, in batches (:388-407)
      for (int64_t j0 = lo[blk]; j0 < hi[blk]; j0 += TCH) {
        const int t = (int)std::min<int64_t>(TCH, hi[blk] - j0);
        std::vector<int64_t> idx(t);
        for (int q = 0; q < t; ++q) idx[q] = j0 + q;
        HIPCHECK(hipMemcpy(didx.p, idx.data(), sizeof(int64_t) * t, hipMemcpyHostToDevice));
        k_extract_f64<<<dim3((unsigned)((n + 255) / 256), (unsigned)t), 256>>>(h->gs, h->ld_gs, didx.p, n, h->gm,
                                                                               Xt.p);
        HIPCHECK(hipGetLastError());
        rb(rocblas_dgemv(blas.h, rocblas_operation_transpose, (rocblas_int)n, t, &one, Xt.p, (rocblas_int)n, dy.p, 1,
                         &zero, dxy.p, 1), "dgemv X'y");
        prepare(Xt.p, t);
        transpose_f64(Xt.p, n, t, XtT.p);
        tall(blas.h, ZbT.p, b, XtT.p, t, Zx.p);
        HIPCHECK(hipMemcpy(Sx.p, Zx.p, sizeof(double) * b * t, hipMemcpyDeviceToDevice));
        rb(rocsolver_dpotrs(blas.h, rocblas_fill_lower, b, t, H.p, b, Sx.p, b), "dpotrs");
        k_window_xx<<<t, 256>>>(Xt.p, n, Zx.p, Sx.p, b, (double)tau, dxx.p);
        HIPCHECK(hipGetLastError());
        std::vector<double> hxy(t), hxx(t);
        HIPCHECK(hipMemcpy(hxy.data(), dxy.p, sizeof(double) * t, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(hxx.data(), dxx.p, sizeof(double) * t, hipMemcpyDeviceToHost));
        for (int q = 0; q < t; ++q) {
          const int64_t j = j0 + q, o = j - test_begin;
          const float s = (float)pc[j];
          if (s <= mac || 2.0f * (float)n - s <= mac) {  // MAC filter (:396-400)
            xy[o] = xx[o] = __builtin_nanf("");
          } else {
            xy[o] = (float)hxy[q];
            xx[o] = (float)hxx[q];
          }
        }
      }
    }
  });
}
