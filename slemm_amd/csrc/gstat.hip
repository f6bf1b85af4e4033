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

}  // namespace
}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_lmm_gstat(slemm_geno* h, const float* X, int64_t ldx, int c, const float* Py, float h2,
                                   const int64_t* snps, int k, int window_size, float tol_L, int precision,
                                   int verbose, float* xPy, float* xPx, float* xBx, float* xSx) {
  return guard([&] {
    REQUIRE(h && X && Py && snps && xPy && xPx && xBx && xSx, "null argument");
    REQUIRE(c >= 1 && ldx >= h->n, "covariates: need c >= 1 and ldx >= n");
    REQUIRE(k >= 1, "need at least one SNP");
    REQUIRE(h2 > 0 && h2 < 1, "h2 must be in (0, 1)");
    HIPCHECK(hipSetDevice(h->device));
    require_model(h);
    const int64_t n = h->n, m = h->m;
    REQUIRE(window_size >= 2 && window_size % 2 == 0, "window_size must be even and positive");
    REQUIRE(window_size < m, "--window_size >= # of SNPs kept for --lmm.");
    for (int i = 0; i < k; ++i) REQUIRE(snps[i] >= 0 && snps[i] < m, "SNP index out of range");
    const float tau1 = (1 - h2) / h2;  // src/reml.cpp:859
    MatD Q;
    covariate_basis(h, X, ldx, c, Q);

    // V = centred dosages of the sampled SNPs; xPy = V' Py before scaling (:936)
    DevT<int64_t> dsnps(k);
    HIPCHECK(hipMemcpy(dsnps.p, snps, sizeof(int64_t) * k, hipMemcpyHostToDevice));
    DevF dV((size_t)n * k), dW((size_t)n * k);
    k_extract_centered<<<dim3((unsigned)((n + 255) / 256), (unsigned)k), 256>>>(h->gs, h->ld_gs, dsnps.p, n, h->gm,
                                                                                dV.p, n);
    HIPCHECK(hipGetLastError());
    std::vector<float> V;
    download(V, dV.p, (size_t)n * k);
    for (int i = 0; i < k; ++i) {
      double s = 0.0;
      for (int64_t r = 0; r < n; ++r) s += (double)V[(size_t)i * n + r] * Py[r];
      xPy[i] = (float)s;
    }
    // v = (I - QQ') R^-1/2 x (:938-939); xSx = ||v||^2 (:942)
    for (int i = 0; i < k; ++i)
      for (int64_t r = 0; r < n; ++r) V[(size_t)i * n + r] *= h->rr_host[r];
    project_out(Q, V, n, k);
    for (int i = 0; i < k; ++i) xSx[i] = (float)sq_norm(&V[(size_t)i * n], n);
    upload(dV.p, V);

    // xPx = v' H^-1 v: one seed over all sampled SNPs at tau1, L_Solve at shift 0 (:940-941)
    {
      SeedGuard sg;
      abi(slemm_gpu_lseed(h, SLEMM_OP_QUADFORM, nullptr, nullptr, tau1, dV.p, n, k, tol_L, 600, k, precision,
                          verbose, nullptr, &sg.s));
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

    // xBx over the window around each SNP (:944-968), FP64 on the device
    const int w = window_size;
    Blas blas;
    DevT<double> zw((size_t)n * w), G((size_t)w * w), qd((size_t)n * std::max(c, 1)), qz((size_t)c * w),
        z((size_t)w);
    DevT<rocblas_int> info(1);
    {
      std::vector<double> qh((size_t)n * c);
      for (size_t e = 0; e < qh.size(); ++e) qh[e] = (double)(float)Q.a[e];  // the handle's FP32 basis
      HIPCHECK(hipMemcpy(qd.p, qh.data(), sizeof(double) * qh.size(), hipMemcpyHostToDevice));
    }
    const double one = 1.0, zero = 0.0, minus_one = -1.0;
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
      rb(rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none, c, w, (rocblas_int)n, &one, qd.p,
                       (rocblas_int)n, zw.p, (rocblas_int)n, &zero, qz.p, c), "dgemm Q'Zw");
      rb(rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, w, c, &minus_one, qd.p,
                       (rocblas_int)n, qz.p, c, &one, zw.p, (rocblas_int)n), "dgemm Zw - Q(Q'Zw)");
      // G = Zw' Zw + tau1 diag(1/gw); z = Zw' v with v the window's column of SNP j
      const int64_t jj = j - wds;
      rb(rocblas_dsyrk(blas.h, rocblas_fill_lower, rocblas_operation_transpose, w, (rocblas_int)n, &one, zw.p,
                       (rocblas_int)n, &zero, G.p, w), "dsyrk");
      k_add_ridge<<<(w + 255) / 256, 256>>>(G.p, w, w, h->gw, wds, (double)tau1);
      HIPCHECK(hipGetLastError());
      rb(rocblas_dgemv(blas.h, rocblas_operation_transpose, (rocblas_int)n, w, &one, zw.p, (rocblas_int)n,
                       zw.p + jj * n, 1, &zero, z.p, 1), "dgemv Zw'v");
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
