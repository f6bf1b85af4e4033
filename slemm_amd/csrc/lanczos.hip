// Krylov engine on the device: batched Lanczos/CG seed systems (L_Seed, src/reml.cpp:246-339),
// shifted-CG replay (L_Solve, :369-431) and the SLQ log-determinant (:440-451, :637-646).
//
// Vectors stay in HBM in FP32 (the operator's input/output format); every dot product, norm
// and recurrence scalar is FP64.  The per-step host work is O(t) scalars: one small
// device->host copy of delta/beta per step drives the CG recurrences and stopping test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "blockprod.hpp"
#include "driver.hpp"
#include "lanczos.hpp"
#include "pipeline.hpp"
#include "tridiag.hpp"

struct slemm_lseed {
  // The device is kept by value: the seed's stored basis and coefficients outlive the genotype
  // handle it was built from (a caller may destroy that handle first), so nothing here may
  // reach through a slemm_geno pointer after the seed has been built.
  int device = 0;
  int t = 0;
  int64_t n = 0;
  int keep_cols = 0;
  int max_steps = 0;
  bool converged = true;
  std::vector<int> steps;        // per column
  std::vector<double> beta;      // t x (max_steps+1), column-major (row k = column k)
  std::vector<double> delta;     // t x max_steps
  std::vector<float*> U;         // kept basis: U[j] is n x keep_cols (ld n), j = 0..max_steps
  ~slemm_lseed() {
    for (float* p : U)
      if (p) (void)hipFree(p);
  }
};

namespace slemm {

constexpr int kRowsPerPart = 4096;

// The per-column passes below read whole float4 groups: a column starts at k ld with ld = n_pad, a
// multiple of kPad, and a part at a multiple of kRowsPerPart, so every group is 16-byte aligned.
// Each thread requests its four groups of a part (one per 256-thread sweep) before using any, so a
// wave keeps four 1 KB loads in flight per array; the rows past the last whole group of a column
// (n not a multiple of 4) are taken one by one.
static_assert(kPad % 4 == 0 && kRowsPerPart % 1024 == 0, "float4 groups of the Lanczos passes");
constexpr int kLzSweeps = 4;  // float4 groups per thread per pass over 1024 groups

// tmp_k -= b_k * Uprev_k (if Uprev); partial dot(Ucur_k, tmp_k).  Inactive columns untouched.
__global__ __launch_bounds__(256) void k_lz_a(const float* __restrict__ Uprev, const float* __restrict__ Ucur,
                                              float* __restrict__ tmp, int64_t ld, int64_t n,
                                              const double* __restrict__ b, const int* __restrict__ active,
                                              double* __restrict__ part, int t) {
  __shared__ double red[256];
  const int k = blockIdx.y;
  const int64_t i0 = (int64_t)blockIdx.x * kRowsPerPart;
  const int64_t i1 = min(n, i0 + kRowsPerPart);
  double s = 0.0;
  if (active[k]) {
    const float bk = Uprev ? (float)b[k] : 0.f;
    const int64_t c0 = (int64_t)k * ld + i0;
    const int64_t nq = (i1 - i0) >> 2;
    float4* t4 = reinterpret_cast<float4*>(tmp + c0);
    const float4* u4 = reinterpret_cast<const float4*>(Ucur + c0);
    const float4* p4 = Uprev ? reinterpret_cast<const float4*>(Uprev + c0) : nullptr;
    for (int64_t q0 = threadIdx.x; q0 < nq; q0 += 256 * kLzSweeps) {
      float4 v[kLzSweeps], u[kLzSweeps], w[kLzSweeps];
#pragma unroll
      for (int r = 0; r < kLzSweeps; ++r) {
        const int64_t q = q0 + 256 * r;
        if (q < nq) {
          v[r] = t4[q];
          u[r] = u4[q];
          if (p4) w[r] = p4[q];
        }
      }
#pragma unroll
      for (int r = 0; r < kLzSweeps; ++r) {
        const int64_t q = q0 + 256 * r;
        if (q < nq) {
          float4 x = v[r];
          if (p4) {
            x.x -= w[r].x * bk;
            x.y -= w[r].y * bk;
            x.z -= w[r].z * bk;
            x.w -= w[r].w * bk;
          }
          t4[q] = x;
          s += (double)u[r].x * (double)x.x;
          s += (double)u[r].y * (double)x.y;
          s += (double)u[r].z * (double)x.z;
          s += (double)u[r].w * (double)x.w;
        }
      }
    }
    for (int64_t i = i0 + 4 * nq + threadIdx.x; i < i1; i += 256) {
      const int64_t o = i + (int64_t)k * ld;
      float v = tmp[o];
      if (Uprev) v -= Uprev[o] * bk;
      tmp[o] = v;
      s += (double)Ucur[o] * (double)v;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.x * t + k] = red[0];
}

// tmp_k -= d_k * Ucur_k; partial ||tmp_k||^2.
__global__ __launch_bounds__(256) void k_lz_b(const float* __restrict__ Ucur, float* __restrict__ tmp, int64_t ld,
                                              int64_t n, const double* __restrict__ d, const int* __restrict__ active,
                                              double* __restrict__ part, int t) {
  __shared__ double red[256];
  const int k = blockIdx.y;
  const int64_t i0 = (int64_t)blockIdx.x * kRowsPerPart;
  const int64_t i1 = min(n, i0 + kRowsPerPart);
  double s = 0.0;
  if (active[k]) {
    const float dk = (float)d[k];
    const int64_t c0 = (int64_t)k * ld + i0;
    const int64_t nq = (i1 - i0) >> 2;
    float4* t4 = reinterpret_cast<float4*>(tmp + c0);
    const float4* u4 = reinterpret_cast<const float4*>(Ucur + c0);
    for (int64_t q0 = threadIdx.x; q0 < nq; q0 += 256 * kLzSweeps) {
      float4 v[kLzSweeps], u[kLzSweeps];
#pragma unroll
      for (int r = 0; r < kLzSweeps; ++r) {
        const int64_t q = q0 + 256 * r;
        if (q < nq) {
          v[r] = t4[q];
          u[r] = u4[q];
        }
      }
#pragma unroll
      for (int r = 0; r < kLzSweeps; ++r) {
        const int64_t q = q0 + 256 * r;
        if (q < nq) {
          float4 x;
          x.x = v[r].x - u[r].x * dk;
          x.y = v[r].y - u[r].y * dk;
          x.z = v[r].z - u[r].z * dk;
          x.w = v[r].w - u[r].w * dk;
          t4[q] = x;
          s += (double)x.x * (double)x.x;
          s += (double)x.y * (double)x.y;
          s += (double)x.z * (double)x.z;
          s += (double)x.w * (double)x.w;
        }
      }
    }
    for (int64_t i = i0 + 4 * nq + threadIdx.x; i < i1; i += 256) {
      const int64_t o = i + (int64_t)k * ld;
      const float v = tmp[o] - Ucur[o] * dk;
      tmp[o] = v;
      s += (double)v * (double)v;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.x * t + k] = red[0];
}

// res[k] = the sum of column k's partials (double), optionally its square root.  One wave per
// column: lane l adds parts l, l + 64, ..., then a butterfly over the wave (at a million rows a
// column has 245 parts, too long a chain of dependent loads for one thread).
__global__ __launch_bounds__(64) void k_lz_sum(const double* __restrict__ part, int64_t nparts, int t,
                                               double* __restrict__ res, int take_sqrt) {
  const int k = blockIdx.x;
  double s = 0.0;
  for (int64_t p = threadIdx.x; p < nparts; p += 64) s += part[p * t + k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (threadIdx.x == 0) res[k] = take_sqrt ? sqrt(s) : s;
}

// Unext_k = tmp_k / b_k (active columns), a division in FP32 like the reference's; thread q takes
// rows 4q .. 4q + 3 (those below n)
__global__ void k_lz_c(const float* __restrict__ tmp, float* __restrict__ Unext, int64_t ld, int64_t n,
                       const double* __restrict__ b, const int* __restrict__ active) {
  const int k = blockIdx.y;
  const int64_t i = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n || !active[k]) return;
  const int64_t o = i + (int64_t)k * ld;
  const float bk = (float)b[k];
  if (i + 4 <= n) {
    const float4 v = *reinterpret_cast<const float4*>(tmp + o);
    *reinterpret_cast<float4*>(Unext + o) = make_float4(v.x / bk, v.y / bk, v.z / bk, v.w / bk);
  } else {
    for (int64_t e = 0; i + e < n; ++e) Unext[o + e] = tmp[o + e] / bk;
  }
}

// copies n x t (ldb) into n x t (ld) and scales column k by 1/bnorm_k
__global__ void k_lz_init(const float* __restrict__ B, int64_t ldb, float* __restrict__ U0, int64_t ld, int64_t n,
                          const double* __restrict__ bnorm) {
  const int k = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  U0[i + (int64_t)k * ld] = B[i + (int64_t)k * ldb] / (float)bnorm[k];
}

// L_Solve replay: for j < J: X += gamma_j P; P = rho_{j+1} U_{j+1} + omega_j P  (per column)
// coefficient arrays are (J x t2) row-major: c[j*t2 + k].
__global__ void k_lsolve(const float* const* __restrict__ Ulist, int64_t ldu, int col0, const float* __restrict__ B,
                         int64_t ldb, float* __restrict__ X, int64_t ldx, int64_t n, int t2, int J,
                         const double* __restrict__ gam, const double* __restrict__ om,
                         const double* __restrict__ rho, const int* __restrict__ Jk) {
  const int k = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = 0.0;
  double p = (double)B[i + (int64_t)k * ldb];
  const int Jc = Jk[k];
  for (int j = 0; j < Jc; ++j) {
    x += gam[(int64_t)j * t2 + k] * p;
    const double u = (double)Ulist[j + 1][i + (int64_t)(col0 + k) * ldu];
    p = rho[(int64_t)(j + 1) * t2 + k] * u + om[(int64_t)j * t2 + k] * p;
  }
  X[i + (int64_t)k * ldx] = (float)x;
}

}  // namespace slemm

using namespace slemm;

namespace slemm {

bool lanczos_graphs() {
  const char* e = std::getenv("SLEMM_LANCZOS_GRAPH");
  return !(e && e[0] == '0');
}

slemm_lseed* lseed_run(slemm_geno* h, const LanczosOp& apply_op, const int* col_group, const float* B, int64_t ldb,
                       int t, float tol, int maxit, int keep_cols, int verbose, hipStream_t st, bool use_graph,
                       const ActiveFn& on_active) {
    REQUIRE(h && B, "null argument");
    REQUIRE(t > 0, "number of columns must be positive");
    REQUIRE(ldb >= h->n, "ldb < n");
    REQUIRE(maxit >= 2, "maxit must be at least 2");
    REQUIRE(keep_cols >= 0 && keep_cols <= t, "keep_cols out of [0, t]");
    REQUIRE(tol > 0, "tolerance must be positive");
    HIPCHECK(hipSetDevice(h->device));
    const int64_t n = h->n, ld = h->n_pad;
    REQUIRE(ld % 4 == 0, "the Lanczos passes take rows in float4 groups");

    std::vector<int> group(t, 0);
    int ngroups = 1;
    if (col_group) {
      ngroups = 0;
      for (int k = 0; k < t; ++k) {
        REQUIRE(col_group[k] >= 0 && col_group[k] < t, "col_group ids must be in [0, t)");
        group[k] = col_group[k];
        ngroups = std::max(ngroups, group[k] + 1);
      }
    }

    auto s = std::make_unique<slemm_lseed>();
    bool eager_next = false;  // the operator's launches changed: run one step eagerly
    s->device = h->device;
    s->t = t;
    s->n = n;
    s->keep_cols = keep_cols;
    s->steps.assign(t, 0);

    DevBuf bufA, bufB, bufC, tmp, part, dvec, dact;
    bufA.reserve(sizeof(float) * ld * t);
    bufB.reserve(sizeof(float) * ld * t);
    bufC.reserve(sizeof(float) * ld * t);
    tmp.reserve(sizeof(float) * ld * t);
    const int64_t nparts = (n + kRowsPerPart - 1) / kRowsPerPart;
    part.reserve(sizeof(double) * nparts * t);
    dvec.reserve(sizeof(double) * 2 * t);  // [0,t): delta or beta scratch, [t,2t): beta_j
    dact.reserve(sizeof(int) * t);
    double* d_dvec = dvec.as<double>();

    std::vector<int> active(t, 1);
    HIPCHECK(hipMemcpyAsync(dact.p, active.data(), sizeof(int) * t, hipMemcpyHostToDevice, st));
    // frozen columns keep whatever their buffers hold; start from zeros so it stays finite
    HIPCHECK(hipMemsetAsync(bufB.p, 0, sizeof(float) * ld * t, st));
    HIPCHECK(hipMemsetAsync(bufC.p, 0, sizeof(float) * ld * t, st));
    HIPCHECK(hipMemsetAsync(tmp.p, 0, sizeof(float) * ld * t, st));

    // ||b_k|| and U_0 = B / ||b||
    std::vector<double> bnorm(t);
    {
      // ||b||^2 via k_lz_a with Uprev = null and Ucur = tmp = a copy of B
      DevBuf bcopy;
      bcopy.reserve(sizeof(float) * ld * t);
      HIPCHECK(hipMemcpy2DAsync(bcopy.p, sizeof(float) * ld, B, sizeof(float) * ldb, sizeof(float) * n, t,
                                hipMemcpyDeviceToDevice, st));
      k_lz_a<<<dim3((unsigned)nparts, (unsigned)t), 256, 0, st>>>(nullptr, bcopy.as<float>(), bcopy.as<float>(), ld,
                                                                 n, d_dvec, dact.as<int>(), part.as<double>(), t);
      HIPCHECK(hipGetLastError());
      k_lz_sum<<<t, 64, 0, st>>>(part.as<double>(), nparts, t, d_dvec, 1);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(bnorm.data(), d_dvec, sizeof(double) * t, hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      for (int k = 0; k < t; ++k) REQUIRE(bnorm[k] > 0, "a right-hand-side column is zero");
      k_lz_init<<<dim3((unsigned)((n + 255) / 256), (unsigned)t), 256, 0, st>>>(bcopy.as<float>(), ld,
                                                                               bufA.as<float>(), ld, n, d_dvec);
      HIPCHECK(hipGetLastError());
    }

    // coefficient storage (host), column-major t x maxit
    std::vector<double> beta((size_t)t * maxit, 0.0), delta((size_t)t * maxit, 0.0);
    std::vector<double> rho(t), gam(t), omg(t);
    for (int k = 0; k < t; ++k) {
      beta[k] = bnorm[k];
      rho[k] = bnorm[k];
    }
    std::vector<char> gdone(ngroups, 0);
    std::vector<int> gsteps(ngroups, 0);
    std::vector<char> gconv(ngroups, 0);

    auto keep_basis = [&](const float* Udev) {
      if (keep_cols == 0) return;
      float* p = nullptr;
      HIPCHECK(hipMalloc(&p, sizeof(float) * n * keep_cols));
      s->U.push_back(p);
      HIPCHECK(hipMemcpy2DAsync(p, sizeof(float) * n, Udev, sizeof(float) * ld, sizeof(float) * n, keep_cols,
                                hipMemcpyDeviceToDevice, st));
    };
    keep_basis(bufA.as<float>());

    float* Uprev = nullptr;
    float* Ucur = bufA.as<float>();
    float* Unext = bufB.as<float>();
    float* spare = bufC.as<float>();
    // per-step scalars through pinned host memory (direct DMA, no staging copy on either side)
    double* pin = static_cast<double*>(pinned_scratch(0, sizeof(double) * 3 * (size_t)t));
    double* hd = pin;          // [0, 2t): delta_j, beta_{j+1} from the device
    double* hb = pin + 2 * t;  // [0, t): beta_j to the device
    // graph replay needs a capturable (non-null) stream: use a private one when none is given
    struct Graphs {
      bool enabled = false;
      hipStream_t own = nullptr;
      hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
      ~Graphs() {
        for (auto& e : exec)
          if (e) (void)hipGraphExecDestroy(e);
        if (own) {
          (void)hipStreamSynchronize(own);
          (void)hipStreamDestroy(own);
        }
      }
    } graphs;
    graphs.enabled = use_graph;
    if (use_graph && st == nullptr) {
      HIPCHECK(hipStreamSynchronize(nullptr));
      HIPCHECK(hipStreamCreate(&graphs.own));
      st = graphs.own;
    }
    int j = 0;
    const auto t_start = std::chrono::steady_clock::now();
    // One Lanczos step on the device: tmp = A U_j, delta_j, beta_{j+1}, U_{j+1}.  The buffers
    // rotate with period 3 after the first step, so steps j >= 1 replay one of three HIP graphs
    // (captured at their first use) instead of launching every kernel again.
    auto step_work = [&](hipStream_t s) {
      apply_op(Ucur, ld, tmp.as<float>(), ld, s);
      k_lz_a<<<dim3((unsigned)nparts, (unsigned)t), 256, 0, s>>>(Uprev, Ucur, tmp.as<float>(), ld, n, d_dvec + t,
                                                                dact.as<int>(), part.as<double>(), t);
      HIPCHECK(hipGetLastError());
      k_lz_sum<<<t, 64, 0, s>>>(part.as<double>(), nparts, t, d_dvec, 0);  // delta_j
      k_lz_b<<<dim3((unsigned)nparts, (unsigned)t), 256, 0, s>>>(Ucur, tmp.as<float>(), ld, n, d_dvec,
                                                                dact.as<int>(), part.as<double>(), t);
      HIPCHECK(hipGetLastError());
      k_lz_sum<<<t, 64, 0, s>>>(part.as<double>(), nparts, t, d_dvec + t, 1);  // beta_{j+1}
      k_lz_c<<<dim3((unsigned)((n + 1023) / 1024), (unsigned)t), 256, 0, s>>>(tmp.as<float>(), Unext, ld, n,
                                                                           d_dvec + t, dact.as<int>());
      HIPCHECK(hipGetLastError());
    };
    while (j < maxit - 1) {
      // beta_j of each column to device slot [t, 2t)
      for (int k = 0; k < t; ++k) hb[k] = beta[(size_t)j * t + k];
      HIPCHECK(hipMemcpyAsync(d_dvec + t, hb, sizeof(double) * t, hipMemcpyHostToDevice, st));
      if (graphs.enabled && j >= 1 && !eager_next) {
        hipGraphExec_t& ge = graphs.exec[(j - 1) % 3];
        if (!ge) {
          hipGraph_t gr = nullptr;
          HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          try {
            step_work(st);
          } catch (...) {
            (void)hipStreamEndCapture(st, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
          }
          HIPCHECK(hipStreamEndCapture(st, &gr));
          const hipError_t ie = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          HIPCHECK(ie);
        }
        HIPCHECK(hipGraphLaunch(ge, st));
      } else {
        step_work(st);
        eager_next = false;
      }
      keep_basis(Unext);
      HIPCHECK(hipMemcpyAsync(hd, d_dvec, sizeof(double) * 2 * t, hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));

      // CG coefficients (src/reml.cpp:304-307) and per-group stopping (:314-321)
      std::vector<double> gres(ngroups, 0.0);
      for (int k = 0; k < t; ++k) {
        const int gk = group[k];
        if (gdone[gk]) continue;
        const double dj = hd[k], bj1 = hd[t + k];
        delta[(size_t)j * t + k] = dj;
        beta[(size_t)(j + 1) * t + k] = bj1;
        gam[k] = j == 0 ? 1.0 / dj : 1.0 / (dj - omg[k] / gam[k]);
        omg[k] = (bj1 * gam[k]) * (bj1 * gam[k]);
        rho[k] = -bj1 * gam[k] * rho[k];
        gres[gk] = std::max(gres[gk], std::fabs(rho[k]) / bnorm[k]);
      }
      ++j;
      bool all_done = true;
      bool changed = false;
      for (int gi = 0; gi < ngroups; ++gi) {
        if (gdone[gi]) continue;
        gsteps[gi] = j;
        if (verbose && j % 5 == 0) printf("Relative error at step %d is %g\n", j, gres[gi]);
        if (gres[gi] <= tol) {
          gdone[gi] = 1;
          gconv[gi] = 1;
          changed = true;
          if (verbose) printf("Converged after %d iterations.\n", j);
        } else {
          all_done = false;
        }
      }
      if (changed) {
        for (int k = 0; k < t; ++k) active[k] = gdone[group[k]] ? 0 : 1;
        HIPCHECK(hipMemcpyAsync(dact.p, active.data(), sizeof(int) * t, hipMemcpyHostToDevice, st));
        if (on_active && !all_done && on_active(active, st)) {
          for (auto& e : graphs.exec)
            if (e) {
              HIPCHECK(hipStreamSynchronize(st));
              (void)hipGraphExecDestroy(e);
              e = nullptr;
            }
          eager_next = true;
        }
      }
      if (all_done) break;
      // rotate buffers: prev <- cur <- next
      float* old_prev = Uprev ? Uprev : spare;
      Uprev = Ucur;
      Ucur = Unext;
      Unext = old_prev;
    }
    HIPCHECK(hipStreamSynchronize(st));
    s->converged = true;
    for (int gi = 0; gi < ngroups; ++gi) {
      if (!gconv[gi]) {
        s->converged = false;
        if (verbose) printf("Failed to converge after %d iterations.\n", maxit);
      }
    }
    for (int k = 0; k < t; ++k) s->steps[k] = gsteps[group[k]];
    s->max_steps = j;
    s->beta.assign((size_t)t * (j + 1), 0.0);
    s->delta.assign((size_t)t * j, 0.0);
    for (int k = 0; k < t; ++k) {
      for (int q = 0; q <= s->steps[k]; ++q) s->beta[(size_t)q * t + k] = beta[(size_t)q * t + k];
      for (int q = 0; q < s->steps[k]; ++q) s->delta[(size_t)q * t + k] = delta[(size_t)q * t + k];
    }
    (void)t_start;
    return s.release();
}

}  // namespace slemm

extern "C" int slemm_gpu_lseed(slemm_geno* h, int kind, const int* col_kind, const int* col_group, float tau0,
                               const float* B, int64_t ldb, int t, float tol, int maxit, int keep_cols,
                               int precision, int verbose, void* stream, slemm_lseed** out) {
  return guard([&] {
    REQUIRE(h && B && out, "null argument");
    *out = nullptr;
    HIPCHECK(hipSetDevice(h->device));
    require_model(h);
    // Columns of converged groups leave the operator's block once that cuts the MFMA column
    // tiles the block product issues (e.g. 102 -> 101 columns: 21 -> 19 int8 digit tiles, or the
    // last unconverged column alone): the active columns are gathered into a compact block, the
    // operator runs on it and the results go back to their columns.  Each column's products do
    // not depend on the other columns of the block, so this changes no column's arithmetic
    // beyond the reduction split the smaller block may choose.
    struct Compact {
      bool on = false;
      int k = 0;
      std::vector<std::pair<int, int>> runs;  // (first column, count) of the active columns
      std::vector<int> kinds;
      DevBuf in, out;
    } cp;
    const LanczosOp op = [&](const float* U, int64_t ldu, float* o, int64_t ldo, hipStream_t s) {
      if (!cp.on) {
        run_op(h, kind, col_kind, tau0, U, ldu, t, o, ldo, precision, s);
        return;
      }
      const int64_t ld = h->n_pad, n = h->n;
      float* ci = cp.in.as<float>();
      float* co = cp.out.as<float>();
      int64_t off = 0;
      for (const auto& r : cp.runs) {
        HIPCHECK(hipMemcpy2DAsync(ci + off * ld, sizeof(float) * ld, U + r.first * ldu, sizeof(float) * ldu,
                                  sizeof(float) * n, r.second, hipMemcpyDeviceToDevice, s));
        off += r.second;
      }
      run_op(h, kind, cp.kinds.data(), tau0, ci, ld, cp.k, co, ld, precision, s);
      off = 0;
      for (const auto& r : cp.runs) {
        HIPCHECK(hipMemcpy2DAsync(o + r.first * ldo, sizeof(float) * ldo, co + off * ld, sizeof(float) * ld,
                                  sizeof(float) * n, r.second, hipMemcpyDeviceToDevice, s));
        off += r.second;
      }
    };
    // SLEMM_LSEED_COMPACT=0 (read per call) keeps every column in the block
    const char* ecp = std::getenv("SLEMM_LSEED_COMPACT");
    const bool compact = !(ecp && ecp[0] == '0');
    const ActiveFn on_active = [&](const std::vector<int>& active, hipStream_t) -> bool {
      if (!compact) return false;
      int k = 0;
      for (int c = 0; c < t; ++c) k += active[c] ? 1 : 0;
      if (k == 0 || (cp.on && k == cp.k) || op_column_tiles(h, k, precision) >= op_column_tiles(h, t, precision))
        return false;
      cp.on = true;
      cp.k = k;
      cp.runs.clear();
      cp.kinds.clear();
      for (int c = 0; c < t; ++c) {
        if (!active[c]) continue;
        if (!cp.runs.empty() && cp.runs.back().first + cp.runs.back().second == c) ++cp.runs.back().second;
        else cp.runs.emplace_back(c, 1);
        cp.kinds.push_back(col_kind ? col_kind[c] : kind);
      }
      cp.in.reserve(sizeof(float) * h->n_pad * k);
      cp.out.reserve(sizeof(float) * h->n_pad * k);
      return true;
    };
    *out = lseed_run(h, op, col_group, B, ldb, t, tol, maxit, keep_cols, verbose, (hipStream_t)stream,
                     lanczos_graphs(), on_active);
  });
}

extern "C" int slemm_gpu_lseed_info(const slemm_lseed* s, int* t, int* steps, int* converged) {
  return guard([&] {
    REQUIRE(s, "null seed");
    if (t) *t = s->t;
    if (steps) *steps = s->max_steps;
    if (converged) *converged = s->converged ? 1 : 0;
  });
}

extern "C" int slemm_gpu_lseed_col_steps(const slemm_lseed* s, int* steps) {
  return guard([&] {
    REQUIRE(s && steps, "null argument");
    for (int k = 0; k < s->t; ++k) steps[k] = s->steps[k];
  });
}

extern "C" int slemm_gpu_lseed_coeffs(const slemm_lseed* s, double* beta, double* delta) {
  return guard([&] {
    REQUIRE(s, "null seed");
    const int t = s->t, J = s->max_steps;
    // internal storage is step-major (q*t + k); the ABI layout is column-major t x (J+1)
    if (beta)
      for (int q = 0; q <= J; ++q)
        for (int k = 0; k < t; ++k) beta[(size_t)q * t + k] = s->beta[(size_t)q * t + k];
    if (delta)
      for (int q = 0; q < J; ++q)
        for (int k = 0; k < t; ++k) delta[(size_t)q * t + k] = s->delta[(size_t)q * t + k];
  });
}

extern "C" int slemm_gpu_lseed_destroy(slemm_lseed* s) {
  return guard([&] {
    if (!s) return;
    (void)hipSetDevice(s->device);
    delete s;
    (void)hipGetLastError();  // nothing of this call may stay pending for the caller's next check
  });
}

extern "C" int slemm_gpu_lsolve(const slemm_lseed* s, int col0, int t2, const float* B, int64_t ldb, float sigma,
                                float tol, float* X, int64_t ldx, void* stream) {
  return guard([&] {
    REQUIRE(s && B && X, "null argument");
    REQUIRE(t2 > 0 && col0 >= 0 && col0 + t2 <= s->keep_cols, "columns must be kept columns of the seed");
    REQUIRE(ldb >= s->n && ldx >= s->n, "leading dimension < n");
    HIPCHECK(hipSetDevice(s->device));
    hipStream_t st = (hipStream_t)stream;
    const int t = s->t;
    // host replay of the scalar recurrences (src/reml.cpp:404-426); ||U_{j+1}|| = 1
    // the solve's own ||b||: beta_0 of the seed is ||b|| of the same vectors
    int J = 0;
    std::vector<int> Jk(t2);
    const int maxJ = s->max_steps;
    std::vector<double> gam((size_t)std::max(maxJ, 1) * t2, 0.0), om((size_t)std::max(maxJ, 1) * t2, 0.0),
        rho((size_t)(maxJ + 1) * t2, 0.0);
    // all columns of one L_Solve call share the stopping rule: max over its columns
    std::vector<double> g(t2), o(t2), r(t2);
    for (int k = 0; k < t2; ++k) r[k] = s->beta[col0 + k];  // rho_0 = ||B|| (= beta_0 of the seed)
    for (int k = 0; k < t2; ++k) rho[k] = r[k];
    int steps_avail = maxJ;
    for (int k = 0; k < t2; ++k) steps_avail = std::min(steps_avail, s->steps[col0 + k]);
    int j = 0;
    while (j < steps_avail) {
      double res = 0.0;
      for (int k = 0; k < t2; ++k) {
        const int kk = col0 + k;
        const double dj = s->delta[(size_t)j * t + kk] + (double)sigma;
        const double bj1 = s->beta[(size_t)(j + 1) * t + kk];
        g[k] = j == 0 ? 1.0 / dj : 1.0 / (dj - o[k] / g[k]);
        o[k] = (bj1 * g[k]) * (bj1 * g[k]);
        r[k] = -bj1 * g[k] * r[k];
        gam[(size_t)j * t2 + k] = g[k];
        om[(size_t)j * t2 + k] = o[k];
        rho[(size_t)(j + 1) * t2 + k] = r[k];
        res = std::max(res, std::fabs(r[k]) / s->beta[kk]);
      }
      ++j;
      if (res <= tol) break;
    }
    J = j;
    for (int k = 0; k < t2; ++k) Jk[k] = J;
    DevBuf dg, dom, drho, djk, dul;
    dg.reserve(sizeof(double) * std::max(J, 1) * t2);
    dom.reserve(sizeof(double) * std::max(J, 1) * t2);
    drho.reserve(sizeof(double) * (J + 1) * t2);
    djk.reserve(sizeof(int) * t2);
    dul.reserve(sizeof(float*) * (J + 1));
    std::vector<const float*> ul(J + 1);
    for (int q = 0; q <= J; ++q) ul[q] = s->U[q];
    HIPCHECK(hipMemcpyAsync(dg.p, gam.data(), sizeof(double) * std::max(J, 1) * t2, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(dom.p, om.data(), sizeof(double) * std::max(J, 1) * t2, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(drho.p, rho.data(), sizeof(double) * (J + 1) * t2, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(djk.p, Jk.data(), sizeof(int) * t2, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(dul.p, ul.data(), sizeof(float*) * (J + 1), hipMemcpyHostToDevice, st));
    k_lsolve<<<dim3((unsigned)((s->n + 255) / 256), (unsigned)t2), 256, 0, st>>>(
        dul.as<const float*>(), s->n, col0, B, ldb, X, ldx, s->n, t2, J, dg.as<double>(), dom.as<double>(),
        drho.as<double>(), djk.as<int>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
  });
}

extern "C" int slemm_gpu_lseed_spectrum(const slemm_lseed* s, int k, double* D, double* W) {
  return guard([&] {
    REQUIRE(s && D && W, "null argument");
    REQUIRE(k >= 0 && k < s->t, "column out of range");
    const int J = s->steps[k];
    std::vector<double> d(J), e(J > 0 ? J - 1 : 0);
    for (int q = 0; q < J; ++q) d[q] = s->delta[(size_t)q * s->t + k];
    for (int q = 0; q + 1 < J; ++q) e[q] = s->beta[(size_t)(q + 1) * s->t + k];
    tridiag_eig_first_row(J, d.data(), e.data(), D, W);
  });
}

extern "C" int slemm_gpu_slq_ldet(const slemm_lseed* s, int col0, int nv, double sigma, double* ldet) {
  return guard([&] {
    REQUIRE(s && ldet, "null argument");
    REQUIRE(nv > 0 && col0 >= 0 && col0 + nv <= s->t, "probe columns out of range");
    double acc = 0.0;
    for (int k = col0; k < col0 + nv; ++k) {
      const int J = s->steps[k];
      std::vector<double> D(J), W(J);
      std::vector<double> d(J), e(J > 0 ? J - 1 : 0);
      for (int q = 0; q < J; ++q) d[q] = s->delta[(size_t)q * s->t + k];
      for (int q = 0; q + 1 < J; ++q) e[q] = s->beta[(size_t)(q + 1) * s->t + k];
      tridiag_eig_first_row(J, d.data(), e.data(), D.data(), W.data());
      for (int q = 0; q < J; ++q) acc += W[q] * W[q] * std::log(D[q] + sigma);
    }
    *ldet = acc * (double)s->n / (double)nv;
  });
}
