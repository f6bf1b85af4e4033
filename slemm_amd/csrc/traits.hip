// Batched traits on individual subsets: T traits share one device-resident bed; every trait
// keeps its own subset, allele frequencies and SNP weights (src/reml.cpp:612-622 over the
// subset), and all their vector columns go through one split-BF16 block product per pass.
// Columns are zero outside their trait's subset, so a product over all individuals equals the
// product over the subset; the per-column epilogues apply the trait's centring and weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "driver.hpp"
#include "lanczos.hpp"
#include "pipeline.hpp"
#include "traits.hpp"

namespace slemm {
namespace {

// masks (T x n host bytes, uploaded) -> float membership and intercept basis (T x n_pad)
__global__ void k_trait_masks(const uint8_t* __restrict__ m8, int64_t n, int64_t ld, const double* __restrict__ n_t,
                              float* __restrict__ mask, float* __restrict__ q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (i >= ld) return;
  const float v = (i < n && m8[(int64_t)t * n + i]) ? 1.f : 0.f;
  mask[(int64_t)t * ld + i] = v;
  q[(int64_t)t * ld + i] = v * (float)(1.0 / sqrt(n_t[t]));
}

// qd[k] = q_{t(k)}' X_k in FP64 (stage 1: per-block partials; stage 2 in k_trait_dots_final)
__global__ __launch_bounds__(256) void k_trait_dots(const float* __restrict__ X, int64_t ldx, const float* __restrict__ q,
                                                    int64_t ldq, const int* __restrict__ col_trait, int64_t n,
                                                    int64_t rpb, int t, double* __restrict__ part) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  const float* qk = q + (int64_t)col_trait[k] * ldq;
  const int64_t i0 = (int64_t)blockIdx.y * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) s += (double)qk[i] * (double)X[i + (int64_t)k * ldx];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * t + k] = red[0];
}

__global__ void k_trait_dots_final(const double* __restrict__ part, int64_t nblk, int t, int ldo,
                                   double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t) return;
  double s = 0.0;
  for (int64_t b = 0; b < nblk; ++b) s += part[b * t + k];
  out[k] = s;  // layout of the c = 1 projection coefficients: row 0 of a c x ldo block
  (void)ldo;
}

// QUADFORM finish per column: out_ik -= q_t(k),i qd_k, then + tau0 v_ik
__global__ void k_trait_quad_finish(float* __restrict__ out, int64_t ldo, const float* __restrict__ V, int64_t ldv,
                                    const float* __restrict__ q, int64_t ldq, const int* __restrict__ col_trait,
                                    const double* __restrict__ qd, const int* __restrict__ kinds, int t, int64_t n,
                                    float tau0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= n || k >= t || kinds[k] != SLEMM_OP_QUADFORM) return;
  const float qi = q[(int64_t)col_trait[k] * ldq + i];
  const float o = out[i + (int64_t)k * ldo];
  out[i + (int64_t)k * ldo] = (float)((double)o - (double)qi * qd[k]) + tau0 * V[i + (int64_t)k * ldv];
}

void trait_dots(slemm_traits* tr, const float* X, int64_t ldx, int t, double* out, hipStream_t st) {
  slemm_geno* h = tr->h;
  const int64_t rpb = 16384;
  const int64_t nblk = (h->n + rpb - 1) / rpb;
  h->redpart.reserve(sizeof(double) * nblk * t);
  k_trait_dots<<<dim3((unsigned)t, (unsigned)nblk), 256, 0, st>>>(X, ldx, tr->q.as<float>(), h->n_pad,
                                                                 tr->col_trait.as<int>(), h->n, rpb, t,
                                                                 h->redpart.as<double>());
  HIPCHECK(hipGetLastError());
  k_trait_dots_final<<<(t + 255) / 256, 256, 0, st>>>(h->redpart.as<double>(), nblk, t, t, out);
  HIPCHECK(hipGetLastError());
}

// Per 16-column tile of the block products: the extent of the tile's subsets (largest member +
// 1) over its live columns (tr->active, empty = all), uploaded for the kernels' work limits.
// The device arrays keep their addresses for a given width, as replayed graphs require.
void upload_tile_limits(slemm_traits* tr, int t, const Layout& L, hipStream_t st) {
  const int ntile = L.KP / 16;
  std::vector<int32_t> tns(ntile, 0);
  std::vector<int64_t> trows(ntile, 0);
  for (int k = 0; k < t; ++k)
    if (tr->active.empty() || tr->active[k]) {
      const int64_t e = tr->extent[tr->col_trait_host[k]];
      trows[k / 16] = std::max(trows[k / 16], e);
      if (L.tail_r > 0 && k >= L.tail0)  // its lo half, column k + tail_r
        trows[(k + L.tail_r) / 16] = std::max(trows[(k + L.tail_r) / 16], e);
    }
  for (int q = 0; q < ntile; ++q) tns[q] = (int32_t)((trows[q] + kPad - 1) / kPad);
  tr->tile_nsuper.reserve(sizeof(int32_t) * ntile);
  tr->tile_rows.reserve(sizeof(int64_t) * ntile);
  HIPCHECK(hipMemcpyAsync(tr->tile_nsuper.p, tns.data(), sizeof(int32_t) * ntile, hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(tr->tile_rows.p, trows.data(), sizeof(int64_t) * ntile, hipMemcpyHostToDevice, st));
  HIPCHECK(hipStreamSynchronize(st));
}

// Uploads the per-column trait ids (cached) and the tile limits of a new batch.
void upload_col_trait(slemm_traits* tr, const int* col_trait, int t, const Layout& L, hipStream_t st) {
  REQUIRE(col_trait, "col_trait is null");
  std::vector<int> ct(col_trait, col_trait + t);
  for (int k = 0; k < t; ++k) REQUIRE(ct[k] >= 0 && ct[k] < tr->T, "col_trait out of range");
  ct.push_back(2 * L.KP + (L.tail_r > 0 ? 1 : 0));  // the tile arrays cover the padded width and layout too
  if (ct != tr->col_trait_host) {
    tr->col_trait.reserve(sizeof(int) * t);
    HIPCHECK(hipMemcpyAsync(tr->col_trait.p, ct.data(), sizeof(int) * t, hipMemcpyHostToDevice, st));
    tr->col_trait_host = ct;
    tr->active.clear();
    upload_tile_limits(tr, t, L, st);
  }
}

}  // namespace

namespace {

// Thread per member i of trait t (blockIdx.y): its row of the y~, q and probe columns.
__global__ __launch_bounds__(256) void k_traits_seed_block(const int64_t* __restrict__ mem, const int64_t* __restrict__ off,
                                                           const float* __restrict__ yq, const uint8_t* __restrict__ bits,
                                                           const float* __restrict__ qscale, int T, int n_V, int64_t n,
                                                           float* __restrict__ B) {
  const int t = blockIdx.y;
  const int64_t o = off[t], nt = off[t + 1] - o;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nt) return;
  const int64_t row = mem[o + i];
  B[(int64_t)t * n + row] = yq[2 * o + i];
  B[(int64_t)(T + t) * n + row] = yq[2 * o + nt + i];
  const float q = qscale[t];
  for (int j = 0; j < n_V; ++j) {
    const int64_t e = (int64_t)j * nt + i;  // draw index
    B[(int64_t)(2 * T + t * n_V + j) * n + row] = (bits[e >> 3] >> (e & 7)) & 1 ? -q : q;
  }
}

}  // namespace

void traits_fill_seed_block(const slemm_traits* tr, int n_V, const std::vector<float>& yq,
                            const std::vector<uint8_t>& bits, float* B) {
  const int T = tr->T;
  const int64_t n = tr->h->n;
  std::vector<int64_t> off(T + 1, 0), mem;
  std::vector<float> qs(T);
  int64_t nmax = 0;
  for (int t = 0; t < T; ++t) {
    const int64_t nt = (int64_t)tr->members[t].size();
    off[t + 1] = off[t] + nt;
    qs[t] = rademacher_scale(nt);
    nmax = std::max(nmax, nt);
  }
  REQUIRE((int64_t)yq.size() == 2 * off[T], "seed block: phenotype/basis values do not match the subsets");
  REQUIRE((int64_t)bits.size() * 8 >= nmax * n_V, "seed block: too few probe signs");
  mem.reserve(off[T]);
  for (int t = 0; t < T; ++t) mem.insert(mem.end(), tr->members[t].begin(), tr->members[t].end());
  DevBuf dmem, doff, dyq, dbits, dqs;
  dmem.reserve(sizeof(int64_t) * mem.size());
  doff.reserve(sizeof(int64_t) * off.size());
  dyq.reserve(sizeof(float) * yq.size());
  dbits.reserve(std::max<size_t>(bits.size(), 1));
  dqs.reserve(sizeof(float) * qs.size());
  HIPCHECK(hipMemcpy(dmem.p, mem.data(), sizeof(int64_t) * mem.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(doff.p, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(dyq.p, yq.data(), sizeof(float) * yq.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(dbits.p, bits.data(), bits.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(dqs.p, qs.data(), sizeof(float) * qs.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemset(B, 0, sizeof(float) * (size_t)n * T * (2 + n_V)));
  k_traits_seed_block<<<dim3((unsigned)((nmax + 255) / 256), (unsigned)T), 256>>>(
      dmem.as<int64_t>(), doff.as<int64_t>(), dyq.as<float>(), dbits.as<uint8_t>(), dqs.as<float>(), T, n_V, n, B);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(nullptr));
}

void run_op_traits(slemm_traits* tr, const int* col_trait, const int* col_kind, float tau0, const float* V,
                   int64_t ldv, int t, float* out, int64_t ldo, int precision, hipStream_t st) {
  slemm_geno* h = tr->h;
  REQUIRE(use_lds(precision), "batched traits run the split-BF16 block product (SLEMM_PREC_BF16X2)");
  check_dense(V, ldv, h->n, t, "V");
  check_dense(out, ldo, h->n, t, "out");
  REQUIRE((const void*)V != (const void*)out, "out must not alias V");
  const Layout L = layout_for(h, t, precision, true);
  bool any_quad = false;
  const int* kinds = upload_kinds(h, SLEMM_OP_LDET, col_kind, t, any_quad, st);
  upload_col_trait(tr, col_trait, t, L, st);
  const TraitCols tc = tr->cols();

  const double* qtv = nullptr;
  if (any_quad) {
    h->qtv.reserve(sizeof(double) * L.KP);
    trait_dots(tr, V, ldv, t, h->qtv.as<double>(), st);
    qtv = h->qtv.as<double>();
  }
  pack_pass1(h, L, V, ldv, t, nullptr, qtv, kinds, precision, st, true, &tc);
  h->x2.reserve(L.x2_bytes);
  const int64_t nparts = h->m_pad / kFin1Rows;
  h->part2.reserve(sizeof(double) * nparts * L.KP);
  {
    int64_t nsp = 0;
    const int64_t S1 = pass_splits(h, 0, L.slot, nsp);
    const int64_t S2 = pass_splits(h, 1, L.slot, nsp);
    h->partbuf.reserve(sizeof(float) * L.KP * std::max<size_t>((size_t)S1 * h->m_pad, (size_t)S2 * h->n_pad));
  }
  pass1_lds(h, L, t, true, h->x2.p, h->part2.as<double>(), nullptr, 0, st, false, false, &tc);
  double* u = h->colsum.as<double>() + L.KP;
  launch_sum_parts(h->part2.as<double>(), nparts, L.KP, u, st);
  pass2_lds(h, L, t, u, nullptr, V, ldv, tau0, kinds, out, ldo, st, &tc);
  if (any_quad) {
    tr->qd.reserve(sizeof(double) * t);
    trait_dots(tr, out, ldo, t, tr->qd.as<double>(), st);
    k_trait_quad_finish<<<dim3((unsigned)((h->n + 255) / 256), (unsigned)t), 256, 0, st>>>(
        out, ldo, V, ldv, tr->q.as<float>(), h->n_pad, tr->col_trait.as<int>(), tr->qd.as<double>(), kinds, t,
        h->n, tau0);
    HIPCHECK(hipGetLastError());
  }
}

}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_traits_create(slemm_geno* h, int T, const uint8_t* masks, slemm_traits** out) {
  return guard([&] {
    REQUIRE(h && masks && out, "null argument");
    *out = nullptr;
    REQUIRE(T >= 1 && T <= 1024, "number of traits out of [1, 1024]");
    HIPCHECK(hipSetDevice(h->device));
    ensure_counts(h, nullptr);
    const int64_t n = h->n, m = h->m, ldn = h->n_pad, ldm = h->m_pad;
    auto tr = std::make_unique<slemm_traits>();
    tr->h = h;
    tr->device = h->device;
    tr->T = T;
    tr->n_host.assign(T, 0);
    tr->members.assign(T, {});
    for (int t = 0; t < T; ++t) {
      for (int64_t i = 0; i < n; ++i)
        if (masks[(size_t)t * n + i]) tr->members[t].push_back(i);
      tr->n_host[t] = (int64_t)tr->members[t].size();
      REQUIRE(tr->n_host[t] >= 2, "every trait needs at least two individuals");
    }
    tr->extent.assign(T, 0);
    for (int t = 0; t < T; ++t) tr->extent[t] = tr->members[t].back() + 1;
    std::vector<double> nd(tr->n_host.begin(), tr->n_host.end());
    tr->n_t.reserve(sizeof(double) * T);
    HIPCHECK(hipMemcpy(tr->n_t.p, nd.data(), sizeof(double) * T, hipMemcpyHostToDevice));
    tr->mask.reserve(sizeof(float) * ldn * T);
    tr->q.reserve(sizeof(float) * ldn * T);
    {
      DevBuf m8;
      m8.reserve((size_t)T * n);
      HIPCHECK(hipMemcpy(m8.p, masks, (size_t)T * n, hipMemcpyHostToDevice));
      k_trait_masks<<<dim3((unsigned)((ldn + 255) / 256), (unsigned)T), 256>>>(
          m8.as<uint8_t>(), n, ldn, tr->n_t.as<double>(), tr->mask.as<float>(), tr->q.as<float>());
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipDeviceSynchronize());
    }
    // dosage sums over each subset: one raw Z' pass with the membership columns (integers below
    // 2^24, exact in BF16 x FP32 accumulation)
    std::vector<float> sums((size_t)m * T);
    {
      DevBuf zs;
      zs.reserve(sizeof(float) * m * T);
      const Layout L = layout_for(h, T, PREC_BF16X2);
      pack_pass1(h, L, tr->mask.as<float>(), ldn, T, nullptr, nullptr, nullptr, PREC_BF16X2, nullptr, false);
      pass1_lds(h, L, T, false, nullptr, nullptr, zs.as<float>(), m, nullptr, true);
      HIPCHECK(hipDeviceSynchronize());
      HIPCHECK(hipMemcpy(sums.data(), zs.p, sizeof(float) * m * T, hipMemcpyDeviceToHost));
    }
    // per-trait gm and gw with the operations of src/reml.cpp:612-622 (float, same order)
    std::vector<int32_t> pc((size_t)ldm * T, 0);
    std::vector<float> gm((size_t)ldm * T, 0.f), gw((size_t)ldm * T, 0.f);
    tr->monomorphic.assign(T, 0);
    for (int t = 0; t < T; ++t) {
      const float nf = (float)tr->n_host[t];
      for (int64_t j = 0; j < m; ++j) {
        const int32_t p = (int32_t)std::lround(sums[(size_t)t * m + j]);
        pc[(size_t)t * ldm + j] = p;
        const float g = (float)p / nf;
        gm[(size_t)t * ldm + j] = g;
        const float v = (g * (2.0f - g)) / 2.0f;
        const float w = (1.0f / v) / (float)m;
        if (std::isfinite(w)) {
          gw[(size_t)t * ldm + j] = w;
        } else {
          ++tr->monomorphic[t];  // no variation within the subset: the SNP drops out of the model
        }
      }
    }
    tr->pc.reserve(sizeof(int32_t) * ldm * T);
    tr->gm.reserve(sizeof(float) * ldm * T);
    tr->gw.reserve(sizeof(float) * ldm * T);
    HIPCHECK(hipMemcpy(tr->pc.p, pc.data(), sizeof(int32_t) * ldm * T, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(tr->gm.p, gm.data(), sizeof(float) * ldm * T, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(tr->gw.p, gw.data(), sizeof(float) * ldm * T, hipMemcpyHostToDevice));
    *out = tr.release();
  });
}

extern "C" int slemm_gpu_traits_destroy(slemm_traits* tr) {
  return guard([&] {
    if (!tr) return;
    (void)hipSetDevice(tr->device);
    (void)hipDeviceSynchronize();
    delete tr;
    (void)hipGetLastError();
  });
}

extern "C" int slemm_gpu_traits_info(const slemm_traits* tr, int64_t* n_t, int64_t* monomorphic) {
  return guard([&] {
    REQUIRE(tr, "null handle");
    for (int t = 0; t < tr->T; ++t) {
      if (n_t) n_t[t] = tr->n_host[t];
      if (monomorphic) monomorphic[t] = tr->monomorphic[t];
    }
  });
}

extern "C" int slemm_gpu_traits_model(const slemm_traits* tr, int trait, float* gm, float* gw, int32_t* popcnt) {
  return guard([&] {
    REQUIRE(tr, "null handle");
    REQUIRE(trait >= 0 && trait < tr->T, "trait out of range");
    HIPCHECK(hipSetDevice(tr->h->device));
    const int64_t m = tr->h->m, off = (int64_t)trait * tr->h->m_pad;
    if (gm) HIPCHECK(hipMemcpy(gm, tr->gm.as<float>() + off, sizeof(float) * m, hipMemcpyDeviceToHost));
    if (gw) HIPCHECK(hipMemcpy(gw, tr->gw.as<float>() + off, sizeof(float) * m, hipMemcpyDeviceToHost));
    if (popcnt) HIPCHECK(hipMemcpy(popcnt, tr->pc.as<int32_t>() + off, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  });
}

extern "C" int slemm_gpu_traits_op(slemm_traits* tr, const int* col_trait, const int* col_kind, float tau0,
                                   const float* V, int64_t ldv, int t, float* out, int64_t ldo, int precision,
                                   void* stream) {
  return guard([&] {
    REQUIRE(tr, "null handle");
    HIPCHECK(hipSetDevice(tr->h->device));
    run_op_traits(tr, col_trait, col_kind, tau0, V, ldv, t, out, ldo, precision, (hipStream_t)stream);
  });
}

extern "C" int slemm_gpu_traits_lseed(slemm_traits* tr, const int* col_trait, const int* col_kind,
                                      const int* col_group, float tau0, const float* B, int64_t ldb, int t, float tol,
                                      int maxit, int keep_cols, int precision, int verbose, void* stream,
                                      slemm_lseed** out) {
  return guard([&] {
    REQUIRE(tr && out, "null argument");
    *out = nullptr;
    HIPCHECK(hipSetDevice(tr->h->device));
    const std::vector<int> ct(col_trait, col_trait + t);
    const std::vector<int> ck = col_kind ? std::vector<int>(col_kind, col_kind + t) : std::vector<int>(t, SLEMM_OP_LDET);
    const LanczosOp op = [&](const float* U, int64_t ldu, float* o, int64_t ldo, hipStream_t s) {
      run_op_traits(tr, ct.data(), ck.data(), tau0, U, ldu, t, o, ldo, precision, s);
    };
    // converged groups stop costing block-product work: their columns leave the tile limits
    const ActiveFn on_active = [&](const std::vector<int>& active, hipStream_t s) {
      tr->active = active;
      upload_tile_limits(tr, t, layout_for(tr->h, t, precision, true), s);
      return false;  // the same launches, with new tile limits in device memory
    };
    try {
      *out = lseed_run(tr->h, op, col_group, B, ldb, t, tol, maxit, keep_cols, verbose, (hipStream_t)stream,
                       lanczos_graphs(), on_active);
    } catch (...) {
      tr->active.clear();
      tr->col_trait_host.clear();
      throw;
    }
    // later calls on this batch see every column again
    tr->active.clear();
    tr->col_trait_host.clear();
  });
}
