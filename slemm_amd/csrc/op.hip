// Operator pipeline (lnr_op_select and the Z'/Z products of src/reml.cpp:6-233) and the
// FP64 GEBV product (src/pred.cpp:112-130).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "blockprod.hpp"

namespace slemm {

namespace {

struct Layout {
  int NT, KP;
  size_t x1_bytes, x2_bytes;
  int64_t lo1_16, lo2_16;  // lo-half offsets (16-byte units)
};

Layout layout_for(const slemm_geno* h, int t, int precision) {
  Layout L;
  L.NT = blockprod_nt(t);
  L.KP = (int)round_up(t, 16 * L.NT);
  const size_t es = precision == PREC_BF16X2 ? 4 : 4;  // hi+lo bf16, or one fp32
  L.x1_bytes = (size_t)h->n_pad * L.KP * es;
  L.x2_bytes = (size_t)h->m_pad * L.KP * es;
  L.lo1_16 = precision == PREC_BF16X2 ? (int64_t)h->n_pad * L.KP / 8 : 0;
  L.lo2_16 = precision == PREC_BF16X2 ? (int64_t)h->m_pad * L.KP / 8 : 0;
  return L;
}

void check_precision(int precision) {
  REQUIRE(precision == PREC_FP32 || precision == PREC_BF16X2, "unknown precision");
}

const int* upload_kinds(slemm_geno* h, int kind, const int* col_kind, int t, bool& any_quad, hipStream_t st) {
  std::vector<int> kinds(t);
  any_quad = false;
  for (int k = 0; k < t; ++k) {
    const int kk = col_kind ? col_kind[k] : kind;
    REQUIRE(kk == SLEMM_OP_LDET || kk == SLEMM_OP_QUADFORM, "Unknown linear operator in lnr_op_select()");
    kinds[k] = kk;
    any_quad |= kk == SLEMM_OP_QUADFORM;
  }
  if (kinds != h->kinds_host) {
    h->kinds.reserve(sizeof(int) * t);
    HIPCHECK(hipMemcpyAsync(h->kinds.p, kinds.data(), sizeof(int) * t, hipMemcpyHostToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
    h->kinds_host = kinds;
  }
  return h->kinds.as<int>();
}

// pass-1 operand from V (n x t): rows scaled by rr (optional), QUADFORM columns projected,
// column means removed.  colsum layout: [0,KP) sums of the represented centred operand,
// [KP,2KP) pass-2 centring sums, [2KP,3KP) sums of the un-centred operand.
void pack_pass1(slemm_geno* h, const Layout& L, const float* V, int64_t ldv, int t, const float* rr,
                const double* qtv, const int* kinds, int precision, hipStream_t st) {
  h->x1.reserve(L.x1_bytes);
  const int64_t nparts = h->n_pad / 256;
  h->part1.reserve(sizeof(double) * nparts * L.KP);
  h->colsum.reserve(sizeof(double) * 3 * L.KP);
  double* xsum = h->colsum.as<double>() + 2 * L.KP;
  PackArgs p;
  p.V = V; p.ldv = ldv; p.t = t; p.KP = L.KP; p.n = h->n; p.rows_pad = h->n_pad;
  p.rr = rr; p.Q = h->Q; p.ldq = h->n_pad; p.c = h->c; p.qtv = qtv; p.kinds = kinds;
  p.mode = precision; p.X = nullptr; p.lo_offset16 = L.lo1_16; p.part = h->part1.as<double>();
  launch_pack(p, st);  // sums of the un-centred operand
  launch_sum_parts(h->part1.as<double>(), nparts, L.KP, xsum, st);
  p.X = h->x1.p;
  p.xsum = xsum;
  launch_pack(p, st);  // centred operand and its represented column sums
  launch_sum_parts(h->part1.as<double>(), nparts, L.KP, h->colsum.as<double>(), st);
}

void centring_args(slemm_geno* h, const Layout& L, BlockProdArgs& a) {
  a.xsum = h->colsum.as<double>() + 2 * L.KP;
  a.xsum_n = (double)h->n;
  a.popcnt = h->popcnt;
}

void check_dense(const float* p, int64_t ld, int64_t rows, int t, const char* what) {
  if (!p) fail(SLEMM_ERR_ARG, std::string(what) + " is null");
  if (ld < rows) fail(SLEMM_ERR_ARG, std::string("leading dimension of ") + what + " is too small");
  if (t <= 0) fail(SLEMM_ERR_ARG, "number of columns must be positive");
}

}  // namespace

void run_op(slemm_geno* h, int kind, const int* col_kind, float tau0, const float* V, int64_t ldv, int t,
            float* out, int64_t ldo, int precision, hipStream_t st) {
  require_model(h);
  check_precision(precision);
  check_dense(V, ldv, h->n, t, "V");
  check_dense(out, ldo, h->n, t, "out");
  REQUIRE((const void*)V != (const void*)out, "out must not alias V");
  const Layout L = layout_for(h, t, precision);
  bool any_quad = false;
  const int* kinds = upload_kinds(h, kind, col_kind, t, any_quad, st);

  const double* qtv = nullptr;
  if (any_quad && h->c > 0) {
    h->qtv.reserve(sizeof(double) * h->c * L.KP);
    HIPCHECK(hipMemsetAsync(h->qtv.p, 0, sizeof(double) * h->c * L.KP, st));
    tn_product(h->Q, h->n_pad, h->c, V, ldv, t, h->n, h->qtv.as<double>(), L.KP, h->redpart, st);
    qtv = h->qtv.as<double>();
  }
  pack_pass1(h, L, V, ldv, t, h->rr, qtv, kinds, precision, st);

  // pass 1: T = W (Z' X1 - gm s'), written as the pass-2 operand
  h->x2.reserve(L.x2_bytes);
  const int64_t nwg1 = h->m_pad / blockprod_rows_per_wg(L.NT);
  h->part2.reserve(sizeof(double) * nwg1 * L.KP);
  BlockProdArgs a;
  a.G = h->gs; a.ldg = h->ld_gs; a.nsuper = h->n_pad / kPad; a.X = h->x1.p; a.lo_offset16 = L.lo1_16;
  a.KP = L.KP; a.t = t; a.rows = h->m; a.colsum = h->colsum.as<double>();
  a.gm = h->gm; a.gw = h->gw; a.out2 = h->x2.p; a.out2_mode = precision;
  a.out2_lo_offset = L.lo2_16 * 8; a.upart = h->part2.as<double>();
  centring_args(h, L, a);
  launch_blockprod(a, precision, EPI_PASS1, L.NT, h->m_pad, st);
  double* u = h->colsum.as<double>() + L.KP;
  launch_sum_parts(h->part2.as<double>(), nwg1, L.KP, u, st);

  // pass 2: out = rr (Z X2 - 1 u') (+ tau0 V for LDET columns)
  BlockProdArgs b;
  b.G = h->gi; b.ldg = h->ld_gi; b.nsuper = h->m_pad / kPad; b.X = h->x2.p; b.lo_offset16 = L.lo2_16;
  b.KP = L.KP; b.t = t; b.rows = h->n; b.colsum = u; b.out = out; b.ldo = ldo;
  b.rr = h->rr; b.v = V; b.ldv = ldv; b.tau0 = tau0; b.kinds = kinds;
  launch_blockprod(b, precision, EPI_PASS2, L.NT, h->n_pad, st);

  if (any_quad) {
    double* bq = nullptr;
    if (h->c > 0) {
      h->qtv.reserve(sizeof(double) * h->c * L.KP);
      tn_product(h->Q, h->n_pad, h->c, out, ldo, t, h->n, h->qtv.as<double>(), L.KP, h->redpart, st);
      bq = h->qtv.as<double>();
    }
    launch_quad_finish(out, ldo, V, ldv, h->Q, h->n_pad, h->c, bq, L.KP, kinds, t, h->n, tau0, st);
  }
}

// --------------------------------------------------------------------------- GEBV ----------
// One wave per 1024 individuals; thread holds 16 FP64 accumulators (4 bytes of codes) and walks
// the SNPs in order, so every individual's sum is accumulated in the same order as the
// reference loop over SNPs.  z*alpha is exact for z in {0,1,2}, so fma == mul+add.
__global__ __launch_bounds__(64) void k_gebv(const uint8_t* __restrict__ gs, int64_t ld_gs, int64_t m,
                                             int64_t n, const double* __restrict__ alpha,
                                             const uint8_t* __restrict__ flip, double* __restrict__ out) {
  const int64_t i0 = (int64_t)blockIdx.x * 1024 + threadIdx.x * 16;
  if (i0 >= n) return;  // n_pad is a multiple of 16, so i0 < n implies i0 + 16 <= n_pad
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  const uint8_t* base = gs + i0 / 4;
  for (int64_t j = 0; j < m; ++j) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(base + j * ld_gs);
    const double aj = alpha[j];
    const bool fj = flip && flip[j];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const uint32_t code = (w >> (2 * q)) & 3u;
      uint32_t d = (code & 1u) + (code >> 1);
      if (fj) d = 2u - d;
      acc[q] = fma((double)d, aj, acc[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q)
    if (i0 + q < n) out[i0 + q] = acc[q];
}

}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_op(slemm_geno* h, int kind, const int* col_kind, float tau0, const float* V,
                            int64_t ldv, int t, float* out, int64_t ldo, int precision, void* stream) {
  return guard([&] {
    REQUIRE(h, "null handle");
    HIPCHECK(hipSetDevice(h->device));
    run_op(h, kind, col_kind, tau0, V, ldv, t, out, ldo, precision, (hipStream_t)stream);
  });
}

extern "C" int slemm_gpu_zt(slemm_geno* h, const float* V, int64_t ldv, int t, float* out, int64_t ldo,
                            int weighted, int scale_rr, int precision, void* stream) {
  return guard([&] {
    REQUIRE(h, "null handle");
    HIPCHECK(hipSetDevice(h->device));
    hipStream_t st = (hipStream_t)stream;
    check_precision(precision);
    check_dense(V, ldv, h->n, t, "V");
    check_dense(out, ldo, h->m, t, "out");
    ensure_counts(h, st);
    if (weighted || scale_rr) require_model(h);
    const Layout L = layout_for(h, t, precision);
    pack_pass1(h, L, V, ldv, t, scale_rr ? h->rr : nullptr, nullptr, nullptr, precision, st);
    BlockProdArgs a;
    a.G = h->gs; a.ldg = h->ld_gs; a.nsuper = h->n_pad / kPad; a.X = h->x1.p; a.lo_offset16 = L.lo1_16;
    a.KP = L.KP; a.t = t; a.rows = h->m; a.colsum = h->colsum.as<double>();
    a.gm = h->gm; a.gw = weighted ? h->gw : nullptr; a.out = out; a.ldo = ldo;
    centring_args(h, L, a);
    launch_blockprod(a, precision, EPI_PASS1, L.NT, h->m_pad, st);
  });
}

extern "C" int slemm_gpu_z(slemm_geno* h, const float* U, int64_t ldu, int t, float* out, int64_t ldo,
                           int precision, void* stream) {
  return guard([&] {
    REQUIRE(h, "null handle");
    HIPCHECK(hipSetDevice(h->device));
    hipStream_t st = (hipStream_t)stream;
    check_precision(precision);
    check_dense(U, ldu, h->m, t, "U");
    check_dense(out, ldo, h->n, t, "out");
    ensure_counts(h, st);
    const Layout L = layout_for(h, t, precision);
    // pack U as the pass-2 operand; gm-weighted column sums give the centring term
    h->x2.reserve(L.x2_bytes);
    const int64_t nparts = h->m_pad / 256;
    h->part2.reserve(sizeof(double) * nparts * L.KP);
    h->colsum.reserve(sizeof(double) * 3 * L.KP);
    PackArgs p;
    p.V = U; p.ldv = ldu; p.t = t; p.KP = L.KP; p.n = h->m; p.rows_pad = h->m_pad;
    p.mode = precision; p.X = h->x2.p; p.lo_offset16 = L.lo2_16; p.part = h->part2.as<double>();
    p.wsum = h->gm;
    launch_pack(p, st);
    double* u = h->colsum.as<double>() + L.KP;
    launch_sum_parts(h->part2.as<double>(), nparts, L.KP, u, st);
    BlockProdArgs b;
    b.G = h->gi; b.ldg = h->ld_gi; b.nsuper = h->m_pad / kPad; b.X = h->x2.p; b.lo_offset16 = L.lo2_16;
    b.KP = L.KP; b.t = t; b.rows = h->n; b.colsum = u; b.out = out; b.ldo = ldo;
    launch_blockprod(b, precision, EPI_PASS2, L.NT, h->n_pad, st);
  });
}

extern "C" int slemm_gpu_gebv(slemm_geno* h, const double* alpha, const uint8_t* flip, double* gebv,
                              void* stream) {
  return guard([&] {
    REQUIRE(h && alpha && gebv, "null argument");
    HIPCHECK(hipSetDevice(h->device));
    hipStream_t st = (hipStream_t)stream;
    DevBuf dflip;
    if (flip) {
      dflip.reserve(h->m);
      HIPCHECK(hipMemcpyAsync(dflip.p, flip, h->m, hipMemcpyHostToDevice, st));
    }
    k_gebv<<<(unsigned)((h->n_pad + 1023) / 1024), 64, 0, st>>>(h->gs, h->ld_gs, h->m, h->n, alpha,
                                                      dflip.as<uint8_t>(), gebv);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
  });
}
