// Operator pipeline (lnr_op_select and the Z'/Z products of src/reml.cpp:6-233) and the
// FP64 GEBV product (src/pred.cpp:112-130).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pipeline.hpp"

namespace slemm {

static bool i8_wide_enabled();

Layout layout_for(const slemm_geno* h, int t, int precision, bool traits, bool allow_i8) {
  Layout L;
  L.i8 = false;
  L.op_mode = precision;
  L.NT = blockprod_nt(t);
  // Wide split-BF16 blocks (t > 80) run as 80-column groups: the pipelined kernels take at most
  // 5 column tiles (two waves per SIMD), and 6-8-tile groups would fall back to the unpipelined
  // register kernel (SLEMM_WIDE_NT overrides).  Each group reads the codes again; these blocks
  // are MFMA-bound, so the extra reads cost no time.
  if (precision == PREC_BF16X2 && use_lds(precision) && t > 80) {
    static const char* e = std::getenv("SLEMM_WIDE_NT");
    const int nt = e ? std::atoi(e) : 5;
    L.NT = nt >= 1 && nt <= 8 ? nt : 5;
  }
  L.KP = (int)round_up(t, 16 * L.NT);
  const char* ehl = std::getenv("SLEMM_HL_COLUMNS");  // "0": hi and lo as separate blocks
  const bool hl_ok = precision == PREC_BF16X2 && use_lds(precision) && !(ehl && ehl[0] == '0');
  L.hl = hl_ok && 2 * t <= 16;
  L.slot = L.hl ? 0 : L.NT;
  L.tail0 = L.tail_r = 0;
  // Blocks of 1 to 5 vectors (hi and lo halves as columns) run on the int8 narrow kernel
  // (SLEMM_I8_NARROW=0: the split-BF16 ones), blocks of 9 or more vectors on the int8 wide kernel
  // (SLEMM_I8_WIDE=0: the split-BF16 ones); 6 to 8 vectors (hi/lo as columns: one MFMA tile per
  // step, streaming faster than three digit tiles) and the batched traits' narrow blocks keep the
  // split-BF16 kernels.
  if (allow_i8 && precision == PREC_BF16X2 && use_lds(precision)) {
    if (t <= 5) L.i8 = !traits && L.hl && i8_narrow_enabled();
    else if (!(t <= 8 && L.hl)) L.i8 = i8_wide_enabled();
  }
  if (L.i8) {
    L.hl = false;
    L.op_mode = PREC_FP32;
    L.x1_bytes = (size_t)h->n_pad * L.KP * 4;
    L.x2_bytes = (size_t)h->m_pad * L.KP * 4;
    L.lo1_16 = L.lo2_16 = 0;
    return L;
  }
  if (L.hl) {
    L.tail_r = t;
    L.x1_bytes = (size_t)h->n_pad * L.KP * 2;
    L.x2_bytes = (size_t)h->m_pad * L.KP * 2;
    L.lo1_16 = L.lo2_16 = t;  // lo half of column k is column t + k
    return L;
  }
  if (hl_ok && t > 16 && t % 16 >= 1 && t % 16 <= 8 && L.NT <= 5) {
    L.tail0 = t / 16 * 16;
    L.tail_r = t % 16;
  }
  const size_t es = 4;  // hi+lo bf16, or one fp32
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
                const double* qtv, const int* kinds, int precision, hipStream_t st, bool center,
                const TraitCols* tc) {
  h->x1.reserve(L.x1_bytes);
  const int64_t nparts = h->n_pad / kPackRows;
  h->part1.reserve(sizeof(double) * nparts * L.KP);
  h->colsum.reserve(sizeof(double) * 3 * L.KP);
  double* xsum = h->colsum.as<double>() + 2 * L.KP;
  PackArgs p;
  p.V = V; p.ldv = ldv; p.t = t; p.KP = L.KP; p.n = h->n; p.rows_pad = h->n_pad;
  p.rr = rr; p.Q = h->Q; p.ldq = h->n_pad; p.c = h->c; p.qtv = qtv; p.kinds = kinds;
  p.mode = L.op_mode; p.X = nullptr; p.lo_offset16 = L.lo1_16; p.part = h->part1.as<double>();
  p.tail0 = L.tail0; p.tail_r = L.tail_r;
  if (tc) {
    p.tc = *tc;
    p.c = qtv ? 1 : 0;  // per-trait intercept basis
  }
  if (center) {
    launch_pack(p, st);  // sums of the un-centred operand
    launch_sum_parts(h->part1.as<double>(), nparts, L.KP, xsum, st);
    p.xsum = xsum;
  }
  p.X = h->x1.p;
  launch_pack(p, st);  // centred operand and its represented column sums
  launch_sum_parts(h->part1.as<double>(), nparts, L.KP, h->colsum.as<double>(), st);
}

void centring_args(slemm_geno* h, const Layout& L, BlockProdArgs& a) {
  a.xsum = h->colsum.as<double>() + 2 * L.KP;
  a.xsum_n = (double)h->n;
  a.popcnt = h->popcnt;
}

// Split-BF16 products run on the split-reduction block-product kernels (variant and grid size
// per shape from the autotuner); SLEMM_BLOCKPROD_V1=1 selects the fused register kernel of
// blockprod.hip instead.  SLEMM_BP_VARIANT / SLEMM_BP_TARGET force a variant / grid target.
bool use_lds(int precision) {
  static const bool v1 = std::getenv("SLEMM_BLOCKPROD_V1") != nullptr;
  return precision == PREC_BF16X2 && !v1;
}

BpChoice bp_choice(const slemm_geno* h, int pass, int slot) {
  const char* ev = std::getenv("SLEMM_BP_VARIANT");  // read per call: tests switch variants in-process
  const char* et = std::getenv("SLEMM_BP_TARGET");
  BpChoice c{h->bp_variant[pass][slot], h->bp_target[pass][slot]};
  if (ev) c.variant = std::atoi(ev);
  if (et) c.target = std::atoll(et);
  return c;
}

// block-product arguments of the layout's hi/lo arrangement: the staged-B narrow kernels (hl),
// the tail tile without lo MFMAs, and the last tile holding live columns
void tail_args(const Layout& L, int t, LdsArgs& la) {
  la.hl = L.hl;
  la.ncols = L.hl ? 2 * t : t;
  la.tail_tile = (L.tail_r > 0 && !L.hl) ? L.tail0 / 16 : -1;
  la.last_tile = L.hl ? 0 : (L.tail_r > 0 ? L.tail0 / 16 : (t + 15) / 16 - 1);
}

// SLEMM_I8_TARGET sets the workgroup target of the int8 reduction split (read per call): by
// default 2048 for narrow blocks (at n = 1M x m = 60k the SNP-major pass needs the split: 3.06 ms
// at a 512 target, 2.64 ms at 2048; at n = 100k x m = 500k 512 to 4096 are within 2%) and 8192
// for wide blocks (the best of 1024 to 16384 at n = 100k x m = 500k).
static bool i8_wide_enabled() {
  const char* e = std::getenv("SLEMM_I8_WIDE");
  return !(e && e[0] == '0');
}

// the int8 settings of a t-vector block for one pass: workgroup target (SLEMM_I8_TARGET, else the
// autotuner's choice on this handle, else 2048 narrow / 8192 wide) and the narrow row tiles
static int i8_pass_splits(const slemm_geno* h, int pass, int t, int64_t rows_pad, int64_t nsuper, bool squared,
                          int64_t& nsp, int& rt) {
  // SLEMM_I8_TARGET (read per call): one workgroup target for both passes, or "pass1,pass2"
  const char* e = std::getenv("SLEMM_I8_TARGET");
  const int cls = i8_class(t);
  const int64_t dflt = h->i8_target[pass][cls] > 0 ? h->i8_target[pass][cls] : (t <= 5 ? 2048 : 8192);
  int64_t target = dflt;
  if (e && *e) {
    const char* c = std::strchr(e, ',');
    target = std::max<int64_t>(1, std::atoll(pass == 1 && c ? c + 1 : e));
  }
  rt = cls == 0 ? h->i8_rt[pass] : 0;
  return i8_splits(t, rows_pad, nsuper, target, squared, rt, nsp);
}

// sums: where the digit column sums of the represented operand go (weighted by sw), null: none
static void bp_i8(slemm_geno* h, const uint8_t* G, int64_t ldg, int64_t red_pad, int64_t rows_pad, const void* X,
                  const Layout& L, int t, bool squared, int S, int64_t nsp, int rt, hipStream_t st, double* sums,
                  I8SumWeights sw, const int32_t* tile_nsuper = nullptr, const int64_t* tile_rows = nullptr) {
  h->i8buf.reserve(i8_scratch_bytes(red_pad, t));
  I8Launch l;
  l.G = G; l.ldg = ldg; l.red_pad = red_pad; l.rows_pad = rows_pad; l.X = X; l.KP = L.KP; l.t = t;
  l.squared = squared; l.nsplit = S; l.ns_per_split = nsp; l.scratch = h->i8buf.p;
  l.part = h->partbuf.as<float>();
  l.tile_nsuper = tile_nsuper; l.tile_rows = tile_rows;
  l.sum_out = sums; l.sum_w = sw; l.rt = rt;
  launch_bp_i8(l, st);
}

int64_t pass_splits(const slemm_geno* h, int pass, int slot, int64_t& nsp) {
  const BpChoice c = bp_choice(h, pass, slot);
  return pass == 0 ? lds_splits(h->m_pad, h->n_pad / kPad, c.target, nsp)
                   : lds_splits(h->n_pad, h->m_pad / kPad, c.target, nsp);
}

// Z'X1 pass, epilogue T = W (Zc'x) -> pass-2 operand x2 / out / upart.
void pass1_lds(slemm_geno* h, const Layout& L, int t, bool weighted, void* x2, double* upart, float* out,
               int64_t ldo, hipStream_t st, bool raw, bool squared, const TraitCols* tc) {
  int64_t nsp = 0;
  const int64_t nsuper = h->n_pad / kPad;
  const bool i8 = L.i8;
  int rt = 0;
  const int S = i8 ? i8_pass_splits(h, 0, t, h->m_pad, nsuper, squared, nsp, rt) : (int)pass_splits(h, 0, L.slot, nsp);
  h->partbuf.reserve(sizeof(float) * (size_t)S * L.KP * h->m_pad);
  LdsArgs la;
  la.G = h->gs; la.ldg = h->ld_gs; la.nsuper = nsuper; la.ns_per_split = nsp; la.X = h->x1.p;
  la.lo_offset16 = L.lo1_16; la.KP = L.KP; la.rows_pad = h->m_pad; la.part = h->partbuf.as<float>();
  la.squared = squared;
  tail_args(L, t, la);
  if (tc) la.tile_nsuper = tc->tile_nsuper;  // reduction over individuals stops at the subsets' end
  TimingToken tok = timing_begin(st);
  if (i8)  // the centring sums s_k of the epilogue become those of the digits
    bp_i8(h, h->gs, h->ld_gs, h->n_pad, h->m_pad, h->x1.p, L, t, squared, S, nsp, rt, st, h->colsum.as<double>(),
          I8SumWeights{}, la.tile_nsuper);
  else launch_bp(la, L.NT, S, bp_choice(h, 0, L.slot).variant, st);
  timing_end(tok, st, EPI_PASS1, PREC_BF16X2);
  Fin1Args f;
  f.part = h->partbuf.as<float>(); f.nsplit = S; f.rows_pad = h->m_pad; f.rows = h->m; f.KP = L.KP; f.t = t;
  f.colsum = h->colsum.as<double>(); f.xsum = h->colsum.as<double>() + 2 * L.KP; f.xsum_n = (double)h->n;
  f.popcnt = h->popcnt; f.gm = h->gm; f.gw = weighted ? h->gw : nullptr; f.X2 = x2; f.lo_offset16 = L.lo2_16;
  f.upart = upart; f.out = out; f.ldo = ldo; f.raw = raw; f.tail0 = L.tail0; f.tail_r = L.tail_r;
  f.x2_fp32 = L.i8;
  if (tc) f.tc = *tc;
  launch_fin1(f, st);
}

// Z X2 pass, epilogue out = rr (P - u) (+ tau0 v for LDET columns).
void pass2_lds(slemm_geno* h, const Layout& L, int t, double* u, const float* rr, const float* V,
               int64_t ldv, float tau0, const int* kinds, float* out, int64_t ldo, hipStream_t st,
               const TraitCols* tc) {
  int64_t nsp = 0;
  const int64_t nsuper = h->m_pad / kPad;
  const bool i8 = L.i8;
  int rt = 0;
  const int S = i8 ? i8_pass_splits(h, 1, t, h->n_pad, nsuper, false, nsp, rt) : (int)pass_splits(h, 1, L.slot, nsp);
  h->partbuf.reserve(sizeof(float) * (size_t)S * L.KP * h->n_pad);
  LdsArgs la;
  la.G = h->gi; la.ldg = h->ld_gi; la.nsuper = nsuper; la.ns_per_split = nsp; la.X = h->x2.p;
  la.lo_offset16 = L.lo2_16; la.KP = L.KP; la.rows_pad = h->n_pad; la.part = h->partbuf.as<float>();
  tail_args(L, t, la);
  if (tc) la.tile_rows = tc->tile_rows;  // output rows (individuals) past the subsets' end are zero
  TimingToken tok = timing_begin(st);
  if (i8) {  // u_k becomes sum_j gm_j (digit operand)_jk, gm of the column's trait for batched traits
    I8SumWeights sw;
    sw.w = tc ? tc->gm : h->gm;
    if (tc) {
      sw.col_row = tc->col_trait;
      sw.ldw = tc->ld_m;
    }
    bp_i8(h, h->gi, h->ld_gi, h->m_pad, h->n_pad, h->x2.p, L, t, false, S, nsp, rt, st, u, sw, nullptr, la.tile_rows);
  } else {
    launch_bp(la, L.NT, S, bp_choice(h, 1, L.slot).variant, st);
  }
  timing_end(tok, st, EPI_PASS2, PREC_BF16X2);
  Fin2Args f;
  f.part = h->partbuf.as<float>(); f.nsplit = S; f.rows_pad = h->n_pad; f.rows = h->n; f.KP = L.KP; f.t = t;
  f.colsum = u; f.rr = rr; f.v = V; f.ldv = ldv; f.tau0 = tau0; f.kinds = kinds; f.out = out; f.ldo = ldo;
  f.tail0 = L.tail0; f.tail_r = L.tail_r;
  if (tc) f.tc = *tc;
  launch_fin2(f, st);
}

void check_dense(const float* p, int64_t ld, int64_t rows, int t, const char* what) {
  if (!p) fail(SLEMM_ERR_ARG, std::string(what) + " is null");
  if (ld < rows) fail(SLEMM_ERR_ARG, std::string("leading dimension of ") + what + " is too small");
  if (t <= 0) fail(SLEMM_ERR_ARG, "number of columns must be positive");
}

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
  // the span [qlo, qlo + tq) of the QUADFORM columns: the projections and the finish run on it
  // alone (LDET columns, the probes, need neither)
  int qlo = 0, tq = 0;
  if (any_quad) {
    int qhi = -1;
    qlo = t;
    for (int k = 0; k < t; ++k)
      if (h->kinds_host[k] == SLEMM_OP_QUADFORM) {
        qlo = std::min(qlo, k);
        qhi = k;
      }
    tq = qhi - qlo + 1;
  }

  const double* qtv = nullptr;
  if (any_quad && h->c > 0) {
    // entries k in the span are written; pack_pass1 reads them for QUADFORM columns only
    h->qtv.reserve(sizeof(double) * h->c * L.KP);
    tn_product(h->Q, h->n_pad, h->c, V + (int64_t)qlo * ldv, ldv, tq, h->n, h->qtv.as<double>() + qlo, L.KP,
               h->redpart, st);
    qtv = h->qtv.as<double>();
  }
  pack_pass1(h, L, V, ldv, t, h->rr, qtv, kinds, precision, st);

  // pass 1: T = W (Z' X1 - gm s'), written as the pass-2 operand
  h->x2.reserve(L.x2_bytes);
  if (use_lds(precision)) {
    const int64_t nparts = h->m_pad / kFin1Rows;
    h->part2.reserve(sizeof(double) * nparts * L.KP);
    {  // size the split-partial buffer for both passes before anything is in flight
      int64_t nsp = 0;
      const int64_t S1 = pass_splits(h, 0, L.slot, nsp);
      const int64_t S2 = pass_splits(h, 1, L.slot, nsp);
      h->partbuf.reserve(sizeof(float) * L.KP *
                         std::max<size_t>((size_t)S1 * h->m_pad, (size_t)S2 * h->n_pad));
    }
    pass1_lds(h, L, t, true, h->x2.p, h->part2.as<double>(), nullptr, 0, st);
    double* u = h->colsum.as<double>() + L.KP;
    launch_sum_parts(h->part2.as<double>(), nparts, L.KP, u, st);
    pass2_lds(h, L, t, u, h->rr, V, ldv, tau0, kinds, out, ldo, st);
  } else {
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
  }

  if (any_quad) {
    double* bq = nullptr;
    if (h->c > 0) {
      h->qtv.reserve(sizeof(double) * h->c * L.KP);
      tn_product(h->Q, h->n_pad, h->c, out + (int64_t)qlo * ldo, ldo, tq, h->n, h->qtv.as<double>() + qlo, L.KP,
                 h->redpart, st);
      bq = h->qtv.as<double>() + qlo;
    }
    launch_quad_finish(out + (int64_t)qlo * ldo, ldo, V + (int64_t)qlo * ldv, ldv, h->Q, h->n_pad, h->c, bq, L.KP,
                       kinds + qlo, tq, h->n, tau0, st);
  }
}

// Upper bound of the device workspace run_op reserves for a t-column call (the sum of every
// reserve it makes, each 256-byte aligned, counting a buffer that grows between the passes at
// both sizes).  The no-allocation entry point checks a call against its scratch with it.
size_t op_workspace_bytes(const slemm_geno* h, int kind, const int* col_kind, int t, int precision) {
  const Layout L = layout_for(h, t, precision);
  bool any_quad = false;
  for (int k = 0; k < t; ++k) {
    const int kk = col_kind ? col_kind[k] : kind;
    REQUIRE(kk == SLEMM_OP_LDET || kk == SLEMM_OP_QUADFORM, "Unknown linear operator in lnr_op_select()");
    any_quad |= kk == SLEMM_OP_QUADFORM;
  }
  size_t total = 0;
  auto add = [&](size_t b) { total += (b + 255) & ~(size_t)255; };
  add(sizeof(int) * t);  // kinds
  if (any_quad && h->c > 0) {
    add(sizeof(double) * h->c * L.KP);                                  // qtv
    add(sizeof(double) * ((h->n + 16383) / 16384) * h->c * t);          // redpart (Q'V, Q'out)
  }
  add(L.x1_bytes);
  add(sizeof(double) * (h->n_pad / kPackRows) * L.KP);                   // part1
  add(sizeof(double) * 3 * L.KP);                                        // colsum
  add(L.x2_bytes);
  if (use_lds(precision)) {
    add(sizeof(double) * (h->m_pad / kFin1Rows) * L.KP);                 // part2
    int64_t nsp = 0;
    const size_t S1 = (size_t)pass_splits(h, 0, L.slot, nsp), S2 = (size_t)pass_splits(h, 1, L.slot, nsp);
    add(sizeof(float) * L.KP * std::max(S1 * h->m_pad, S2 * h->n_pad));  // partbuf, sized up front
    if (L.i8) {
      int rt = 0;
      const size_t I1 = (size_t)i8_pass_splits(h, 0, t, h->m_pad, h->n_pad / kPad, false, nsp, rt);
      const size_t I2 = (size_t)i8_pass_splits(h, 1, t, h->n_pad, h->m_pad / kPad, false, nsp, rt);
      add(sizeof(float) * L.KP * I1 * h->m_pad);                          // partbuf grown by the int8 splits,
      add(sizeof(float) * L.KP * I2 * h->n_pad);                          // once per pass at most
      add(i8_scratch_bytes(h->n_pad, t));                                 // digits, pass 1
      add(i8_scratch_bytes(h->m_pad, t));                                 // digits, pass 2
    }
  } else {
    add(sizeof(double) * (h->m_pad / blockprod_rows_per_wg(L.NT)) * L.KP);  // part2
  }
  return total;
}

// --------------------------------------------------------------------------- GEBV ----------
// One thread per individual over the individual-major copy: each thread walks its own row of
// codes SNP by SNP, so every individual's FP64 sum is accumulated in exactly the reference's
// order (tgv += z*alpha_j, src/pred.cpp:128).  The chain of dependent FP64 adds is the bound (n
// chains of m steps), so a step is kept to a bit-field extract of the 2-bit code, its LDS
// address, one 8-byte LDS read and the add: a table of the four values z(code) * alpha_j per SNP
// (z = popcount(code), 2 - z for a flipped SNP; the products are exact and 0 * alpha keeps its
// sign, as the reference's z*alpha does) is staged in LDS, 128 SNPs (4 KB) per chunk,
// double-buffered and filled one chunk ahead with in-order vector loads.  Codes come in chunks
// of 128 SNPs (32 bytes of the row per thread): each iteration takes the chunk loaded during
// the previous one and at once issues the load of the next (clamped to the last chunk, so it is
// unconditional).  Table padding past m is zero, so padding SNPs add +0 (the sum never becomes
// -0: it starts at +0 and x + (-x) is +0).  Staging the alphas in LDS rather than reading them
// with scalar loads matters at low occupancy (100k individuals = 1.5 waves per SIMD): scalar
// loads can only be waited for all together, which left an L2 round trip per few steps.
__global__ __launch_bounds__(256) void k_gebv(const uint8_t* __restrict__ gi, int64_t ld_gi, int64_t nchunk,
                                              int64_t n, const double* __restrict__ table,
                                              double* __restrict__ out) {
  __shared__ double tab[2][128 * 4];
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * 256 + tid;  // < n_pad: rows exist in gi
  const uint4* row = reinterpret_cast<const uint4*>(gi + i * ld_gi);
  const uint4* tsrc = reinterpret_cast<const uint4*>(table);  // 4 KB = 256 x 16 bytes per chunk
  const int64_t last = nchunk - 1;
  double acc = 0.0;
  reinterpret_cast<uint4*>(tab[0])[tid] = tsrc[tid];
  uint4 n0 = row[0], n1 = row[1];
  uint4 tn = tsrc[min<int64_t>(1, last) * 256 + tid];
  __syncthreads();
  for (int64_t c = 0; c < nchunk; ++c) {
    const uint4 c0 = n0, c1 = n1;
    const uint4 tc = tn;
    const int64_t cn = min(c + 1, last), cn2 = min(c + 2, last);
    n0 = row[cn * 2];
    n1 = row[cn * 2 + 1];
    __builtin_amdgcn_sched_barrier(0);
    // chunk c+1's table, loaded during chunk c-1, into the buffer chunk c-1 used
    if (c + 1 < nchunk) reinterpret_cast<uint4*>(tab[(c + 1) & 1])[tid] = tc;
    tn = tsrc[cn2 * 256 + tid];
    const double* t = tab[c & 1];
    const uint32_t w[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
#pragma unroll
      for (int k = 0; k < 16; ++k) acc += t[(q * 16 + k) * 4 + __builtin_amdgcn_ubfe(w[q], 2 * k, 2)];
    }
    __builtin_amdgcn_sched_barrier(0);
    // no instruction: a use of the next chunk here keeps its load in this iteration (left to
    // itself the compiler moves it into the next one, right before the first step)
    asm volatile("" ::"v"(n0.x), "v"(n0.y), "v"(n0.z), "v"(n0.w), "v"(n1.x), "v"(n1.y), "v"(n1.z), "v"(n1.w));
    __syncthreads();
  }
  if (i < n) out[i] = acc;
}

// table[4j + code] = z(code) * alpha_j with z = popcount(code), 2 - z for a flipped SNP (flip
// words: bits 2k, 2k+1 of word j/16); zeros past m
__global__ void k_gebv_table(const double* __restrict__ alpha, const uint32_t* __restrict__ flipw, int64_t m,
                             int64_t len4, double* __restrict__ table) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len4) return;
  const int64_t j = e >> 2;
  const uint32_t code = (uint32_t)(e & 3);
  if (j >= m) {
    table[e] = 0.0;
    return;
  }
  const uint32_t f = (flipw[j >> 4] >> (2 * (j & 15))) & 3u;
  table[e] = (double)__builtin_popcount(code ^ f) * alpha[j];
}

// GRAMMAR-gamma score columns: [Py, rr o Q_1..c, rr^2] (n x (2+c), column-major, ld n).
__global__ void k_gwas_cols(const float* __restrict__ Py, const float* __restrict__ rr, const float* __restrict__ Q,
                            int64_t ldq, int c, int64_t n, float* __restrict__ V) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float r = rr[i];
  V[i] = Py[i];
  for (int q = 0; q < c; ++q) V[i + (int64_t)(1 + q) * n] = r * Q[i + (int64_t)q * ldq];
  V[i + (int64_t)(1 + c) * n] = r * r;
}

// xy_j = x_c'Py and xx_j = ||(I-QQ')R^-1/2 x_c||^2 = sum rr^2 (d-mu)^2 - ||Q'(rr o x_c)||^2 with
// sum rr^2 (d-mu)^2 = S2 - 2 mu C1 - mu^2 R2  (S2 = sum rr^2 d^2, C1 = Zc'(rr^2), R2 = sum rr^2).
// Without error-variance weights (S2 null) S2 is the exact count sum popcnt + 2 n2.
__global__ void k_gwas_combine(const float* __restrict__ A, int64_t lda, int c, const float* __restrict__ S2,
                               const int32_t* __restrict__ cnt2, const float* __restrict__ gm,
                               const int32_t* __restrict__ pc, int64_t m, int64_t n, double R2, float mac,
                               float* __restrict__ xy, float* __restrict__ xx) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const float s = (float)pc[j];
  if (s <= mac || 2.0f * (float)n - s <= mac) {
    xy[j] = __builtin_nanf("");
    xx[j] = __builtin_nanf("");
    return;
  }
  const double mu = gm[j];
  double proj = 0.0;
  for (int q = 0; q < c; ++q) {
    const double a = A[j + (int64_t)(1 + q) * lda];
    proj += a * a;
  }
  const double C1 = A[j + (int64_t)(1 + c) * lda];
  const double s2 = S2 ? (double)S2[j] : (double)pc[j] + 2.0 * (double)cnt2[j];
  xy[j] = A[j];
  xx[j] = (float)(s2 - 2.0 * mu * C1 - mu * mu * R2 - proj);
}

// Column dots, stage 1: block (k, b) sums A[i,k] B[i,k] over rows [b*rpb, (b+1)*rpb) in FP64.
__global__ __launch_bounds__(256) void k_col_dots(const float* __restrict__ A, int64_t lda,
                                                  const float* __restrict__ B, int64_t ldb, int64_t n,
                                                  int64_t rpb, int t, double* __restrict__ part) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  const int64_t i0 = (int64_t)blockIdx.y * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256)
    s += (double)A[i + (int64_t)k * lda] * (double)B[i + (int64_t)k * ldb];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * t + k] = red[0];
}

// stage 2: res[k] = sum_b part[b][k], fixed order (deterministic).
__global__ void k_col_dots_final(const double* __restrict__ part, int64_t nblk, int t, double* __restrict__ res) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t) return;
  double s = 0.0;
  for (int64_t b = 0; b < nblk; ++b) s += part[b * t + k];
  res[k] = s;
}

void col_dots(slemm_geno* h, const float* A, int64_t lda, const float* B, int64_t ldb, int64_t n, int t,
              double* res_dev, hipStream_t st) {
  if (t <= 0) return;
  const int64_t rpb = 16384;
  const int64_t nblk = (n + rpb - 1) / rpb;
  h->redpart.reserve(sizeof(double) * nblk * t);
  k_col_dots<<<dim3((unsigned)t, (unsigned)nblk), 256, 0, st>>>(A, lda, B, ldb, n, rpb, t, h->redpart.as<double>());
  HIPCHECK(hipGetLastError());
  k_col_dots_final<<<(t + 255) / 256, 256, 0, st>>>(h->redpart.as<double>(), nblk, t, res_dev);
  HIPCHECK(hipGetLastError());
}

}  // namespace slemm

using namespace slemm;

static_assert(2 + 2 * BP_NVARIANTS * 3 == SLEMM_AUTOTUNE_MS, "autotune timing table size");

// Pseudo-random BF16 pairs for autotuning: each 32-bit word holds two BF16 values in (-1, 1)
// (hi halves of the operand) or about 2^-9 of that (lo halves: the upper half of the buffer).
static __global__ void k_fill_split(uint32_t* __restrict__ w, size_t nwords, uint32_t salt) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x) {
    uint32_t x = (uint32_t)i * 0x9E3779B9u ^ (salt * 0x85EBCA6Bu);
    x ^= x >> 15; x *= 0x2C1B3C6Du; x ^= x >> 12; x *= 0x297A2D39u; x ^= x >> 15;
    const uint32_t ex = i < nwords / 2 ? 0x3F00u : 0x3B00u;  // BF16 exponents of [0.5,1) / 2^-9 scale
    const uint32_t a = ex | (x & 0x7Fu) | ((x & 0x100u) << 7);
    const uint32_t b = ex | ((x >> 16) & 0x7Fu) | ((x & 0x200u) << 6);
    w[i] = a | (b << 16);
  }
}

// FP32 operand values in (-1, 1) for autotuning the int8 kernels
static __global__ void k_fill_f32(float* __restrict__ x, size_t n, uint32_t salt) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint32_t v = (uint32_t)i * 0x9E3779B9u ^ (salt * 0x85EBCA6Bu);
    v ^= v >> 15; v *= 0x2C1B3C6Du; v ^= v >> 12; v *= 0x297A2D39u; v ^= v >> 15;
    x[i] = (float)((int32_t)v) * 4.656613e-10f;
  }
}

// Autotuning of a t-vector block that runs on the int8 kernels: per pass, the workgroup target of
// the reduction split (1024 ... 16384) and, for narrow blocks, the row tiles per wave (4 or 8),
// each timed over full passes (digit preparation + block product) on random FP32 operands; the
// fastest is kept on the handle for its kernel class (i8_class) and pass.
static void autotune_i8(slemm_geno* h, const Layout& L, int t, int reps, int* variant, int* target, double* best_ms) {
  h->x1.reserve(L.x1_bytes);
  h->x2.reserve(L.x2_bytes);
  k_fill_f32<<<4096, 256>>>(h->x1.as<float>(), L.x1_bytes / 4, 1);
  k_fill_f32<<<4096, 256>>>(h->x2.as<float>(), L.x2_bytes / 4, 2);
  HIPCHECK(hipGetLastError());
  const int cls = i8_class(t);
  const int64_t targets[] = {1024, 2048, 4096, 8192, 16384};
  const int rts[] = {4, 8};
  const int nrt = cls == 0 ? 2 : 1;
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  for (int pass = 0; pass < 2; ++pass) {
    const int64_t rows_pad = pass == 0 ? h->m_pad : h->n_pad, red_pad = pass == 0 ? h->n_pad : h->m_pad;
    double best = 1e300;
    int brt = 0;
    int64_t bt = 0;
    for (int ri = 0; ri < nrt; ++ri)
      for (int ti = 0; ti < 5; ++ti) {
        const int rt = cls == 0 ? rts[ri] : 0;
        int64_t nsp = 0;
        const int S = i8_splits(t, rows_pad, red_pad / kPad, targets[ti], false, rt, nsp);
        h->partbuf.reserve(sizeof(float) * (size_t)S * L.KP * rows_pad);
        h->i8buf.reserve(i8_scratch_bytes(red_pad, t));
        I8Launch l;
        l.G = pass == 0 ? h->gs : h->gi; l.ldg = pass == 0 ? h->ld_gs : h->ld_gi;
        l.red_pad = red_pad; l.rows_pad = rows_pad; l.X = pass == 0 ? h->x1.p : h->x2.p; l.KP = L.KP; l.t = t;
        l.nsplit = S; l.ns_per_split = nsp; l.scratch = h->i8buf.p; l.part = h->partbuf.as<float>(); l.rt = rt;
        launch_bp_i8(l, nullptr);  // warm-up
        HIPCHECK(hipEventRecord(e0, nullptr));
        for (int r = 0; r < reps; ++r) launch_bp_i8(l, nullptr);
        HIPCHECK(hipEventRecord(e1, nullptr));
        HIPCHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        const double avg = ms / reps;
        if (best_ms) best_ms[2 + pass * (BP_NVARIANTS * 3) + ri * 5 + ti] = avg;
        if (avg < best) {
          best = avg;
          brt = rt;
          bt = targets[ti];
        }
      }
    h->i8_target[pass][cls] = (int)bt;
    if (cls == 0) h->i8_rt[pass] = brt;
    if (variant) variant[pass] = brt;
    if (target) target[pass] = (int)bt;
    if (best_ms) best_ms[pass] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

// MFMA instructions a t-column block's product issues per 16 x 64 tile of the output rows and
// reduction (int8 16x16x64: one per digit tile; split-BF16 16x16x32: two per 16-column tile with
// hi and lo as columns, four with separate hi and lo tiles; FP32: one unit per 16-column tile,
// compared only with FP32 blocks): what dropping columns from a block can save.
int slemm::op_column_tiles(const slemm_geno* h, int t, int precision) {
  const Layout L = layout_for(h, t, precision);
  if (L.i8) return (i8_digits(t).ndc + 15) / 16;
  const int tiles = (int)(L.KP / 16);
  if (precision != PREC_BF16X2) return tiles;
  return (L.hl ? 2 : 4) * tiles;
}

extern "C" int slemm_gpu_block_kernel(const slemm_geno* h, int t, int precision, int* kernel) {
  return guard([&] {
    REQUIRE(h && kernel, "null argument");
    REQUIRE(t > 0, "t must be positive");
    check_precision(precision);
    const Layout L = layout_for(h, t, precision);
    *kernel = L.i8 ? SLEMM_KERNEL_I8_NARROW + i8_class(t) : precision == PREC_BF16X2 ? SLEMM_KERNEL_BF16X2 : SLEMM_KERNEL_FP32;
  });
}

extern "C" int slemm_gpu_autotune(slemm_geno* h, int t, int reps, int* variant, int* target, double* best_ms) {
  return guard([&] {
    REQUIRE(h, "null handle");
    REQUIRE(t > 0 && reps > 0, "t and reps must be positive");
    HIPCHECK(hipSetDevice(h->device));
    if (best_ms)
      for (int e = 0; e < SLEMM_AUTOTUNE_MS; ++e) best_ms[e] = -1.0;
    {
      const Layout Li = layout_for(h, t, PREC_BF16X2);
      if (Li.i8) {
        autotune_i8(h, Li, t, reps, variant, target, best_ms);
        return;
      }
    }
    const Layout L = layout_for(h, t, PREC_BF16X2, false, false);  // the split-BF16 kernels only
    h->x1.reserve(L.x1_bytes);
    h->x2.reserve(L.x2_bytes);
    // operands with the value spread of real split-BF16 blocks (hi ~ U(-1,1), lo ~ 2^-9 hi): zero
    // operands draw less power and would favour variants that only win at boosted clocks
    k_fill_split<<<4096, 256>>>(reinterpret_cast<uint32_t*>(h->x1.p), L.x1_bytes / 4, 1);
    k_fill_split<<<4096, 256>>>(reinterpret_cast<uint32_t*>(h->x2.p), L.x2_bytes / 4, 2);
    HIPCHECK(hipGetLastError());
    const int64_t targets[] = {512, 1024, 2048};
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    for (int pass = 0; pass < 2; ++pass) {
      const int64_t rows_pad = pass == 0 ? h->m_pad : h->n_pad;
      const int64_t nsuper = (pass == 0 ? h->n_pad : h->m_pad) / kPad;
      double best = 1e300;
      int bv = 0;
      int64_t bt = 1024;
      for (int v = 0; v < BP_NVARIANTS; ++v)
        for (int ti = 0; ti < 3; ++ti) {
          const int64_t tg = targets[ti];
          if (L.tail_r > 0 && !bp_hl_variant(v)) {  // not implemented for hi/lo-in-columns blocks
            if (best_ms) best_ms[2 + pass * (BP_NVARIANTS * 3) + v * 3 + ti] = -1.0;
            continue;
          }
          int64_t nsp = 0;
          const int S = lds_splits(rows_pad, nsuper, tg, nsp);
          h->partbuf.reserve(sizeof(float) * (size_t)S * L.KP * rows_pad);
          LdsArgs la;
          la.G = pass == 0 ? h->gs : h->gi;
          la.ldg = pass == 0 ? h->ld_gs : h->ld_gi;
          la.nsuper = nsuper; la.ns_per_split = nsp; la.X = pass == 0 ? h->x1.p : h->x2.p;
          la.lo_offset16 = pass == 0 ? L.lo1_16 : L.lo2_16; la.KP = L.KP; la.rows_pad = rows_pad;
          la.part = h->partbuf.as<float>();
          tail_args(L, t, la);
          launch_bp(la, L.NT, S, v, nullptr);  // warm-up
          HIPCHECK(hipEventRecord(e0, nullptr));
          for (int r = 0; r < reps; ++r) launch_bp(la, L.NT, S, v, nullptr);
          HIPCHECK(hipEventRecord(e1, nullptr));
          HIPCHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
          const double avg = ms / reps;
          if (best_ms) best_ms[2 + pass * (BP_NVARIANTS * 3) + v * 3 + ti] = avg;
          if (avg < best) {
            best = avg;
            bv = v;
            bt = tg;
          }
        }
      h->bp_variant[pass][L.slot] = bv;
      h->bp_target[pass][L.slot] = (int)bt;
      if (variant) variant[pass] = bv;
      if (target) target[pass] = (int)bt;
      if (best_ms) best_ms[pass] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

extern "C" int slemm_gpu_gwas_score(slemm_geno* h, const float* Py, float skip_mac, float* xy, float* xx,
                                    void* stream) {
  return guard([&] {
    REQUIRE(h && Py && xy && xx, "null argument");
    HIPCHECK(hipSetDevice(h->device));
    require_model(h);
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = h->n, m = h->m;
    const int c = h->c;
    const int t = 2 + c;
    DevBuf cols, A, B;
    cols.reserve(sizeof(float) * n * t);
    A.reserve(sizeof(float) * m * t);
    B.reserve(sizeof(float) * m);
    k_gwas_cols<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(Py, h->rr, h->Q, h->n_pad, c, n, cols.as<float>());
    HIPCHECK(hipGetLastError());
    // centred products Zc'[Py, rr o Q, rr^2]
    const Layout LA = layout_for(h, t, PREC_BF16X2);
    pack_pass1(h, LA, cols.as<float>(), n, t, nullptr, nullptr, nullptr, PREC_BF16X2, st);
    pass1_lds(h, LA, t, false, nullptr, nullptr, A.as<float>(), m, st);
    // raw squared-dosage product Z2'(rr^2): a second pass over the bed, needed only with
    // error-variance weights (otherwise the dosage counts give it exactly)
    if (!h->rr_unit) {
      const Layout LB = layout_for(h, 1, PREC_BF16X2);
      pack_pass1(h, LB, cols.as<float>() + (size_t)(1 + c) * n, n, 1, nullptr, nullptr, nullptr, PREC_BF16X2, st,
                 false);
      pass1_lds(h, LB, 1, false, nullptr, nullptr, B.as<float>(), m, st, true, true);
    }
    double R2 = 0.0;
    for (int64_t i = 0; i < n; ++i) R2 += (double)h->rr_host[i] * h->rr_host[i];
    k_gwas_combine<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(A.as<float>(), m, c,
                                                                h->rr_unit ? nullptr : B.as<float>(), h->cnt2,
                                                                h->gm, h->popcnt, m, n, R2, skip_mac, xy, xx);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
  });
}

extern "C" int slemm_gpu_op(slemm_geno* h, int kind, const int* col_kind, float tau0, const float* V,
                            int64_t ldv, int t, float* out, int64_t ldo, int precision, void* stream) {
  return guard([&] {
    REQUIRE(h, "null handle");
    HIPCHECK(hipSetDevice(h->device));
    run_op(h, kind, col_kind, tau0, V, ldv, t, out, ldo, precision, (hipStream_t)stream);
  });
}

namespace slemm {

// (W) Zc' (R^-1/2) V -> out (m x t), one pass over the SNP-major copy.
void run_zt(slemm_geno* h, const float* V, int64_t ldv, int t, float* out, int64_t ldo, bool weighted,
            bool scale_rr, int precision, hipStream_t st) {
  check_precision(precision);
  check_dense(V, ldv, h->n, t, "V");
  check_dense(out, ldo, h->m, t, "out");
  ensure_counts(h, st);
  if (weighted || scale_rr) require_model(h);
  const Layout L = layout_for(h, t, precision);
  pack_pass1(h, L, V, ldv, t, scale_rr ? h->rr : nullptr, nullptr, nullptr, precision, st);
  if (use_lds(precision)) {
    pass1_lds(h, L, t, weighted, nullptr, nullptr, out, ldo, st);
    return;
  }
  BlockProdArgs a;
  a.G = h->gs; a.ldg = h->ld_gs; a.nsuper = h->n_pad / kPad; a.X = h->x1.p; a.lo_offset16 = L.lo1_16;
  a.KP = L.KP; a.t = t; a.rows = h->m; a.colsum = h->colsum.as<double>();
  a.gm = h->gm; a.gw = weighted ? h->gw : nullptr; a.out = out; a.ldo = ldo;
  centring_args(h, L, a);
  launch_blockprod(a, precision, EPI_PASS1, L.NT, h->m_pad, st);
}

namespace {

// out_j = sum_k coef_k T_jk T_j(t+k), FP64 accumulation over the k columns.
__global__ void k_adjoint_combine(const float* __restrict__ T, int64_t ldt, int t, const double* __restrict__ coef,
                                  int64_t m, double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  double acc = 0.0;
  for (int k = 0; k < t; ++k) acc += coef[k] * ((double)T[j + k * ldt] * (double)T[j + (k + t) * ldt]);
  out[j] = acc;
}

}  // namespace
}  // namespace slemm

extern "C" int slemm_gpu_zt(slemm_geno* h, const float* V, int64_t ldv, int t, float* out, int64_t ldo,
                            int weighted, int scale_rr, int precision, void* stream) {
  return guard([&] {
    REQUIRE(h, "null handle");
    HIPCHECK(hipSetDevice(h->device));
    run_zt(h, V, ldv, t, out, ldo, weighted != 0, scale_rr != 0, precision, (hipStream_t)stream);
  });
}

extern "C" int slemm_gpu_adjoint(slemm_geno* h, const float* A, int64_t lda, const float* B, int64_t ldb, int t,
                                 const double* coef, int scale_rr, double* out, int precision, void* stream) {
  return guard([&] {
    REQUIRE(h && coef && out, "null argument");
    HIPCHECK(hipSetDevice(h->device));
    hipStream_t st = (hipStream_t)stream;
    check_dense(A, lda, h->n, t, "A");
    check_dense(B, ldb, h->n, t, "B");
    const int64_t n = h->n, m = h->m;
    // [A | B] as one n x 2t operand: both sides of every product share the pass over the bed
    DevBuf ab, zt, dc;
    ab.reserve(sizeof(float) * n * 2 * t);
    HIPCHECK(hipMemcpy2DAsync(ab.p, sizeof(float) * n, A, sizeof(float) * lda, sizeof(float) * n, t,
                              hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpy2DAsync(ab.as<float>() + (size_t)n * t, sizeof(float) * n, B, sizeof(float) * ldb,
                              sizeof(float) * n, t, hipMemcpyDeviceToDevice, st));
    zt.reserve(sizeof(float) * m * 2 * t);
    run_zt(h, ab.as<float>(), n, 2 * t, zt.as<float>(), m, false, scale_rr != 0, precision, st);
    dc.reserve(sizeof(double) * t);
    HIPCHECK(hipMemcpyAsync(dc.p, coef, sizeof(double) * t, hipMemcpyHostToDevice, st));
    k_adjoint_combine<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(zt.as<float>(), m, t, dc.as<double>(), m, out);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
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
    const int64_t nparts = h->m_pad / kPackRows;
    h->part2.reserve(sizeof(double) * nparts * L.KP);
    h->colsum.reserve(sizeof(double) * 3 * L.KP);
    PackArgs p;
    p.V = U; p.ldv = ldu; p.t = t; p.KP = L.KP; p.n = h->m; p.rows_pad = h->m_pad;
    p.mode = L.op_mode; p.X = h->x2.p; p.lo_offset16 = L.lo2_16; p.part = h->part2.as<double>();
    p.wsum = h->gm;
    p.tail0 = L.tail0; p.tail_r = L.tail_r;
    launch_pack(p, st);
    double* u = h->colsum.as<double>() + L.KP;
    launch_sum_parts(h->part2.as<double>(), nparts, L.KP, u, st);
    if (use_lds(precision)) {
      pass2_lds(h, L, t, u, nullptr, nullptr, 0, 0.f, nullptr, out, ldo, st);
      return;
    }
    BlockProdArgs b;
    b.G = h->gi; b.ldg = h->ld_gi; b.nsuper = h->m_pad / kPad; b.X = h->x2.p; b.lo_offset16 = L.lo2_16;
    b.KP = L.KP; b.t = t; b.rows = h->n; b.colsum = u; b.out = out; b.ldo = ldo;
    launch_blockprod(b, precision, EPI_PASS2, L.NT, h->n_pad, st);
  });
}

// Operand of the fast GEBV: u_j = -alpha_j for a flipped SNP, alpha_j otherwise (FP32; zero past
// m), and per block the FP64 sum of 2 alpha_j over the flipped SNPs (2 - z = 2 + (-1) z).
static __global__ void k_gebv_operand(const double* __restrict__ alpha, const uint8_t* __restrict__ flip,
                                      int64_t m, int64_t m_pad, float* __restrict__ u, double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < m_pad; j += (int64_t)gridDim.x * 256) {
    const double aj = j < m ? alpha[j] : 0.0;
    const bool f = j < m && flip && flip[j];
    u[j] = (float)(f ? -aj : aj);
    if (f) acc += 2.0 * aj;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// gebv_i = (Zc u)_i + sum_j gm_j (represented u)_j + sum of the flipped 2 alpha_j, in FP64
static __global__ void k_gebv_finish(const float* __restrict__ zc, int64_t n, const double* __restrict__ gm_u,
                                     const double* __restrict__ part, int nblk, double* __restrict__ gebv) {
  __shared__ double c;
  if (threadIdx.x == 0) {
    double s = gm_u[0];
    for (int b = 0; b < nblk; ++b) s += part[b];
    c = s;
  }
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    gebv[i] = (double)zc[i] + c;
}

// Fast GEBV (no reference counterpart; src/pred.cpp:112-130 is what slemm_gpu_gebv reproduces bit
// for bit): with s_j = -1 for a flipped SNP, z -> 2 - z is 2 + s_j z, so GEBV = Z u + 2 sum_flipped
// alpha with u_j = s_j alpha_j.  Z u is one pass-2 product of the 1-vector block on the block
// kernels (FP32 operand; on the default path int8 digits of its 23-bit rendering with exact
// sums), returned centred, Zc u = Z u - sum_j gm_j u_j: adding back the centring sum the kernels
// used (of the operand as represented) gives Z u of that operand.  Accuracy: FP32 rendering of
// alpha and one FP32 rounding of Zc u, against one HBM pass instead of the reference order's
// dependent FP64 chain per individual.
extern "C" int slemm_gpu_gebv_fast(slemm_geno* h, const double* alpha, const uint8_t* flip, double* gebv,
                                   void* stream) {
  return guard([&] {
    REQUIRE(h && alpha && gebv, "null argument");
    HIPCHECK(hipSetDevice(h->device));
    hipStream_t st = (hipStream_t)stream;
    DevBuf du, dflip, dpart, dzc;
    constexpr int nblk = 256;
    du.reserve(sizeof(float) * h->m_pad);
    dpart.reserve(sizeof(double) * nblk);
    dzc.reserve(sizeof(float) * h->n);
    if (flip) {
      dflip.reserve(h->m);
      HIPCHECK(hipMemcpyAsync(dflip.p, flip, h->m, hipMemcpyHostToDevice, st));
    }
    k_gebv_operand<<<nblk, 256, 0, st>>>(alpha, dflip.as<uint8_t>(), h->m, h->m_pad, du.as<float>(),
                                         dpart.as<double>());
    HIPCHECK(hipGetLastError());
    const int rc = slemm_gpu_z(h, du.as<float>(), h->m, 1, dzc.as<float>(), h->n, PREC_BF16X2, stream);
    if (rc != SLEMM_OK) fail(rc, slemm_gpu_last_error());
    const Layout L = layout_for(h, 1, PREC_BF16X2);
    k_gebv_finish<<<(unsigned)std::min<int64_t>(4096, (h->n + 255) / 256), 256, 0, st>>>(
        dzc.as<float>(), h->n, h->colsum.as<double>() + L.KP, dpart.as<double>(), nblk, gebv);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
  });
}

extern "C" int slemm_gpu_gebv(slemm_geno* h, const double* alpha, const uint8_t* flip, double* gebv,
                              void* stream) {
  return guard([&] {
    REQUIRE(h && alpha && gebv, "null argument");
    HIPCHECK(hipSetDevice(h->device));
    hipStream_t st = (hipStream_t)stream;
    REQUIRE(h->m_pad % 512 == 0 && h->ld_gi % 16 == 0 && h->ld_gi * 4 >= h->m_pad && h->n_pad % 256 == 0,
            "GEBV: unexpected padding");
    const int64_t ngroup = h->m_pad / 16;
    // flip words: bits 2k, 2k+1 of word j/16 set when SNP j (< m) is flipped
    std::vector<uint32_t> fw((size_t)ngroup, 0u);
    if (flip)
      for (int64_t j = 0; j < h->m; ++j)
        if (flip[j]) fw[(size_t)(j >> 4)] |= 3u << (2 * (j & 15));
    DevBuf dfw, dtab;
    dfw.reserve(sizeof(uint32_t) * fw.size());
    HIPCHECK(hipMemcpyAsync(dfw.p, fw.data(), sizeof(uint32_t) * fw.size(), hipMemcpyHostToDevice, st));
    const int64_t len4 = h->m_pad * 4;
    dtab.reserve(sizeof(double) * (size_t)len4);
    k_gebv_table<<<(unsigned)((len4 + 255) / 256), 256, 0, st>>>(alpha, dfw.as<uint32_t>(), h->m, len4,
                                                                   dtab.as<double>());
    HIPCHECK(hipGetLastError());
    k_gebv<<<(unsigned)(h->n_pad / 256), 256, 0, st>>>(h->gi, h->ld_gi, h->m_pad / 128, h->n, dtab.as<double>(),
                                                      gebv);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
  });
}
