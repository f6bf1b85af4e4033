// Block products on the int8 matrix cores: the Z'x / Zx passes of src/reml.cpp:6-233 for the
// probe blocks, Lanczos / CG vectors, single right-hand sides and the GWAS score vectors.  A
// narrow kernel (1 to 5 vectors: one MFMA tile, digits in registers, no LDS) and a wide kernel
// (16-vector tiles of three digit tiles each plus a tail tile, digits staged in LDS) share the
// digit preparation below.
//
// The operand is FP32 (element (row, k) at row KP + k), as the pack and pass-1 epilogue kernels
// write it for these kernels.  Per column it is rendered as a 23-bit fixed-point integer
// X = rint(x 2^22 / s) with s the power of two at or above the column's largest magnitude, split
// into three balanced int8 digits X = 2^16 d0 + 2^8 d1 + d2; per element the rendering error is at
// most 2^-23 s.  The column sums of the represented operand (plain for pass 1, gm-weighted for
// pass 2) are taken from the digits themselves, so the exact-centring terms of the pass
// epilogues match what the MFMAs multiply.  The genotype codes decode in registers to int8
// dosages {0,1,1,2} ({0,1,1,4} for the squared pass) with one shift, one mask and one byte permute
// per four codes, so no table and no LDS are involved.  Each digit is one column of a
// 16 x 16 x 64 int8 MFMA tile, whose int32 sums are exact: |digit| <= 128, so each reduction split
// is kept to at most 2^23 rows (2^22 for the squared pass) and |sum| <= 2^31.  The epilogue
// combines the three digit columns of a vector exactly in 64-bit integers before one rounding to
// FP32.  The split-BF16 kernels instead keep a 16-bit hi + lo operand and round every product sum
// in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <type_traits>

#include "blockprod.hpp"

namespace slemm {
namespace {

typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

constexpr int kI8Blocks = 512;  // column-maximum partials per vector (row ranges of k_i8_colmax)
static_assert(kI8Blocks % 64 == 0, "k_i8_scale gives each lane kI8Blocks / 64 partials");

// Digit columns: vectors [0, 16 nvt) in tiles of 16, digit p of vector k in column
// 16 (3 (k / 16) + p) + k % 16 (the three digits of a vector in the same lane of three tiles);
// the r8 <= 5 vectors after them in one tail tile, digit p of vector 16 nvt + k' in column
// 48 nvt + 3 k' + p.  ndc columns in all (blocks of at most 5 vectors: nvt = 0, ndc = 3 t).
__device__ __forceinline__ int digit_col(int k, int p, int nvt) {
  return k < 16 * nvt ? 16 * (3 * (k >> 4) + p) + (k & 15) : 48 * nvt + 3 * (k - 16 * nvt) + p;
}

// Partial maxima of |x| over the FP32 operand X[row KP + k] (rows [0, rows)): block (x, y) covers
// columns 64 y .. 64 y + 63 (one per lane, so each wave reads 256 contiguous bytes of a row) and
// the x-th of kI8Blocks contiguous row ranges, 4 rows at a time, eight of them in flight per
// thread (independent maxima: the loop is bound by memory latency, not by the compares).  A NaN or
// infinite element makes the partial +inf (fmaxf alone would drop NaN).  pmax[k kI8Blocks + x].
__global__ __launch_bounds__(256) void k_i8_colmax(const float* __restrict__ X, int KP, int t, int64_t rows,
                                                   float* __restrict__ pmax) {
  __shared__ float red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int k = blockIdx.y * 64 + tx;
  const int64_t per = (rows + kI8Blocks - 1) / kI8Blocks;
  const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(rows, r0 + per);
  auto absx = [](float v) { const float a = fabsf(v); return a <= FLT_MAX ? a : INFINITY; };
  float mx = 0.f;
  if (k < t) {
    float m8[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int64_t r = r0 + ty;
    for (; r + 28 < r1; r += 32) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = X[(r + 4 * u) * KP + k];
#pragma unroll
      for (int u = 0; u < 8; ++u) m8[u] = fmaxf(m8[u], absx(v[u]));
    }
    for (; r < r1; r += 4) m8[0] = fmaxf(m8[0], absx(X[r * KP + k]));
#pragma unroll
    for (int u = 0; u < 8; ++u) mx = fmaxf(mx, m8[u]);
  }
  red[ty][tx] = mx;
  __syncthreads();
  if (ty == 0 && k < t) pmax[k * kI8Blocks + blockIdx.x] = fmaxf(fmaxf(red[0][tx], red[1][tx]), fmaxf(red[2][tx], red[3][tx]));
}

// Per-vector scale from the kI8Blocks partial maxima: s_k = 2^e_k, the power of two at or above the
// column's largest magnitude (s = 1 for an all-zero column); ex[k] = e_k and sc[k] = 2^(e_k - 22)
// in double (a float would underflow for columns below 2^-127).  A column with a NaN or infinite
// element (partial +inf) gets ex[k] = INT32_MIN and a NaN scale: zero digits, and all its
// products come out NaN, as the floating-point kernels' do.
// One wave per column (the maximum is exact in any order).
__global__ __launch_bounds__(64) void k_i8_scale(const float* __restrict__ pmax, int t, double* __restrict__ sc,
                                                 int* __restrict__ ex) {
  const int k = blockIdx.x, l = threadIdx.x;
  float mx = 0.f;
#pragma unroll
  for (int u = 0; u < kI8Blocks / 64; ++u) mx = fmaxf(mx, pmax[(int64_t)k * kI8Blocks + l + 64 * u]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  if (l != 0) return;
  if (!(mx <= FLT_MAX)) {
    sc[k] = __builtin_nan("");
    ex[k] = INT32_MIN;
    return;
  }
  int e = 0;
  if (mx > 0.f) {
    (void)frexpf(mx, &e);               // mx = f 2^e with 0.5 <= f < 1, so mx < 2^e
    if (mx == ldexpf(0.5f, e)) e -= 1;  // a power of two: s = mx
  }
  sc[k] = ldexp(1.0, e - 22);
  ex[k] = e;
}

// Digits in MFMA lane order.  Step T (64 reduction rows) of lane group g covers rows
// 512 (T / 8) + 128 g + 16 (T % 8) + j, j < 16, the rows whose codes the block-product kernels'
// lane group g decodes at that step; byte 4 m + b of a lane's 16 bytes holds row j = 4 b + m (the
// order in which the codes decode).  B[(4 T + g) ndc + digit_col(k, p)] holds digit p of vector k:
// X = rint(x 2^(22 - e_k)) = 2^16 d0 + 2^8 d1 + d2 with balanced digits, from the FP32 element x
// (the product x 2^(22 - e_k) is formed in double: exact, and no overflow for tiny columns).
// Thread v: vector k = v mod t of the pair (T, g) = (v / t) / 4, (v / t) mod 4, so that
// neighbouring threads read neighbouring columns of a row and write neighbouring digit columns.
// Each thread also leaves psum[v] = sum_j w_j X_j of its 16 rows (w = 1 without weights, else
// w[row] of the vector's weight row), for the digit column sums.
__global__ __launch_bounds__(256) void k_i8_quant(const float* __restrict__ X, int KP, int t, int nvt, int ndc,
                                                  int64_t nstep, const int* __restrict__ ex, I8SumWeights sw,
                                                  i32x4_t* __restrict__ B, double* __restrict__ psum) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nstep * 4 * t) return;
  const int k = (int)(v % t);
  const int64_t tg = v / t;
  const int64_t T = tg >> 2;
  const int g = (int)(tg & 3);
  const int e = ex[k];
  const bool finite = e != INT32_MIN;
  const double mult = finite ? ldexp(1.0, 22 - e) : 0.0;  // |X| <= 2^22
  const int64_t r0 = (T >> 3) * 512 + g * 128 + (T & 7) * 16;  // first of 16 rows
  const float* w = sw.w ? sw.w + (sw.col_row ? (int64_t)sw.col_row[k] * sw.ldw : 0) : nullptr;
  int X16[16];
  double ws = 0.0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    X16[j] = finite ? (int)rint((double)X[(r0 + j) * KP + k] * mult) : 0;
    ws += w ? (double)w[r0 + j] * (double)X16[j] : (double)X16[j];
  }
  psum[v] = ws;
  uint32_t d[3][4] = {};
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int x = X16[j];
    const int d2 = ((x + 128) & 255) - 128;
    const int x1 = (x - d2) >> 8;
    const int d1 = ((x1 + 128) & 255) - 128;
    const int d0 = (x1 - d1) >> 8;  // |x| <= 2^22: d0 in [-64, 64]
    const int m = j & 3, bb = j >> 2;  // byte 4 m + bb: dword m, byte bb
    d[0][m] |= (uint32_t)(d0 & 255) << (8 * bb);
    d[1][m] |= (uint32_t)(d1 & 255) << (8 * bb);
    d[2][m] |= (uint32_t)(d2 & 255) << (8 * bb);
  }
#pragma unroll
  for (int p = 0; p < 3; ++p)
    B[(4 * T + g) * ndc + digit_col(k, p, nvt)] = i32x4_t{(int)d[p][0], (int)d[p][1], (int)d[p][2], (int)d[p][3]};
}

// Column sums of the represented operand, out[k] = 2^(e_k - 22) sum_q psum[q t + k] over the nq
// per-thread sums of the quantization (the sum over the digits themselves, so the exact-centring
// terms of the pass epilogues match what the MFMAs multiply), NaN for a non-finite column.  Two
// stages in fixed order: block (x, y) sums the y-th of kI8SumBlocks contiguous ranges of q for
// columns 64 x .. 64 x + 63 (each wave reads 512 contiguous bytes), then one wave per column.
constexpr int kI8SumBlocks = 256;
static_assert(kI8SumBlocks % 64 == 0, "the second stage gives each lane kI8SumBlocks / 64 stage sums");

__global__ __launch_bounds__(256) void k_i8_digit_sums1(const double* __restrict__ psum, int64_t nq, int t,
                                                        double* __restrict__ part) {
  __shared__ double red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + tx;
  const int64_t per = (nq + kI8SumBlocks - 1) / kI8SumBlocks;
  const int64_t q0 = (int64_t)blockIdx.y * per, q1 = min(nq, q0 + per);
  double s = 0.0;
  if (k < t) {
    // four independent sums, so that four loads are in flight per thread
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t q = q0 + ty;
    for (; q + 12 < q1; q += 16) {
#pragma unroll
      for (int u = 0; u < 4; ++u) s4[u] += psum[(q + 4 * u) * t + k];
    }
    for (; q < q1; q += 4) s4[0] += psum[q * t + k];
    s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
  }
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && k < t) part[(int64_t)blockIdx.y * t + k] = (red[0][tx] + red[1][tx]) + (red[2][tx] + red[3][tx]);
}

// one wave per column: lane l sums stage sums l, l + 64, ..., then a fixed butterfly
__global__ __launch_bounds__(64) void k_i8_digit_sums2(const double* __restrict__ part, int t,
                                                       const int* __restrict__ ex, double* __restrict__ out) {
  const int k = blockIdx.x, l = threadIdx.x;
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < kI8SumBlocks / 64; ++u) s += part[(int64_t)(l + 64 * u) * t + k];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (l == 0) out[k] = ex[k] == INT32_MIN ? __builtin_nan("") : ldexp(s, ex[k] - 22);
}

struct I8Args {
  const uint8_t* G = nullptr;
  int64_t ldg = 0;
  int64_t nsuper = 0, ns_per_split = 0;
  const i32x4_t* B = nullptr;
  int t = 0;
  const double* sc = nullptr;
  float* part = nullptr;
  int KP = 0;
  int64_t rows_pad = 0;
};

// One workgroup: 4 waves, each on 16 RT rows of one reduction slice (blockIdx.z).  Lane (r, g)
// holds, per super-chunk s (512 reduction rows), bytes [128 s + 32 g, 128 s + 32 g + 32) of its
// rows: word w is step w.  Digits and codes of super-chunk s + 1 are requested before the
// steps of s (digits first, so waiting for them never waits for the codes).
template <int RT, bool SQ>
__global__ __launch_bounds__(256) void k_bp_i8(I8Args a) {
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * (64 * RT) + wave * (16 * RT);
  const int64_t s0 = blockIdx.z * a.ns_per_split;
  const int64_t s1 = min(a.nsuper, s0 + a.ns_per_split);
  const int nc = 3 * a.t;
  const uint8_t* lbase = a.G + (row0 + r) * a.ldg + g * 32;
  const i32x4_t* bbase = a.B + g * nc + (r % nc);  // lanes of dead columns read a live one
  const uint32_t tab = SQ ? 0x04010100u : 0x02010100u;
  i32x4_t acc[RT];
#pragma unroll
  for (int i = 0; i < RT; ++i) acc[i] = i32x4_t{0, 0, 0, 0};
  uint32_t cA[RT][8], cB[RT][8];
  i32x4_t bA[8], bB[8];
  auto ld = [&](int64_t s, uint32_t (&cc)[RT][8], i32x4_t (&bb)[8]) {
    s = s < s1 ? s : s1 - 1;  // past the slice: re-read its last super-chunk (never used)
#pragma unroll
    for (int w = 0; w < 8; ++w) bb[w] = bbase[(s * 8 + w) * 4 * nc];
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const u32x4_t* q = reinterpret_cast<const u32x4_t*>(lbase + (int64_t)i * 16 * a.ldg + s * 128);
      const u32x4_t q0 = q[0], q1 = q[1];
      cc[i][0] = q0.x; cc[i][1] = q0.y; cc[i][2] = q0.z; cc[i][3] = q0.w;
      cc[i][4] = q1.x; cc[i][5] = q1.y; cc[i][6] = q1.z; cc[i][7] = q1.w;
    }
  };
  auto steps = [&](const uint32_t (&cc)[RT][8], const i32x4_t (&bb)[8]) {
#pragma unroll
    for (int w = 0; w < 8; ++w) {
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        const uint32_t W = cc[i][w];
        // byte b of dword m: code 4 b + m of the word, mapped to its dosage
        const i32x4_t av = {(int)__builtin_amdgcn_perm(0u, tab, W & 0x03030303u),
                            (int)__builtin_amdgcn_perm(0u, tab, (W >> 2) & 0x03030303u),
                            (int)__builtin_amdgcn_perm(0u, tab, (W >> 4) & 0x03030303u),
                            (int)__builtin_amdgcn_perm(0u, tab, (W >> 6) & 0x03030303u)};
        acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bb[w], acc[i], 0, 0, 0);
      }
    }
  };
  // the scheduling barriers keep each super-chunk's requests ahead of the previous one's steps
  // (left to itself the compiler sinks them among the steps to save registers, and the in-order
  // memory counter then makes the steps wait for the prefetch)
  if (s0 < s1) ld(s0, cA, bA);
  for (int64_t s = s0; s < s1; s += 2) {
    ld(s + 1, cB, bB);
    __builtin_amdgcn_sched_barrier(0);
    steps(cA, bA);
    __builtin_amdgcn_sched_barrier(0);
    if (s + 1 < s1) {
      ld(s + 2, cA, bA);
      __builtin_amdgcn_sched_barrier(0);
      steps(cB, bB);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // lane (r, g) holds rows 4 g .. 4 g + 3 of column r; column 3 k + p is digit p of vector k.
  // Lanes of column 3 k combine the three digits.
  const int k = r / 3, p = r % 3;
  const bool live = r < nc;
  const double scale = live ? (double)a.sc[k] : 0.0;
#pragma unroll
  for (int i = 0; i < RT; ++i) {
    float o[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int v0 = acc[i][q];
      const int v1 = __shfl(v0, lane + 1);
      const int v2 = __shfl(v0, lane + 2);
      const int64_t tot = (int64_t)v0 * 65536 + (int64_t)v1 * 256 + (int64_t)v2;
      o[q] = (float)((double)tot * scale);
    }
    const int64_t row = row0 + 16 * i + 4 * g;
    if (live && p == 0)
      *reinterpret_cast<float4*>(a.part + (blockIdx.z * (int64_t)a.KP + k) * a.rows_pad + row) =
          make_float4(o[0], o[1], o[2], o[3]);
  }
}

struct I8WArgs {
  const uint8_t* G = nullptr;
  int64_t ldg = 0;
  int64_t nsuper = 0, ns_per_split = 0;
  const i32x4_t* B = nullptr;
  int ndc = 0, nvt = 0, r8 = 0, t = 0;
  const double* sc = nullptr;
  float* part = nullptr;
  int KP = 0;
  int64_t rows_pad = 0;
  int tail0 = 0, tail_r = 0;  // operand layout: vectors k >= tail0 have a lo column k + tail_r
  // optional per 16-vector tile limits (batched traits): reduction super-chunks that can be
  // nonzero, and rows from which the tile's products are zero (both skipped)
  const int32_t* tile_nsuper = nullptr;
  const int64_t* tile_rows = nullptr;
  int vt_base = 0;         // first vector tile of group 0 (second form: a block split over two launches)
  bool xcd_pairs = false;  // dispatch the column groups of a row block back to back on one XCD
  int prio = 1;            // issue priority of a wave while it issues a step's MFMAs (0: none)
};

// Wide blocks: column group blockIdx.y holds vector tiles [NVG y, NVG y + NVG) (three digit tiles
// each, the digits of a vector in one lane) and, in the last group, the tail tile of up to 5
// vectors.  One workgroup: 4 waves, each on 16 RT rows of one reduction slice (blockIdx.z).
// Codes as in the narrow kernel (whole 128-byte lines per row and super-chunk, a super-chunk
// ahead); the group's digits of each half super-chunk (4 steps) are staged in LDS, requested a
// half ahead and before the codes, with one barrier per half.
// TAIL = false (blocks without a tail tile, one vector tile per group): 3 digit tiles, few enough
// registers for three waves per SIMD.
// (An instantiation without the per-tile branches, for groups of NVG full vector tiles without
// tile limits, was tried: the scheduler then hoisted LDS reads and decodes, the kernel spilled and
// ran 20% slower at K=65.)
template <int RT, int NVG, bool TAIL = true>
__global__ __launch_bounds__(256, TAIL ? 2 : 3) void k_bp_i8w(I8WArgs a) {
  constexpr int NDT = 3 * NVG + (TAIL ? 1 : 0);  // digit tiles of a group; with TAIL the last is the tail tile
  __shared__ i32x4_t sb[2][NDT * 16 * 16];  // [half parity][tile j][4 w4 + g][column]
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  // (row block bx, column group by) of this workgroup.  Workgroups are dispatched round-robin
  // over the 8 XCDs in linear order; with xcd_pairs the column groups of one row block take
  // consecutive slots of the same XCD, so all but the first read the codes from that XCD's L2.
  int64_t bx = blockIdx.x;
  int by = blockIdx.y;
  {
    const int64_t gx = gridDim.x, gy = gridDim.y, gx8 = gx - gx % 8;
    const int64_t L = blockIdx.x + gx * blockIdx.y;
    if (a.xcd_pairs && gy > 1 && L < gx8 * gy) {
      const int64_t slot = L / 8;
      bx = (slot / gy) * 8 + L % 8;
      by = (int)(slot % gy);
    } else if (a.xcd_pairs && gy > 1) {  // the last gx % 8 row blocks, in any order
      const int64_t L2 = L - gx8 * gy, rem = gx - gx8;
      bx = gx8 + L2 % rem;
      by = (int)(L2 / rem);
    }
  }
  const int64_t row0 = bx * (64 * RT) + wave * (16 * RT);
  const int64_t wg_row0 = bx * (64 * RT);
  const int vt0 = by * NVG;
  const int nv = min(NVG, a.nvt - vt0);                      // vector tiles of this group
  const bool tail = TAIL && a.r8 > 0 && by == (int)gridDim.y - 1;  // the group holds the tail tile
  // super-chunk indices in 32 bits (the host checks nsuper < 2^31): the scalar unit compares
  // 32-bit integers, 64-bit ones would go through the VALU in every tile's liveness test
  const int s0 = (int)blockIdx.z * (int)a.ns_per_split;
  int s1 = min((int)a.nsuper, s0 + (int)a.ns_per_split);
  const uint32_t tab = 0x02010100u;
  // per-tile limits: digit tile j of vector tile vt0 + j / 3 (the tail tile: vector tile nvt)
  // runs MFMAs for super-chunks s < lim[j]; the staged loop bound is the largest over the
  // workgroup's rows, since all its waves meet at a barrier every half super-chunk
  int32_t lim[NDT];
  int32_t limmax = 0;
#pragma unroll
  for (int j = 0; j < NDT; ++j) {
    const bool tslot = TAIL && j == NDT - 1;
    const int vtile = !tslot ? vt0 + j / 3 : a.nvt;
    const bool on = !tslot ? j < 3 * nv : tail;
    const int32_t ln = !on ? 0 : (a.tile_nsuper ? a.tile_nsuper[vtile] : INT32_MAX);
    lim[j] = (a.tile_rows && on && row0 >= a.tile_rows[vtile]) ? 0 : ln;
    limmax = max(limmax, (a.tile_rows && on && wg_row0 >= a.tile_rows[vtile]) ? 0 : ln);
  }
  s1 = min(s1, (int)limmax);
  i32x4_t acc[RT][NDT];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NDT; ++j) acc[i][j] = i32x4_t{0, 0, 0, 0};
  auto live = [&](int j) { return j < 3 * nv || (TAIL && j == NDT - 1 && tail); };
  const uint8_t* lbase = a.G + (row0 + r) * a.ldg + g * 32;
  auto ld = [&](int sc, uint32_t (&dst)[RT][8]) {
    sc = sc < s1 ? sc : s1 - 1;  // past the slice: re-read its last super-chunk (never used)
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const u32x4_t* q = reinterpret_cast<const u32x4_t*>(lbase + (int64_t)i * 16 * a.ldg + (int64_t)sc * 128);
      const u32x4_t q0 = q[0], q1 = q[1];
      dst[i][0] = q0.x; dst[i][1] = q0.y; dst[i][2] = q0.z; dst[i][3] = q0.w;
      dst[i][4] = q1.x; dst[i][5] = q1.y; dst[i][6] = q1.z; dst[i][7] = q1.w;
    }
  };
  // staging: thread piece q is entry 256 q + threadIdx.x = (16 q + 4 w4 + g') 16 + c of the
  // half's buffer, digit column c of tile q for lane group g' at step 4 hh + w4 (global step
  // numbering); tail-tile columns past 3 r8 repeat a live one.  The tile of a piece is uniform
  // across the workgroup, so dead tiles cost a scalar branch.
  const int hlast = 2 * s1 - 1;
  const int wg = threadIdx.x >> 4, cc16 = threadIdx.x & 15;
  i32x4_t stg[NDT];
  auto bload = [&](int hh) {
    hh = hh <= hlast ? hh : hlast;
    const int64_t brow = (int64_t)(4 * (4 * hh + (wg >> 2)) + (wg & 3)) * a.ndc;
#pragma unroll
    for (int q = 0; q < 3 * NVG; ++q)
      if (q < 3 * nv) stg[q] = a.B[brow + 16 * (3 * vt0 + q) + cc16];
    if (TAIL && tail) stg[NDT - 1] = a.B[brow + 48 * a.nvt + cc16 % (3 * a.r8)];
  };
  auto bstore = [&](int hh) {
    i32x4_t* dst = &sb[hh & 1][threadIdx.x];
#pragma unroll
    for (int q = 0; q < 3 * NVG; ++q)
      if (q < 3 * nv) dst[256 * q] = stg[q];
    if (TAIL && tail) dst[256 * (NDT - 1)] = stg[NDT - 1];
  };
  // lm: bit j set when digit tile j runs its MFMAs in this super-chunk (uniform; tested as bits of
  // a scalar integer, so that no per-tile condition is carried in vector registers)
  auto step = [&](const uint32_t (&cc)[RT][8], int w, int buf, uint32_t lm) {
    i32x4_t av[RT];
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const uint32_t W = cc[i][w];
      av[i] = i32x4_t{(int)__builtin_amdgcn_perm(0u, tab, W & 0x03030303u),
                      (int)__builtin_amdgcn_perm(0u, tab, (W >> 2) & 0x03030303u),
                      (int)__builtin_amdgcn_perm(0u, tab, (W >> 4) & 0x03030303u),
                      (int)__builtin_amdgcn_perm(0u, tab, (W >> 6) & 0x03030303u)};
    }
    const i32x4_t* bp = &sb[buf][(4 * (w & 3) + g) * 16 + r];
    // each tile's digits are read one tile ahead, outside the branch around its MFMAs, so that
    // the read overlaps the previous tile's MFMAs (reads of dead tiles stay inside the buffer and
    // are not used)
    i32x4_t bcur = bp[0];
    if (a.prio) __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int j = 0; j < NDT; ++j) {
      const i32x4_t bv = bcur;
      if (j + 1 < NDT) bcur = bp[256 * (j + 1)];
      if ((lm >> j) & 1u) {
#pragma unroll
        for (int i = 0; i < RT; ++i)
          acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[i], bv, acc[i][j], 0, 0, 0);
      }
    }
    if (a.prio != 0) __builtin_amdgcn_s_setprio(0);
  };
  uint32_t cA[RT][8], cB[RT][8];
  if (s0 < s1) {
    bload(2 * s0);
    ld(s0, cA);
    bstore(2 * s0);
    __syncthreads();
  }
  // half H of super-chunk s from the codes cc; its first half requests super-chunk s + 1 into nx.
  // The buffer of half hh + 1 is written after every wave has finished half hh - 1 (the barrier
  // that ended it), so one barrier per half suffices.
  auto half = [&](auto hconst, const uint32_t (&cc)[RT][8], uint32_t (&nx)[RT][8], int s) {
    constexpr int H = decltype(hconst)::value;
    uint32_t lm = 0;
#pragma unroll
    for (int j = 0; j < NDT; ++j)
      if (live(j) && s < lim[j]) lm |= 1u << j;
    lm = __builtin_amdgcn_readfirstlane(lm);
    bload(2 * s + H + 1);
    if constexpr (H == 0) ld(s + 1, nx);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int w4 = 0; w4 < 4; ++w4) step(cc, 4 * H + w4, H, lm);
    __builtin_amdgcn_sched_barrier(0);
    bstore(2 * s + H + 1);
    __syncthreads();
  };
  using H0 = std::integral_constant<int, 0>;
  using H1 = std::integral_constant<int, 1>;
  for (int s = s0; s < s1; s += 2) {
    half(H0{}, cA, cB, s);
    half(H1{}, cA, cB, s);
    if (s + 1 < s1) {
      half(H0{}, cB, cA, s + 1);
      half(H1{}, cB, cA, s + 1);
    }
  }
  // lane (r, g) holds rows 4 g .. 4 g + 3 of column r of every tile
  auto put = [&](int k, int64_t row, const float (&o)[4]) {
    float* base = a.part + (blockIdx.z * (int64_t)a.KP + k) * a.rows_pad + row;
    *reinterpret_cast<float4*>(base) = make_float4(o[0], o[1], o[2], o[3]);
    if (a.tail_r > 0 && k >= a.tail0)  // the operand's lo column, which the epilogues add
      *reinterpret_cast<float4*>(base + (int64_t)a.tail_r * a.rows_pad) = make_float4(0.f, 0.f, 0.f, 0.f);
  };
#pragma unroll
  for (int i = 0; i < RT; ++i) {
    const int64_t row = row0 + 16 * i + 4 * g;
#pragma unroll
    for (int jv = 0; jv < NVG; ++jv) {
      const int k = 16 * (vt0 + jv) + r;
      if (jv >= nv || k >= a.t) continue;
      const double scale = (double)a.sc[k];
      float o[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        o[q] = (float)((double)((int64_t)acc[i][3 * jv][q] * 65536 + (int64_t)acc[i][3 * jv + 1][q] * 256 +
                                (int64_t)acc[i][3 * jv + 2][q]) * scale);
      put(k, row, o);
    }
    if (TAIL && tail) {
      const int k = 16 * a.nvt + r / 3;
      const bool first = r % 3 == 0 && r < 3 * a.r8;
      const double scale = first ? (double)a.sc[k] : 0.0;
      float o[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int v0 = acc[i][NDT - 1][q];
        const int v1 = __shfl(v0, lane + 1);
        const int v2 = __shfl(v0, lane + 2);
        o[q] = (float)((double)((int64_t)v0 * 65536 + (int64_t)v1 * 256 + (int64_t)v2) * scale);
      }
      if (first) put(k, row, o);
    }
  }
}

// s_waitcnt immediate (gfx9 encoding) for "at most n vector-memory operations outstanding", the
// other counters unconstrained
constexpr int vmcnt_imm(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }

// Wide blocks, second form, for blocks whose column groups all hold NVG whole vector tiles (blocks
// of an odd number of vector tiles keep the first form): the tiling, digit layout and epilogue of
// k_bp_i8w, with per-group rather than per-tile limits for the batched traits, and with
//  - the digits of each half super-chunk loaded straight into LDS (global_load_lds_dwordx4: each
//    lane's 16 bytes land at the wave's base + 16 lane, which is the staging layout), so that no
//    registers stage them and no store pass is needed; the full tiles two halves ahead in a ring
//    of three buffers, the tail tile one half ahead in two, all separate LDS objects, so that the
//    reads of one never wait for the loads into another, and each half ends with an explicit
//    wait for the next half's digits before the barrier;
//  - each step run row tile by row tile: row tile i's MFMAs on all digit tiles (their digits held
//    in registers, each refilled with the next step's right after its last use), while row tile
//    i + 1 (or the next step's row tile 0) is decoded between them, so that a wave issues its
//    MFMAs back to back even when the other wave of its SIMD waits, with one decoded row tile live
//    beside the one being decoded; only the tail tile's MFMAs sit behind a (uniform) branch;
//  - codes loaded per half super-chunk into a ring of three, each requested two halves ahead of
//    its use (the registers of one whole super-chunk ahead, twice the lead).
template <int RT, int NVG, bool TAIL = true>
__global__ __launch_bounds__(256, TAIL ? 2 : 3) void k_bp_i8x(I8WArgs a) {
  constexpr int NDT = 3 * NVG + (TAIL ? 1 : 0);  // digit tiles of a group; with TAIL the last is the tail tile
  constexpr int NF = 3 * NVG;                    // full digit tiles
  // digit staging, [tile j][4 w4 + g][column] per half super-chunk: the full tiles in a ring of
  // three buffers (loaded two halves ahead), the tail tile in two (one half ahead); five separate
  // LDS objects, so that reads of one never wait for loads into another.  80 KB with the tail tile,
  // which two workgroups per CU still fit.
  __shared__ i32x4_t fb0[NF * 256], fb1[NF * 256], fb2[NF * 256];
  __shared__ i32x4_t tb0[TAIL ? 256 : 1], tb1[TAIL ? 256 : 1];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  // (row block bx, column group by), remapped with xcd_pairs as in k_bp_i8w
  int64_t bx = blockIdx.x;
  int by = blockIdx.y;
  {
    const int64_t gx = gridDim.x, gy = gridDim.y, gx8 = gx - gx % 8;
    const int64_t L = blockIdx.x + gx * blockIdx.y;
    if (a.xcd_pairs && gy > 1 && L < gx8 * gy) {
      const int64_t slot = L / 8;
      bx = (slot / gy) * 8 + L % 8;
      by = (int)(slot % gy);
    } else if (a.xcd_pairs && gy > 1) {
      const int64_t L2 = L - gx8 * gy, rem = gx - gx8;
      bx = gx8 + L2 % rem;
      by = (int)(L2 / rem);
    }
  }
  const int64_t row0 = bx * (64 * RT) + wave * (16 * RT);
  const int vt0 = a.vt_base + by * NVG;
  const bool tail = TAIL && a.r8 > 0 && by == (int)gridDim.y - 1;  // the group holds the tail tile
  const int s0 = (int)blockIdx.z * (int)a.ns_per_split;
  int s1 = min((int)a.nsuper, s0 + (int)a.ns_per_split);
  if (a.tile_nsuper || a.tile_rows) {
    // batched traits: the group's limits, the largest over its tiles.  The operand is zero past
    // each tile's own reduction limit and the epilogues mask the rows outside every subset, so the
    // limits only skip work, and a group's tiles (one trait's probes, as a rule) share them.
    const int64_t wg_row0 = bx * (64 * RT);
    int32_t gl = 0;
#pragma unroll
    for (int jv = 0; jv <= NVG; ++jv) {
      if (jv == NVG && !tail) continue;
      const int vtile = jv < NVG ? vt0 + jv : a.nvt;
      int32_t ln = a.tile_nsuper ? a.tile_nsuper[vtile] : INT32_MAX;
      if (a.tile_rows && wg_row0 >= a.tile_rows[vtile]) ln = 0;
      gl = max(gl, ln);
    }
    s1 = min(s1, (int)gl);
  }
  i32x4_t acc[RT][NDT];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NDT; ++j) acc[i][j] = i32x4_t{0, 0, 0, 0};
  const uint8_t* lbase = a.G + (row0 + r) * a.ldg + g * 32;
  const int hlast = 2 * s1 - 1;
  // codes of half super-chunk hh (one 16-byte load per row tile): dst[i][w4] is the code word of
  // step w4 of that half; past the slice the last half is re-read (never used)
  auto ldh = [&](int hh, uint32_t (&dst)[RT][4]) __attribute__((always_inline)) {
    hh = hh <= hlast ? hh : hlast;
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const u32x4_t q = *reinterpret_cast<const u32x4_t*>(lbase + (int64_t)i * 16 * a.ldg + (int64_t)(hh >> 1) * 128 +
                                                          (hh & 1) * 16);
      dst[i][0] = q.x; dst[i][1] = q.y; dst[i][2] = q.z; dst[i][3] = q.w;
    }
  };
  const int wg = threadIdx.x >> 4, cc16 = threadIdx.x & 15;
  // digits of half hh: thread piece q is entry 256 q + threadIdx.x of the buffer (tail-tile
  // columns past 3 r8 repeat a live one)
  auto brow = [&](int hh) __attribute__((always_inline)) {
    hh = hh <= hlast ? hh : hlast;
    return a.B + (int64_t)(4 * (4 * hh + (wg >> 2)) + (wg & 3)) * a.ndc;
  };
  auto fbuf = [&](auto fc) __attribute__((always_inline)) {
    constexpr int F = decltype(fc)::value;
    return F == 0 ? fb0 : (F == 1 ? fb1 : fb2);
  };
  auto load_full = [&](auto fc, int hh) __attribute__((always_inline)) {
    i32x4_t* dst = fbuf(fc) + 64 * wave;
    const i32x4_t* src = brow(hh) + cc16;
#pragma unroll
    for (int q = 0; q < NF; ++q) __builtin_amdgcn_global_load_lds(src + 16 * (3 * vt0 + q), dst + 256 * q, 16, 0, 0);
  };
  auto tbuf = [&](auto tc) __attribute__((always_inline)) { return decltype(tc)::value == 0 ? tb0 : tb1; };
  auto load_tail = [&](auto tc, int hh) __attribute__((always_inline)) {
    i32x4_t* dst = tbuf(tc) + 64 * wave;
    const i32x4_t* src = brow(hh) + 48 * a.nvt + cc16 % (3 * a.r8);
    if (TAIL && tail) __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
  };
  const uint32_t tab = 0x02010100u;
  auto decode1 = [&](uint32_t W) __attribute__((always_inline)) {
    return i32x4_t{(int)__builtin_amdgcn_perm(0u, tab, W & 0x03030303u),
                   (int)__builtin_amdgcn_perm(0u, tab, (W >> 2) & 0x03030303u),
                   (int)__builtin_amdgcn_perm(0u, tab, (W >> 4) & 0x03030303u),
                   (int)__builtin_amdgcn_perm(0u, tab, (W >> 6) & 0x03030303u)};
  };
  // codes in a ring of three half super-chunks, each requested two halves ahead of its use
  uint32_t c0[RT][4], c1[RT][4], c2[RT][4];
  i32x4_t avc;  // the next row tile to multiply, decoded
  if (s0 < s1) {
    load_tail(std::integral_constant<int, 0>{}, 2 * s0);
    load_full(std::integral_constant<int, 0>{}, 2 * s0);
    load_full(std::integral_constant<int, 1>{}, 2 * s0 + 1);
    ldh(2 * s0, c0);
    ldh(2 * s0 + 1, c1);
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
    __syncthreads();
    avc = decode1(c0[0][0]);
  }
  // half hh (full-tile buffer F = hh mod 3, tail buffer T = hh mod 2, counting from 2 s0) from
  // its codes cc and the next half's codes cn (whose first step the last step here decodes).  It
  // requests, in this order, the tail digits of half hh + 1, the full-tile digits of half hh + 2
  // and the codes of half hh + 2 into cl; each buffer written was last read in half hh - 1, which
  // every wave finished at the barrier that ended it, so one barrier per half suffices.
  auto half = [&](auto fc, auto tc, int hh, const uint32_t (&cc)[RT][4], const uint32_t (&cn)[RT][4],
                  uint32_t (&cl)[RT][4]) __attribute__((always_inline)) {
    constexpr int F = decltype(fc)::value, T = decltype(tc)::value;
    load_tail(std::integral_constant<int, 1 - T>{}, hh + 1);
    load_full(std::integral_constant<int, (F + 2) % 3>{}, hh + 2);
    ldh(hh + 2, cl);
    const i32x4_t* bp = fbuf(fc) + g * 16 + r;                   // step w4 at + 64 w4, tile j at + 256 j
    const i32x4_t* tp = tbuf(tc) + g * 16 + r;                   // the tail tile's buffer
    i32x4_t B[NDT];
#pragma unroll
    for (int j = 0; j < NF; ++j) B[j] = bp[256 * j];
    if (TAIL) B[NDT - 1] = tp[0];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int w4 = 0; w4 < 4; ++w4) {
      if (a.prio) __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        const uint32_t nwd = i + 1 < RT ? cc[i + 1][w4] : (w4 + 1 < 4 ? cc[0][w4 + 1] : cn[0][0]);
        const i32x4_t avn = decode1(nwd);
        // the tail tile first (behind its branch), then the other tiles in one block, where the
        // decode's 11 VALU are interleaved with the MFMAs: a wave issues in order, so VALU placed
        // after all the MFMAs would wait behind each of them
        if (TAIL && tail) acc[i][NDT - 1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(avc, B[NDT - 1], acc[i][NDT - 1], 0, 0, 0);
        if (TAIL && i == RT - 1 && w4 < 3) B[NDT - 1] = tp[64 * (w4 + 1)];
#pragma unroll
        for (int j = 0; j < NF; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(avc, B[j], acc[i][j], 0, 0, 0);
          if (i == RT - 1 && w4 < 3) B[j] = bp[256 * j + 64 * (w4 + 1)];
        }
        avc = avn;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                                  // MFMA
          if (i == RT - 1 && w4 < 3) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);       // next step's digits
          if (j < 11 - NF + 1) __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);             // decode VALU
          else __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);
        }
        __builtin_amdgcn_sched_barrier(0);  // one row tile per scheduling region
      }
      if (a.prio) __builtin_amdgcn_s_setprio(0);
    }
    // the digits of half hh + 1 are in (its full tiles were requested a half earlier, its tail
    // tile first thing here; only this half's full-tile and code requests, issued after, may still
    // be in flight), then the barrier that hands them to every wave
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(NF + RT));
    __syncthreads();
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  const int hend = 2 * s1;  // halves [2 s0, hend), an even number
  for (int hh = 2 * s0; hh < hend; hh += 6) {
    half(I0{}, I0{}, hh, c0, c1, c2);
    half(I1{}, I1{}, hh + 1, c1, c2, c0);
    if (hh + 2 < hend) {
      half(I2{}, I0{}, hh + 2, c2, c0, c1);
      half(I0{}, I1{}, hh + 3, c0, c1, c2);
    }
    if (hh + 4 < hend) {
      half(I1{}, I0{}, hh + 4, c1, c2, c0);
      half(I2{}, I1{}, hh + 5, c2, c0, c1);
    }
  }
  auto put = [&](int k, int64_t row, const float (&o)[4]) {
    float* base = a.part + (blockIdx.z * (int64_t)a.KP + k) * a.rows_pad + row;
    *reinterpret_cast<float4*>(base) = make_float4(o[0], o[1], o[2], o[3]);
  };
#pragma unroll
  for (int i = 0; i < RT; ++i) {
    const int64_t row = row0 + 16 * i + 4 * g;
#pragma unroll
    for (int jv = 0; jv < NVG; ++jv) {
      const int k = 16 * (vt0 + jv) + r;
      if (k >= a.t) continue;
      const double scale = (double)a.sc[k];
      float o[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        o[q] = (float)((double)((int64_t)acc[i][3 * jv][q] * 65536 + (int64_t)acc[i][3 * jv + 1][q] * 256 +
                                (int64_t)acc[i][3 * jv + 2][q]) * scale);
      put(k, row, o);
    }
    if (TAIL && tail) {
      const int k = 16 * a.nvt + r / 3;
      const bool first = r % 3 == 0 && r < 3 * a.r8;
      const double scale = first ? (double)a.sc[k] : 0.0;
      float o[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int v0 = acc[i][NDT - 1][q];
        const int v1 = __shfl(v0, lane + 1);
        const int v2 = __shfl(v0, lane + 2);
        o[q] = (float)((double)((int64_t)v0 * 65536 + (int64_t)v1 * 256 + (int64_t)v2) * scale);
      }
      if (first) put(k, row, o);
    }
  }
}

}  // namespace

bool i8_narrow_enabled() {
  const char* e = std::getenv("SLEMM_I8_NARROW");  // read per call: "0" keeps the BF16 kernels
  return !(e && e[0] == '0');
}

// 16-row tiles per wave (2, 4 or 8; SLEMM_I8_RT, read per call).  Eight measured fastest at
// n = 100,000 x m = 500,000 for 1 to 5 vectors: each digit load then serves 8 MFMAs.
static int i8_rt(int rt) {
  const char* e = std::getenv("SLEMM_I8_RT");
  if (e) rt = std::atoi(e);
  else if (rt <= 0) rt = kI8RT;
  return rt == 2 || rt == 8 ? rt : 4;
}

int i8_class(int t) {
  const I8Digits d = i8_digits(t);
  return d.nvt == 0 ? 0 : (d.nvt == 1 && d.r8 == 0 ? 1 : 2);
}

int i8_splits(int t, int64_t rows_pad, int64_t nsuper, int64_t target_wgs, bool squared, int rt, int64_t& ns_per) {
  const bool narrow = i8_digits(t).nvt == 0;
  const int64_t wgs = rows_pad / (64 * (narrow ? i8_rt(rt) : kI8WRT)) * i8_groups(t);
  int64_t S = (target_wgs + wgs - 1) / wgs;
  // every split's int32 sums stay exact: |digit| <= 128 times dosages <= 2 (<= 4 squared) over at
  // most 2^23 (2^22) rows, at most 2^31 in magnitude
  const int64_t max_ns = (squared ? (int64_t)1 << 22 : (int64_t)1 << 23) / 512;
  S = std::max<int64_t>(S, (nsuper + max_ns - 1) / max_ns);
  S = std::max<int64_t>(1, std::min<int64_t>(S, nsuper));
  ns_per = (nsuper + S - 1) / S;
  return (int)((nsuper + ns_per - 1) / ns_per);
}

I8Digits i8_digits(int t) {
  I8Digits d;
  if (t <= 5) {
    d.nvt = 0; d.r8 = t; d.ndc = 3 * t;
  } else if (t % 16 >= 1 && t % 16 <= 5) {
    d.nvt = t / 16; d.r8 = t % 16; d.ndc = 16 * (3 * d.nvt + 1);
  } else {
    d.nvt = (t + 15) / 16; d.r8 = 0; d.ndc = 48 * d.nvt;
  }
  return d;
}

// scratch: sc[t] (doubles) and ex[t] (ints), each rounded up to 64 entries, the kI8Blocks partial
// maxima per vector (floats), the per-thread digit sums of the quantization (4 per step and
// vector, doubles), their kI8SumBlocks stage sums per vector (doubles), then the digits
// (16-byte aligned)
size_t i8_scratch_bytes(int64_t red_pad, int t) {
  const size_t t64 = 64 * (size_t)((t + 63) / 64);
  return 12 * t64 + sizeof(float) * (size_t)t * kI8Blocks + sizeof(double) * 4 * (size_t)(red_pad / 64) * t +
         sizeof(double) * (size_t)kI8SumBlocks * t + (size_t)red_pad * i8_digits(t).ndc;
}

int i8_groups(int t) {
  const I8Digits d = i8_digits(t);
  return d.nvt == 0 ? 1 : (d.nvt + kI8NVG - 1) / kI8NVG;
}

void launch_bp_i8(const I8Launch& l, hipStream_t st) {
  if (l.t < 1) fail(SLEMM_ERR_ARG, "int8 blocks take at least one vector");
  const I8Digits dl = i8_digits(l.t);
  const bool narrow = dl.nvt == 0;
  if (narrow && l.KP != 16) fail(SLEMM_ERR_ARG, "int8 narrow blocks have one 16-column tile");
  if (l.KP < l.t || l.KP % 16 != 0) fail(SLEMM_ERR_ARG, "int8 block: column padding mismatch");
  if (l.rows_pad % 512 != 0 || l.red_pad % 512 != 0 || l.ldg * 4 < l.red_pad || l.ldg % 16 != 0)
    fail(SLEMM_ERR_ARG, "int8 block: padding mismatch");
  // int32 exactness of every split's sums (i8_splits keeps to this)
  if (l.ns_per_split * 512 > (l.squared ? (int64_t)1 << 22 : (int64_t)1 << 23))
    fail(SLEMM_ERR_ARG, "int8 block: reduction split too long for exact int32 sums");
  if (l.squared && l.t != 1) fail(SLEMM_ERR_ARG, "squared-dosage products take one vector");
  if (l.tail_r != 0) fail(SLEMM_ERR_ARG, "int8 blocks take the FP32 operand layout (no lo columns)");
  const size_t t64 = 64 * (size_t)((l.t + 63) / 64);
  double* sc = static_cast<double*>(l.scratch);
  int* ex = reinterpret_cast<int*>(sc + t64);
  float* pmax = reinterpret_cast<float*>(ex + t64);
  const int64_t nstep = l.red_pad / 64;
  double* psum = reinterpret_cast<double*>(pmax + (size_t)l.t * kI8Blocks);
  double* spart = psum + 4 * (size_t)nstep * l.t;
  i32x4_t* B = reinterpret_cast<i32x4_t*>(spart + (size_t)kI8SumBlocks * l.t);
  if (reinterpret_cast<uintptr_t>(psum) % 8 != 0 || reinterpret_cast<uintptr_t>(B) % 16 != 0)
    fail(SLEMM_ERR_ARG, "int8 block: scratch alignment");
  const float* X = static_cast<const float*>(l.X);
  k_i8_colmax<<<dim3(kI8Blocks, (unsigned)((l.t + 63) / 64)), 256, 0, st>>>(X, l.KP, l.t, l.red_pad, pmax);
  k_i8_scale<<<(unsigned)l.t, 64, 0, st>>>(pmax, l.t, sc, ex);
  k_i8_quant<<<(unsigned)((nstep * 4 * l.t + 255) / 256), 256, 0, st>>>(X, l.KP, l.t, dl.nvt, dl.ndc, nstep, ex,
                                                                         l.sum_w, B, psum);
  if (l.sum_out) {
    k_i8_digit_sums1<<<dim3((unsigned)((l.t + 63) / 64), kI8SumBlocks), 256, 0, st>>>(psum, 4 * nstep, l.t, spart);
    k_i8_digit_sums2<<<(unsigned)l.t, 64, 0, st>>>(spart, l.t, ex, l.sum_out);
  }
  if (narrow) {
    I8Args a;
    a.G = l.G; a.ldg = l.ldg; a.nsuper = l.red_pad / 512; a.ns_per_split = l.ns_per_split; a.B = B; a.t = l.t;
    a.sc = sc; a.part = l.part; a.KP = l.KP; a.rows_pad = l.rows_pad;
    const int rt = i8_rt(l.rt);
    const dim3 grid((unsigned)(l.rows_pad / (64 * rt)), 1, (unsigned)l.nsplit);
    if (rt == 2) {
      if (l.squared) k_bp_i8<2, true><<<grid, 256, 0, st>>>(a);
      else k_bp_i8<2, false><<<grid, 256, 0, st>>>(a);
    } else if (rt == 8) {
      if (l.squared) k_bp_i8<8, true><<<grid, 256, 0, st>>>(a);
      else k_bp_i8<8, false><<<grid, 256, 0, st>>>(a);
    } else {
      if (l.squared) k_bp_i8<4, true><<<grid, 256, 0, st>>>(a);
      else k_bp_i8<4, false><<<grid, 256, 0, st>>>(a);
    }
  } else {
    I8WArgs a;
    a.G = l.G; a.ldg = l.ldg; a.nsuper = l.red_pad / 512; a.ns_per_split = l.ns_per_split; a.B = B;
    a.ndc = dl.ndc; a.nvt = dl.nvt; a.r8 = dl.r8; a.t = l.t; a.sc = sc; a.part = l.part; a.KP = l.KP;
    a.rows_pad = l.rows_pad; a.tail0 = l.tail0; a.tail_r = l.tail_r;
    a.tile_nsuper = l.tile_nsuper; a.tile_rows = l.tile_rows;
    // the second form takes groups of whole vector tiles without lo columns (SLEMM_I8_WIDE_FORM=1:
    // the first form for every block, read per call)
    // Blocks whose vector tiles do not pair up (an odd number, the tail tile aside) run as two
    // launches of the second form: the pairs, then the last vector tile with the tail tile as one
    // group of its own (each group reads the codes once either way, and the two write disjoint
    // columns of the partials).
    const char* e4 = std::getenv("SLEMM_I8_WIDE_FORM");
    const bool first_form = (e4 && e4[0] == '1') || a.tail_r != 0;
    // SLEMM_I8_XCD (0/1, read per call): the column groups of a row block dispatched back to back on
    // one XCD, so that all but the first read the codes from that XCD's L2.  SLEMM_I8_PRIO (0/1):
    // issue priority 1 while a wave issues a step's MFMAs.  Defaults as measured at n = 100,000 x
    // m = 500,000, K = 65: the second form is fastest with the pairing and without the priority
    // (operator 16.94 -> 15.77 ms; pairing alone 16.61, no priority alone 16.93); the first form
    // with the priority and without the pairing (20.4 -> 20.1 ms from the priority; the pairing cut
    // 8-14% of its fetched bytes and no time).
    const char* ex = std::getenv("SLEMM_I8_XCD");
    a.xcd_pairs = ex ? ex[0] == '1' : !first_form;
    const char* ep = std::getenv("SLEMM_I8_PRIO");
    a.prio = ep ? (ep[0] == '1' ? 1 : 0) : (first_form ? 1 : 0);
    if (a.nsuper >= INT32_MAX / 2) fail(SLEMM_ERR_ARG, "int8 block: reduction too long");
    // (one-vector-tile blocks measured slower with 2 row tiles per wave at four waves per SIMD:
    // each digit load then serves half as many MFMAs)
    const unsigned nrb = (unsigned)(l.rows_pad / (64 * kI8WRT));
    const dim3 grid(nrb, (unsigned)i8_groups(l.t), (unsigned)l.nsplit);
    const char* e3 = std::getenv("SLEMM_I8_ONE_TILE");  // "0": the general kernel (read per call)
    if (dl.nvt == 1 && dl.r8 == 0 && !(e3 && e3[0] == '0')) {
      if (first_form) k_bp_i8w<kI8WRT, 1, false><<<grid, 256, 0, st>>>(a);
      else k_bp_i8x<kI8WRT, 1, false><<<grid, 256, 0, st>>>(a);
    } else if (first_form) {
      k_bp_i8w<kI8WRT, kI8NVG><<<grid, 256, 0, st>>>(a);
    } else if (dl.nvt % kI8NVG == 0) {
      k_bp_i8x<kI8WRT, kI8NVG><<<grid, 256, 0, st>>>(a);
    } else {
      const int pairs = dl.nvt / kI8NVG;
      if (pairs > 0) {
        I8WArgs a2 = a;
        a2.r8 = 0;  // the tail tile goes with the last vector tile
        k_bp_i8x<kI8WRT, kI8NVG><<<dim3(nrb, (unsigned)pairs, (unsigned)l.nsplit), 256, 0, st>>>(a2);
        HIPCHECK(hipGetLastError());
      }
      I8WArgs a1 = a;
      a1.vt_base = pairs * kI8NVG;
      a1.xcd_pairs = false;
      const dim3 g1(nrb, 1, (unsigned)l.nsplit);
      if (dl.r8 > 0) k_bp_i8x<kI8WRT, 1, true><<<g1, 256, 0, st>>>(a1);
      else k_bp_i8x<kI8WRT, 1, false><<<g1, 256, 0, st>>>(a1);
    }
  }
  HIPCHECK(hipGetLastError());
}

}  // namespace slemm
