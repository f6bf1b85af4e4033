// Narrow block products (at most 5 vectors, hi and lo halves as columns of one tile) on the int8
// matrix cores: the Z'x / Zx passes of src/reml.cpp:6-233 for the Lanczos / CG vectors, the
// single right-hand sides and the GWAS score vectors.
//
// The operand is the split-BF16 one the pack and pass-1 epilogue kernels write (value hi + lo per
// element; their column sums feed the exact-centring terms).  Per column it is rendered as a
// 23-bit fixed-point integer X = rint((hi + lo) 2^22 / s) with s the power of two at or above the
// column's largest magnitude, split into three balanced int8 digits X = 2^16 d0 + 2^8 d1 + d2.
// The genotype codes decode in registers to int8 dosages {0,1,1,2} ({0,1,1,4} for the squared
// pass) with one shift, one mask and one byte permute per four codes, so no table and no LDS
// are involved.  Each digit is one column of a 16 x 16 x 64 int8 MFMA tile, whose int32 sums
// are exact (|sum| <= 2 * 128 * reduction length < 2^31 for reductions below 8.3M), and the
// epilogue combines the three digit columns of a vector exactly in 64-bit integers before one
// rounding to FP32.  Per element the rendering error is at most 2^-23 s; the split-BF16 kernels
// instead round every product sum in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "blockprod.hpp"

namespace slemm {
namespace {

typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

constexpr int kI8Blocks = 64;  // column-maximum partials per vector

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

// Partial maxima of |hi + lo| over 8-row blocks: block (x, k) covers blocks x, x + 64, ... of
// column k.  X holds 8 rows of one column per 16 bytes: hi halves in columns [0, t), lo halves in
// columns [t, 2t) of each KP-wide row block.
__global__ __launch_bounds__(256) void k_i8_colmax(const uint4* __restrict__ X, int KP, int t, int64_t nrb,
                                                   float* __restrict__ pmax) {
  __shared__ float red[4];
  const int k = blockIdx.y;
  float mx = 0.f;
  for (int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x; b < nrb; b += (int64_t)kI8Blocks * 256) {
    const uint4 h = X[b * KP + k], l = X[b * KP + t + k];
    const uint32_t hw[4] = {h.x, h.y, h.z, h.w}, lw[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      mx = fmaxf(mx, fabsf(bf16_lo(hw[q]) + bf16_lo(lw[q])));
      mx = fmaxf(mx, fabsf(bf16_hi(hw[q]) + bf16_hi(lw[q])));
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) pmax[k * kI8Blocks + blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// Digits in MFMA lane order.  Step T (64 reduction rows) of lane group g covers rows
// 512 (T / 8) + 128 g + 16 (T % 8) + j, j < 16, the rows whose codes the block-product kernel's
// lane group g decodes at that step; byte 4 m + b of a lane's 16 bytes holds row j = 4 b + m (the
// order in which the codes decode).  B[(4 T + g) nc + 3 k + p] holds digit p of vector k.
// Thread (x, k): one (T, g) pair of vector k.  sc[k] = s_k 2^-22 (written by block 0).
__global__ __launch_bounds__(256) void k_i8_quant(const uint4* __restrict__ X, int KP, int t, int64_t nstep,
                                                  const float* __restrict__ pmax, float* __restrict__ sc,
                                                  i32x4_t* __restrict__ B) {
  __shared__ float smax;
  const int k = blockIdx.y;
  if (threadIdx.x < 64) {
    float mx = pmax[k * kI8Blocks + threadIdx.x];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (threadIdx.x == 0) smax = mx;
  }
  __syncthreads();
  const float mx = smax;
  // s = 2^e >= mx (s = 1 for an all-zero column); mult = 2^22 / s scales exactly
  int e = 0;
  if (mx > 0.f) {
    (void)frexpf(mx, &e);                      // mx = f 2^e with 0.5 <= f < 1, so mx < 2^e
    if (mx == ldexpf(0.5f, e)) e -= 1;         // a power of two: s = mx
  }
  const float mult = ldexpf(1.0f, 22 - e);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[k] = ldexpf(1.0f, e - 22);
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nstep * 4) return;
  const int64_t T = v >> 2;
  const int g = (int)(v & 3);
  const int64_t r0 = (T >> 3) * 512 + g * 128 + (T & 7) * 16;  // first of 16 rows
  int X16[16];
#pragma unroll
  for (int h8 = 0; h8 < 2; ++h8) {
    const int64_t b = (r0 >> 3) + h8;
    const uint4 hv = X[b * KP + k], lv = X[b * KP + t + k];
    const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w}, lw[4] = {lv.x, lv.y, lv.z, lv.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      X16[h8 * 8 + 2 * q] = (int)rintf((bf16_lo(hw[q]) + bf16_lo(lw[q])) * mult);
      X16[h8 * 8 + 2 * q + 1] = (int)rintf((bf16_hi(hw[q]) + bf16_hi(lw[q])) * mult);
    }
  }
  uint32_t d[3][4] = {};
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int x = X16[j];
    const int d2 = ((x + 128) & 255) - 128;
    const int x1 = (x - d2) >> 8;
    const int d1 = ((x1 + 128) & 255) - 128;
    const int d0 = (x1 - d1) >> 8;  // |x| <= 2^22: d0 in [-64, 64]
    const int m = j & 3, b = j >> 2;  // byte 4 m + b: dword m, byte b
    d[0][m] |= (uint32_t)(d0 & 255) << (8 * b);
    d[1][m] |= (uint32_t)(d1 & 255) << (8 * b);
    d[2][m] |= (uint32_t)(d2 & 255) << (8 * b);
  }
  const int nc = 3 * t;
#pragma unroll
  for (int p = 0; p < 3; ++p)
    B[(4 * T + g) * nc + 3 * k + p] = i32x4_t{(int)d[p][0], (int)d[p][1], (int)d[p][2], (int)d[p][3]};
}

struct I8Args {
  const uint8_t* G = nullptr;
  int64_t ldg = 0;
  int64_t nsuper = 0, ns_per_split = 0;
  const i32x4_t* B = nullptr;
  int t = 0;
  const float* sc = nullptr;
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
  // Lanes of column 3 k combine the three digits; lanes of column 3 k + 1 clear the lo-half
  // column t + k, which the pass epilogues add to column k.
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
    else if (live && p == 1)
      *reinterpret_cast<float4*>(a.part + (blockIdx.z * (int64_t)a.KP + a.t + k) * a.rows_pad + row) =
          make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

}  // namespace

bool i8_narrow_enabled() {
  const char* e = std::getenv("SLEMM_I8_NARROW");  // read per call: "0" keeps the BF16 kernels
  return !(e && e[0] == '0');
}

size_t i8_scratch_bytes(int64_t red_pad, int t) {
  return sizeof(float) * (64 + (size_t)t * kI8Blocks) + (size_t)red_pad * 3 * t;
}

// 16-row tiles per wave (2, 4 or 8; SLEMM_I8_RT, read per call).  Eight measured fastest at
// n = 100,000 x m = 500,000 for 1 to 5 vectors: each digit load then serves 8 MFMAs.
static int i8_rt() {
  const char* e = std::getenv("SLEMM_I8_RT");
  const int rt = e ? std::atoi(e) : kI8RT;
  return rt == 2 || rt == 8 ? rt : 4;
}

int i8_splits(int64_t rows_pad, int64_t nsuper, int64_t target_wgs, int64_t& ns_per) {
  const int64_t wgs = rows_pad / (64 * i8_rt());
  int64_t S = (target_wgs + wgs - 1) / wgs;
  S = std::max<int64_t>(1, std::min<int64_t>(S, nsuper));
  ns_per = (nsuper + S - 1) / S;
  return (int)((nsuper + ns_per - 1) / ns_per);
}

void launch_bp_i8(const I8Launch& l, hipStream_t st) {
  if (l.t < 1 || 3 * l.t > 16) fail(SLEMM_ERR_ARG, "int8 narrow blocks take 1 to 5 vectors");
  if (l.KP != 16) fail(SLEMM_ERR_ARG, "int8 narrow blocks have one 16-column tile");
  if (l.rows_pad % 512 != 0 || l.red_pad % 512 != 0 || l.ldg * 4 < l.red_pad || l.ldg % 16 != 0)
    fail(SLEMM_ERR_ARG, "int8 narrow block: padding mismatch");
  if (l.red_pad > 8 * 1024 * 1024) fail(SLEMM_ERR_ARG, "int8 narrow block: reduction too long for int32 sums");
  if (l.squared && l.t != 1) fail(SLEMM_ERR_ARG, "squared-dosage products take one vector");
  float* sc = static_cast<float*>(l.scratch);
  float* pmax = sc + 64;
  i32x4_t* B = reinterpret_cast<i32x4_t*>(pmax + (size_t)l.t * kI8Blocks);
  if (reinterpret_cast<uintptr_t>(B) % 16 != 0) fail(SLEMM_ERR_ARG, "int8 narrow block: scratch alignment");
  const uint4* X = static_cast<const uint4*>(l.X);
  k_i8_colmax<<<dim3(kI8Blocks, l.t), 256, 0, st>>>(X, l.KP, l.t, l.red_pad / 8, pmax);
  const int64_t nstep = l.red_pad / 64;
  k_i8_quant<<<dim3((unsigned)((nstep * 4 + 255) / 256), l.t), 256, 0, st>>>(X, l.KP, l.t, nstep, pmax, sc, B);
  I8Args a;
  a.G = l.G; a.ldg = l.ldg; a.nsuper = l.red_pad / 512; a.ns_per_split = l.ns_per_split; a.B = B; a.t = l.t;
  a.sc = sc; a.part = l.part; a.KP = l.KP; a.rows_pad = l.rows_pad;
  const int rt = i8_rt();
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
  HIPCHECK(hipGetLastError());
}

}  // namespace slemm
