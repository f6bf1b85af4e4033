// Split-BF16 genotype block product with the vector block staged in LDS.
//
// Same product as k_blockprod (P = G X with G the 2-bit codes unpacked to BF16 dosages and X the
// hi/lo BF16 split of the vector block), restructured for gfx950's data paths:
//  * the four waves of a workgroup share every B-operand fragment through LDS: a step group of
//    TS MFMA steps (TS*32 reduction elements x KB columns x hi/lo) is copied once per workgroup
//    with 16-byte loads and read back with conflict-free ds_read_b128, double-buffered so the
//    copy of group q+1 overlaps the MFMAs of group q;
//  * the codes of a wave's rows stay in registers for a whole 512-element super-chunk and the
//    next super-chunk's codes are fetched one chunk ahead;
//  * the reduction can be split over gridDim.z so that passes with few rows (Z U over the
//    individuals, small SNP sets) still fill all 256 CUs; each split writes raw FP32 partial
//    products and the epilogue kernels below sum the splits in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <climits>
#include <cstdlib>

#include "blockprod.hpp"

namespace slemm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ uint32_t bf16_bits_rn(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_to_float(uint32_t b) { return __uint_as_float(b << 16); }

// 8 two-bit codes -> 8 BF16 dosages with one byte-permute per pair of codes.  Table indexed by
// the raw code c: low byte {00,80,80,00}[c], high byte {00,3F,3F,40}[c] (BF16 of 0,1,1,2).
// SQ selects squared dosages {0,1,1,4}: low byte {00,80,80,80}[c] (BF16(4) = 0x4080).
template <bool SQ = false>
__device__ __forceinline__ bf16x8_t unpack8(uint32_t p16) {
  constexpr uint32_t lo_tab = SQ ? 0x80808000u : 0x00808000u;
  // sp holds code 2q at bits 4q..4q+1 and code 2q+1 at bits 4q+16..4q+17 (bits 14-15 mix two
  // codes but are never read), so each pair needs one shift and one mask
  const uint32_t sp = p16 | (p16 << 14);
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t x = (sp >> (4 * q)) & 0x00030003u;  // c_2q -> byte 0, c_2q+1 -> byte 2
    w[q] = __builtin_amdgcn_perm(0x403F3F00u, lo_tab, x * 0x0101u + 0x04000400u);
  }
  return __builtin_bit_cast(bf16x8_t, make_uint4(w[0], w[1], w[2], w[3]));
}

}  // namespace

template <int RT, int NT, int TS, bool PREF, bool SQ>
__global__ __launch_bounds__(256) void k_bp_lds(LdsArgs a) {
  constexpr int KB = 16 * NT;
  constexpr int GPS = 16 / TS;                 // step groups per super-chunk
  constexpr int NSTAGE = TS * 8 * KB / 256;    // 16-byte loads per thread per step group
  __shared__ uint4 sB[2][TS * 2 * 4 * KB];      // [buf][tt][piece][g][col]

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * (64 * RT) + wave * (16 * RT);
  const int col0 = blockIdx.y * KB;
  const int64_t s0 = (int64_t)blockIdx.z * a.ns_per_split;
  const int64_t s1 = min(a.nsuper, s0 + a.ns_per_split);
  const int64_t ngrp = (s1 - s0) * GPS;
  const uint4* X = reinterpret_cast<const uint4*>(a.X);

  f32x4_t acc[RT][NT];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  uint4 stage[NSTAGE];
  auto load_group = [&](int64_t q) {
#pragma unroll
    for (int e = 0; e < NSTAGE; ++e) {
      const int L = tid + 256 * e;
      const int col = L % KB;
      int rest = L / KB;
      const int gg = rest & 3;
      rest >>= 2;
      const int piece = rest & 1;
      const int tt = rest >> 1;
      const int64_t qs = q * TS + tt;
      const int64_t s = s0 + qs / 16;
      const int64_t blk = s * 64 + gg * 16 + (qs % 16);
      stage[e] = X[piece * a.lo_offset16 + blk * a.KP + col0 + col];
    }
  };
  auto store_group = [&](int buf) {
#pragma unroll
    for (int e = 0; e < NSTAGE; ++e) sB[buf][tid + 256 * e] = stage[e];
  };

  const uint8_t* gbase = a.G + (row0 + r) * a.ldg + g * 32;
  uint32_t pk[RT][8], pn[PREF ? RT : 1][8];
  auto load_codes = [&](int64_t s, auto& dst) {
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const u32x4_t* src = reinterpret_cast<const u32x4_t*>(gbase + (int64_t)i * 16 * a.ldg + s * 128);
      const u32x4_t q0 = __builtin_nontemporal_load(src);
      const u32x4_t q1 = __builtin_nontemporal_load(src + 1);
      dst[i][0] = q0.x; dst[i][1] = q0.y; dst[i][2] = q0.z; dst[i][3] = q0.w;
      dst[i][4] = q1.x; dst[i][5] = q1.y; dst[i][6] = q1.z; dst[i][7] = q1.w;
    }
  };

  if (ngrp > 0) {
    load_group(0);
    store_group(0);
    load_codes(s0, pk);
  }
  for (int64_t q = 0; q < ngrp; ++q) {
    __syncthreads();
    const int buf = (int)(q & 1);
    if (q + 1 < ngrp) load_group(q + 1);
    const int gq = (int)(q % GPS);
    if constexpr (PREF) {
      if (gq == 0 && q + GPS < ngrp) load_codes(s0 + q / GPS + 1, pn);
    } else {
      if (gq == 0 && q > 0) load_codes(s0 + q / GPS, pk);
    }
    const uint4* sb = sB[buf];
#pragma unroll
    for (int tt = 0; tt < TS; ++tt) {
      const int t = gq * TS + tt;
      bf16x8_t av[RT];
#pragma unroll
      for (int i = 0; i < RT; ++i) av[i] = unpack8<SQ>((pk[i][t >> 1] >> (16 * (t & 1))) & 0xFFFFu);
      // all hi products of the step, then all lo products: RT*NT independent MFMAs separate
      // the two updates of one accumulator
#pragma unroll
      for (int piece = 0; piece < 2; ++piece)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          const bf16x8_t b = __builtin_bit_cast(bf16x8_t, sb[((tt * 2 + piece) * 4 + g) * KB + j * 16 + r]);
#pragma unroll
          for (int i = 0; i < RT; ++i)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], b, acc[i][j], 0, 0, 0);
        }
    }
    if constexpr (PREF) {
      if (gq == GPS - 1 && q + 1 < ngrp) {
#pragma unroll
        for (int i = 0; i < RT; ++i)
#pragma unroll
          for (int w = 0; w < 8; ++w) pk[i][w] = pn[i][w];
      }
    }
    if (q + 1 < ngrp) store_group(buf ^ 1);
  }

  // raw partial products, column-major per split: part[(z*KP + col) * rows_pad + row]
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int64_t col = col0 + 16 * j + r;
      const int64_t row = row0 + 16 * i + 4 * g;
      float4* dst = reinterpret_cast<float4*>(a.part + ((int64_t)blockIdx.z * a.KP + col) * a.rows_pad + row);
      *dst = make_float4(acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
    }
}

// Register-only variant: every wave reads its B fragments straight from global memory (L2/MALL
// hits after the first wave of a workgroup wave-front); no LDS, no barriers, lower register
// use, so two waves fit per SIMD.  Same split-K partial output as k_bp_lds.
template <int RT, int NT, bool SQ, bool LUT = false>
__global__ __launch_bounds__(256) void k_bp_reg(LdsArgs a) {
  // LUT: codes are unpacked through a 16-entry LDS table (nibble = two codes -> two BF16
  // dosages), one broadcast-friendly ds_read_b32 per pair instead of the permute sequence
  __shared__ uint32_t lut[16];
  if constexpr (LUT) {
    if (threadIdx.x < 16) {
      const uint32_t bf[4] = {0x0000u, 0x3F80u, 0x3F80u, SQ ? 0x4080u : 0x4000u};  // BF16 of d or d^2
      lut[threadIdx.x] = bf[threadIdx.x & 3] | (bf[threadIdx.x >> 2] << 16);
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * (64 * RT) + wave * (16 * RT);
  const int col0 = blockIdx.y * (16 * NT);
  const int64_t s0 = (int64_t)blockIdx.z * a.ns_per_split;
  int64_t s1 = min(a.nsuper, s0 + a.ns_per_split);
  // per-tile work limits (batched traits): MFMAs of tile j run only for super-chunks s < lim[j]
  int32_t lim[NT];
  int32_t limmax = 0;
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int64_t tile = (int64_t)blockIdx.y * NT + j;
    int32_t l = a.tile_nsuper ? a.tile_nsuper[tile] : INT32_MAX;
    if (a.tile_rows && row0 >= a.tile_rows[tile]) l = 0;  // this wave's rows of the tile are zero
    lim[j] = l;
    limmax = max(limmax, l);
  }
  s1 = min(s1, (int64_t)limmax);
  const uint4* xh = reinterpret_cast<const uint4*>(a.X);
  const uint4* xl = xh + a.lo_offset16;
  f32x4_t acc[RT][NT];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const uint8_t* gbase = a.G + (row0 + r) * a.ldg + g * 32;
  for (int64_t s = s0; s < s1; ++s) {
    uint32_t pk[RT][8];
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const u32x4_t* src = reinterpret_cast<const u32x4_t*>(gbase + (int64_t)i * 16 * a.ldg + s * 128);
      const u32x4_t q0 = __builtin_nontemporal_load(src);
      const u32x4_t q1 = __builtin_nontemporal_load(src + 1);
      pk[i][0] = q0.x; pk[i][1] = q0.y; pk[i][2] = q0.z; pk[i][3] = q0.w;
      pk[i][4] = q1.x; pk[i][5] = q1.y; pk[i][6] = q1.z; pk[i][7] = q1.w;
    }
    const int64_t blk0 = s * 64 + g * 16;
#pragma unroll 4
    for (int t = 0; t < 16; ++t) {
      bf16x8_t av[RT];
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        const uint32_t p16 = (pk[i][t >> 1] >> (16 * (t & 1))) & 0xFFFFu;
        if constexpr (LUT) {
          av[i] = __builtin_bit_cast(bf16x8_t, make_uint4(lut[p16 & 0xF], lut[(p16 >> 4) & 0xF],
                                                          lut[(p16 >> 8) & 0xF], lut[p16 >> 12]));
        } else {
          av[i] = unpack8<SQ>(p16);
        }
      }
      const int64_t boff = (blk0 + t) * a.KP + col0 + r;
      bf16x8_t bh[NT], bl[NT];
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        bh[j] = __builtin_bit_cast(bf16x8_t, xh[boff + j * 16]);
        bl[j] = __builtin_bit_cast(bf16x8_t, xl[boff + j * 16]);
      }
      // all hi products of the step, then all lo products (independent MFMAs in between)
#pragma unroll
      for (int j = 0; j < NT; ++j)
        if (s < lim[j])
#pragma unroll
          for (int i = 0; i < RT; ++i)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < NT; ++j)
        if (s < lim[j])
#pragma unroll
          for (int i = 0; i < RT; ++i)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bl[j], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int64_t col = col0 + 16 * j + r;
      const int64_t row = row0 + 16 * i + 4 * g;
      float4* dst = reinterpret_cast<float4*>(a.part + ((int64_t)blockIdx.z * a.KP + col) * a.rows_pad + row);
      *dst = make_float4(acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
    }
}

// Software-pipelined register variant.  Same tiles and partial output as k_bp_reg, but loads
// are requested well before their data is used, without a second register set for B:
//  * B: the fragment register of column tile j (hi, then lo) is refilled with step T+1's
//    fragment as soon as the RT MFMAs of step T that read it have been issued, so each load has
//    about one step of matrix work (2*RT*NT MFMAs) to arrive;
//  * codes: the 512-element super-chunk is consumed in two halves of 8 steps (16 bytes of codes
//    per lane and row tile) with two register sets; the next half is requested when the current
//    one starts.
// Loads past the split's last step re-read that step (valid addresses, no branches).
// 256-entry LDS table for the table-driven unpack: entry b holds the four BF16 dosages (or
// squared dosages) of the four codes of byte b, two per 32-bit word.  R copies of each entry
// are stored side by side (index b * R + copy) so that lanes reading different copies hit
// different banks.
template <bool SQ, int R = 1, int NTH = 256>
__device__ __forceinline__ void fill_byte_table(uint2* lut8) {
  static_assert((256 * R) % NTH == 0, "table size must be a multiple of the workgroup size");
  const uint32_t bf[4] = {0x0000u, 0x3F80u, 0x3F80u, SQ ? 0x4080u : 0x4000u};
#pragma unroll
  for (int k = 0; k < 256 * R / NTH; ++k) {
    const uint32_t idx = threadIdx.x + (uint32_t)NTH * k;  // NTH threads
    const uint32_t t = idx / R;
    lut8[idx] = make_uint2(bf[t & 3] | (bf[(t >> 2) & 3] << 16), bf[(t >> 4) & 3] | (bf[t >> 6] << 16));
  }
}

// One (row block bx, reduction split bz) tile of the pipelined product for column group blockIdx.y.
// LUT > 0: codes are unpacked by two 8-byte LDS reads per 8 codes (lut8, LUT copies per entry,
// filled by the caller) instead of the permute sequence: a third of the VALU work, for blocks
// too narrow to hide it.  HL: hi/lo-in-columns block, one MFMA per row tile and step.  WPG waves
// per workgroup, each on its own 16 * RT rows.  FULL: each row's 128 code bytes of a super-chunk
// are requested at once (lane group g reads bytes [32 g, 32 g + 32)), a super-chunk ahead, so
// that the DRAM sees whole lines per row instead of two 64-byte halves eight steps apart.
// SBC > 0 (hi/lo-in-columns blocks with at most SBC live columns, FULL): the workgroup stages
// the live B columns of each half super-chunk in LDS (sb, 2 x 32 x SBC 16-byte pieces), loaded
// a half ahead and before the codes.  Per-step B loads from global memory would share the
// in-order vector-memory counter with the code loads, so every wait for a B fragment would also
// wait for the next super-chunk's codes and the code prefetch would shrink to a few steps.
template <int RT, int NT, bool SQ, int LUT = 0, bool HL = false, int WPG = 4, bool FULL = false, int SBC = 0>
__device__ __forceinline__ void bp_pipe_tile(const LdsArgs& a, int64_t bx, int64_t bz, const uint2* lut8 = nullptr,
                                             uint4* sb = nullptr) {
  static_assert(SBC == 0 || (FULL && (HL || NT == 1)), "staged B is for one-tile blocks with whole-line codes");
  static_assert(!HL || NT == 1, "hi/lo-in-columns blocks have one column tile");
  // the wave index is uniform across the wave: keep it (and the tile limits) in scalar registers
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  const int64_t row0 = bx * (WPG * 16 * RT) + wave * (16 * RT);
  const int col0 = blockIdx.y * (16 * NT);
  const int64_t s0 = bz * a.ns_per_split;
  int64_t s1 = min(a.nsuper, s0 + a.ns_per_split);
  // per-tile work limits (batched traits): MFMAs of tile j run only for super-chunks s < lim[j]
  // (with staged B the loop bound must be the same for all waves of the workgroup, which meet at
  // a barrier every super-chunk: it is taken over the workgroup's rows)
  // Tiles past last_tile hold only zero columns and run no MFMA; the tail tile holds hi and lo
  // halves as columns and runs no lo MFMA (both flags are uniform across the workgroup).
  int32_t lim[NT];
  bool lo_on[NT];
  int32_t limmax = 0;
  const int64_t wg_row0 = bx * (WPG * 16 * RT);
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int64_t tile = (int64_t)blockIdx.y * NT + j;
    lo_on[j] = tile != a.tail_tile;
    const int32_t ln =
        (a.last_tile >= 0 && tile > a.last_tile) ? 0 : (a.tile_nsuper ? a.tile_nsuper[tile] : INT32_MAX);
    // this wave's (workgroup's) rows of the tile are zero from tile_rows on
    lim[j] = (a.tile_rows && row0 >= a.tile_rows[tile]) ? 0 : ln;
    limmax = max(limmax, SBC > 0 ? ((a.tile_rows && wg_row0 >= a.tile_rows[tile]) ? 0 : ln) : lim[j]);
  }
  s1 = min(s1, (int64_t)limmax);
  const uint4* xh = reinterpret_cast<const uint4*>(a.X);
  const uint4* xl = xh + a.lo_offset16;
  const uint8_t* gbase = a.G + (row0 + r) * a.ldg + g * 16;
  f32x4_t acc[RT][NT];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  // codes of half-chunk hc = 2s + h (8 steps): a super-chunk s spans row bytes [128 s, 128 s + 128)
  // and half h its bytes [64 h, 64 h + 64), of which lane group g reads [16 g, 16 g + 16).  Each
  // half is one contiguous 64-byte sector per row, so no sector is fetched twice (two halves
  // interleaved within a line would make the second load re-fetch the streamed line).  Element
  // (h, g, step t8, e) of the super-chunk is 256 h + 64 g + 8 t8 + e, i.e. vector block row
  // 32 h + 8 g + t8 (boff_of below).
  auto load_codes = [&](int64_t hc, uint32_t (&dst)[RT][4]) {
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      // a plain (not streaming) load: the line must survive in L2 until the other half is read
      const u32x4_t q = *reinterpret_cast<const u32x4_t*>(
          gbase + (int64_t)i * 16 * a.ldg + (hc >> 1) * 128 + (hc & 1) * 64);
      dst[i][0] = q.x; dst[i][1] = q.y; dst[i][2] = q.z; dst[i][3] = q.w;
    }
  };
  // B fragments of global step T = 16 s + 8 h + t8 sit in block row 64 s + 32 h + 8 g + t8
  // (FULL: step T = 16 s + t uses codes 128 g + 8 t + e of the super-chunk, block row 64 s + 16 g + t)
  auto boff_of = [&](int64_t T) {
    if constexpr (FULL) return ((T >> 4) * 64 + g * 16 + (T & 15)) * a.KP + col0 + r;
    return ((T >> 4) * 64 + ((T >> 3) & 1) * 32 + g * 8 + (T & 7)) * a.KP + col0 + r;
  };
  const int64_t Tlast = s1 * 16 - 1;
  // B prefetch distance in steps: a step with few column tiles is short, so narrow blocks keep
  // D register sets in flight (D divides 8, so step t8 of a half always uses set t8 % D)
  // (4 with whole-line code sets, which take the room)
  constexpr int D = NT == 1 ? (FULL ? 4 : 8) : (NT == 2 ? 4 : (NT == 3 ? 2 : 1));
  bf16x8_t bh[D][NT], bl[D][NT];
  constexpr int LS = LUT > 1 ? LUT : 1;                    // entry stride of the table
  const uint2* lutl = LUT > 1 ? lut8 + (lane % LS) : lut8;  // this lane's copy
  // 32 copies: entry b of this lane's copy is at byte b * 256 + (lane % 32) * 8, so each lookup
  // address is one mask-and-or of the code word (bits 8-15 select the entry)
  const char* lutb = reinterpret_cast<const char*>(lut8);
  const uint32_t lane_off = (uint32_t)(lane % 32) * 8u;
  // step t8 (0..7 of a half, or 0..15 of a whole line) of the code set cc
  auto step = [&](const auto& cc, int t8, int64_t T) {
    bf16x8_t av[RT];
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      if constexpr (LUT == 32) {
        const uint32_t W = cc[i][t8 >> 1];
        const uint32_t olo = ((t8 & 1) ? (W >> 8) : (W << 8)) & 0xFF00u;
        const uint32_t ohi = ((t8 & 1) ? (W >> 16) : W) & 0xFF00u;
        const uint2 lo = *reinterpret_cast<const uint2*>(lutb + (olo | lane_off));
        const uint2 hi = *reinterpret_cast<const uint2*>(lutb + (ohi | lane_off));
        av[i] = __builtin_bit_cast(bf16x8_t, make_uint4(lo.x, lo.y, hi.x, hi.y));
      } else if constexpr (LUT > 0) {
        const uint32_t w = cc[i][t8 >> 1] >> (16 * (t8 & 1));
        const uint2 lo = lutl[(w & 0xFFu) * LS], hi = lutl[((w >> 8) & 0xFFu) * LS];
        av[i] = __builtin_bit_cast(bf16x8_t, make_uint4(lo.x, lo.y, hi.x, hi.y));
      } else {
        av[i] = unpack8<SQ>((cc[i][t8 >> 1] >> (16 * (t8 & 1))) & 0xFFFFu);
      }
    }
    const int64_t sc = T >> 4;  // super-chunk of this step
    if constexpr (SBC > 0) {
      // staged B: piece (g, step t8 & 7, live column r % SBC) of this half's buffer; lanes of
      // dead columns read a live column's piece, which only reaches unread outputs.  Blocks with
      // hi and lo as two tiles keep the lo pieces 2 x 32 x SBC pieces further on.
      const int pi = ((T >> 3) & 1) * 32 * SBC + (g * 8 + (T & 7)) * SBC + r % a.ncols;
      const bf16x8_t bv = __builtin_bit_cast(bf16x8_t, sb[pi]);
      if (sc < lim[0]) {
#pragma unroll
        for (int i = 0; i < RT; ++i) acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv, acc[i][0], 0, 0, 0);
        if constexpr (!HL) {
          const bf16x8_t bl2 = __builtin_bit_cast(bf16x8_t, sb[2 * 32 * SBC + pi]);
#pragma unroll
          for (int i = 0; i < RT; ++i)
            acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bl2, acc[i][0], 0, 0, 0);
        }
      }
      return;
    }
    const int slot = t8 % D;
    const int64_t bn = boff_of(T + D <= Tlast ? T + D : Tlast);  // refill with step T + D
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      if (sc < lim[j])
#pragma unroll
        for (int i = 0; i < RT; ++i)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bh[slot][j], acc[i][j], 0, 0, 0);
      bh[slot][j] = __builtin_bit_cast(bf16x8_t, xh[bn + j * 16]);
      __builtin_amdgcn_sched_barrier(0);  // keep the refill after the MFMAs that read the set
    }
    if constexpr (!HL) {
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        if (!lo_on[j]) continue;
        if (sc < lim[j])
#pragma unroll
          for (int i = 0; i < RT; ++i)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bl[slot][j], acc[i][j], 0, 0, 0);
        bl[slot][j] = __builtin_bit_cast(bf16x8_t, xl[bn + j * 16]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  };

  if constexpr (FULL) {
    // two whole-line sets; the loop walks pairs of super-chunks so that set indices are constant
    uint32_t cA[RT][8], cB[RT][8];
    const uint8_t* lbase = a.G + (row0 + r) * a.ldg + g * 32;
    auto ld = [&](int64_t sc, uint32_t (&dst)[RT][8]) {
      sc = sc < s1 ? sc : s1 - 1;  // loads past the range re-read the last line
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        const u32x4_t* q = reinterpret_cast<const u32x4_t*>(lbase + (int64_t)i * 16 * a.ldg + sc * 128);
        const u32x4_t q0 = q[0], q1 = q[1];
        dst[i][0] = q0.x; dst[i][1] = q0.y; dst[i][2] = q0.z; dst[i][3] = q0.w;
        dst[i][4] = q1.x; dst[i][5] = q1.y; dst[i][6] = q1.z; dst[i][7] = q1.w;
      }
    };
    // staged B (SBC > 0): piece p < 32 * ncols of half hc = 2 s + h is column p % ncols of block
    // row 64 s + 16 g + 8 h + u with (g, u) = (p / ncols) / 8, % 8; thread threadIdx.x holds
    // pieces threadIdx.x + 64 * WPG * q
    constexpr int NP = SBC > 0 ? (32 * SBC + 64 * WPG - 1) / (64 * WPG) : 1;
    const int nc = a.ncols;
    const int64_t hlast = 2 * s1 - 1;
    uint4 stg[NP], stl[HL ? 1 : NP];
    auto bload = [&](int64_t hc) {
      hc = hc <= hlast ? hc : hlast;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const int p = threadIdx.x + 64 * WPG * q;
        const int row = p / nc;
        if (p < 32 * nc) {
          const int64_t o = ((hc >> 1) * 64 + (row >> 3) * 16 + (hc & 1) * 8 + (row & 7)) * a.KP + p % nc;
          stg[q] = xh[o];
          if constexpr (!HL) stl[q] = xl[o];
        }
      }
    };
    auto bstore = [&](int64_t hc) {
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const int p = threadIdx.x + 64 * WPG * q;
        if (p < 32 * nc) {
          const int o = (hc & 1) * 32 * SBC + (p / nc) * SBC + p % nc;
          sb[o] = stg[q];
          if constexpr (!HL) sb[2 * 32 * SBC + o] = stl[q];
        }
      }
    };
    if (s0 < s1) {
      if constexpr (SBC > 0) bload(2 * s0);
      ld(s0, cA);
      if constexpr (SBC > 0) {
        bstore(2 * s0);
        __syncthreads();
      } else {
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const int64_t b0 = boff_of(s0 * 16 + d);
#pragma unroll
          for (int j = 0; j < NT; ++j) {
            bh[d][j] = __builtin_bit_cast(bf16x8_t, xh[b0 + j * 16]);
            if constexpr (!HL) if (lo_on[j]) bl[d][j] = __builtin_bit_cast(bf16x8_t, xl[b0 + j * 16]);
          }
        }
      }
    }
    // staged B: the buffer of half hc + 1 is written after every wave has finished half hc - 1
    // (the barrier that ended it), so one barrier per half suffices; B of the next half is
    // requested before the codes, so waiting for it never waits for the codes
    // half H of super-chunk s from the codes cc; its first half requests super-chunk s + 1 into nx
    auto half = [&](auto hconst, const uint32_t (&cc)[RT][8], uint32_t (&nx)[RT][8], int64_t s) {
      constexpr int H = decltype(hconst)::value;
      if constexpr (SBC > 0) bload(2 * s + H + 1);
      if constexpr (H == 0) ld(s + 1, nx);
#pragma unroll
      for (int t = 0; t < 8; ++t) step(cc, 8 * H + t, s * 16 + 8 * H + t);
      if constexpr (SBC > 0) {
        bstore(2 * s + H + 1);
        __syncthreads();
      }
    };
    using H0 = std::integral_constant<int, 0>;
    using H1 = std::integral_constant<int, 1>;
    for (int64_t s = s0; s < s1; s += 2) {
      half(H0{}, cA, cB, s);
      half(H1{}, cA, cB, s);
      if (s + 1 < s1) {
        half(H0{}, cB, cA, s + 1);
        half(H1{}, cB, cA, s + 1);
      }
    }
  } else {
  uint32_t cA[RT][4], cB[RT][4];
  if (s0 < s1) {
    load_codes(2 * s0, cA);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int64_t b0 = boff_of(s0 * 16 + d);
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        bh[d][j] = __builtin_bit_cast(bf16x8_t, xh[b0 + j * 16]);
        if constexpr (!HL) if (lo_on[j]) bl[d][j] = __builtin_bit_cast(bf16x8_t, xl[b0 + j * 16]);
      }
    }
  }
  for (int64_t s = s0; s < s1; ++s) {
    load_codes(2 * s + 1, cB);
#pragma unroll
    for (int t8 = 0; t8 < 8; ++t8) step(cA, t8, s * 16 + t8);
    load_codes(s + 1 < s1 ? 2 * s + 2 : 2 * s + 1, cA);
#pragma unroll
    for (int t8 = 0; t8 < 8; ++t8) step(cB, t8, s * 16 + 8 + t8);
  }
  }
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int64_t col = col0 + 16 * j + r;
      const int64_t row = row0 + 16 * i + 4 * g;
      float4* dst = reinterpret_cast<float4*>(a.part + (bz * a.KP + col) * a.rows_pad + row);
      *dst = make_float4(acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
    }
}

// hi/lo-in-columns blocks (HL) run with whole-line codes and B staged in LDS (SBC live columns)
template <int RT, int NT, bool SQ, bool HL = false, int SBC = 0>
__global__ __launch_bounds__(256, 2) void k_bp_pipe(LdsArgs a) {
  __shared__ uint4 sb[SBC > 0 ? (HL ? 2 : 4) * 32 * SBC : 1];
  bp_pipe_tile<RT, NT, SQ, 0, HL, 4, (SBC > 0), SBC>(a, blockIdx.x, blockIdx.z, nullptr, sb);
}

// byte-table unpack with R copies of the table (R = 1: one shared 2 KB table)
template <int RT, int NT, bool SQ, int R = 1, bool HL = false, int SBC = 0>
__global__ __launch_bounds__(256, 2) void k_bp_pipel(LdsArgs a) {
  __shared__ uint2 lut8[256 * R];
  __shared__ uint4 sb[SBC > 0 ? (HL ? 2 : 4) * 32 * SBC : 1];
  fill_byte_table<SQ, R>(lut8);
  __syncthreads();
  bp_pipe_tile<RT, NT, SQ, R, HL, 4, NT == 1, SBC>(a, blockIdx.x, blockIdx.z, lut8, sb);
}

// 32 table copies (64 KB) shared by a 512-thread workgroup of 8 waves on 512 rows: two such
// workgroups fit a CU's LDS, i.e. four waves per SIMD
template <int RT, int NT, bool SQ, bool HL = false, int SBC = 0>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(HL ? 4 : 2))) void k_bp_pipew(LdsArgs a) {
  __shared__ uint2 lut8[256 * 32];
  __shared__ uint4 sb[SBC > 0 ? (HL ? 2 : 4) * 32 * SBC : 1];
  fill_byte_table<SQ, 32, 512>(lut8);
  __syncthreads();
  bp_pipe_tile<RT, NT, SQ, 32, HL, 8, (SBC > 0), SBC>(a, blockIdx.x, blockIdx.z, lut8, sb);
}

// Persistent form: a grid sized to what the GPU holds at once (two workgroups per CU) walks the
// (row block, split) tiles in order, tile v = split * nbx + row block.  All resident workgroups
// then sweep the same reduction range at the same pace, so each vector-block line fetched into
// an XCD's L2 serves every workgroup of that XCD instead of being re-fetched by late starters.
template <int RT, int NT, bool SQ, bool HL = false>
__global__ __launch_bounds__(256, 2) void k_bp_persist(LdsArgs a, int64_t nbx, int64_t ntiles) {
  for (int64_t v = blockIdx.x; v < ntiles; v += gridDim.x) bp_pipe_tile<RT, NT, SQ, 0, HL>(a, v % nbx, v / nbx);
}

// Pass-1 epilogue from the split partials: T_jk = w_j (P_jk - gm_j s_k + xbar_k (pc_j - n gm_j));
// writes the BF16 hi/lo pass-2 operand, optionally the column-major result, and per-workgroup
// partial sums of gm_j * represented(T_jk).  A workgroup covers 64 columns x kFin1Rows (64) rows:
// the partials are summed over the splits with row-contiguous reads (16 threads per column, 256
// contiguous bytes) into LDS, then thread (tx = column, ty) finishes rows 16 ty .. 16 ty + 15, so
// that the operand's writes stay contiguous across the columns.  Every global load of a thread
// (its partials and the 16 rows' gm, w and popcounts) is issued before the one barrier, so a
// workgroup waits on one memory round trip; small workgroups keep enough of them resident.  Tail
// hi/lo layouts: P_jk of a tail vector is the sum of the partials of columns k and k + tail_r,
// and threads of the columns holding lo halves only write their (zero) workgroup sums; pad columns
// past them write their hi halves (zero) only.
__global__ __launch_bounds__(256) void k_fin1(Fin1Args a) {
  __shared__ double red[4][64];
  __shared__ float pt[64][65];  // [column][row of the tile], padded against bank conflicts
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + tx;
  // batched traits: this column's per-SNP model terms and subset size
  const int trait = (a.tc.col_trait && k < a.t) ? a.tc.col_trait[k] : -1;
  const float* gmk = trait >= 0 ? a.tc.gm + (int64_t)trait * a.tc.ld_m : a.gm;
  const float* gwk = trait >= 0 ? (a.gw ? a.tc.gw + (int64_t)trait * a.tc.ld_m : nullptr) : a.gw;
  const int32_t* pck = trait >= 0 ? a.tc.pc + (int64_t)trait * a.tc.ld_m : a.popcnt;
  const double nk = trait >= 0 ? a.tc.n_t[trait] : a.xsum_n;
  const bool tail = a.tail_r > 0 && k >= a.tail0 && k < a.t;
  const bool lo_col = a.tail_r > 0 && k >= a.t && k < a.tail0 + 2 * a.tail_r;
  const bool active = k < a.KP && !lo_col;
  const int64_t trow0 = (int64_t)blockIdx.y * kFin1Rows;
  const int rl0 = ty * 16;
  // the rows' model terms, requested with the partials below
  float gmv[16], wv[16];
  int32_t pcv[16];
  if (active) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int64_t row = trow0 + rl0 + e;
      gmv[e] = gmk[row];
      wv[e] = gwk ? gwk[row] : 1.0f;
      pcv[e] = pck[row];
    }
  }
  const double sk = active ? a.colsum[k] : 0.0;
  const double xb = (active && a.xsum && k < a.t) ? (double)(float)(a.xsum[k] / nk) : 0.0;
  // sums of the partials: piece q = column q / 16, rows 4 (q % 16) .. + 3 of the tile.  Padding
  // columns (k >= t) are zero; their partials are not read (the block products leave them
  // undefined).
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    const int q = threadIdx.x + 256 * it;
    const int c = q >> 4, r4 = (q & 15) * 4;
    const int kk = blockIdx.x * 64 + c;
    float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
    if (kk < a.t) {
      const bool tl = a.tail_r > 0 && kk >= a.tail0;
      for (int h = 0; h < (tl ? 2 : 1); ++h)
        for (int z = 0; z < a.nsplit; ++z) {
          const float4 v = *reinterpret_cast<const float4*>(
              a.part + ((int64_t)z * a.KP + kk + h * a.tail_r) * a.rows_pad + trow0 + r4);
          sum.x += v.x; sum.y += v.y; sum.z += v.z; sum.w += v.w;
        }
    }
    pt[c][r4] = sum.x; pt[c][r4 + 1] = sum.y; pt[c][r4 + 2] = sum.z; pt[c][r4 + 3] = sum.w;
  }
  __syncthreads();
  double usum = 0.0;
  if (active) {
#pragma unroll
    for (int g8 = 0; g8 < 2; ++g8) {
      const int rl = rl0 + g8 * 8;
      const int64_t i0 = trow0 + rl;
      uint32_t hb[8], lb[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t row = i0 + e;
        const float P = pt[tx][rl + e];
        const float gm1 = gmv[g8 * 8 + e];
        const double rem = xb * ((double)pcv[g8 * 8 + e] - nk * (double)gm1);
        const float tv = a.raw ? P : wv[g8 * 8 + e] * (float)((double)P - (double)gm1 * sk + rem);
        if (a.x2_fp32) {  // FP32 operand for the int8 kernels, which sum its digits themselves
          if (a.X2) reinterpret_cast<float*>(a.X2)[row * a.KP + k] = tv;
          usum += (double)gm1 * (double)tv;
        } else {
          hb[e] = bf16_bits_rn(tv);
          lb[e] = bf16_bits_rn(tv - bf16_to_float(hb[e]));
          usum += (double)gm1 * ((double)bf16_to_float(hb[e]) + (double)bf16_to_float(lb[e]));
        }
        if (a.out && k < a.t && row < a.rows) a.out[row + (int64_t)k * a.ldo] = tv;
      }
      if (a.X2 && !a.x2_fp32) {
        uint4* x2 = reinterpret_cast<uint4*>(a.X2);
        const int64_t idx = (i0 >> 3) * a.KP + k;
        x2[idx] = make_uint4(hb[0] | (hb[1] << 16), hb[2] | (hb[3] << 16), hb[4] | (hb[5] << 16), hb[6] | (hb[7] << 16));
        const uint4 lo4 =
            make_uint4(lb[0] | (lb[1] << 16), lb[2] | (lb[3] << 16), lb[4] | (lb[5] << 16), lb[6] | (lb[7] << 16));
        if (tail) x2[idx + a.tail_r] = lo4;
        else if (a.tail_r == 0 || k < a.tail0) x2[idx + a.lo_offset16] = lo4;
      }
    }
  }
  red[ty][tx] = usum;
  __syncthreads();
  if (ty == 0 && k < a.KP && a.upart)
    a.upart[(int64_t)blockIdx.y * a.KP + k] = (red[0][tx] + red[1][tx]) + (red[2][tx] + red[3][tx]);
}

// Pass-2 epilogue: out_ik = rr_i (sum_z P_ik - u_k) (+ tau0 v_ik for LDET columns).
__global__ __launch_bounds__(256) void k_fin2(Fin2Args a) {
  const int k = blockIdx.y;
  const int64_t i = 4 * ((int64_t)blockIdx.x * 256 + threadIdx.x);
  if (i >= a.rows || k >= a.t) return;
  float4 P = make_float4(0.f, 0.f, 0.f, 0.f);
  const bool tail = a.tail_r > 0 && k >= a.tail0;  // tail hi/lo: lo partials in column k + tail_r
  for (int h = 0; h < (tail ? 2 : 1); ++h)
    for (int z = 0; z < a.nsplit; ++z) {
      const float4 p =
          *reinterpret_cast<const float4*>(a.part + ((int64_t)z * a.KP + k + h * a.tail_r) * a.rows_pad + i);
      P.x += p.x; P.y += p.y; P.z += p.z; P.w += p.w;
    }
  const float Pv[4] = {P.x, P.y, P.z, P.w};
  const double uk = a.colsum[k];
  const bool add_tau = a.v && (a.kinds == nullptr || a.kinds[k] == 0);
  const float* tmask = a.tc.col_trait ? a.tc.mask + (int64_t)a.tc.col_trait[k] * a.tc.ld_n : nullptr;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int64_t row = i + e;
    if (row >= a.rows) break;
    float zv = (float)((double)Pv[e] - uk);
    if (a.rr) zv *= a.rr[row];
    if (tmask) zv *= tmask[row];
    if (add_tau) zv += a.tau0 * a.v[row + (int64_t)k * a.ldv];
    a.out[row + (int64_t)k * a.ldo] = zv;
  }
}

// ---------------------------------------------------------------------------- launchers -----
int lds_nt(int t) {
  const int nt = (t + 15) / 16;
  return nt < 8 ? nt : 8;
}

int lds_splits(int64_t rows_pad, int64_t nsuper, int64_t target_wgs, int64_t& ns_per) {
  const int64_t wgs = rows_pad / (64 * kLdsRT);  // 256-row workgroups
  int64_t S = (target_wgs + wgs - 1) / wgs;
  if (S < 1) S = 1;
  if (S > nsuper) S = nsuper;
  ns_per = (nsuper + S - 1) / S;
  return (int)((nsuper + ns_per - 1) / ns_per);
}

template <int NT, bool SQ>
static void launch_variant(const LdsArgs& a, int variant, dim3 grid, hipStream_t st) {
  switch (variant) {
    case BP_REG: k_bp_reg<kLdsRT, NT, SQ><<<grid, 256, 0, st>>>(a); break;
    case BP_REGL: k_bp_reg<kLdsRT, NT, SQ, true><<<grid, 256, 0, st>>>(a); break;
    case BP_REG8:
      grid.x /= 2;  // 8 row tiles per wave: 512 rows per workgroup
      if constexpr (NT <= 5) k_bp_reg<8, NT, SQ><<<grid, 256, 0, st>>>(a);
      else k_bp_reg<kLdsRT, NT, SQ><<<dim3(grid.x * 2, grid.y, grid.z), 256, 0, st>>>(a);
      break;
    case BP_PIPE:  // wider column tiles exceed two waves' register budget: plain register variant
      if constexpr (NT <= 5) k_bp_pipe<kLdsRT, NT, SQ><<<grid, 256, 0, st>>>(a);
      else k_bp_reg<kLdsRT, NT, SQ><<<grid, 256, 0, st>>>(a);
      break;
    case BP_PIPEL:
      if constexpr (NT <= 5) k_bp_pipel<kLdsRT, NT, SQ><<<grid, 256, 0, st>>>(a);
      else k_bp_reg<kLdsRT, NT, SQ, true><<<grid, 256, 0, st>>>(a);
      break;
    case BP_PERSIST:
      if constexpr (NT <= 5) {
        const int64_t nbx = grid.x, ntiles = (int64_t)grid.x * grid.z;
        const int64_t resident = std::max<int64_t>(64, 512 / (int64_t)grid.y);  // 2 workgroups per CU
        k_bp_persist<kLdsRT, NT, SQ><<<dim3((unsigned)std::min(ntiles, resident), grid.y, 1), 256, 0, st>>>(
            a, nbx, ntiles);
      } else {
        k_bp_reg<kLdsRT, NT, SQ><<<grid, 256, 0, st>>>(a);
      }
      break;
    case BP_PIPER16:
      if constexpr (NT <= 5) k_bp_pipel<kLdsRT, NT, SQ, 16><<<grid, 256, 0, st>>>(a);
      else k_bp_reg<kLdsRT, NT, SQ, true><<<grid, 256, 0, st>>>(a);
      break;
    case BP_PIPER32:
      if constexpr (NT <= 5) k_bp_pipew<kLdsRT, NT, SQ><<<dim3(grid.x / 2, grid.y, grid.z), 512, 0, st>>>(a);
      else k_bp_reg<kLdsRT, NT, SQ, true><<<grid, 256, 0, st>>>(a);
      break;
    case BP_LDS4: k_bp_lds<kLdsRT, NT, 4, true, SQ><<<grid, 256, 0, st>>>(a); break;
    case BP_LDS2: k_bp_lds<kLdsRT, NT, 2, false, SQ><<<grid, 256, 0, st>>>(a); break;
    default: fail(SLEMM_ERR_ARG, "unknown block-product variant");
  }
}

// hi/lo-in-columns blocks (one column tile, at most SBC live columns): pipelined variants only
template <bool SQ, int SBC>
static void launch_variant_hl(const LdsArgs& a, int variant, dim3 grid, hipStream_t st) {
  switch (variant) {
    case BP_PIPEL: k_bp_pipel<kLdsRT, 1, SQ, 1, true, SBC><<<grid, 256, 0, st>>>(a); break;
    case BP_PIPER16: k_bp_pipel<kLdsRT, 1, SQ, 16, true, SBC><<<grid, 256, 0, st>>>(a); break;
    case BP_PIPER32: k_bp_pipew<kLdsRT, 1, SQ, true, SBC><<<dim3(grid.x / 2, grid.y, grid.z), 512, 0, st>>>(a); break;
    case BP_PERSIST: {  // B from global memory, per step
      const int64_t nbx = grid.x, ntiles = (int64_t)grid.x * grid.z;
      const int64_t resident = std::max<int64_t>(64, 512 / (int64_t)grid.y);
      k_bp_persist<kLdsRT, 1, SQ, true><<<dim3((unsigned)std::min(ntiles, resident), grid.y, 1), 256, 0, st>>>(
          a, nbx, ntiles);
      break;
    }
    default:
      if (variant < 0 || variant >= BP_NVARIANTS) fail(SLEMM_ERR_ARG, "unknown block-product variant");
      k_bp_pipe<kLdsRT, 1, SQ, true, SBC><<<grid, 256, 0, st>>>(a);
  }
}

// SLEMM_STAGED_ONE_TILE=0 keeps one-tile hi/lo blocks on the per-step B loads (read per call)
static bool staged_one_tile() {
  const char* e = std::getenv("SLEMM_STAGED_ONE_TILE");
  return !(e && e[0] == '0');
}

// one-tile blocks with hi and lo as two tiles (up to 16 live columns): the pipelined variants
// with the block's live B columns staged in LDS per half super-chunk, as for hi/lo-in-columns
// blocks; the other variants run their usual kernels
template <int SBC>
static void launch_variant_nt1s(const LdsArgs& a, int variant, dim3 grid, hipStream_t st) {
  switch (variant) {
    case BP_PIPE: k_bp_pipe<kLdsRT, 1, false, false, SBC><<<grid, 256, 0, st>>>(a); break;
    case BP_PIPEL: k_bp_pipel<kLdsRT, 1, false, 1, false, SBC><<<grid, 256, 0, st>>>(a); break;
    case BP_PIPER16: k_bp_pipel<kLdsRT, 1, false, 16, false, SBC><<<grid, 256, 0, st>>>(a); break;
    case BP_PIPER32:
      k_bp_pipew<kLdsRT, 1, false, false, SBC><<<dim3(grid.x / 2, grid.y, grid.z), 512, 0, st>>>(a);
      break;
    default: launch_variant<1, false>(a, variant, grid, st);
  }
}

void launch_bp(const LdsArgs& a, int NT, int nsplit, int variant, hipStream_t st) {
  if (a.rows_pad % (64 * kLdsRT) != 0) fail(SLEMM_ERR_ARG, "row padding mismatch");
  // a tail hi/lo tile is only implemented by the pipelined kernels (NT <= 5)
  if (a.tail_tile >= 0) {
    if (NT > 5) fail(SLEMM_ERR_ARG, "tail hi/lo blocks need at most 5 column tiles per group");
    if (!bp_hl_variant(variant)) variant = BP_PIPE;
  }
  if (a.KP % (16 * NT) != 0) fail(SLEMM_ERR_ARG, "column padding mismatch");
  dim3 grid((unsigned)(a.rows_pad / (64 * kLdsRT)), (unsigned)(a.KP / (16 * NT)), (unsigned)nsplit);
  if (a.hl) {
    if (NT != 1) fail(SLEMM_ERR_ARG, "hi/lo-in-columns blocks have one column tile");
    if (a.ncols < 1 || a.ncols > 16) fail(SLEMM_ERR_ARG, "hi/lo-in-columns blocks have 1 to 16 live columns");
    if (a.squared) {
      if (a.ncols > 2) fail(SLEMM_ERR_ARG, "squared-dosage products take one vector");
      launch_variant_hl<true, 2>(a, variant, grid, st);
    } else if (a.ncols <= 2) {
      launch_variant_hl<false, 2>(a, variant, grid, st);
    } else if (a.ncols <= 8) {
      launch_variant_hl<false, 8>(a, variant, grid, st);
    } else {
      launch_variant_hl<false, 16>(a, variant, grid, st);
    }
  } else if (a.squared) {
    if (NT != 1) fail(SLEMM_ERR_ARG, "squared-dosage products support one column tile");
    launch_variant<1, true>(a, variant, grid, st);
  } else if (NT == 1 && a.tail_tile < 0 && a.ncols >= 1 && a.ncols <= 16 && staged_one_tile()) {
    if (a.ncols <= 8) launch_variant_nt1s<8>(a, variant, grid, st);
    else launch_variant_nt1s<16>(a, variant, grid, st);
  } else {
    switch (NT) {
      case 1: launch_variant<1, false>(a, variant, grid, st); break;
      case 2: launch_variant<2, false>(a, variant, grid, st); break;
      case 3: launch_variant<3, false>(a, variant, grid, st); break;
      case 4: launch_variant<4, false>(a, variant, grid, st); break;
      case 5: launch_variant<5, false>(a, variant, grid, st); break;
      case 6: launch_variant<6, false>(a, variant, grid, st); break;
      case 7: launch_variant<7, false>(a, variant, grid, st); break;
      case 8: launch_variant<8, false>(a, variant, grid, st); break;
      default: fail(SLEMM_ERR_ARG, "unsupported column tile count");
    }
  }
  HIPCHECK(hipGetLastError());
}

void launch_fin1(const Fin1Args& a, hipStream_t st) {
  dim3 grid((unsigned)((a.KP + 63) / 64), (unsigned)(a.rows_pad / kFin1Rows));
  k_fin1<<<grid, 256, 0, st>>>(a);
  HIPCHECK(hipGetLastError());
}

void launch_fin2(const Fin2Args& a, hipStream_t st) {
  dim3 grid((unsigned)((a.rows + 1023) / 1024), (unsigned)a.t);
  k_fin2<<<grid, 256, 0, st>>>(a);
  HIPCHECK(hipGetLastError());
}

}  // namespace slemm
