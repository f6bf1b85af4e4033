// Genotype block products on gfx950 MFMA.
//
// Both halves of the operator Zc W Zc' V are the same shape: a packed 2-bit matrix G
// (rows x C, C = reduction length, codes unpacked on the fly to dosages 0/1/2) times a
// skinny dense block X (C x K):
//   pass 1 (Z'V): rows = SNPs,        G = SNP-major copy,        X = R^-1/2-scaled V
//   pass 2 (Z U): rows = individuals, G = individual-major copy, X = W-weighted pass-1 result
// Centring is exact and never enters the MFMA: Zc'v = Z'v - gm (1'v) and Zc u = Z u - 1 (gm'u),
// applied in the epilogues with column sums of the *represented* X, so the MFMA sees only the
// exact integers {0,1,2}.
//
// Work decomposition: a workgroup is 4 waves; each wave owns RT row tiles of 16 rows and NT
// column tiles of 16 columns and streams the whole reduction in super-chunks of 512 elements
// (128 bytes of codes per row).  Within a super-chunk lane (r = l%16, g = l/16) loads the 32
// bytes of codes g*32..g*32+31 of each of its rows once and keeps them in registers; the
// 16x16 MFMA tiles consume them 8 (BF16) or 1 (FP32) element(s) at a time.
//
// Operand formats of X (written by the pack kernels / the pass-1 epilogue):
//   FP32:   float  X[C][KP]                    (B operand of v_mfma_f32_16x16x4_f32)
//   BF16X2: bf16   Xh[C/8][KP][8], Xl[C/8][KP][8] = hi and lo parts, v = hi + lo
//           (B operand of v_mfma_f32_16x16x32_bf16; 16 contiguous bytes per lane)
#include <hip/hip_runtime.h>

#include <vector>

#include "blockprod.hpp"

namespace slemm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t bf16_bits_rn(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;  // round to nearest even (finite inputs)
}
__device__ __forceinline__ float bf16_to_float(uint32_t b) { return __uint_as_float(b << 16); }

// 8 two-bit codes (16 bits) -> 8 BF16 dosages.  d = low bit + high bit; BF16(d) built with a
// byte permute: low byte = {00,80,00}[d], high byte = {00,3F,40}[d].
__device__ __forceinline__ bf16x8_t unpack8_bf16(uint32_t p16) {
  const uint32_t s = (p16 & 0x5555u) + ((p16 >> 1) & 0x5555u);
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t t = (s >> (4 * q)) & 0xFu;
    const uint32_t x = (t | (t << 14)) & 0x00030003u;  // d0 -> bits 0-1, d1 -> bits 16-17
    const uint32_t sel = x * 0x0101u + 0x04000400u;
    w[q] = __builtin_amdgcn_perm(0x00403F00u, 0x00008000u, sel);
  }
  uint4 v = make_uint4(w[0], w[1], w[2], w[3]);
  return __builtin_bit_cast(bf16x8_t, v);
}

template <int MODE, int RT, int NT, int EPI>
__global__ __launch_bounds__(256) void k_blockprod(BlockProdArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * (64 * RT) + wave * (16 * RT);
  const int col0 = blockIdx.y * (16 * NT);
  const int KP = a.KP;

  f32x4_t acc[RT][NT];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  const uint8_t* gbase = a.G + (row0 + r) * a.ldg + g * 32;

  for (int64_t s = 0; s < a.nsuper; ++s) {
    uint32_t pk[RT][8];
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const u32x4_t* src = reinterpret_cast<const u32x4_t*>(gbase + (int64_t)i * 16 * a.ldg + s * 128);
      const u32x4_t q0 = __builtin_nontemporal_load(src);
      const u32x4_t q1 = __builtin_nontemporal_load(src + 1);
      pk[i][0] = q0.x; pk[i][1] = q0.y; pk[i][2] = q0.z; pk[i][3] = q0.w;
      pk[i][4] = q1.x; pk[i][5] = q1.y; pk[i][6] = q1.z; pk[i][7] = q1.w;
    }
    if constexpr (MODE == PREC_BF16X2) {
      const uint4* xh = reinterpret_cast<const uint4*>(a.X);
      const uint4* xl = xh + a.lo_offset16;
      const int64_t blk0 = s * 64 + g * 16;
#pragma unroll 4
      for (int t = 0; t < 16; ++t) {
        bf16x8_t av[RT];
#pragma unroll
        for (int i = 0; i < RT; ++i) av[i] = unpack8_bf16((pk[i][t >> 1] >> (16 * (t & 1))) & 0xFFFFu);
        const int64_t boff = (blk0 + t) * KP + col0 + r;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          const bf16x8_t bh = __builtin_bit_cast(bf16x8_t, xh[boff + j * 16]);
          const bf16x8_t bl = __builtin_bit_cast(bf16x8_t, xl[boff + j * 16]);
#pragma unroll
          for (int i = 0; i < RT; ++i) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bh, acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bl, acc[i][j], 0, 0, 0);
          }
        }
      }
    } else {
      const float* xf = reinterpret_cast<const float*>(a.X);
      const int64_t e0 = s * 512 + g * 128;
#pragma unroll 8
      for (int t = 0; t < 128; ++t) {
        float av[RT];
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          const uint32_t code = (pk[i][t >> 4] >> (2 * (t & 15))) & 3u;
          av[i] = (float)((code & 1u) + (code >> 1));
        }
        const int64_t boff = (e0 + t) * KP + col0 + r;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          const float b = xf[boff + j * 16];
#pragma unroll
          for (int i = 0; i < RT; ++i)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], b, acc[i][j], 0, 0, 0);
        }
      }
    }
  }

  // ---------------------------------------------------------------- epilogues
  // acc[i][j][v] = P[row0 + 16i + 4g + v][col0 + 16j + r]
  if constexpr (EPI == EPI_PASS1) {
    __shared__ double red[4][16 * NT];
    double upart[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) upart[j] = 0.0;
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int k = col0 + 16 * j + r;
      const double sk = a.colsum[k];
      const double xb = (a.xsum && k < a.t) ? (double)(float)(a.xsum[k] / a.xsum_n) : 0.0;
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        const int64_t rowb = row0 + 16 * i + 4 * g;
        float tv[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t row = rowb + v;
          const float gmv = a.gm[row];
          const float wv = a.gw ? a.gw[row] : 1.0f;
          // Zc'x = Z'x~ - gm (1'x~) + xbar (popcnt - n gm),  x = x~ + xbar
          const double rem = xb * ((double)a.popcnt[row] - a.xsum_n * (double)gmv);
          const float cen = (float)((double)acc[i][j][v] - (double)gmv * sk + rem);
          tv[v] = wv * cen;
        }
        // represented value (what pass 2 will consume) and its gm-weighted column sum
        if (a.out2) {
          if (a.out2_mode == PREC_BF16X2) {
            uint32_t hb[4], lb[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              hb[v] = bf16_bits_rn(tv[v]);
              lb[v] = bf16_bits_rn(tv[v] - bf16_to_float(hb[v]));
              upart[j] += (double)a.gm[rowb + v] * ((double)bf16_to_float(hb[v]) + (double)bf16_to_float(lb[v]));
            }
            const int64_t off = ((rowb >> 3) * KP + k) * 8 + (rowb & 7);  // in bf16 elements
            uint16_t* o2 = reinterpret_cast<uint16_t*>(a.out2);
            *reinterpret_cast<uint2*>(o2 + off) = make_uint2(hb[0] | (hb[1] << 16), hb[2] | (hb[3] << 16));
            *reinterpret_cast<uint2*>(o2 + off + a.out2_lo_offset) =
                make_uint2(lb[0] | (lb[1] << 16), lb[2] | (lb[3] << 16));
          } else {
            float* o2 = reinterpret_cast<float*>(a.out2);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              o2[(rowb + v) * KP + k] = tv[v];
              upart[j] += (double)a.gm[rowb + v] * (double)tv[v];
            }
          }
        }
        if (a.out && k < a.t) {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (rowb + v < a.rows) a.out[(rowb + v) + (int64_t)k * a.ldo] = tv[v];
        }
      }
    }
    if (a.upart) {
      // reduce upart over the 4 lane groups, then over the 4 waves (fixed order)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        upart[j] += __shfl_xor(upart[j], 16);
        upart[j] += __shfl_xor(upart[j], 32);
      }
      if (g == 0)
#pragma unroll
        for (int j = 0; j < NT; ++j) red[wave][16 * j + r] = upart[j];
      __syncthreads();
      if (threadIdx.x < 16 * NT) {
        const int cc = threadIdx.x;
        const double sum = (red[0][cc] + red[1][cc]) + (red[2][cc] + red[3][cc]);
        a.upart[(int64_t)blockIdx.x * KP + col0 + cc] = sum;
      }
    }
  } else {
    // pass 2: out[i,k] = rr_i (P - u_k) (+ tau0 v_ik for LDET columns)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int k = col0 + 16 * j + r;
      if (k >= a.t) continue;
      const double uk = a.colsum[k];
      const bool add_tau = a.v && (a.kinds == nullptr || a.kinds[k] == 0);
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        const int64_t rowb = row0 + 16 * i + 4 * g;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t row = rowb + v;
          if (row < a.rows) {
            float z = (float)((double)acc[i][j][v] - uk);
            if (a.rr) z *= a.rr[row];
            if (add_tau) z += a.tau0 * a.v[row + (int64_t)k * a.ldv];
            a.out[row + (int64_t)k * a.ldo] = z;
          }
        }
      }
    }
  }
}

template <int MODE, int EPI>
static void launch_nt(const BlockProdArgs& a, int NT, int64_t rows_pad, hipStream_t st) {
#define SLEMM_LAUNCH(NTV, RTV)                                                                \
  case NTV: {                                                                                 \
    dim3 grid((unsigned)(rows_pad / (64 * RTV)), (unsigned)(a.KP / (16 * NTV)));              \
    k_blockprod<MODE, RTV, NTV, EPI><<<grid, 256, 0, st>>>(a);                                \
    break;                                                                                    \
  }
  switch (NT) {
    SLEMM_LAUNCH(1, 8)
    SLEMM_LAUNCH(2, 8)
    SLEMM_LAUNCH(3, 4)
    SLEMM_LAUNCH(4, 4)
    SLEMM_LAUNCH(5, 4)
    SLEMM_LAUNCH(6, 4)
    SLEMM_LAUNCH(7, 4)
    SLEMM_LAUNCH(8, 4)
    default:
      fail(SLEMM_ERR_ARG, "unsupported column tile count");
  }
#undef SLEMM_LAUNCH
  HIPCHECK(hipGetLastError());
}

int blockprod_nt(int t) {
  const int nt = (t + 15) / 16;
  return nt < 8 ? nt : 8;
}

int64_t blockprod_rows_per_wg(int NT) { return NT <= 2 ? 512 : 256; }

// Optional per-launch timing of the block-product kernels with HIP events recorded on the
// launch stream (slemm_gpu_timing_*).  Events are created per launch and read back on demand.
namespace {
struct TimingRec {
  hipEvent_t beg, end;
  int pass, precision;
};
bool g_timing = false;
std::vector<TimingRec> g_timing_recs;
}  // namespace

TimingToken timing_begin(hipStream_t st) {
  TimingToken tok;
  if (!g_timing) return tok;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (st && hipStreamIsCapturing(st, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone) return tok;
  HIPCHECK(hipEventCreate(&tok.beg));
  HIPCHECK(hipEventRecord(tok.beg, st));
  tok.on = true;
  return tok;
}

void timing_end(TimingToken& tok, hipStream_t st, int pass, int precision) {
  if (!tok.on) return;
  TimingRec rec{};
  rec.beg = tok.beg;
  HIPCHECK(hipEventCreate(&rec.end));
  HIPCHECK(hipEventRecord(rec.end, st));
  rec.pass = pass;
  rec.precision = precision;
  g_timing_recs.push_back(rec);
  tok.on = false;
}

void launch_blockprod(const BlockProdArgs& a, int precision, int epi, int NT, int64_t rows_pad,
                      hipStream_t st) {
  if (rows_pad % blockprod_rows_per_wg(NT) != 0) fail(SLEMM_ERR_ARG, "row padding mismatch");
  if (a.KP % (16 * NT) != 0) fail(SLEMM_ERR_ARG, "column padding mismatch");
  TimingToken tok = timing_begin(st);
  if (precision == PREC_BF16X2) {
    if (epi == EPI_PASS1) launch_nt<PREC_BF16X2, EPI_PASS1>(a, NT, rows_pad, st);
    else launch_nt<PREC_BF16X2, EPI_PASS2>(a, NT, rows_pad, st);
  } else {
    if (epi == EPI_PASS1) launch_nt<PREC_FP32, EPI_PASS1>(a, NT, rows_pad, st);
    else launch_nt<PREC_FP32, EPI_PASS2>(a, NT, rows_pad, st);
  }
  timing_end(tok, st, epi, precision);
}

}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_timing(int enable) {
  return guard([&] { g_timing = enable != 0; });
}

extern "C" int slemm_gpu_timing_read(int pass, int precision, double* total_ms, int* launches) {
  return guard([&] {
    REQUIRE(total_ms && launches, "null argument");
    double tot = 0.0;
    int cnt = 0;
    for (auto& r : g_timing_recs) {
      if (r.pass != pass || r.precision != precision) continue;
      HIPCHECK(hipEventSynchronize(r.end));
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, r.beg, r.end));
      tot += ms;
      ++cnt;
    }
    *total_ms = tot;
    *launches = cnt;
  });
}

extern "C" int slemm_gpu_timing_reset(void) {
  return guard([&] {
    for (auto& r : g_timing_recs) {
      (void)hipEventSynchronize(r.end);
      (void)hipEventDestroy(r.beg);
      (void)hipEventDestroy(r.end);
    }
    g_timing_recs.clear();
  });
}

namespace slemm {

// ------------------------------------------------------------------ pack (pass-1 operand) ----
// x_ik = rr_i * (v_ik - sum_q Q_iq a_qk)   (projection only for QUADFORM columns, a = Q'V)
// A block covers kPackRows = 64 rows x 64 columns (17 KB of LDS, so that eight blocks fit a CU
// and keep enough loads in flight).  The block's V tile is first staged in LDS with
// row-contiguous reads (lane = row, wave ty loading columns ty, ty + 4, ...), then thread
// (tx = column within 64, ty) packs 2 blocks of 8 rows of its column, so the packed writes stay
// column-adjacent too.
__global__ __launch_bounds__(256) void k_pack(PackArgs a) {
  __shared__ double red[4][64];
  __shared__ float vt[64][kPackRows + 1];  // [column][row]; the pad makes the column-wise reads conflict-free
  __shared__ float rrs[kPackRows];         // the rows' r_rsqrt
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + tx;
  const int64_t rblk = (int64_t)blockIdx.y * kPackRows;
  const int64_t rbase = rblk + ty * (kPackRows / 4);
  {
    const int64_t i = rblk + tx;
    const int kmax = min(64, a.t - (int)blockIdx.x * 64);
    if (a.rr && ty == 0) rrs[tx] = i < a.n ? a.rr[i] : 0.f;
#pragma unroll
    for (int q = 0; q < 16; ++q) {  // 16 loads in flight per thread
      const int kk = ty + 4 * q;
      vt[kk][tx] = (kk < kmax && i < a.n) ? a.V[i + (int64_t)(blockIdx.x * 64 + kk) * a.ldv] : 0.f;
    }
  }
  __syncthreads();
  double csum = 0.0;
  const bool live = k < a.t;
  const bool proj = live && a.qtv && (a.kinds == nullptr || a.kinds[k] == 1) && a.c > 0;
  // batched traits: the column's subset mask, intercept basis and subset size
  const int trait = (live && a.tc.col_trait) ? a.tc.col_trait[k] : -1;
  const float* tmask = trait >= 0 ? a.tc.mask + (int64_t)trait * a.tc.ld_n : nullptr;
  const float* Qk = trait >= 0 ? a.tc.q + (int64_t)trait * a.tc.ld_n : a.Q;
  const double nk = trait >= 0 ? a.tc.n_t[trait] : (double)a.n;
  // column mean removed before the product (the exact remainder is added back in the pass-1
  // epilogue), so the MFMA accumulates a centred sum and a constant column costs no accuracy
  const float xbar = (live && a.xsum) ? (float)(a.xsum[k] / nk) : 0.f;
  if (k < a.KP) {
    for (int b8 = 0; b8 < kPackRows / 32; ++b8) {
      const int64_t i0 = rbase + b8 * 8;
      float x[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t i = i0 + e;
        float val = 0.f;
        if (live && i < a.n && (!tmask || tmask[i] != 0.f)) {
          float v = vt[tx][i - rblk];
          if (proj) {
            double acc = 0.0;
            for (int q = 0; q < a.c; ++q) acc += (double)Qk[i + (int64_t)q * a.ldq] * a.qtv[q * a.KP + k];
            v = (float)((double)v - acc);
          }
          val = a.rr ? rrs[i - rblk] * v : v;
          val -= xbar;
        }
        x[e] = val;
      }
      if (!a.X) {  // sums-only sweep (column sums of the un-centred operand)
#pragma unroll
        for (int e = 0; e < 8; ++e) csum += (double)x[e];
        continue;
      }
      if (a.mode == PREC_BF16X2) {
        uint32_t hb[8], lb[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          hb[e] = bf16_bits_rn(x[e]);
          lb[e] = bf16_bits_rn(x[e] - bf16_to_float(hb[e]));
          const double rep = (double)bf16_to_float(hb[e]) + (double)bf16_to_float(lb[e]);
          csum += a.wsum ? (i0 + e < a.n ? (double)a.wsum[i0 + e] * rep : 0.0) : rep;
        }
        // tail hi/lo: columns [t, tail0 + 2 tail_r) hold lo halves of the tail vectors
        if (a.tail_r > 0 && !live && k < a.tail0 + 2 * a.tail_r) continue;
        uint4* xh = reinterpret_cast<uint4*>(a.X);
        const int64_t idx = (i0 >> 3) * a.KP + k;
        xh[idx] = make_uint4(hb[0] | (hb[1] << 16), hb[2] | (hb[3] << 16), hb[4] | (hb[5] << 16), hb[6] | (hb[7] << 16));
        const uint4 lo4 =
            make_uint4(lb[0] | (lb[1] << 16), lb[2] | (lb[3] << 16), lb[4] | (lb[5] << 16), lb[6] | (lb[7] << 16));
        if (a.tail_r > 0 && live && k >= a.tail0) xh[idx + a.tail_r] = lo4;
        else if (a.tail_r == 0) xh[idx + a.lo_offset16] = lo4;
        else if (k < a.tail0) xh[idx + a.lo_offset16] = lo4;  // pad columns past the tail: hi only
      } else {
        float* xf = reinterpret_cast<float*>(a.X);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          xf[(i0 + e) * a.KP + k] = x[e];
          csum += a.wsum ? (i0 + e < a.n ? (double)a.wsum[i0 + e] * (double)x[e] : 0.0) : (double)x[e];
        }
      }
    }
  }
  red[ty][tx] = csum;
  __syncthreads();
  if (ty == 0 && k < a.KP)
    a.part[(int64_t)blockIdx.y * a.KP + k] = (red[0][tx] + red[1][tx]) + (red[2][tx] + red[3][tx]);
}

void launch_pack(const PackArgs& a, hipStream_t st) {
  dim3 grid((unsigned)((a.KP + 63) / 64), (unsigned)(a.rows_pad / kPackRows));
  k_pack<<<grid, 256, 0, st>>>(a);
  HIPCHECK(hipGetLastError());
}

// sum over nparts partial rows: out[k] = sum_p part[p*KP + k] in fixed order (double)
__global__ void k_sum_parts(const double* __restrict__ part, int64_t nparts, int KP, double* __restrict__ out) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int64_t p = threadIdx.x; p < nparts; p += 256) s += part[p * KP + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

void launch_sum_parts(const double* part, int64_t nparts, int KP, double* out, hipStream_t st) {
  k_sum_parts<<<KP, 256, 0, st>>>(part, nparts, KP, out);
  HIPCHECK(hipGetLastError());
}

// ------------------------------------------------------------------ small dense helpers ----
// part[b][q*t + k] = sum over rows of chunk b of A[i,q] * B[i,k]  (q < c, k < t), FP64.
__global__ __launch_bounds__(256) void k_tn_part(const float* __restrict__ A, int64_t lda, int c,
                                                 const float* __restrict__ B, int64_t ldb, int t,
                                                 int64_t n, int64_t rows_per_block, double* __restrict__ part) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  const int q = blockIdx.z;
  const int64_t i0 = (int64_t)blockIdx.y * rows_per_block;
  const int64_t i1 = min(n, i0 + rows_per_block);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256)
    s += (double)A[i + (int64_t)q * lda] * (double)B[i + (int64_t)k * ldb];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * c * t + (int64_t)q * t + k] = red[0];
}

// out[q*ldo + k] = sum_b part[b][q*t+k]
__global__ void k_tn_final(const double* __restrict__ part, int64_t nblk, int ct, int t, int ldo,
                           double* __restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ct) return;
  double s = 0.0;
  for (int64_t b = 0; b < nblk; ++b) s += part[b * ct + idx];
  out[(idx / t) * ldo + idx % t] = s;
}

void tn_product(const float* A, int64_t lda, int c, const float* B, int64_t ldb, int t, int64_t n,
                double* out, int ldo, DevBuf& scratch, hipStream_t st) {
  if (c <= 0 || t <= 0) return;
  const int64_t rpb = 16384;
  const int64_t nblk = (n + rpb - 1) / rpb;
  scratch.reserve(sizeof(double) * nblk * c * t);
  dim3 grid((unsigned)t, (unsigned)nblk, (unsigned)c);
  k_tn_part<<<grid, 256, 0, st>>>(A, lda, c, B, ldb, t, n, rpb, scratch.as<double>());
  HIPCHECK(hipGetLastError());
  k_tn_final<<<(c * t + 255) / 256, 256, 0, st>>>(scratch.as<double>(), nblk, c * t, t, ldo, out);
  HIPCHECK(hipGetLastError());
}

// QUADFORM finish: out_ik = out_ik - sum_q Q_iq b_qk + tau0 v_ik for columns of kind 1.
__global__ void k_quad_finish(float* __restrict__ out, int64_t ldo, const float* __restrict__ V, int64_t ldv,
                              const float* __restrict__ Q, int64_t ldq, int c, const double* __restrict__ b,
                              int KP, const int* __restrict__ kinds, int t, int64_t n, float tau0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= n || k >= t) return;
  if (kinds && kinds[k] != 1) return;
  double acc = 0.0;
  for (int q = 0; q < c; ++q) acc += (double)Q[i + (int64_t)q * ldq] * b[q * KP + k];
  const float o = out[i + (int64_t)k * ldo];
  out[i + (int64_t)k * ldo] = (float)((double)o - acc) + tau0 * V[i + (int64_t)k * ldv];
}

void launch_quad_finish(float* out, int64_t ldo, const float* V, int64_t ldv, const float* Q, int64_t ldq,
                        int c, const double* b, int KP, const int* kinds, int t, int64_t n, float tau0,
                        hipStream_t st) {
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)t);
  k_quad_finish<<<grid, 256, 0, st>>>(out, ldo, V, ldv, Q, ldq, c, b, KP, kinds, t, n, tau0);
  HIPCHECK(hipGetLastError());
}

}  // namespace slemm
