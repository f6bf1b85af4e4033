// Matrix-core issue ceiling of this GPU: a register-only loop of independent MFMA accumulator
// chains, the ceiling against which the block-product kernels' issued rates are read.  No
// reference counterpart (instrumentation, like the kernel timing of blockprod.hip).
//
// Every lane holds operands derived from its own index (so nothing folds) and `chains`
// independent accumulators; each loop iteration issues one MFMA per chain.  Two operand sets: the
// same small operands for every chain, or (instruction ids + 4) the block products' kind of data
// -- per chain its own full-range random B operand (int8 digits; BF16 values with random
// mantissas), the A operand (dosages 0/1/2 for int8) shared by four chains -- so that, as in the
// block products, consecutive MFMAs see different operand bits; switching bits costs power, and
// under a power limit the clock, so this set gives the ceiling for real data.  A grid of
// 256 x waves_per_simd workgroups of four waves puts waves_per_simd waves on every SIMD of the
// 256 CUs.  Wave 0 of workgroup 0 reads the shader clock counter (s_memtime) and the 100 MHz
// real-time counter (s_memrealtime) before and after its loop, which gives the clock the matrix
// cores ran at under this load.
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"

namespace slemm {
namespace {

typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef int i32x16_t __attribute__((ext_vector_type(16)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;

// instruction ids of slemm_gpu_mfma_ceiling
enum { I8_16x16x64 = 0, I8_32x32x32 = 1, BF16_16x16x32 = 2, BF16_32x32x16 = 3 };
constexpr double kOps[4] = {16.0 * 16 * 64 * 2, 32.0 * 32 * 32 * 2, 16.0 * 16 * 32 * 2, 32.0 * 32 * 16 * 2};

template <int I> struct Acc;
template <> struct Acc<I8_16x16x64> { using T = i32x4_t; };
template <> struct Acc<I8_32x32x32> { using T = i32x16_t; };
template <> struct Acc<BF16_16x16x32> { using T = f32x4_t; };
template <> struct Acc<BF16_32x32x16> { using T = f32x16_t; };

__device__ __forceinline__ bf16x8_t bf16_operand(uint32_t seed) {
  // small values (|x| < 2^-6) with varying mantissas, so the sums grow slowly and stay finite
  bf16x8_t v;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const uint32_t x = (seed * 2654435761u + e * 40503u) >> 25;  // 7 mantissa bits
    const uint16_t bits = (uint16_t)(0x3C00u | x | ((x & 1u) << 15));
    v[e] = __builtin_bit_cast(__bf16, bits);
  }
  return v;
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// four random bytes, each 0, 1 or 2 (genotype dosages)
__device__ __forceinline__ int dosage_word(uint32_t seed) {
  const uint32_t h = mix32(seed);
  uint32_t w = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) w |= (((h >> (8 * b)) & 0xffu) % 3u) << (8 * b);
  return (int)w;
}

template <int I, int NACC, bool VARY>
__global__ __launch_bounds__(256) void k_mfma_probe(int iters, long long* clocks, float* sink) {
  using T = typename Acc<I>::T;
  T acc[NACC];
  const uint32_t tid = blockIdx.x * 256 + threadIdx.x;
#pragma unroll
  for (int j = 0; j < NACC; ++j) {  // distinct starting values: identical chains invite merging
    acc[j] = T{};
    acc[j][0] = j + 1;
    acc[j][1] = (int)(tid & 7);
  }
  constexpr int NA = VARY ? (NACC + 3) / 4 : 1, NB = VARY ? NACC : 1;
  i32x4_t ia[NA], ib[NB];
  bf16x8_t fa[NA], fb[NB];
  if constexpr (VARY) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      ia[j] = i32x4_t{dosage_word(tid * 64 + j * 4), dosage_word(tid * 64 + j * 4 + 1),
                      dosage_word(tid * 64 + j * 4 + 2), dosage_word(tid * 64 + j * 4 + 3)};
      fa[j] = bf16_operand(mix32(tid * 64 + j));
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      ib[j] = i32x4_t{(int)mix32(tid * 97 + j * 4 + 11), (int)mix32(tid * 97 + j * 4 + 12),
                      (int)mix32(tid * 97 + j * 4 + 13), (int)mix32(tid * 97 + j * 4 + 14)};
      fb[j] = bf16_operand(mix32(tid * 97 + j + 0x5bd1u));
    }
  } else {
    ia[0] = i32x4_t{(int)((tid & 3) + 1), (int)((tid >> 2) & 3), 1, (int)(tid & 1)};
    ib[0] = i32x4_t{1, (int)((tid >> 1) & 3), (int)((tid & 1) + 1), 2};
    fa[0] = bf16_operand(tid);
    fb[0] = bf16_operand(tid ^ 0x5bd1u);
  }
  const bool timer = blockIdx.x == 0 && threadIdx.x == 0;
  long long t0 = 0, r0 = 0;
  if (timer) {
    t0 = __builtin_amdgcn_s_memtime();
    r0 = __builtin_amdgcn_s_memrealtime();
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < NACC; ++j) {
      const int ja = VARY ? j / 4 : 0, jb = VARY ? j : 0;
      if constexpr (I == I8_16x16x64) acc[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ia[ja], ib[jb], acc[j], 0, 0, 0);
      else if constexpr (I == I8_32x32x32) acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ia[ja], ib[jb], acc[j], 0, 0, 0);
      else if constexpr (I == BF16_16x16x32) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[ja], fb[jb], acc[j], 0, 0, 0);
      else acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[ja], fb[jb], acc[j], 0, 0, 0);
    }
  }
  if (timer) {
    clocks[0] = __builtin_amdgcn_s_memtime() - t0;
    clocks[1] = __builtin_amdgcn_s_memrealtime() - r0;
  }
  // every element of every accumulator stays live (a partial read lets the compiler merge chains)
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < NACC; ++j)
#pragma unroll
    for (int e = 0; e < (int)(sizeof(T) / 4); ++e) s += (float)acc[j][e];
  sink[tid] = s;
}

template <int I, bool VARY>
void launch_probe(int chains, dim3 grid, int iters, long long* clocks, float* sink, hipStream_t st) {
  switch (chains) {
    case 1: k_mfma_probe<I, 1, VARY><<<grid, 256, 0, st>>>(iters, clocks, sink); break;
    case 2: k_mfma_probe<I, 2, VARY><<<grid, 256, 0, st>>>(iters, clocks, sink); break;
    case 4: k_mfma_probe<I, 4, VARY><<<grid, 256, 0, st>>>(iters, clocks, sink); break;
    default: k_mfma_probe<I, 8, VARY><<<grid, 256, 0, st>>>(iters, clocks, sink); break;
  }
}

template <bool VARY>
void launch_probe_instr(int instr, int chains, dim3 grid, int iters, long long* clocks, float* sink) {
  switch (instr) {
    case 0: launch_probe<I8_16x16x64, VARY>(chains, grid, iters, clocks, sink, nullptr); break;
    case 1: launch_probe<I8_32x32x32, VARY>(chains, grid, iters, clocks, sink, nullptr); break;
    case 2: launch_probe<BF16_16x16x32, VARY>(chains, grid, iters, clocks, sink, nullptr); break;
    default: launch_probe<BF16_32x32x16, VARY>(chains, grid, iters, clocks, sink, nullptr); break;
  }
}

}  // namespace
}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_mfma_ceiling(int instr, int chains, int waves_per_simd, int iters, double* tops,
                                      double* clock_ratio, double* ms) {
  return guard([&] {
    REQUIRE(instr >= 0 && instr <= 7,
            "instr: 0 i8 16x16x64, 1 i8 32x32x32, 2 bf16 16x16x32, 3 bf16 32x32x16; + 4 with random operands");
    REQUIRE(chains == 1 || chains == 2 || chains == 4 || chains == 8, "chains must be 1, 2, 4 or 8");
    REQUIRE(waves_per_simd >= 1 && waves_per_simd <= 8, "waves_per_simd must be in [1, 8]");
    REQUIRE(iters >= 1 && iters <= (1 << 20), "iters must be in [1, 2^20]");
    REQUIRE(tops, "null argument");
    int dev = 0, cus = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const dim3 grid((unsigned)(cus * waves_per_simd));
    DevBuf clk, sink;
    clk.reserve(sizeof(long long) * 2);
    sink.reserve(sizeof(float) * grid.x * 256);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    auto run = [&](int it) {
      if (instr < 4) launch_probe_instr<false>(instr, chains, grid, it, clk.as<long long>(), sink.as<float>());
      else launch_probe_instr<true>(instr - 4, chains, grid, it, clk.as<long long>(), sink.as<float>());
      HIPCHECK(hipGetLastError());
    };
    run(std::max(1, iters / 10));  // warm up (clocks ramp, code object loaded)
    HIPCHECK(hipEventRecord(e0, nullptr));
    run(iters);
    HIPCHECK(hipEventRecord(e1, nullptr));
    HIPCHECK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHECK(hipEventElapsedTime(&t, e0, e1));
    long long c[2] = {0, 0};
    HIPCHECK(hipMemcpy(c, clk.p, sizeof(c), hipMemcpyDeviceToHost));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double ops = kOps[instr % 4] * chains * (double)iters * 4.0 * grid.x;
    *tops = ops / (t * 1e-3) / 1e12;
    if (ms) *ms = t;
    // shader-clock ticks per 100 MHz tick: the core clock in units of 100 MHz
    if (clock_ratio) *clock_ratio = c[1] > 0 ? (double)c[0] / (double)c[1] : 0.0;
  });
}
