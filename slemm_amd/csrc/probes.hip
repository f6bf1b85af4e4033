// Rademacher probes of SLDF_REML (src/reml.cpp:597-605): n_V columns of n draws from
// mt19937(seed + 54321), entry -1 when uniform_real_distribution<float>(0, 1) < 0.5 else +1, each
// column then divided by its float norm.  The norm is closed-form: the reference sums v_i^2 = 1 in
// float, which counts exactly up to 2^24 and stays there (2^24 + 1 rounds back to 2^24), so
// ss = min(n, 2^24) and every entry is -q or +q with q = 1 / sqrt(ss) rounded once.  The draws
// come from whole twisted MT19937 blocks (one engine word per draw, the only sequential part);
// on the device path only one byte per draw crosses to the GPU, where a kernel writes the floats.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "driver.hpp"
#include "rng.hpp"

namespace slemm {
namespace {

// 1 where the draw is -1: uniform_real_distribution<float>(0, 1) is canonical24(u) * (1 - 0) + 0
uint8_t negative(uint32_t u) { return canonical24(u) * 1.0f + 0.0f < 0.5f ? 1 : 0; }

}  // namespace

// the signs of the first `total` draws in draw order (probe column j of an n-row block is draws
// j n .. j n + n - 1)
std::vector<uint8_t> rademacher_signs(int rng_seed, int64_t total) {
  std::vector<uint8_t> neg((size_t)std::max<int64_t>(total, 0));
  Mt19937Blocks mt((uint32_t)(rng_seed + 54321));
  uint32_t w[Mt19937Blocks::kN];
  for (int64_t e = 0; e < total; e += Mt19937Blocks::kN) {
    mt.block(w);
    const int64_t k1 = std::min<int64_t>(Mt19937Blocks::kN, total - e);
    for (int64_t k = 0; k < k1; ++k) neg[(size_t)(e + k)] = negative(w[k]);
  }
  return neg;
}

// 1 / sqrt of the float sum of n unit squares
float rademacher_scale(int64_t n) {
  const float ss = n < (1 << 24) ? (float)n : 16777216.0f;
  return 1.0f / std::sqrt(ss);
}

namespace {

__global__ __launch_bounds__(256) void k_probe_cols(const uint8_t* __restrict__ neg, int64_t n, int n_V, float q,
                                                    float* __restrict__ B, int64_t ldb) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * n_V) return;
  const int64_t j = e / n, i = e - j * n;
  B[i + j * ldb] = neg[e] ? -q : q;
}

}  // namespace

void rademacher_probes(int rng_seed, int64_t n, int n_V, float* B) {
  const std::vector<uint8_t> neg = rademacher_signs(rng_seed, n * n_V);
  const float q = rademacher_scale(n);
  parallel_parts(n_V, host_threads(), [&](int64_t j) {
    for (int64_t i = 0; i < n; ++i) B[(size_t)j * n + i] = neg[(size_t)j * n + i] ? -q : q;
  });
}

// Stream-ordered on the null stream (allocation, copy, kernel, release), waiting only for that
// stream.
void rademacher_probes_device(int rng_seed, int64_t n, int n_V, float* dB, int64_t ldb) {
  if (n <= 0 || n_V <= 0) return;
  const std::vector<uint8_t> neg = rademacher_signs(rng_seed, n * n_V);
  uint8_t* dneg = nullptr;
  HIPCHECK(hipMallocAsync(reinterpret_cast<void**>(&dneg), neg.size(), nullptr));
  hipError_t e = hipMemcpyAsync(dneg, neg.data(), neg.size(), hipMemcpyHostToDevice, nullptr);
  if (e == hipSuccess) {
    const int64_t total = n * n_V;
    k_probe_cols<<<(unsigned)((total + 255) / 256), 256>>>(dneg, n, n_V, rademacher_scale(n), dB, ldb);
    e = hipGetLastError();
  }
  const hipError_t fe = hipFreeAsync(dneg, nullptr);
  const hipError_t se = hipStreamSynchronize(nullptr);  // the signs' host buffer is released on return
  HIPCHECK(e);
  HIPCHECK(fe);
  HIPCHECK(se);
}

}  // namespace slemm

using namespace slemm;

extern "C" int slemm_gpu_rademacher_probes(int rng_seed, int64_t n, int n_V, float* B) {
  return guard([&] {
    REQUIRE(n >= 0 && n_V >= 0, "sizes must be non-negative");
    REQUIRE(B || n * n_V == 0, "null argument");
    rademacher_probes(rng_seed, n, n_V, B);
  });
}
