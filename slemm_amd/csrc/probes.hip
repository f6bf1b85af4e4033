// Rademacher probes of SLDF_REML (src/reml.cpp:597-605): n_V columns of n draws from
// mt19937(seed + 54321), entry -1 when uniform_real_distribution<float>(0, 1) < 0.5 else +1, each
// column then divided by its float norm.  The norm is closed-form: the reference sums v_i^2 = 1 in
// float, which counts exactly up to 2^24 and stays there (2^24 + 1 rounds back to 2^24), so
// ss = min(n, 2^24) and every entry is -q or +q with q = 1 / sqrt(ss) rounded once.  The signs
// come from rademacher_sign_bits (one engine word per draw); on the device path only one bit per
// draw crosses to the GPU, where a kernel writes the floats.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "driver.hpp"
#include "rng.hpp"

namespace slemm {

namespace {

__global__ __launch_bounds__(256) void k_probe_cols(const uint8_t* __restrict__ bits, int64_t n, int n_V, float q,
                                                    float* __restrict__ B, int64_t ldb) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;  // draw index
  if (e >= n * n_V) return;
  const int64_t j = e / n, i = e - j * n;
  B[i + j * ldb] = (bits[e >> 3] >> (e & 7)) & 1 ? -q : q;
}

}  // namespace

void rademacher_probes(int rng_seed, int64_t n, int n_V, float* B) {
  const std::vector<uint8_t> bits = rademacher_sign_bits(rng_seed, n * n_V);
  const float q = rademacher_scale(n);
  parallel_parts(n_V, host_threads(), [&](int64_t j) {
    for (int64_t i = 0; i < n; ++i) {
      const int64_t e = j * n + i;
      B[(size_t)e] = (bits[(size_t)(e >> 3)] >> (e & 7)) & 1 ? -q : q;
    }
  });
}

// Stream-ordered on the null stream (allocation, copy, kernel, release), waiting only for that
// stream.
void rademacher_probes_device(int rng_seed, int64_t n, int n_V, float* dB, int64_t ldb) {
  if (n <= 0 || n_V <= 0) return;
  const size_t nbytes = (size_t)((n * n_V + 7) / 8);
  uint8_t* bits = static_cast<uint8_t*>(pinned_scratch(2, nbytes));
  rademacher_sign_bits_into(rng_seed, n * n_V, bits);
  uint8_t* dbits = nullptr;
  HIPCHECK(hipMallocAsync(reinterpret_cast<void**>(&dbits), nbytes, nullptr));
  hipError_t e = hipMemcpyAsync(dbits, bits, nbytes, hipMemcpyHostToDevice, nullptr);
  if (e == hipSuccess) {
    const int64_t total = n * n_V;
    k_probe_cols<<<(unsigned)((total + 255) / 256), 256>>>(dbits, n, n_V, rademacher_scale(n), dB, ldb);
    e = hipGetLastError();
  }
  const hipError_t fe = hipFreeAsync(dbits, nullptr);
  const hipError_t se = hipStreamSynchronize(nullptr);  // the bits' host buffer is released on return
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
