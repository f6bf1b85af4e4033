// Host random streams of the reference's drivers, value for value: std::mt19937 generated a
// whole 624-word block at a time, libstdc++'s generate_canonical<float, 24> and the coordinates of
// normal_distribution<float>'s polar method, and a small host thread pool.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace slemm {

// generate_canonical<float, 24>(mt19937) as libstdc++ evaluates it: one engine value u,
// float(u) / 2^32 (exact but for the rounding of float(u)), and the largest float below 1 where
// that rounds up to 1.
inline float canonical24(uint32_t u) {
  float r = (float)u * 0x1p-32f;
  if (r >= 1.0f) r = 0x1.fffffep-1f;
  return r;
}

// 2 U - 1 as normal_distribution evaluates it: (float(2) * U) - 1.0 in double, stored as float.
inline float polar_coord(uint32_t u) { return (float)((double)(2.0f * canonical24(u)) - 1.0); }

// MT19937 (Matsumoto and Nishimura, ACM TOMACS 8(1), 1998) producing its output 624 words at a
// time: the state initialised as std::mt19937(seed) does (x_i = 1812433253 (x_{i-1} ^
// (x_{i-1} >> 30)) + i), each block one twist of the whole state followed by the tempering of
// every word, in plain loops the compiler vectorises.  The words come out in the order
// std::mt19937's operator() returns them.
class Mt19937Blocks {
 public:
  static constexpr int kN = 624, kM = 397;
  // always inlined, so that a caller compiled for wider vectors (target_clones) vectorises them too
  __attribute__((always_inline)) explicit Mt19937Blocks(uint32_t seed) {
    x_[0] = seed;
    for (int i = 1; i < kN; ++i) x_[i] = 1812433253u * (x_[i - 1] ^ (x_[i - 1] >> 30)) + (uint32_t)i;
  }
  // The state as the window of the last kN untempered words, oldest first: the next block's
  // outputs are those that follow them.
  const uint32_t* window() const { return x_; }
  void set_window(const uint32_t* w) {
    for (int i = 0; i < kN; ++i) x_[i] = w[i];
  }
  // the next kN outputs into dst
  __attribute__((always_inline)) void block(uint32_t* dst) {
    int i = 0;
    for (; i < kN - kM; ++i) x_[i] = step(x_[i], x_[i + 1], x_[i + kM]);
    for (; i < kN - 1; ++i) x_[i] = step(x_[i], x_[i + 1], x_[i + kM - kN]);
    x_[kN - 1] = step(x_[kN - 1], x_[0], x_[kM - 1]);
    for (int k = 0; k < kN; ++k) {
      uint32_t y = x_[k];
      y ^= y >> 11;
      y ^= (y << 7) & 0x9d2c5680u;
      y ^= (y << 15) & 0xefc60000u;
      y ^= y >> 18;
      dst[k] = y;
    }
  }

 private:
  __attribute__((always_inline)) static uint32_t step(uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return c ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  }
  uint32_t x_[kN];
};

// The next nblk blocks of mt into dst (nblk * 624 words), with AVX2 where the CPU has it
// (signs.cpp).
void mt_fill_blocks(Mt19937Blocks& mt, int64_t nblk, uint32_t* dst);

// Runs f(k) for k < parts on up to `threads` host threads (contiguous ranges).
template <typename F>
void parallel_parts(int64_t parts, int threads, const F& f) {
  if (threads <= 1 || parts <= 1) {
    for (int64_t k = 0; k < parts; ++k) f(k);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t per = (parts + threads - 1) / threads;
  for (int64_t k0 = 0; k0 < parts; k0 += per)
    pool.emplace_back([&f, k0, parts, per] {
      for (int64_t k = k0; k < std::min(parts, k0 + per); ++k) f(k);
    });
  for (auto& th : pool) th.join();
}

// Host threads for the parallel parts of the draws: the hardware's, at most 16.
inline int host_threads() { return (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency())); }

}  // namespace slemm
