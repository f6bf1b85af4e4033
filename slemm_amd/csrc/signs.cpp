// Signs of the Rademacher probes of SLDF_REML (src/reml.cpp:597-605) and of the batched traits'
// seed blocks: draw e is -1 when uniform_real_distribution<float>(0, 1) applied to word e of
// mt19937(seed + 54321) is below 0.5.  That distribution evaluates generate_canonical<float, 24>
// (float(u) / 2^32) times 1 plus 0, which is below 0.5 exactly when float(u) < 2^31, that is for
// u < 2^31 - 64: words in [2^31 - 128, 2^31) lie between the neighbouring floats 2^31 - 128 and
// 2^31, and from the midpoint 2^31 - 64 on they round up (the tie to the even 2^31).  So each
// draw is one integer compare of the tempered word, kept as one bit.  With AVX2 the twist and
// tempering run eight words to an instruction and each compare of eight words gives one byte of
// bits.  Streams of at least two chunks (kMtJumpChunk words) are cut into up to 16 chunks that
// host threads produce at once, each starting from the engine state moved ahead by its offset
// (jump polynomials of gen_mt_jump.py, generated at build time).  mt_fill_blocks gives the same
// choice of instruction set to callers that take the engine's words themselves.
#include <immintrin.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "driver.hpp"
#include "mt_jump.inc"  // kMtJumpChunk, kMtJumpPolys, kMtJumpWords, kMtJumpPoly (build time)
#include "rng.hpp"

namespace slemm {
namespace {

static_assert(kMtJumpChunk % Mt19937Blocks::kN == 0 && kMtJumpChunk % 8 == 0,
              "chunks start on engine blocks and on whole bytes of sign bits");

// out = P(A) w0: the window after e more words, for P = x^e mod phi (phi the minimal polynomial of
// the one-word step A).  Horner's rule on a linear buffer: each step appends the next word of the
// accumulator window; where P has a term the start window is added.  The result can differ from
// the true window only in the 31 low bits of its oldest word, which no later word depends on.
__attribute__((target("avx2"))) void mt_jump_avx2(const uint32_t* w0, const uint64_t* P, uint32_t* out) {
  constexpr int N = Mt19937Blocks::kN, M = Mt19937Blocks::kM;
  int deg = kMtJumpWords * 64 - 1;
  while (deg > 0 && !((P[deg >> 6] >> (deg & 63)) & 1)) --deg;
  std::vector<uint32_t> a((size_t)N + deg + 1, 0u);
  size_t base = 0;
  for (int i = deg; i >= 0; --i) {
    if (i != deg) {
      const uint32_t y = (a[base] & 0x80000000u) | (a[base + 1] & 0x7fffffffu);
      a[base + N] = a[base + M] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
      ++base;
    }
    if ((P[i >> 6] >> (i & 63)) & 1) {
      uint32_t* d = a.data() + base;
      for (int k = 0; k < N; ++k) d[k] ^= w0[k];
    }
  }
  std::copy(a.begin() + (int64_t)base, a.begin() + (int64_t)base + N, out);
}

constexpr uint32_t kNegBelow = 0x7fffffc0u;  // 2^31 - 64

// Bits of draws [e, e + 624) of one block w into o (bytes ceil(k1 / 8), k1 live draws): bit k of
// byte b is draw 8 b + k; bits past k1 are zero.
inline __attribute__((always_inline)) void tail_bits(const uint32_t* w, int64_t k1, uint8_t* o) {
  for (int64_t b = 0; 8 * b < k1; ++b) {
    uint32_t byte = 0;
    for (int k = 0; k < 8 && 8 * b + k < k1; ++k) byte |= (uint32_t)(w[8 * b + k] < kNegBelow) << k;
    o[b] = (uint8_t)byte;
  }
}

// The engine is passed in (constructed by the caller): its inlined twist and tempering are then
// compiled with this function's instruction set.
__attribute__((target("avx2"))) void sign_bits_avx2(Mt19937Blocks& mt, int64_t total, uint8_t* bits) {
  alignas(32) uint32_t w[Mt19937Blocks::kN];
  const __m256i flip = _mm256_set1_epi32(INT32_MIN);
  const __m256i lim = _mm256_set1_epi32((int)(kNegBelow ^ 0x80000000u));  // unsigned order as signed
  for (int64_t e = 0; e < total; e += Mt19937Blocks::kN) {
    mt.block(w);
    uint8_t* o = bits + e / 8;  // kN = 624 is a multiple of 8
    const int64_t k1 = std::min<int64_t>(Mt19937Blocks::kN, total - e);
    if (k1 < Mt19937Blocks::kN) {
      tail_bits(w, k1, o);
      break;
    }
    for (int b = 0; b < Mt19937Blocks::kN / 8; ++b) {
      const __m256i v = _mm256_xor_si256(_mm256_load_si256(reinterpret_cast<const __m256i*>(w + 8 * b)), flip);
      o[b] = (uint8_t)_mm256_movemask_ps(_mm256_castsi256_ps(_mm256_cmpgt_epi32(lim, v)));
    }
  }
}

void sign_bits_plain(Mt19937Blocks& mt, int64_t total, uint8_t* bits) {
  uint32_t w[Mt19937Blocks::kN];
  for (int64_t e = 0; e < total; e += Mt19937Blocks::kN) {
    mt.block(w);
    tail_bits(w, std::min<int64_t>(Mt19937Blocks::kN, total - e), bits + e / 8);
  }
}

__attribute__((target("avx2"))) void fill_blocks_avx2(Mt19937Blocks& mt, int64_t nblk, uint32_t* dst) {
  for (int64_t b = 0; b < nblk; ++b) mt.block(dst + b * Mt19937Blocks::kN);
}

void fill_blocks_plain(Mt19937Blocks& mt, int64_t nblk, uint32_t* dst) {
  for (int64_t b = 0; b < nblk; ++b) mt.block(dst + b * Mt19937Blocks::kN);
}

}  // namespace

void mt_fill_blocks(Mt19937Blocks& mt, int64_t nblk, uint32_t* dst) {
  if (__builtin_cpu_supports("avx2")) fill_blocks_avx2(mt, nblk, dst);
  else fill_blocks_plain(mt, nblk, dst);
}

std::vector<uint8_t> rademacher_sign_bits(int rng_seed, int64_t total) {
  std::vector<uint8_t> bits((size_t)((std::max<int64_t>(total, 0) + 7) / 8));
  rademacher_sign_bits_into(rng_seed, total, bits.data());
  return bits;
}

void rademacher_sign_bits_into(int rng_seed, int64_t total, uint8_t* bits) {
  if (total <= 0) return;
  const Mt19937Blocks mt0((uint32_t)(rng_seed + 54321));
  const bool avx2 = __builtin_cpu_supports("avx2");
  auto run = [&](Mt19937Blocks& mt, int64_t count, uint8_t* out) {
    if (avx2) sign_bits_avx2(mt, count, out);
    else sign_bits_plain(mt, count, out);
  };
  const int threads = host_threads();
  const int64_t L = kMtJumpChunk;
  if (!avx2 || threads < 2 || total < 2 * L) {
    Mt19937Blocks mt = mt0;
    run(mt, total, bits);
    return;
  }
  // chunk c: draws [c L, (c + 1) L), the last one to the end of the stream
  const int nch = (int)std::min<int64_t>(kMtJumpPolys + 1, total / L);
  parallel_parts(nch, threads, [&](int64_t c) {
    Mt19937Blocks mt = mt0;
    if (c > 0) {
      uint32_t w[Mt19937Blocks::kN];
      mt_jump_avx2(mt0.window(), kMtJumpPoly[c - 1], w);
      mt.set_window(w);
    }
    const int64_t e0 = c * L, e1 = c + 1 == nch ? total : e0 + L;
    run(mt, e1 - e0, bits + e0 / 8);
  });
}

// 1 / sqrt of the float sum of n unit squares
float rademacher_scale(int64_t n) {
  const float ss = n < (1 << 24) ? (float)n : 16777216.0f;
  return 1.0f / std::sqrt(ss);
}

}  // namespace slemm
