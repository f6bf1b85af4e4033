// Packed genotype storage on the device: upload/compaction of PLINK bed rows, the
// individual-major transposed copy, bit counts, synthetic beds and the model terms.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "common.hpp"

namespace slemm {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e == hipSuccess) return;
  std::string msg = std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" +
                    std::to_string(line) + ")";
  fail(e == hipErrorOutOfMemory ? SLEMM_ERR_ALLOC : SLEMM_ERR_HIP, msg);
}

ScratchArena*& thread_arena() {
  static thread_local ScratchArena* a = nullptr;
  return a;
}

void* ScratchArena::take(size_t nbytes) {
  const size_t off = (used + 255) & ~(size_t)255;
  if (off > cap || nbytes > cap - off)
    fail(SLEMM_ERR_ARG, "workspace of this call exceeds the no-allocation scratch (" +
                            std::to_string(cap >> 20) + " MiB)");
  used = off + nbytes;
  return base + off;
}

void DevBuf::reserve(size_t nbytes) {
  if (nbytes <= bytes) return;
  release();
  if (ScratchArena* a = thread_arena()) {
    p = a->take(nbytes);
    bytes = nbytes;
    borrowed = true;
    return;
  }
  hipError_t e = hipMalloc(&p, nbytes);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    hip_check(e, "hipMalloc(workspace)", __FILE__, __LINE__);
  }
  bytes = nbytes;
}

void DevBuf::release() {
  if (p && !borrowed) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  borrowed = false;
}

// ------------------------------------------------------------------------------ kernels ----

// One thread per output byte: gathers the 4 kept individuals of that byte from a raw bed row.
// keep == nullptr means "all individuals"; bits beyond n are forced to code 00.
__global__ void k_compact(const uint8_t* __restrict__ raw, int64_t raw_ld, int rows,
                          const int32_t* __restrict__ keep, int64_t n, uint8_t* __restrict__ gs,
                          int64_t ld_gs) {
  const int64_t ob = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (r >= rows || ob >= ld_gs) return;
  const uint8_t* src = raw + (int64_t)r * raw_ld;
  uint32_t out = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t kk = ob * 4 + q;
    if (kk < n) {
      const int64_t i = keep ? (int64_t)keep[kk] : kk;
      const uint32_t code = (src[i >> 2] >> (2 * (i & 3))) & 3u;
      out |= code << (2 * q);
    }
  }
  gs[(int64_t)r * ld_gs + ob] = (uint8_t)out;
}

// gs (SNP-major) -> gi (individual-major).  Tile: 512 SNPs x 256 individuals; 256 threads.
// Thread t assembles the 128 bytes (512 SNPs, one whole cache line) of individual i0 + t from the
// staged tile, so that no output line is shared between workgroups (which may sit on different
// XCDs and L2s); the lines go back through the same LDS so that eight neighbouring lanes store
// each line's eight 16-byte pieces (whole lines per store instead of 64 lines touched 16 bytes
// at a time).  The transpose of packed 2-bit codes is its own inverse: launched on gi (rows =
// individuals) over a (m_pad / 256, n_pad / 512) grid it writes gs.  With a row list (rows
// n_rows and beyond read as zero) input row j is row rows[j] of gs, and the rows as read are also
// written to copy (leading dimension ld_gs): one pass then gathers a subset of individuals into
// both layouts.
__global__ __launch_bounds__(256) void k_transpose(const uint8_t* __restrict__ gs, int64_t ld_gs,
                                                   uint8_t* __restrict__ gi, int64_t ld_gi,
                                                   const int64_t* __restrict__ rows = nullptr, int64_t n_rows = 0,
                                                   uint8_t* __restrict__ copy = nullptr) {
  __shared__ uint32_t tile[512][64 / 4 + 1];  // 512 SNP rows x 64 bytes (+pad against conflicts)
  const int64_t i0 = (int64_t)blockIdx.x * 256;
  const int64_t j0 = (int64_t)blockIdx.y * 512;
  const int t = threadIdx.x;
  // cooperative load: 512 rows x 4 uint4 = 2048 uint4, 8 per thread
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int idx = q * 256 + t;
    const int row = idx >> 2, c4 = idx & 3;
    uint4 v;
    if (rows) {
      const int64_t j = j0 + row;
      v = j < n_rows ? *reinterpret_cast<const uint4*>(gs + rows[j] * ld_gs + i0 / 4 + c4 * 16)
                     : make_uint4(0u, 0u, 0u, 0u);
      *reinterpret_cast<uint4*>(copy + j * ld_gs + i0 / 4 + c4 * 16) = v;
    } else {
      v = *reinterpret_cast<const uint4*>(gs + (j0 + row) * ld_gs + i0 / 4 + c4 * 16);
    }
    tile[row][c4 * 4 + 0] = v.x;
    tile[row][c4 * 4 + 1] = v.y;
    tile[row][c4 * 4 + 2] = v.z;
    tile[row][c4 * 4 + 3] = v.w;
  }
  __syncthreads();
  const int dw = t >> 4;          // dword holding individual t within a row
  const int sh = 2 * (t & 15);    // bit offset of its code in that dword
  uint32_t outw[32];              // the individual's 512 codes, 16 to a word
#pragma unroll
  for (int w = 0; w < 32; ++w) {
    uint32_t acc = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc |= ((tile[w * 16 + q][dw] >> sh) & 3u) << (2 * q);
    outw[w] = acc;
  }
  __syncthreads();  // the tile is read; its memory now holds the output lines
  // output staging: line t's word w at obuf[t * 33 + w] (the odd stride keeps both the row-wise
  // writes and the column-wise reads free of bank conflicts); 256 x 33 words fit the tile
  uint32_t* obuf = &tile[0][0];
  static_assert(256 * 33 <= 512 * 17, "output staging fits the tile");
#pragma unroll
  for (int w = 0; w < 32; ++w) obuf[t * 33 + w] = outw[w];
  __syncthreads();
  // stores: piece idx = 256 q + t is 16-byte piece idx & 7 of line idx >> 3
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int idx = q * 256 + t;
    const int line = idx >> 3, pc = idx & 7;
    const uint32_t* src = obuf + line * 33 + pc * 4;
    *reinterpret_cast<uint4*>(gi + (i0 + line) * ld_gi + j0 / 4 + pc * 16) = make_uint4(src[0], src[1], src[2], src[3]);
  }
}

// One wave per SNP row: popcount, count of 11 codes, count of nonzero codes; gm = pc/n.
__global__ __launch_bounds__(256) void k_counts(const uint8_t* __restrict__ gs, int64_t ld_gs,
                                                int64_t m, int64_t n, int32_t* __restrict__ pc,
                                                int32_t* __restrict__ c2, int32_t* __restrict__ cany,
                                                float* __restrict__ gm) {
  const int lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= m) return;
  const uint4* row = reinterpret_cast<const uint4*>(gs + j * ld_gs);
  const int64_t nv = ld_gs / 16;
  int s_pc = 0, s_2 = 0, s_any = 0;
  for (int64_t v = lane; v < nv; v += 64) {
    const uint4 w4 = row[v];
    const uint32_t w[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t lo = w[q] & 0x55555555u, hi = (w[q] >> 1) & 0x55555555u;
      s_pc += __popc(w[q]);
      s_2 += __popc(lo & hi);
      s_any += __popc(lo | hi);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s_pc += __shfl_xor(s_pc, off);
    s_2 += __shfl_xor(s_2, off);
    s_any += __shfl_xor(s_any, off);
  }
  if (lane == 0) {
    pc[j] = s_pc;
    c2[j] = s_2;
    cany[j] = s_any;
    gm[j] = __fdiv_rn((float)s_pc, (float)n);  // src/reml.cpp:620, IEEE round-to-nearest
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Synthetic bed: thread per output byte of gs.  p_j ~ U(lo,hi); code 01 with prob `missing`,
// else A2 count ~ Bin(2,p_j) coded 00/10/11.
__global__ void k_synth(uint8_t* __restrict__ gs, int64_t ld_gs, int64_t n, int64_t m,
                        uint64_t seed, float lo, float hi, float missing, const float* __restrict__ pj) {
  const int64_t ob = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (j >= m || ob >= ld_gs) return;
  const uint64_t hj = mix64(seed ^ mix64((uint64_t)j + 0x1234567ull));
  const float p = pj ? pj[j] : lo + (hi - lo) * (float)((hj >> 40) * (1.0 / 16777216.0));
  const uint32_t pthr = (uint32_t)(p * 2097152.0f);            // 21-bit threshold
  const uint32_t mthr = (uint32_t)(missing * 2097152.0f);
  uint32_t out = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t i = ob * 4 + q;
    if (i < n) {
      const uint64_t r = mix64(hj ^ ((uint64_t)i * 0xD1B54A32D192ED03ull));
      const uint32_t u1 = (uint32_t)(r & 0x1FFFFF), u2 = (uint32_t)((r >> 21) & 0x1FFFFF),
                     u3 = (uint32_t)((r >> 42) & 0x1FFFFF);
      uint32_t code;
      if (u3 < mthr) {
        code = 1u;  // missing
      } else {
        const uint32_t d = (u1 < pthr) + (u2 < pthr);
        code = d == 0 ? 0u : (d == 1 ? 2u : 3u);
      }
      out |= code << (2 * q);
    }
  }
  gs[j * ld_gs + ob] = (uint8_t)out;
}

// Gathers whole SNP-major rows (subset / reorder on the device).
__global__ void k_gather_rows(const uint8_t* __restrict__ src, int64_t ld, const int64_t* __restrict__ idx,
                              int64_t rows, uint8_t* __restrict__ dst) {
  const int64_t r = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (r >= rows) return;
  const uint4* s = reinterpret_cast<const uint4*>(src + idx[r] * ld);
  uint4* d = reinterpret_cast<uint4*>(dst + r * ld);
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < ld / 16;
       v += (int64_t)gridDim.x * blockDim.x)
    d[v] = s[v];
}

// ------------------------------------------------------------------------------ helpers ----

static slemm_geno* alloc_geno(int64_t n, int64_t m) {
  auto* h = new slemm_geno();
  HIPCHECK(hipGetDevice(&h->device));
  h->n = n;
  h->m = m;
  h->n_pad = round_up(n, kPad);
  h->cap_n_pad = h->n_pad;
  h->m_pad = round_up(m, kPad);
  h->ld_gs = h->n_pad / 4;
  h->ld_gi = h->m_pad / 4;
  try {
    HIPCHECK(hipMalloc(&h->gs, (size_t)h->m_pad * h->ld_gs));
    HIPCHECK(hipMalloc(&h->gi, (size_t)h->n_pad * h->ld_gi));
    HIPCHECK(hipMemset(h->gs, 0, (size_t)h->m_pad * h->ld_gs));
    HIPCHECK(hipMalloc(&h->popcnt, sizeof(int32_t) * h->m_pad));
    HIPCHECK(hipMalloc(&h->cnt2, sizeof(int32_t) * h->m_pad));
    HIPCHECK(hipMalloc(&h->gm, sizeof(float) * h->m_pad));
    HIPCHECK(hipMalloc(&h->gw, sizeof(float) * h->m_pad));
    HIPCHECK(hipMalloc(&h->rr, sizeof(float) * h->n_pad));
    HIPCHECK(hipMemset(h->popcnt, 0, sizeof(int32_t) * h->m_pad));
    HIPCHECK(hipMemset(h->cnt2, 0, sizeof(int32_t) * h->m_pad));
    HIPCHECK(hipMemset(h->gm, 0, sizeof(float) * h->m_pad));
    HIPCHECK(hipMemset(h->gw, 0, sizeof(float) * h->m_pad));
    HIPCHECK(hipMemset(h->rr, 0, sizeof(float) * h->n_pad));
  } catch (...) {
    slemm_gpu_geno_destroy(h);
    throw;
  }
  return h;
}

static void build_transpose(slemm_geno* h, hipStream_t st) {
  static_assert(kPad % 512 == 0, "the transpose tiles 512 SNPs");
  dim3 grid((unsigned)(h->n_pad / 256), (unsigned)(h->m_pad / 512));
  k_transpose<<<grid, 256, 0, st>>>(h->gs, h->ld_gs, h->gi, h->ld_gi);
  HIPCHECK(hipGetLastError());
}

void ensure_counts(slemm_geno* h, hipStream_t st) {
  if (h->counts_ready) return;
  DevBuf cany;
  cany.reserve(sizeof(int32_t) * h->m);
  k_counts<<<(unsigned)((h->m + 3) / 4), 256, 0, st>>>(h->gs, h->ld_gs, h->m, h->n, h->popcnt, h->cnt2,
                                                      cany.as<int32_t>(), h->gm);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(st));
  h->counts_ready = true;
}

// The allocation is kept while it is large enough (a handle reused for many fits, as the
// one-trait-at-a-time plan does, then frees and allocates nothing here).
void set_Q(slemm_geno* h, const float* Q, int c) {
  const size_t bytes = sizeof(float) * h->n_pad * (size_t)c;
  if (bytes > h->q_bytes) {
    if (h->Q) (void)hipFree(h->Q);
    h->Q = nullptr;
    h->q_bytes = 0;
    HIPCHECK(hipMalloc(&h->Q, bytes));
    h->q_bytes = bytes;
  }
  h->c = c;
  if (c > 0) {
    HIPCHECK(hipMemset(h->Q, 0, bytes));
    HIPCHECK(hipMemcpy2D(h->Q, sizeof(float) * h->n_pad, Q, sizeof(float) * h->n, sizeof(float) * h->n, c,
                         hipMemcpyHostToDevice));
  }
}

void set_model_weights(slemm_geno* h, const float* gw_dev) {
  HIPCHECK(hipSetDevice(h->device));
  ensure_counts(h, nullptr);
  HIPCHECK(hipMemcpy(h->gw, gw_dev, sizeof(float) * h->m, hipMemcpyDeviceToDevice));
  std::vector<float> ones(h->n, 1.0f);
  HIPCHECK(hipMemcpy(h->rr, ones.data(), sizeof(float) * h->n, hipMemcpyHostToDevice));
  h->rr_host = ones;
  h->rr_unit = true;
  h->rvec_host = ones;
  set_Q(h, nullptr, 0);
  h->model_ready = true;
}

void require_model(const slemm_geno* h) {
  if (!h->model_ready) fail(SLEMM_ERR_STATE, "model terms not set: call slemm_gpu_set_model first");
}

}  // namespace slemm

using namespace slemm;

// ------------------------------------------------------------------------------ C ABI ------

extern "C" const char* slemm_gpu_last_error(void) { return g_last_error.c_str(); }

extern "C" int slemm_gpu_version(void) { return 1; }

extern "C" int slemm_gpu_device_count(int* count) {
  return guard([&] {
    REQUIRE(count, "count is null");
    HIPCHECK(hipGetDeviceCount(count));
  });
}

// validated individual keep list (strictly increasing, in range) as int32; returns n
static int64_t check_keep(const int64_t* keep_idx, int64_t n_keep, int64_t n_raw, std::vector<int32_t>& keep32) {
  const int64_t n = keep_idx ? n_keep : n_raw;
  REQUIRE(n > 0, "individual set retained for analysis is empty");
  if (keep_idx) {
    keep32.resize(n);
    for (int64_t k = 0; k < n; ++k) {
      REQUIRE(keep_idx[k] >= 0 && keep_idx[k] < n_raw, "keep_idx out of range");
      REQUIRE(k == 0 || keep_idx[k] > keep_idx[k - 1], "keep_idx must be strictly increasing");
      keep32[k] = (int32_t)keep_idx[k];
    }
  }
  return n;
}

// writes raw bed rows [j0, j0 + rows) of the output order into dst (rows x row_bytes)
using RowSource = std::function<void(int64_t j0, int64_t rows, uint8_t* dst)>;

// Fills h->gs from raw rows in chunks of at most ~128 MB through two pinned host buffers and
// two device staging buffers: while the GPU copies and compacts chunk c (individual keep list
// applied, padding codes zeroed) the host fills chunk c + 1.  Then builds the individual-major
// copy.  All work is on one private stream.
static void stream_rows(slemm_geno* h, int64_t row_bytes, const std::vector<int32_t>& keep32,
                        const RowSource& src) {
  const int64_t n = h->n, m = h->m;
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>({m, (128ll << 20) / row_bytes, (int64_t)65535}));
  const size_t cbytes = (size_t)chunk * row_bytes;
  hipStream_t st = nullptr;
  uint8_t* pin[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  DevBuf dkeep, stage[2];
  auto cleanup = [&] {
    if (st) (void)hipStreamSynchronize(st);
    for (int b = 0; b < 2; ++b) {
      if (pin[b]) (void)hipHostFree(pin[b]);
      if (done[b]) (void)hipEventDestroy(done[b]);
    }
    if (st) (void)hipStreamDestroy(st);
  };
  try {
    HIPCHECK(hipStreamCreate(&st));
    for (int b = 0; b < 2; ++b) {
      HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&pin[b]), cbytes));
      HIPCHECK(hipEventCreate(&done[b]));
      stage[b].reserve(cbytes);
    }
    if (!keep32.empty()) {
      dkeep.reserve(sizeof(int32_t) * n);
      HIPCHECK(hipMemcpyAsync(dkeep.p, keep32.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    }
    int64_t c = 0;
    for (int64_t j0 = 0; j0 < m; j0 += chunk, ++c) {
      const int b = (int)(c & 1);
      const int64_t rows = std::min(chunk, m - j0);
      if (c >= 2) HIPCHECK(hipEventSynchronize(done[b]));  // chunk c - 2 has left pin[b] and stage[b]
      src(j0, rows, pin[b]);
      HIPCHECK(hipMemcpyAsync(stage[b].p, pin[b], (size_t)rows * row_bytes, hipMemcpyHostToDevice, st));
      dim3 grid((unsigned)((h->ld_gs + 255) / 256), (unsigned)rows);
      k_compact<<<grid, 256, 0, st>>>(stage[b].as<uint8_t>(), row_bytes, (int)rows, dkeep.as<int32_t>(), n,
                                      h->gs + j0 * h->ld_gs, h->ld_gs);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipEventRecord(done[b], st));
    }
    build_transpose(h, st);
    HIPCHECK(hipStreamSynchronize(st));
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

extern "C" int slemm_gpu_geno_create(const uint8_t* bed_rows, int64_t n_raw, int64_t m_raw,
                                     int64_t row_bytes, const int64_t* keep_idx, int64_t n_keep,
                                     const int64_t* snp_idx, int64_t m, slemm_geno** out) {
  return guard([&] {
    REQUIRE(out, "out is null");
    *out = nullptr;
    REQUIRE(bed_rows, "bed_rows is null");
    REQUIRE(n_raw > 0 && m_raw > 0, "empty genotype matrix");
    REQUIRE(row_bytes >= (n_raw + 3) / 4, "row_bytes < ceil(n_raw/4)");
    REQUIRE(n_raw < (int64_t)1 << 31, "n_raw too large");
    std::vector<int32_t> keep32;
    const int64_t n = check_keep(keep_idx, n_keep, n_raw, keep32);
    if (!snp_idx) m = m_raw;
    REQUIRE(m > 0, "marker set retained for analysis is empty");
    if (snp_idx)
      for (int64_t j = 0; j < m; ++j)
        REQUIRE(snp_idx[j] >= 0 && snp_idx[j] < m_raw, "snp_idx out of range");

    slemm_geno* h = alloc_geno(n, m);
    try {
      // rows of the output order, from the caller's array (gathered when snp_idx reorders)
      const RowSource src = [&](int64_t j0, int64_t rows, uint8_t* dst) {
        if (!snp_idx) {
          std::memcpy(dst, bed_rows + j0 * row_bytes, (size_t)rows * row_bytes);
          return;
        }
        for (int64_t r = 0; r < rows; ++r)
          std::memcpy(dst + r * row_bytes, bed_rows + snp_idx[j0 + r] * row_bytes, row_bytes);
      };
      stream_rows(h, row_bytes, keep32, src);
    } catch (...) {
      slemm_gpu_geno_destroy(h);
      throw;
    }
    *out = h;
  });
}

// .bed file -> handle: the body is read with pread straight into pinned buffers (runs of
// consecutive wanted SNPs in one read), overlapped with the copies and compaction of the
// previous chunk; the file is never held in host memory as a whole.
extern "C" int slemm_gpu_geno_load_bed(const char* bed_path, int64_t n_raw, int64_t m_raw, const int64_t* keep_idx,
                                       int64_t n_keep, const int64_t* snp_idx, int64_t m, slemm_geno** out) {
  return guard([&] {
    REQUIRE(out && bed_path, "null argument");
    *out = nullptr;
    REQUIRE(n_raw > 0 && m_raw > 0, "empty genotype matrix");
    REQUIRE(n_raw < (int64_t)1 << 31, "n_raw too large");
    const int64_t row_bytes = (n_raw + 3) / 4;
    std::vector<int32_t> keep32;
    const int64_t n = check_keep(keep_idx, n_keep, n_raw, keep32);
    if (!snp_idx) m = m_raw;
    REQUIRE(m > 0, "marker set retained for analysis is empty");
    if (snp_idx)
      for (int64_t j = 0; j < m; ++j)
        REQUIRE(snp_idx[j] >= 0 && snp_idx[j] < m_raw, "snp_idx out of range");
    const int fd = open(bed_path, O_RDONLY);
    if (fd < 0) fail(SLEMM_ERR_DATA, std::string("Cannot open ") + bed_path);
    struct FdGuard {
      int fd;
      ~FdGuard() { close(fd); }
    } fg{fd};
    // PLINK magic 0x6c 0x1b and 0x01 for SNP-major order (src/input.cpp:352-358)
    uint8_t magic[3] = {0, 0, 0};
    if (pread(fd, magic, 3, 0) != 3 || magic[0] != 0x6c || magic[1] != 0x1b || magic[2] != 0x01)
      fail(SLEMM_ERR_DATA, "Error: problem with the BED file: not a SNP-major PLINK .bed");
    // the body must reach the last SNP read (the reference stops reading there, src/pred.cpp:129)
    int64_t last = m_raw - 1;
    if (snp_idx) last = *std::max_element(snp_idx, snp_idx + m);
    struct stat sb;
    if (fstat(fd, &sb) != 0 || (int64_t)sb.st_size < 3 + (last + 1) * row_bytes)
      fail(SLEMM_ERR_DATA, "Error: problem with the BED file ... has the FAM/BIM file been changed?");
    slemm_geno* h = alloc_geno(n, m);
    try {
      const RowSource src = [&](int64_t j0, int64_t rows, uint8_t* dst) {
        int64_t r = 0;
        while (r < rows) {
          const int64_t first = snp_idx ? snp_idx[j0 + r] : j0 + r;
          int64_t run = 1;
          while (r + run < rows && (snp_idx ? snp_idx[j0 + r + run] : j0 + r + run) == first + run) ++run;
          const size_t want = (size_t)(run * row_bytes);
          size_t got = 0;
          while (got < want) {
            const ssize_t k = pread(fd, dst + r * row_bytes + got, want - got, (off_t)(3 + first * row_bytes + (int64_t)got));
            if (k <= 0) fail(SLEMM_ERR_DATA, "Error: problem with the BED file ... has the FAM/BIM file been changed?");
            got += (size_t)k;
          }
          r += run;
        }
      };
      stream_rows(h, row_bytes, keep32, src);
    } catch (...) {
      slemm_gpu_geno_destroy(h);
      throw;
    }
    *out = h;
  });
}

extern "C" int slemm_gpu_geno_subset(const slemm_geno* hs, const int64_t* snp_idx, int64_t m,
                                     slemm_geno** out) {
  return guard([&] {
    REQUIRE(hs && snp_idx && out, "null argument");
    *out = nullptr;
    REQUIRE(m > 0, "marker set retained for analysis is empty");
    for (int64_t j = 0; j < m; ++j) REQUIRE(snp_idx[j] >= 0 && snp_idx[j] < hs->m, "snp_idx out of range");
    HIPCHECK(hipSetDevice(hs->device));
    slemm_geno* h = alloc_geno(hs->n, m);
    try {
      DevBuf didx;
      didx.reserve(sizeof(int64_t) * m);
      HIPCHECK(hipMemcpy(didx.p, snp_idx, sizeof(int64_t) * m, hipMemcpyHostToDevice));
      dim3 grid(8, (unsigned)std::min<int64_t>(m, 65535), (unsigned)((m + 65534) / 65535));
      k_gather_rows<<<grid, 256>>>(hs->gs, hs->ld_gs, didx.as<int64_t>(), m, h->gs);
      HIPCHECK(hipGetLastError());
      build_transpose(h, nullptr);
      HIPCHECK(hipDeviceSynchronize());
    } catch (...) {
      slemm_gpu_geno_destroy(h);
      throw;
    }
    *out = h;
  });
}

namespace slemm {

slemm_geno* geno_alloc_like(const slemm_geno* src, int64_t cap_n) {
  HIPCHECK(hipSetDevice(src->device));
  return alloc_geno(cap_n, src->m);
}

void geno_refill(slemm_geno* h, const slemm_geno* hs, const int64_t* keep_idx, int64_t n_keep) {
  REQUIRE(h && hs && keep_idx, "null argument");
  std::vector<int32_t> keep32;
  check_keep(keep_idx, n_keep, hs->n, keep32);
  REQUIRE(h->m == hs->m && h->device == hs->device, "handles of different SNP sets or devices");
  REQUIRE(round_up(n_keep, kPad) <= h->cap_n_pad, "more individuals than the handle holds");
  HIPCHECK(hipSetDevice(hs->device));
  h->n = n_keep;
  h->n_pad = round_up(n_keep, kPad);
  h->ld_gs = h->n_pad / 4;
  h->counts_ready = h->model_ready = false;
  set_Q(h, nullptr, 0);
  // the error-variance weights beyond n and the workspaces read as zero, as after allocation (a
  // shorter layout reuses the same memory)
  HIPCHECK(hipMemset(h->rr, 0, sizeof(float) * h->cap_n_pad));
  for (DevBuf* b : {&h->x1, &h->x2, &h->part1, &h->part2, &h->colsum, &h->qtv, &h->tmp, &h->kinds, &h->redpart,
                    &h->partbuf, &h->i8buf})
    if (b->p) HIPCHECK(hipMemset(b->p, 0, b->bytes));
  h->kinds_host.clear();  // the per-column kinds on the device are gone: upload them again
  // individuals are rows of the individual-major copy: gathered and transposed in one pass
  REQUIRE(h->ld_gi == hs->ld_gi, "handles of different SNP layouts");
  DevBuf dkeep;
  dkeep.reserve(sizeof(int64_t) * n_keep);
  HIPCHECK(hipMemcpy(dkeep.p, keep_idx, sizeof(int64_t) * n_keep, hipMemcpyHostToDevice));
  // one pass: the kept rows of the source's individual-major copy (zero rows past n_keep) into
  // this handle's individual-major copy and, transposed, its SNP-major copy
  k_transpose<<<dim3((unsigned)(h->m_pad / 256), (unsigned)(h->n_pad / 512)), 256>>>(
      hs->gi, h->ld_gi, h->gs, h->ld_gs, dkeep.as<int64_t>(), n_keep, h->gi);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipDeviceSynchronize());
}

}  // namespace slemm

extern "C" int slemm_gpu_geno_keep(const slemm_geno* hs, const int64_t* keep_idx, int64_t n_keep,
                                   slemm_geno** out) {
  return guard([&] {
    REQUIRE(hs && keep_idx && out, "null argument");
    *out = nullptr;
    REQUIRE(n_keep > 0, "individual set retained for analysis is empty");
    std::vector<int32_t> keep32;
    check_keep(keep_idx, n_keep, hs->n, keep32);
    slemm_geno* h = geno_alloc_like(hs, n_keep);
    try {
      geno_refill(h, hs, keep_idx, n_keep);
    } catch (...) {
      slemm_gpu_geno_destroy(h);
      throw;
    }
    *out = h;
  });
}

static void synthesize(int64_t n, int64_t m, uint64_t seed, float maf_lo, float maf_hi, float missing,
                       const float* p_host, slemm_geno** out) {
  REQUIRE(out, "out is null");
  *out = nullptr;
  REQUIRE(n > 0 && m > 0, "empty genotype matrix");
  REQUIRE(n < (int64_t)1 << 31, "n too large");
  REQUIRE(maf_lo >= 0 && maf_hi <= 1 && maf_lo <= maf_hi, "bad allele-frequency range");
  REQUIRE(missing >= 0 && missing < 1, "bad missing rate");
  if (p_host)
    for (int64_t j = 0; j < m; ++j) REQUIRE(p_host[j] >= 0 && p_host[j] <= 1, "allele frequencies must be in [0,1]");
  slemm_geno* h = alloc_geno(n, m);
  try {
    DevBuf dp;
    if (p_host) {
      dp.reserve(sizeof(float) * m);
      HIPCHECK(hipMemcpy(dp.p, p_host, sizeof(float) * m, hipMemcpyHostToDevice));
    }
    dim3 grid((unsigned)((h->ld_gs + 255) / 256), (unsigned)std::min<int64_t>(m, 65535),
              (unsigned)((m + 65534) / 65535));
    k_synth<<<grid, 256>>>(h->gs, h->ld_gs, n, m, seed, maf_lo, maf_hi, missing, dp.as<float>());
    HIPCHECK(hipGetLastError());
    build_transpose(h, nullptr);
    HIPCHECK(hipDeviceSynchronize());
  } catch (...) {
    slemm_gpu_geno_destroy(h);
    throw;
  }
  *out = h;
}

extern "C" int slemm_gpu_geno_synthesize_p(int64_t n, int64_t m, uint64_t seed, const float* p, float missing,
                                           slemm_geno** out) {
  return guard([&] {
    REQUIRE(p, "p is null");
    synthesize(n, m, seed, 0.f, 1.f, missing, p, out);
  });
}

extern "C" int slemm_gpu_geno_synthesize(int64_t n, int64_t m, uint64_t seed, float maf_lo,
                                         float maf_hi, float missing, slemm_geno** out) {
  return guard([&] { synthesize(n, m, seed, maf_lo, maf_hi, missing, nullptr, out); });
}

extern "C" int slemm_gpu_geno_destroy(slemm_geno* h) {
  return guard([&] {
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
    if (h->gs) (void)hipFree(h->gs);
    if (h->gi) (void)hipFree(h->gi);
    if (h->popcnt) (void)hipFree(h->popcnt);
    if (h->cnt2) (void)hipFree(h->cnt2);
    if (h->gm) (void)hipFree(h->gm);
    if (h->gw) (void)hipFree(h->gw);
    if (h->rr) (void)hipFree(h->rr);
    if (h->Q) (void)hipFree(h->Q);
    delete h;
    (void)hipGetLastError();  // nothing of this call may stay pending for the caller's next check
  });
}

extern "C" int slemm_gpu_geno_dims(const slemm_geno* h, int64_t* n, int64_t* m) {
  return guard([&] {
    REQUIRE(h, "null handle");
    if (n) *n = h->n;
    if (m) *m = h->m;
  });
}

extern "C" int slemm_gpu_geno_download(const slemm_geno* h, int64_t j0, int64_t cnt, uint8_t* rows) {
  return guard([&] {
    REQUIRE(h && rows, "null argument");
    REQUIRE(j0 >= 0 && cnt >= 0 && j0 + cnt <= h->m, "row range out of bounds");
    HIPCHECK(hipSetDevice(h->device));
    const int64_t rb = (h->n + 3) / 4;
    HIPCHECK(hipMemcpy2D(rows, rb, h->gs + j0 * h->ld_gs, h->ld_gs, rb, cnt, hipMemcpyDeviceToHost));
  });
}

extern "C" int slemm_gpu_geno_counts(slemm_geno* h, int32_t* popcnt, int32_t* n0, int32_t* n1,
                                     int32_t* n2, float* gm) {
  return guard([&] {
    REQUIRE(h, "null handle");
    HIPCHECK(hipSetDevice(h->device));
    DevBuf cany;
    cany.reserve(sizeof(int32_t) * h->m);
    k_counts<<<(unsigned)((h->m + 3) / 4), 256>>>(h->gs, h->ld_gs, h->m, h->n, h->popcnt, h->cnt2,
                                                  cany.as<int32_t>(), h->gm);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
    h->counts_ready = true;
    std::vector<int32_t> v2(h->m), vany(h->m);
    HIPCHECK(hipMemcpy(v2.data(), h->cnt2, sizeof(int32_t) * h->m, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(vany.data(), cany.p, sizeof(int32_t) * h->m, hipMemcpyDeviceToHost));
    if (popcnt) HIPCHECK(hipMemcpy(popcnt, h->popcnt, sizeof(int32_t) * h->m, hipMemcpyDeviceToHost));
    if (gm) HIPCHECK(hipMemcpy(gm, h->gm, sizeof(float) * h->m, hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < h->m; ++j) {
      if (n2) n2[j] = v2[j];
      if (n1) n1[j] = vany[j] - v2[j];
      if (n0) n0[j] = (int32_t)(h->n - vany[j]);
    }
  });
}

extern "C" int slemm_gpu_set_model(slemm_geno* h, const float* persnp, const float* rvec,
                                   const float* Q, int64_t ldq, int c) {
  return guard([&] {
    REQUIRE(h, "null handle");
    REQUIRE(c >= 0 && c <= 64, "number of covariates out of [0, 64]");
    REQUIRE(c == 0 || (Q && ldq >= h->n), "Q is null or ldq < n");
    HIPCHECK(hipSetDevice(h->device));
    ensure_counts(h, nullptr);
    const int64_t n = h->n, m = h->m;
    std::vector<float> gm(m), gw(m), rr(n);
    HIPCHECK(hipMemcpy(gm.data(), h->gm, sizeof(float) * m, hipMemcpyDeviceToHost));
    // mean of the per-SNP weights (src/reml.cpp:622 persnp.mean()); exact for unit weights
    float mean = 1.0f;
    if (persnp) {
      double s = 0;
      for (int64_t j = 0; j < m; ++j) s += persnp[j];
      mean = (float)(s / (double)m);
      REQUIRE(mean > 0, "per-SNP weights must have a positive mean");
    }
    for (int64_t j = 0; j < m; ++j) {
      const float g = gm[j];
      const float v = (g * (2.0f - g)) / 2.0f;           // src/reml.cpp:622, float arithmetic
      const float wj = persnp ? persnp[j] / mean : 1.0f;
      gw[j] = ((1.0f / v) / (float)m) * wj;
    }
    for (int64_t i = 0; i < n; ++i) {
      if (rvec) {
        REQUIRE(rvec[i] > 0, "error-variance weights must be positive");
        rr[i] = 1.0f / std::sqrt(rvec[i]);                // src/reml.cpp:574
      } else {
        rr[i] = 1.0f;
      }
    }
    for (int64_t j = 0; j < m; ++j)
      REQUIRE(std::isfinite(gw[j]), "monomorphic SNP (allele frequency 0 or 1) gives an infinite weight");
    HIPCHECK(hipMemcpy(h->gw, gw.data(), sizeof(float) * m, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h->rr, rr.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    h->rr_host = rr;
    h->rr_unit = std::all_of(rr.begin(), rr.end(), [](float v) { return v == 1.0f; });
    h->rvec_host.assign(n, 1.0f);
    if (rvec) std::copy(rvec, rvec + n, h->rvec_host.begin());
    std::vector<float> qc((size_t)n * c);
    for (int q = 0; q < c; ++q) std::copy(Q + (size_t)q * ldq, Q + (size_t)q * ldq + n, qc.begin() + (size_t)q * n);
    set_Q(h, qc.data(), c);
    h->model_ready = true;
  });
}

extern "C" int slemm_gpu_set_gw(slemm_geno* h, const float* gw) {
  return guard([&] {
    REQUIRE(h && gw, "null argument");
    require_model(h);
    for (int64_t j = 0; j < h->m; ++j)
      REQUIRE(std::isfinite(gw[j]) && gw[j] >= 0, "SNP weights must be finite and non-negative");
    HIPCHECK(hipSetDevice(h->device));
    HIPCHECK(hipMemcpy(h->gw, gw, sizeof(float) * h->m, hipMemcpyHostToDevice));
  });
}

extern "C" int slemm_gpu_get_model(const slemm_geno* h, float* gm, float* gw, float* r_rsqrt) {
  return guard([&] {
    REQUIRE(h, "null handle");
    require_model(h);
    HIPCHECK(hipSetDevice(h->device));
    if (gm) HIPCHECK(hipMemcpy(gm, h->gm, sizeof(float) * h->m, hipMemcpyDeviceToHost));
    if (gw) HIPCHECK(hipMemcpy(gw, h->gw, sizeof(float) * h->m, hipMemcpyDeviceToHost));
    if (r_rsqrt) HIPCHECK(hipMemcpy(r_rsqrt, h->rr, sizeof(float) * h->n, hipMemcpyDeviceToHost));
  });
}
