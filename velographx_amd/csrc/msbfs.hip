// Multi-source BFS: up to 64 roots traversed together (batched BFS).
//
// The reference has no multi-source traversal; its closest analogue runs
// independent single-root BFS on threads (benchmarks/thread_scaling_bfs.cpp
// :78-108).  Here one adjacency pass serves every root: vertex v carries a
// 64-bit mask seen[v] (bit r = root r reached it) and each level stores the
// bits that became new, newbits[L][v].  Bit r of newbits[L][v] is set exactly
// when v is at distance L from root r, so every root's distances equal its
// single-source BFS.
//
// A level pulls: an unfinished vertex ORs the level's frontier masks of its
// in-neighbours and stops as soon as every root is accounted for.  Rows of
// up to 32 entries are scanned by one lane, longer rows by the whole wave.
// Per-root uint8 levels (255 = unreachable) are written at the end by a
// transpose that keeps the byte stores coalesced.
#include <algorithm>
#include <vector>

#include "vgx_internal.hpp"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);
extern thread_local std::string t_last_error;

constexpr int kMsBlock = 256;
constexpr uint32_t kMsShortRow = 32;

struct MsCounters {
  unsigned long long new_pairs;  // (vertex, root) pairs reached this level
};

}  // namespace vgx

struct vgx_msbfs {
  vgx_graph* g = nullptr;
  std::vector<uint32_t> roots;
  uint32_t k = 0;
  uint64_t n = 0;
  uint32_t levels = 0;  // number of stored levels (newbits[0..levels))
  bool ran = false;
  vgx::DevBuf<unsigned long long> seen;
  vgx::DevBuf<unsigned long long> newbits;  // [levels_cap][n]
  uint64_t levels_cap = 0;
  vgx::MsCounters* d_ctr = nullptr;
  vgx::MsCounters* h_ctr = nullptr;
  std::vector<uint64_t> reached;  // per root
};

namespace vgx {

__global__ void k_ms_init(unsigned long long* seen, unsigned long long* nb0, uint64_t n, const uint32_t* roots,
                          uint32_t k) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    unsigned long long m = 0;
    for (uint32_t r = 0; r < k; ++r)
      if (roots[r] == v) m |= 1ull << r;
    seen[v] = m;
    nb0[v] = m;
  }
}

// One pull level: frontier = newbits[L], output newbits[L+1].
__global__ __launch_bounds__(kMsBlock) void k_ms_pull(const uint2* __restrict__ in_row,
                                                       const uint32_t* __restrict__ in_arena, uint32_t n,
                                                       unsigned long long all, unsigned long long* __restrict__ seen,
                                                       const unsigned long long* __restrict__ frontier,
                                                       unsigned long long* __restrict__ next, MsCounters* ctr) {
  const uint32_t W = gridDim.x * (kMsBlock / kWave);
  const uint32_t gw = blockIdx.x * (kMsBlock / kWave) + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  unsigned long long pairs = 0;
  const uint32_t groups = (n + kWave - 1) / kWave;
  for (uint32_t grp = gw; grp < groups; grp += W) {
    const uint32_t v = grp * kWave + lane;
    const unsigned long long s = v < n ? seen[v] : all;
    const bool open = s != all;
    uint2 rb = make_uint2(0u, 0u);
    if (open) rb = in_row[v];
    unsigned long long acc = 0;
    const bool long_row = open && rb.y > kMsShortRow;
    if (open && !long_row) {
      const uint32_t* p = in_arena + rb.x;
      uint32_t j = 0;
      for (; j + 4 <= rb.y; j += 4) {
        const uint32_t x0 = p[j], x1 = p[j + 1], x2 = p[j + 2], x3 = p[j + 3];
        acc |= (x0 & kTomb ? 0ull : frontier[x0]) | (x1 & kTomb ? 0ull : frontier[x1]) |
               (x2 & kTomb ? 0ull : frontier[x2]) | (x3 & kTomb ? 0ull : frontier[x3]);
        if ((acc | s) == all) break;
      }
      if ((acc | s) != all)
        for (; j < rb.y; ++j) {
          const uint32_t x = p[j];
          if (!(x & kTomb)) acc |= frontier[x];
        }
    }
    unsigned long long lm = __ballot(long_row);
    while (lm) {
      const int l = __ffsll(static_cast<long long>(lm)) - 1;
      lm &= lm - 1;
      const uint32_t rx = __shfl(rb.x, l, kWave), ry = __shfl(rb.y, l, kWave);
      const unsigned long long sl = __shfl(s, l, kWave);
      unsigned long long a = 0;
      for (uint32_t j = 0; j < ry; j += kWave) {
        const uint32_t x = j + lane < ry ? in_arena[rx + j + lane] : kTomb;
        a |= x & kTomb ? 0ull : frontier[x];
        unsigned long long w = a;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) w |= __shfl_xor(w, d, kWave);
        a = w;
        if ((a | sl) == all) break;
      }
      if (static_cast<int>(lane) == l) acc = a;
    }
    const unsigned long long nw = open ? (acc & ~s) : 0ull;
    if (nw) {
      seen[v] = s | nw;
      next[v] = nw;
      pairs += __popcll(nw);
    }
  }
  pairs = wave_sum(pairs);
  if (lane == 0 && pairs) atomicAdd(&ctr->new_pairs, pairs);
}

// Per-root uint8 levels, root-major: out[r * n + v].  Each lane builds the
// 64 levels of its vertex in registers, then every root's row is written
// with consecutive lanes on consecutive bytes.
__global__ void k_ms_levels_u8(const unsigned long long* __restrict__ newbits, uint32_t levels, uint64_t n,
                               uint32_t k, uint8_t* __restrict__ out) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t packed[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) packed[i] = 0xFFFFFFFFu;
    for (uint32_t L = 0; L < levels; ++L) {
      unsigned long long m = newbits[static_cast<uint64_t>(L) * n + v];
      while (m) {
        const int r = __ffsll(static_cast<long long>(m)) - 1;
        m &= m - 1;
        const int w = r >> 2, sh = (r & 3) * 8;
#pragma unroll
        for (int i = 0; i < 16; ++i)
          if (i == w) packed[i] = (packed[i] & ~(0xFFu << sh)) | (L << sh);
      }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const uint32_t r = i * 4 + b;
        if (r < k) out[static_cast<uint64_t>(r) * n + v] = static_cast<uint8_t>(packed[i] >> (b * 8));
      }
    }
  }
}

// uint32 distances of one root from the per-level masks.
__global__ void k_ms_distances(const unsigned long long* __restrict__ newbits, uint32_t levels, uint64_t n,
                               uint32_t r, uint32_t* __restrict__ out) {
  const unsigned long long bit = 1ull << r;
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t d = kInf;
    for (uint32_t L = 0; L < levels; ++L)
      if (newbits[static_cast<uint64_t>(L) * n + v] & bit) {
        d = L;
        break;
      }
    out[v] = d;
  }
}

static unsigned ms_blocks(uint64_t items, unsigned per = kMsBlock, unsigned cap = 4096) {
  uint64_t b = (items + per - 1) / per;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

static void ms_ensure_levels(vgx_msbfs* m, uint64_t need) {
  if (need <= m->levels_cap) return;
  uint64_t cap = std::max<uint64_t>(need, m->levels_cap ? m->levels_cap * 2 : 16);
  m->newbits.grow_copy(cap * m->n + 1, m->levels_cap * m->n, m->g->stream);
  m->levels_cap = cap;
}

static void ms_run(vgx_msbfs* m) {
  vgx_graph* g = m->g;
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  m->n = n;
  m->levels_cap = 0;
  m->newbits.release();
  m->seen.ensure(n + 1);
  ms_ensure_levels(m, 16);
  const unsigned long long all = m->k == 64 ? ~0ull : ((1ull << m->k) - 1);
  DevBuf<uint32_t> droots;
  droots.ensure(m->k + 1);
  VGX_HIP(hipMemcpyAsync(droots.p, m->roots.data(), m->k * 4, hipMemcpyHostToDevice, s));
  m->levels = 0;
  if (n) {
    k_ms_init<<<ms_blocks(n), kMsBlock, 0, s>>>(m->seen.p, m->newbits.p, n, droots.p, m->k);
    VGX_LAUNCH_CHECK();
  }
  m->levels = 1;
  const uint2* in_row = g->in().row.p;
  const uint32_t* in_arena = g->in().arena.p;
  const unsigned blocks = ms_blocks((n + kWave - 1) / kWave, kMsBlock / kWave, 2048);
  while (n) {
    ms_ensure_levels(m, m->levels + 1);
    unsigned long long* nxt = m->newbits.p + m->levels * n;
    VGX_HIP(hipMemsetAsync(nxt, 0, n * 8, s));
    VGX_HIP(hipMemsetAsync(m->d_ctr, 0, sizeof(MsCounters), s));
    k_ms_pull<<<blocks, kMsBlock, 0, s>>>(in_row, in_arena, static_cast<uint32_t>(n), all, m->seen.p,
                                          m->newbits.p + (m->levels - 1) * n, nxt, m->d_ctr);
    VGX_LAUNCH_CHECK();
    VGX_HIP(hipMemcpyAsync(m->h_ctr, m->d_ctr, sizeof(MsCounters), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    if (m->h_ctr->new_pairs == 0) break;
    m->levels += 1;
  }
  m->ran = true;
  m->reached.assign(m->k, 0);
}

}  // namespace vgx

using namespace vgx;

#define VGX_API_BEGIN try {
#define VGX_API_END                                                   \
  }                                                                   \
  catch (const vgx::Error& e) {                                       \
    return vgx::set_error(e.code, e.what());                          \
  }                                                                   \
  catch (const std::bad_alloc&) {                                     \
    return vgx::set_error(VGX_OUT_OF_MEMORY, "host allocation failed"); \
  }                                                                   \
  catch (const std::exception& e) {                                   \
    return vgx::set_error(VGX_INVALID_ARGUMENT, e.what());            \
  }                                                                   \
  vgx::t_last_error.clear();                                          \
  return VGX_OK;

extern "C" {

vgx_status vgx_msbfs_create(vgx_graph* g, const uint32_t* roots, uint32_t k, vgx_msbfs** out) {
  VGX_API_BEGIN
  if (!g || !out || (k && !roots)) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (k == 0 || k > 64) throw Error(VGX_INVALID_ARGUMENT, "between 1 and 64 roots");
  VGX_HIP(hipSetDevice(g->device));
  auto* m = new vgx_msbfs();
  try {
    m->g = g;
    m->roots.assign(roots, roots + k);
    m->k = k;
    m->d_ctr = static_cast<MsCounters*>(dev_alloc(sizeof(MsCounters)));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&m->h_ctr), sizeof(MsCounters), hipHostMallocDefault));
  } catch (...) {
    vgx_msbfs_destroy(m);
    throw;
  }
  *out = m;
  VGX_API_END
}

vgx_status vgx_msbfs_destroy(vgx_msbfs* m) {
  VGX_API_BEGIN
  if (!m) return VGX_OK;
  if (m->g) (void)hipStreamSynchronize(m->g->stream);
  if (m->d_ctr) dev_free(m->d_ctr);
  if (m->h_ctr) (void)hipHostFree(m->h_ctr);
  delete m;
  VGX_API_END
}

vgx_status vgx_msbfs_run(vgx_msbfs* m) {
  VGX_API_BEGIN
  if (!m) throw Error(VGX_INVALID_ARGUMENT, "null msbfs");
  VGX_HIP(hipSetDevice(m->g->device));
  ms_run(m);
  VGX_API_END
}

vgx_status vgx_msbfs_distances(vgx_msbfs* m, uint32_t r, uint32_t* out, uint64_t cap) {
  VGX_API_BEGIN
  if (!m || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (!m->ran) throw Error(VGX_INVALID_ARGUMENT, "run() first");
  if (r >= m->k) throw Error(VGX_OUT_OF_RANGE, "root index out of range");
  if (cap < m->n) throw Error(VGX_INVALID_ARGUMENT, "output buffer too small");
  VGX_HIP(hipSetDevice(m->g->device));
  hipStream_t s = m->g->stream;
  DevBuf<uint32_t> tmp;
  tmp.ensure(m->n + 1);
  if (m->n) {
    k_ms_distances<<<ms_blocks(m->n), kMsBlock, 0, s>>>(m->newbits.p, m->levels, m->n, r, tmp.p);
    VGX_LAUNCH_CHECK();
    VGX_HIP(hipMemcpyAsync(out, tmp.p, m->n * 4, hipMemcpyDeviceToHost, s));
  }
  VGX_HIP(hipStreamSynchronize(s));
  VGX_API_END
}

vgx_status vgx_msbfs_levels_u8(vgx_msbfs* m, uint8_t* out, uint64_t cap) {
  VGX_API_BEGIN
  if (!m || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (!m->ran) throw Error(VGX_INVALID_ARGUMENT, "run() first");
  if (cap < static_cast<uint64_t>(m->k) * m->n) throw Error(VGX_INVALID_ARGUMENT, "output buffer too small");
  if (m->levels > 255) throw Error(VGX_UNSUPPORTED, "depth exceeds 254: use vgx_msbfs_distances");
  VGX_HIP(hipSetDevice(m->g->device));
  hipStream_t s = m->g->stream;
  DevBuf<uint8_t> tmp;
  tmp.ensure(static_cast<uint64_t>(m->k) * m->n + 1);
  if (m->n) {
    k_ms_levels_u8<<<ms_blocks(m->n), kMsBlock, 0, s>>>(m->newbits.p, m->levels, m->n, m->k, tmp.p);
    VGX_LAUNCH_CHECK();
    VGX_HIP(hipMemcpyAsync(out, tmp.p, static_cast<uint64_t>(m->k) * m->n, hipMemcpyDeviceToHost, s));
  }
  VGX_HIP(hipStreamSynchronize(s));
  VGX_API_END
}

vgx_status vgx_msbfs_reached(vgx_msbfs* m, uint64_t* reached_per_root, uint32_t cap) {
  VGX_API_BEGIN
  if (!m || !reached_per_root) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (!m->ran) throw Error(VGX_INVALID_ARGUMENT, "run() first");
  VGX_HIP(hipSetDevice(m->g->device));
  std::vector<unsigned long long> seen(m->n);
  if (m->n) {
    VGX_HIP(hipMemcpyAsync(seen.data(), m->seen.p, m->n * 8, hipMemcpyDeviceToHost, m->g->stream));
    VGX_HIP(hipStreamSynchronize(m->g->stream));
  }
  for (uint32_t r = 0; r < m->k && r < cap; ++r) {
    uint64_t c = 0;
    for (uint64_t v = 0; v < m->n; ++v) c += (seen[v] >> r) & 1ull;
    reached_per_root[r] = c;
  }
  VGX_API_END
}

// Number of levels of the last run (the deepest distance + 1).
vgx_status vgx_msbfs_level_count(vgx_msbfs* m, uint32_t* levels) {
  VGX_API_BEGIN
  if (!m || !levels) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  *levels = m->levels;
  VGX_API_END
}

}  // extern "C"
