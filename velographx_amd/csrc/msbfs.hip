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
// Each level runs in one of two directions over the packed CSR snapshots:
//   push: every frontier vertex u ORs its new bits, minus what the target
//         has seen, into next[v] of its out-neighbours (64-bit atomicOr);
//   pull: every unfinished vertex ORs the frontier masks of its
//         in-neighbours and stops once every root is accounted for.
// Push is used while the frontier's out-arcs are fewer than 1/alpha of all
// arcs.  Rows of up to 32 arcs are handled by one lane (64 consecutive
// vertices per wave); longer rows are cut into chunks of 1024 arcs, one wave
// each, so R-MAT hubs (which cluster at low ids) spread over many waves.
// OR is commutative and idempotent, so the result does not depend on the
// order of the atomics.  A separate pass folds next into seen and counts the
// new (vertex, root) pairs and the next frontier's out-arcs.
// Per-root uint8 levels (255 = unreachable) are written at the end by a
// transpose that keeps the byte stores coalesced.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <vector>

#include "vgx_internal.hpp"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);
extern thread_local std::string t_last_error;

constexpr int kMsBlock = 256;
constexpr uint32_t kMsWaves = kMsBlock / kWave;
constexpr uint32_t kMsShortRow = 32;
constexpr uint32_t kMsChunk = 1024;

// Level counters, striped over kMsStripes cache lines (workgroup b adds to
// stripe b % kMsStripes; the host sums the stripes), so the finalize pass
// does not serialise on one address.
constexpr uint32_t kMsStripes = 32;
struct alignas(128) MsStripe {
  unsigned long long new_pairs;      // (vertex, root) pairs reached this level
  unsigned long long frontier_arcs;  // out-arcs of the vertices reached this level
};
struct MsCounters {
  MsStripe s[kMsStripes];
  unsigned long long pairs() const {
    unsigned long long t = 0;
    for (const auto& x : s) t += x.new_pairs;
    return t;
  }
  unsigned long long arcs() const {
    unsigned long long t = 0;
    for (const auto& x : s) t += x.frontier_arcs;
    return t;
  }
};

// Chunks of the rows longer than kMsShortRow in one snapshot: row v owns
// chunks [row_chunk[v], row_chunk[v+1]).
struct MsPlan {
  DevBuf<uint32_t> row_chunk;
  DevBuf<uint32_t> chunk_row;
  uint64_t nchunks = 0;
  uint64_t version = ~0ull;
  uint64_t n = 0;
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
  uint32_t blocks = 0;            // short-row kernel workgroups, 0 = auto
  double alpha = 4.0;             // push while frontier arcs < arcs / alpha
  vgx::MsPlan plan[2];            // [0] out-rows (push), [1] in-rows (pull)
  vgx::DevBuf<uint8_t> scan;
  uint32_t push_levels = 0;
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

__global__ void k_ms_chunk_count(const uint32_t* __restrict__ off, uint32_t n, uint32_t* __restrict__ cnt) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t len = off[v + 1] - off[v];
    cnt[v] = len > kMsShortRow ? (len + kMsChunk - 1) / kMsChunk : 0u;
  }
}

__global__ void k_ms_chunk_fill(const uint32_t* __restrict__ row_chunk, uint32_t n, uint32_t* __restrict__ chunk_row) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    for (uint32_t c = row_chunk[v]; c < row_chunk[v + 1]; ++c) chunk_row[c] = v;
}

// push, short rows: lane per frontier vertex u.
__global__ __launch_bounds__(kMsBlock) void k_ms_push_short(const uint32_t* __restrict__ off,
                                                             const uint32_t* __restrict__ adj, uint32_t n,
                                                             const unsigned long long* __restrict__ frontier,
                                                             const unsigned long long* __restrict__ seen,
                                                             unsigned long long* next) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const unsigned long long f = frontier[u];
    if (!f) continue;
    const uint32_t b = off[u], e = off[u + 1];
    if (e - b > kMsShortRow) continue;
    for (uint32_t i = b; i < e; ++i) {
      const uint32_t v = adj[i];
      const unsigned long long m = f & ~seen[v];
      // bits only ever get added to next[v]: skip the atomic when they are there
      if (m & ~next[v]) atomicOr(&next[v], m);
    }
  }
}

// push, long rows: wave per chunk, lane per arc.
__global__ __launch_bounds__(kMsBlock) void k_ms_push_long(const uint32_t* __restrict__ off,
                                                            const uint32_t* __restrict__ adj,
                                                            const uint32_t* __restrict__ row_chunk,
                                                            const uint32_t* __restrict__ chunk_row, uint32_t nchunks,
                                                            const unsigned long long* __restrict__ frontier,
                                                            const unsigned long long* __restrict__ seen,
                                                            unsigned long long* next) {
  for (uint32_t c = blockIdx.x * kMsWaves + threadIdx.x / kWave; c < nchunks; c += gridDim.x * kMsWaves) {
    const uint32_t u = chunk_row[c];
    const unsigned long long f = frontier[u];
    if (!f) continue;
    const uint32_t b = off[u] + (c - row_chunk[u]) * kMsChunk;
    const uint32_t e = min(b + kMsChunk, off[u + 1]);
    for (uint32_t i = b + lane_id(); i < e; i += kWave) {
      const uint32_t v = adj[i];
      const unsigned long long m = f & ~seen[v];
      if (m & ~next[v]) atomicOr(&next[v], m);
    }
  }
}

// pull, short rows: lane per unfinished vertex v, early exit once every
// root is accounted for.
__global__ __launch_bounds__(kMsBlock) void k_ms_pull_short(const uint32_t* __restrict__ off,
                                                             const uint32_t* __restrict__ adj, uint32_t n,
                                                             unsigned long long all,
                                                             const unsigned long long* __restrict__ frontier,
                                                             const unsigned long long* __restrict__ seen,
                                                             unsigned long long* __restrict__ next) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const unsigned long long s = seen[v];
    if (s == all) continue;
    const uint32_t b = off[v], e = off[v + 1];
    if (e - b > kMsShortRow) continue;
    unsigned long long acc = 0;
    uint32_t i = b;
    for (; i + 4 <= e; i += 4) {
      acc |= frontier[adj[i]] | frontier[adj[i + 1]] | frontier[adj[i + 2]] | frontier[adj[i + 3]];
      if ((acc | s) == all) break;
    }
    if ((acc | s) != all)
      for (; i < e; ++i) acc |= frontier[adj[i]];
    const unsigned long long m = acc & ~s;
    if (m) next[v] = m;  // the only writer of a short row's entry
  }
}

__device__ __forceinline__ unsigned long long wave_or(unsigned long long x) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) x |= __shfl_xor(x, d, kWave);
  return x;
}

// pull, long rows: wave per chunk; the early-exit test runs every 4 steps.
__global__ __launch_bounds__(kMsBlock) void k_ms_pull_long(const uint32_t* __restrict__ off,
                                                            const uint32_t* __restrict__ adj,
                                                            const uint32_t* __restrict__ row_chunk,
                                                            const uint32_t* __restrict__ chunk_row, uint32_t nchunks,
                                                            unsigned long long all,
                                                            const unsigned long long* __restrict__ frontier,
                                                            const unsigned long long* __restrict__ seen,
                                                            unsigned long long* __restrict__ next) {
  for (uint32_t c = blockIdx.x * kMsWaves + threadIdx.x / kWave; c < nchunks; c += gridDim.x * kMsWaves) {
    const uint32_t v = chunk_row[c];
    const unsigned long long s = seen[v];
    if (s == all) continue;
    const uint32_t b = off[v] + (c - row_chunk[v]) * kMsChunk;
    const uint32_t e = min(b + kMsChunk, off[v + 1]);
    unsigned long long acc = 0;
    for (uint32_t i0 = b; i0 < e; i0 += 4 * kWave) {
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t i = i0 + j * kWave + lane_id();
        if (i < e) acc |= frontier[adj[i]];
      }
      acc = wave_or(acc);
      if ((acc | s) == all) break;
    }
    const unsigned long long m = acc & ~s;
    if (lane_id() == 0 && m) atomicOr(&next[v], m);
  }
}

// Fold the level into seen; count new pairs and the next frontier's arcs.
__global__ __launch_bounds__(kMsBlock) void k_ms_finalize(const uint32_t* __restrict__ out_off, uint32_t n,
                                                           unsigned long long* __restrict__ seen,
                                                           const unsigned long long* __restrict__ next,
                                                           MsCounters* ctr) {
  unsigned long long pairs = 0, arcs = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const unsigned long long nw = next[v];
    if (!nw) continue;
    seen[v] |= nw;
    pairs += __popcll(nw);
    arcs += out_off[v + 1] - out_off[v];
  }
  __shared__ unsigned long long sp[kMsWaves], sa[kMsWaves];
  pairs = wave_sum(pairs);
  arcs = wave_sum(arcs);
  if (lane_id() == 0) {
    sp[threadIdx.x / kWave] = pairs;
    sa[threadIdx.x / kWave] = arcs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < kMsWaves; ++w) {
      pairs += sp[w];
      arcs += sa[w];
    }
    MsStripe& st = ctr->s[blockIdx.x % kMsStripes];
    if (pairs) atomicAdd(&st.new_pairs, pairs);
    if (arcs) atomicAdd(&st.frontier_arcs, arcs);
  }
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

// Long-row chunks of one snapshot, rebuilt when the snapshot changes.
static const MsPlan& ms_plan(vgx_msbfs* m, int dir, const CsrSnapshot& sn, hipStream_t s) {
  MsPlan& pl = m->plan[dir];
  if (pl.version == sn.version && pl.n == sn.n && pl.row_chunk.p) return pl;
  const uint32_t n = static_cast<uint32_t>(sn.n);
  pl.row_chunk.ensure(sn.n + 2);
  VGX_HIP(hipMemsetAsync(pl.row_chunk.p, 0, sizeof(uint32_t), s));
  uint32_t total = 0;
  if (n) {
    DevBuf<uint32_t> cnt;
    cnt.ensure(sn.n + 1);
    k_ms_chunk_count<<<ms_blocks(n), kMsBlock, 0, s>>>(sn.off.p, n, cnt.p);
    VGX_LAUNCH_CHECK();
    size_t bytes = 0;
    VGX_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, cnt.p, pl.row_chunk.p + 1, static_cast<int>(n), s));
    m->scan.ensure(bytes + 16);
    VGX_HIP(hipcub::DeviceScan::InclusiveSum(m->scan.p, bytes, cnt.p, pl.row_chunk.p + 1, static_cast<int>(n), s));
    VGX_HIP(hipMemcpyAsync(&total, pl.row_chunk.p + n, sizeof(total), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    pl.chunk_row.ensure(static_cast<uint64_t>(total) + 1);
    k_ms_chunk_fill<<<ms_blocks(n), kMsBlock, 0, s>>>(pl.row_chunk.p, n, pl.chunk_row.p);
    VGX_LAUNCH_CHECK();
  }
  pl.nchunks = total;
  pl.version = sn.version;
  pl.n = sn.n;
  return pl;
}

static void ms_run(vgx_msbfs* m) {
  vgx_graph* g = m->g;
  hipStream_t s = g->stream;
  const CsrSnapshot& out = graph_snapshot(g, 0);
  const CsrSnapshot& in = graph_snapshot(g, 1);
  const MsPlan& po = ms_plan(m, 0, out, s);
  const MsPlan& pi = ms_plan(m, 1, in, s);
  const uint64_t n = g->n;
  if (n != m->n) {  // the per-level arrays are laid out with stride n
    m->levels_cap = 0;
    m->newbits.release();
  }
  m->n = n;
  m->seen.ensure(n + 1);
  ms_ensure_levels(m, 16);
  const unsigned long long all = m->k == 64 ? ~0ull : ((1ull << m->k) - 1);
  DevBuf<uint32_t> droots;
  droots.ensure(m->k + 1);
  VGX_HIP(hipMemcpyAsync(droots.p, m->roots.data(), m->k * 4, hipMemcpyHostToDevice, s));
  m->levels = 0;
  m->push_levels = 0;
  if (n) {
    k_ms_init<<<ms_blocks(n), kMsBlock, 0, s>>>(m->seen.p, m->newbits.p, n, droots.p, m->k);
    VGX_LAUNCH_CHECK();
  }
  m->levels = 1;
  const uint32_t nn = static_cast<uint32_t>(n);
  const unsigned bshort = m->blocks ? m->blocks : ms_blocks(n, kMsBlock, 4096);
  const unsigned bout = ms_blocks(po.nchunks, kMsWaves, 8192);
  const unsigned bin = ms_blocks(pi.nchunks, kMsWaves, 8192);
  const double arcs = static_cast<double>(out.arcs);
  unsigned long long frontier_arcs = 0;  // unknown for the roots: push
  bool first = true;
  while (n) {
    ms_ensure_levels(m, m->levels + 1);
    const unsigned long long* cur = m->newbits.p + (m->levels - 1) * n;
    unsigned long long* nxt = m->newbits.p + m->levels * n;
    VGX_HIP(hipMemsetAsync(nxt, 0, n * 8, s));
    VGX_HIP(hipMemsetAsync(m->d_ctr, 0, sizeof(MsCounters), s));
    const bool push = first || static_cast<double>(frontier_arcs) * m->alpha < arcs;
    if (push) {
      k_ms_push_short<<<bshort, kMsBlock, 0, s>>>(out.off.p, out.adj.p, nn, cur, m->seen.p, nxt);
      VGX_LAUNCH_CHECK();
      if (po.nchunks) {
        k_ms_push_long<<<bout, kMsBlock, 0, s>>>(out.off.p, out.adj.p, po.row_chunk.p, po.chunk_row.p,
                                                 static_cast<uint32_t>(po.nchunks), cur, m->seen.p, nxt);
        VGX_LAUNCH_CHECK();
      }
      m->push_levels += 1;
    } else {
      k_ms_pull_short<<<bshort, kMsBlock, 0, s>>>(in.off.p, in.adj.p, nn, all, cur, m->seen.p, nxt);
      VGX_LAUNCH_CHECK();
      if (pi.nchunks) {
        k_ms_pull_long<<<bin, kMsBlock, 0, s>>>(in.off.p, in.adj.p, pi.row_chunk.p, pi.chunk_row.p,
                                                static_cast<uint32_t>(pi.nchunks), all, cur, m->seen.p, nxt);
        VGX_LAUNCH_CHECK();
      }
    }
    k_ms_finalize<<<ms_blocks(n, kMsBlock, 1024), kMsBlock, 0, s>>>(out.off.p, nn, m->seen.p, nxt, m->d_ctr);
    VGX_LAUNCH_CHECK();
    VGX_HIP(hipMemcpyAsync(m->h_ctr, m->d_ctr, sizeof(MsCounters), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    first = false;
    if (m->h_ctr->pairs() == 0) break;
    frontier_arcs = m->h_ctr->arcs();
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

vgx_status vgx_msbfs_set_tuning(vgx_msbfs* m, uint32_t blocks, double alpha) {
  VGX_API_BEGIN
  if (!m) throw Error(VGX_INVALID_ARGUMENT, "null msbfs");
  if (blocks > 65535) throw Error(VGX_INVALID_ARGUMENT, "too many workgroups");
  if (!(alpha > 0.0)) throw Error(VGX_INVALID_ARGUMENT, "alpha must be positive");
  m->blocks = blocks;
  m->alpha = alpha;
  VGX_API_END
}

vgx_status vgx_msbfs_autotune(vgx_msbfs* m, uint32_t reps, uint32_t* best_blocks, double* best_alpha,
                              double* best_ms) {
  VGX_API_BEGIN
  if (!m) throw Error(VGX_INVALID_ARGUMENT, "null msbfs");
  if (reps == 0) reps = 1;
  VGX_HIP(hipSetDevice(m->g->device));
  auto time_run = [&]() {
    ms_run(m);  // allocations and plans settle
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t r = 0; r < reps; ++r) ms_run(m);
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
  };
  double best_t = time_run();
  uint32_t bb = m->blocks;
  double ba = m->alpha;
  for (double a : {2.0, 4.0, 8.0, 16.0, 32.0}) {
    m->alpha = a;
    const double t = time_run();
    if (t < best_t) {
      best_t = t;
      ba = a;
    }
  }
  m->alpha = ba;
  for (uint32_t b : {0u, 512u, 1024u, 2048u, 4096u, 8192u}) {
    m->blocks = b;
    const double t = time_run();
    if (t < best_t) {
      best_t = t;
      bb = b;
    }
  }
  m->blocks = bb;
  ms_run(m);
  if (best_blocks) *best_blocks = bb;
  if (best_alpha) *best_alpha = ba;
  if (best_ms) *best_ms = best_t;
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
