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
// Each level runs in one of two directions over the graph's live row tables
// (sorted rows, deleted entries tombstoned, so an update batch needs no
// repacked copy):
//   push: every frontier vertex u ORs its new bits, minus what the target
//         has seen, into next[v] of its out-neighbours (64-bit atomicOr);
//   pull: every unfinished vertex ORs the frontier masks of its
//         in-neighbours and stops once every root is accounted for.
// Push is used while the frontier's out-arcs are fewer than 1/alpha of all
// arcs.  Rows of up to 32 arcs are handled by one lane (64 consecutive
// vertices per wave); longer rows are cut into chunks of 1024 arcs, one wave
// each, so R-MAT hubs (which cluster at low ids) spread over many waves.
// OR is commutative and idempotent, so the result does not depend on the
// order of the atomics.
//
// The pass is a fixed program of launches that the device steers by
// itself: every level runs push/pull kernels that return at once unless the
// level's direction (decided on the device at the end of the previous level)
// is theirs, a finalize kernel that folds the level into seen and counts new
// pairs and next-frontier arcs, and a one-wave decide kernel that picks the
// next direction or marks the pass done.  Kernel arguments are one device
// block rewritten before each pass, so the same launch program can be
// captured once in a hipGraph and replayed for every pass while the graph
// changes underneath (only its size class forces a new capture).
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
constexpr uint32_t kMsStripes = 32;
// Levels one launch program covers; deeper traversals continue with
// further programs from where the previous one stopped.
constexpr uint32_t kMsProgram = 16;

// Level counters, striped over kMsStripes cache lines (workgroup b adds to
// stripe b % kMsStripes; the decide kernel sums the stripes).
struct alignas(128) MsStripe {
  unsigned long long new_pairs;      // (vertex, root) pairs reached this level
  unsigned long long frontier_arcs;  // out-arcs of the vertices reached this level
};
struct MsLevel {
  MsStripe s[kMsStripes];
};

// Device-side control of one pass.
struct MsCtl {
  uint32_t dir[kMsProgram + 1];  // per program level: 0 push, 1 pull
  uint32_t done;                 // set once a level found nothing new
  uint32_t levels;               // stored levels (valid once done)
  uint32_t push_levels;
  uint32_t pad;
  MsLevel ctr[kMsProgram];
};

// Everything a level kernel needs, in device memory (rewritten per pass).
// The passes read the graph's live row tables and arenas (rows sorted,
// deleted entries tombstoned), so an update batch needs no repacked copy.
struct MsArgs {
  const uint2* out_row;
  const uint32_t* out_arena;
  const uint32_t* out_deg;  // live out-degrees
  const uint32_t* out_rc;   // long-row chunk plan of the out-rows
  const uint32_t* out_cr;
  const uint2* in_row;
  const uint32_t* in_arena;
  const uint32_t* in_rc;
  const uint32_t* in_cr;
  uint32_t out_nch, in_nch;
  unsigned long long* seen;
  unsigned long long* newbits;  // level l of this program at newbits + (base + l) * n
  const uint32_t* roots;
  uint32_t n, k;
  uint32_t base;  // absolute level of program level 0
  uint32_t pad;
  unsigned long long all;
  double arcs;
  double alpha;
  MsCtl* ctl;
};

// Chunks of the rows longer than kMsShortRow in one direction's row table:
// row v owns chunks [row_chunk[v], row_chunk[v+1]).
struct MsPlan {
  DevBuf<uint32_t> row_chunk;
  DevBuf<uint32_t> chunk_row;
  DevBuf<uint32_t> cnt;  // chunks per row (scratch of the rebuild)
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
  vgx::DevBuf<uint32_t> droots;
  vgx::MsArgs* d_args = nullptr;
  vgx::MsCtl* d_ctl = nullptr;
  vgx::MsCtl* h_ctl = nullptr;  // pinned
  std::vector<uint64_t> reached;  // per root
  uint32_t blocks = 0;            // short-row kernel workgroups, 0 = auto
  double alpha = 4.0;             // push while frontier arcs < arcs / alpha
  vgx::MsPlan plan[2];            // [0] out-rows (push), [1] in-rows (pull)
  vgx::DevBuf<uint8_t> scan;
  uint32_t push_levels = 0;
  // captured launch program
  bool use_graph = false;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  uint64_t graph_key[5] = {0, 0, 0, 0, 0};
  uint64_t replays = 0;
};

namespace vgx {

__device__ __forceinline__ unsigned long long* level_ptr(const MsArgs& a, uint32_t l) {
  return a.newbits + static_cast<uint64_t>(a.base + l) * a.n;
}

// Program start: at absolute level 0 the roots' bits; later programs
// continue from the last level of the previous one.  Clears the level
// counters and the first `next` array, and sets the first direction.
__global__ void k_ms_start(const MsArgs* __restrict__ pa, int fresh, uint32_t first_dir) {
  const MsArgs a = *pa;
  const uint64_t n = a.n;
  unsigned long long* nb0 = level_ptr(a, 0);
  unsigned long long* nb1 = level_ptr(a, 1);
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (fresh) {
      unsigned long long m = 0;
      for (uint32_t r = 0; r < a.k; ++r)
        if (a.roots[r] == v) m |= 1ull << r;
      a.seen[v] = m;
      nb0[v] = m;
    }
    nb1[v] = 0;
  }
  if (blockIdx.x == 0) {
    MsCtl* c = a.ctl;
    for (uint32_t i = threadIdx.x; i < kMsProgram * kMsStripes; i += blockDim.x) {
      c->ctr[i / kMsStripes].s[i % kMsStripes].new_pairs = 0;
      c->ctr[i / kMsStripes].s[i % kMsStripes].frontier_arcs = 0;
    }
    if (threadIdx.x == 0) {
      c->dir[0] = first_dir;
      c->done = 0;
      c->levels = 0;
      c->push_levels = 0;
    }
  }
}

__global__ void k_ms_chunk_count(const uint2* __restrict__ row, uint32_t n, uint32_t* __restrict__ cnt) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t len = row[v].y;
    cnt[v] = len > kMsShortRow ? (len + kMsChunk - 1) / kMsChunk : 0u;
  }
}

__global__ void k_ms_chunk_fill(const uint32_t* __restrict__ row_chunk, uint32_t n, uint32_t* __restrict__ chunk_row) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    for (uint32_t c = row_chunk[v]; c < row_chunk[v + 1]; ++c) chunk_row[c] = v;
}

// Whether program level l runs in direction `dir`.
__device__ __forceinline__ bool level_runs(const MsArgs& a, uint32_t l, uint32_t dir) {
  const MsCtl* c = a.ctl;
  return !c->done && c->dir[l] == dir;
}

// push, short rows: lane per frontier vertex u.
__global__ __launch_bounds__(kMsBlock) void k_ms_push_short(const MsArgs* __restrict__ pa, uint32_t l) {
  const MsArgs a = *pa;
  if (!level_runs(a, l, 0)) return;
  const unsigned long long* frontier = level_ptr(a, l);
  unsigned long long* next = level_ptr(a, l + 1);
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < a.n; u += gridDim.x * blockDim.x) {
    const unsigned long long f = frontier[u];
    if (!f) continue;
    const uint2 rb = a.out_row[u];
    if (rb.y > kMsShortRow) continue;
    for (uint32_t i = rb.x; i < rb.x + rb.y; ++i) {
      const uint32_t v = a.out_arena[i];
      if (v & kTomb) continue;
      const unsigned long long m = f & ~a.seen[v];
      // bits only ever get added to next[v]: skip the atomic when they are there
      if (m & ~next[v]) atomicOr(&next[v], m);
    }
  }
}

// push, long rows: wave per chunk, lane per arc.
__global__ __launch_bounds__(kMsBlock) void k_ms_push_long(const MsArgs* __restrict__ pa, uint32_t l) {
  const MsArgs a = *pa;
  if (!level_runs(a, l, 0)) return;
  const unsigned long long* frontier = level_ptr(a, l);
  unsigned long long* next = level_ptr(a, l + 1);
  for (uint32_t c = blockIdx.x * kMsWaves + threadIdx.x / kWave; c < a.out_nch; c += gridDim.x * kMsWaves) {
    const uint32_t u = a.out_cr[c];
    const unsigned long long f = frontier[u];
    if (!f) continue;
    const uint2 rb = a.out_row[u];
    const uint32_t b = rb.x + (c - a.out_rc[u]) * kMsChunk;
    const uint32_t e = min(b + kMsChunk, rb.x + rb.y);
    for (uint32_t i = b + lane_id(); i < e; i += kWave) {
      const uint32_t v = a.out_arena[i];
      if (v & kTomb) continue;
      const unsigned long long m = f & ~a.seen[v];
      if (m & ~next[v]) atomicOr(&next[v], m);
    }
  }
}

// Frontier mask of parent entry x (nothing for a deleted entry, whose id
// bits still name a vertex).
__device__ __forceinline__ unsigned long long in_mask(const unsigned long long* frontier, uint32_t x) {
  const unsigned long long f = frontier[x & kIdMask];
  return (x & kTomb) ? 0ull : f;
}

// pull, short rows: lane per unfinished vertex v, early exit once every
// root is accounted for.
__global__ __launch_bounds__(kMsBlock) void k_ms_pull_short(const MsArgs* __restrict__ pa, uint32_t l) {
  const MsArgs a = *pa;
  if (!level_runs(a, l, 1)) return;
  const unsigned long long* frontier = level_ptr(a, l);
  unsigned long long* next = level_ptr(a, l + 1);
  const unsigned long long all = a.all;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += gridDim.x * blockDim.x) {
    const unsigned long long s = a.seen[v];
    if (s == all) continue;
    const uint2 rb = a.in_row[v];
    if (rb.y > kMsShortRow) continue;
    const uint32_t* adj = a.in_arena;
    const uint32_t b = rb.x, e = rb.x + rb.y;
    unsigned long long acc = 0;
    uint32_t i = b;
    for (; i + 4 <= e; i += 4) {
      acc |= in_mask(frontier, adj[i]) | in_mask(frontier, adj[i + 1]) | in_mask(frontier, adj[i + 2]) |
             in_mask(frontier, adj[i + 3]);
      if ((acc | s) == all) break;
    }
    if ((acc | s) != all)
      for (; i < e; ++i) acc |= in_mask(frontier, adj[i]);
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
__global__ __launch_bounds__(kMsBlock) void k_ms_pull_long(const MsArgs* __restrict__ pa, uint32_t l) {
  const MsArgs a = *pa;
  if (!level_runs(a, l, 1)) return;
  const unsigned long long* frontier = level_ptr(a, l);
  unsigned long long* next = level_ptr(a, l + 1);
  const unsigned long long all = a.all;
  for (uint32_t c = blockIdx.x * kMsWaves + threadIdx.x / kWave; c < a.in_nch; c += gridDim.x * kMsWaves) {
    const uint32_t v = a.in_cr[c];
    const unsigned long long s = a.seen[v];
    if (s == all) continue;
    const uint2 rb = a.in_row[v];
    const uint32_t b = rb.x + (c - a.in_rc[v]) * kMsChunk;
    const uint32_t e = min(b + kMsChunk, rb.x + rb.y);
    unsigned long long acc = 0;
    for (uint32_t i0 = b; i0 < e; i0 += 4 * kWave) {
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t i = i0 + j * kWave + lane_id();
        if (i < e) acc |= in_mask(frontier, a.in_arena[i]);
      }
      acc = wave_or(acc);
      if ((acc | s) == all) break;
    }
    const unsigned long long m = acc & ~s;
    if (lane_id() == 0 && m) atomicOr(&next[v], m);
  }
}

// Fold level l into seen; count new pairs and the next frontier's arcs;
// clear the array the next level will fill.
__global__ __launch_bounds__(kMsBlock) void k_ms_finalize(const MsArgs* __restrict__ pa, uint32_t l) {
  const MsArgs a = *pa;
  if (a.ctl->done) return;
  const unsigned long long* next = level_ptr(a, l + 1);
  unsigned long long* after = level_ptr(a, l + 2);
  unsigned long long pairs = 0, arcs = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += gridDim.x * blockDim.x) {
    after[v] = 0;
    const unsigned long long nw = next[v];
    if (!nw) continue;
    a.seen[v] |= nw;
    pairs += __popcll(nw);
    arcs += a.out_deg[v];
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
    MsStripe& st = a.ctl->ctr[l].s[blockIdx.x % kMsStripes];
    if (pairs) atomicAdd(&st.new_pairs, pairs);
    if (arcs) atomicAdd(&st.frontier_arcs, arcs);
  }
}

// One wave: sum level l's stripes; finish the pass or set the direction of
// level l + 1.
__global__ void k_ms_decide(const MsArgs* __restrict__ pa, uint32_t l) {
  const MsArgs a = *pa;
  MsCtl* c = a.ctl;
  if (c->done) return;
  const unsigned lane = lane_id();
  unsigned long long pairs = lane < kMsStripes ? c->ctr[l].s[lane].new_pairs : 0ull;
  unsigned long long arcs = lane < kMsStripes ? c->ctr[l].s[lane].frontier_arcs : 0ull;
  pairs = wave_sum(pairs);
  arcs = wave_sum(arcs);
  if (lane == 0) {
    if (c->dir[l] == 0) c->push_levels += 1;
    if (pairs == 0) {
      c->done = 1;
      c->levels = a.base + l + 1;
    } else {
      c->levels = a.base + l + 2;
      c->dir[l + 1] = static_cast<double>(arcs) * a.alpha < a.arcs ? 0u : 1u;
    }
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

// Grid of a chunk kernel: the next power of two of the waves it needs, so
// small changes of the chunk count keep the launch program unchanged.
static unsigned ms_chunk_blocks(uint64_t nchunks) {
  uint64_t b = 1;
  while (b * kMsWaves < nchunks && b < 8192) b <<= 1;
  return static_cast<unsigned>(b);
}

static void ms_ensure_levels(vgx_msbfs* m, uint64_t need) {
  if (need <= m->levels_cap) return;
  uint64_t cap = std::max<uint64_t>(need, m->levels_cap ? m->levels_cap * 2 : 32);
  m->newbits.grow_copy(cap * m->n + 1, m->levels_cap * m->n, m->g->stream);
  m->levels_cap = cap;
}

// Long-row chunks of one direction's row table, rebuilt when the graph
// changes (chunks cover stored entries, tombstones included).
static const MsPlan& ms_plan(vgx_msbfs* m, int dir, const Adjacency& adj, uint64_t version, uint64_t nv,
                             hipStream_t s) {
  MsPlan& pl = m->plan[dir];
  if (pl.version == version && pl.n == nv && pl.row_chunk.p) return pl;
  const uint32_t n = static_cast<uint32_t>(nv);
  pl.row_chunk.ensure(nv + 2);
  VGX_HIP(hipMemsetAsync(pl.row_chunk.p, 0, sizeof(uint32_t), s));
  uint32_t total = 0;
  if (n) {
    DevBuf<uint32_t>& cnt = pl.cnt;
    cnt.ensure(nv + 1);
    k_ms_chunk_count<<<ms_blocks(n), kMsBlock, 0, s>>>(adj.row.p, n, cnt.p);
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
  pl.version = version;
  pl.n = nv;
  return pl;
}

struct MsGrids {
  unsigned start, shrt, out_long, in_long, fin;
};

// The launch program of `prog` (at most kMsProgram) levels.  Only kernel
// launches: it is also what gets captured into the hipGraph.
static void ms_program(const vgx_msbfs* m, const MsGrids& gr, uint32_t prog, int fresh, uint32_t first_dir,
                       hipStream_t s) {
  const MsArgs* pa = m->d_args;
  k_ms_start<<<gr.start, kMsBlock, 0, s>>>(pa, fresh, first_dir);
  VGX_LAUNCH_CHECK();
  for (uint32_t l = 0; l < prog; ++l) {
    k_ms_push_short<<<gr.shrt, kMsBlock, 0, s>>>(pa, l);
    VGX_LAUNCH_CHECK();
    k_ms_push_long<<<gr.out_long, kMsBlock, 0, s>>>(pa, l);
    VGX_LAUNCH_CHECK();
    k_ms_pull_short<<<gr.shrt, kMsBlock, 0, s>>>(pa, l);
    VGX_LAUNCH_CHECK();
    k_ms_pull_long<<<gr.in_long, kMsBlock, 0, s>>>(pa, l);
    VGX_LAUNCH_CHECK();
    k_ms_finalize<<<gr.fin, kMsBlock, 0, s>>>(pa, l);
    VGX_LAUNCH_CHECK();
    k_ms_decide<<<1, kWave, 0, s>>>(pa, l);
    VGX_LAUNCH_CHECK();
  }
}

static void ms_drop_graph(vgx_msbfs* m) {
  if (m->exec) (void)hipGraphExecDestroy(m->exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->exec = nullptr;
  m->graph = nullptr;
}

static void ms_run(vgx_msbfs* m) {
  vgx_graph* g = m->g;
  hipStream_t s = g->stream;
  const Adjacency& out = g->adj[0];
  const Adjacency& in = g->in();
  const MsPlan& po = ms_plan(m, 0, out, g->rows_epoch, g->n, s);
  const MsPlan& pi = g->directed ? ms_plan(m, 1, in, g->rows_epoch, g->n, s) : po;  // undirected: one table, one plan
  const uint64_t n = g->n;
  if (n != m->n) {  // the per-level arrays are laid out with stride n
    m->levels_cap = 0;
    m->newbits.release();
    ms_drop_graph(m);
  }
  m->n = n;
  m->seen.ensure(n + 1);
  m->droots.ensure(m->k + 1);
  VGX_HIP(hipMemcpyAsync(m->droots.p, m->roots.data(), m->k * 4, hipMemcpyHostToDevice, s));
  // Levels per program: the previous pass's depth plus two (levels past
  // the end of a pass still cost a launch each; the slack keeps a pass one
  // level deeper within one program), between 4 and kMsProgram; a deeper
  // pass continues with another program.
  const uint32_t prog =
      m->levels ? std::min<uint32_t>(kMsProgram, std::max<uint32_t>(4, m->levels + 2)) : kMsProgram;
  m->levels = 0;
  m->push_levels = 0;
  if (n == 0) {
    m->ran = true;
    return;
  }
  MsGrids gr;
  gr.start = ms_blocks(n, kMsBlock, 4096);
  gr.shrt = m->blocks ? m->blocks : ms_blocks(n, kMsBlock, 4096);
  gr.out_long = ms_chunk_blocks(po.nchunks);
  gr.in_long = ms_chunk_blocks(pi.nchunks);
  gr.fin = ms_blocks(n, kMsBlock, 1024);
  // Programs of `prog` levels; the first starts from the roots, the next
  // ones from the last level of the previous program.
  uint32_t base = 0;
  uint32_t first_dir = 0;
  for (;;) {
    ms_ensure_levels(m, base + prog + 2);
    MsArgs a{};
    a.out_row = out.row.p;
    a.out_arena = out.arena.p;
    a.out_deg = out.deg.p;
    a.out_rc = po.row_chunk.p;
    a.out_cr = po.chunk_row.p;
    a.out_nch = static_cast<uint32_t>(po.nchunks);
    a.in_row = in.row.p;
    a.in_arena = in.arena.p;
    a.in_rc = pi.row_chunk.p;
    a.in_cr = pi.chunk_row.p;
    a.in_nch = static_cast<uint32_t>(pi.nchunks);
    a.seen = m->seen.p;
    a.newbits = m->newbits.p;
    a.roots = m->droots.p;
    a.n = static_cast<uint32_t>(n);
    a.k = m->k;
    a.base = base;
    a.all = m->k == 64 ? ~0ull : ((1ull << m->k) - 1);
    a.arcs = static_cast<double>(out.live);
    a.alpha = m->alpha;
    a.ctl = m->d_ctl;
    VGX_HIP(hipMemcpyAsync(m->d_args, &a, sizeof(a), hipMemcpyHostToDevice, s));
    const bool fresh = base == 0;
    if (m->use_graph && fresh) {
      const uint64_t key[5] = {n, (static_cast<uint64_t>(gr.shrt) << 32) | gr.fin,
                               (static_cast<uint64_t>(gr.out_long) << 32) | gr.in_long, gr.start, prog};
      if (!m->exec || !std::equal(key, key + 5, m->graph_key)) {
        ms_drop_graph(m);
        VGX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        ms_program(m, gr, prog, 1, 0, s);
        VGX_HIP(hipStreamEndCapture(s, &m->graph));
        VGX_HIP(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
        std::copy(key, key + 5, m->graph_key);
      }
      VGX_HIP(hipGraphLaunch(m->exec, s));
      m->replays += 1;
    } else {
      ms_program(m, gr, prog, fresh ? 1 : 0, first_dir, s);
    }
    VGX_HIP(hipMemcpyAsync(m->h_ctl, m->d_ctl, sizeof(MsCtl), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    m->push_levels += m->h_ctl->push_levels;
    if (m->h_ctl->done) {
      m->levels = m->h_ctl->levels;
      break;
    }
    // deeper than one program: continue from the last level written
    first_dir = m->h_ctl->dir[prog];
    base += prog;
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
    m->d_args = static_cast<MsArgs*>(dev_alloc(sizeof(MsArgs)));
    m->d_ctl = static_cast<MsCtl*>(dev_alloc(sizeof(MsCtl)));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&m->h_ctl), sizeof(MsCtl), hipHostMallocDefault));
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
  ms_drop_graph(m);
  if (m->d_args) dev_free(m->d_args);
  if (m->d_ctl) dev_free(m->d_ctl);
  if (m->h_ctl) (void)hipHostFree(m->h_ctl);
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

vgx_status vgx_msbfs_set_graph(vgx_msbfs* m, int on) {
  VGX_API_BEGIN
  if (!m) throw Error(VGX_INVALID_ARGUMENT, "null msbfs");
  m->use_graph = on != 0;
  if (!m->use_graph) ms_drop_graph(m);
  VGX_API_END
}

vgx_status vgx_msbfs_stats(vgx_msbfs* m, uint32_t* push_levels, uint64_t* graph_replays) {
  VGX_API_BEGIN
  if (!m) throw Error(VGX_INVALID_ARGUMENT, "null msbfs");
  if (push_levels) *push_levels = m->push_levels;
  if (graph_replays) *graph_replays = m->replays;
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
    ms_run(m);  // allocations, plans and the captured program settle
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
