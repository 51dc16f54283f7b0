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
// itself: every level runs one level kernel that takes the direction the
// previous level decided on the device (its workgroups split between short
// rows and long-row chunks), a finalize kernel that folds the level into
// seen and counts new pairs and next-frontier arcs, and a one-wave decide
// kernel that picks the next direction or marks the pass done.  Kernel arguments are one device
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
// Defaults of the tunable degree buckets: rows of up to kMsShortRow arcs are
// handled by one lane, longer rows in chunks of kMsChunk arcs, one wave each.
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
  double beta;          // a pulling pass turns back to push only when new pairs * beta < n * k (0: off)
  uint32_t short_row;   // rows of at most this many arcs: one lane
  uint32_t chunk;       // longer rows: chunks of this many arcs, one wave each
  MsCtl* ctl;
};

// Incremental repair of all roots after an update batch.  Distances are kept
// vertex-major, d[v * 64 + r] (kInf = unreachable), so a wave handles one
// vertex with one lane per root and every neighbour costs one coalesced
// 256-byte load.  Per-vertex 64-bit root masks carry the state of the
// invalidation (cand, aff, tm, sup, nm) and of the propagation (pend, pm).
// Lists come in two parities: round i fills parity i & 1 and its first
// kernel clears the counters of parity (i + 1) & 1, which the previous round
// has finished with.  Plain lists hold vertices; packed lists (vertex + edge
// prefix, see wave_push) drive the edge-balanced loops.
struct MsrState {
  unsigned long long tq[2];  // plain: vertices given candidate roots (invalidation input)
  unsigned long long vq[2];  // plain: of those, the ones with roots left to evaluate
  unsigned long long sq[2];  // packed: of those, the ones with in-arcs (support scan of the row heads)
  unsigned long long rq;     // packed: vertices still unsupported after the head scan (rest of the rows)
  unsigned long long eq[2];  // packed: newly affected vertices with out-arcs (emit)
  unsigned long long aall;   // plain: affected vertices (any root)
  unsigned long long apairs; // affected (vertex, root) pairs
  unsigned long long seedq;  // packed: affected vertices with in-arcs
  unsigned long long fq[2];  // plain: vertices with lowered roots (propagation input)
  unsigned long long pq[2];  // packed: of those, the ones with out-arcs
  unsigned long long candidates;  // (vertex, root) candidate pairs
  unsigned long long lowered;     // pairs lowered by seeds, additions and propagation
  uint32_t fallback;
  uint32_t pad;
};

struct MsrArgs {
  const uint2* out_row;
  const uint32_t* out_arena;
  const uint2* in_row;
  const uint32_t* in_arena;
  uint32_t* d;  // [n][64]
  unsigned long long* cand;
  unsigned long long* aff;
  unsigned long long* tm;
  unsigned long long* sup;
  unsigned long long* nm;
  unsigned long long* pend;
  unsigned long long* pm;
  uint32_t* tl[2];
  uint32_t* vl[2];
  uint32_t* sl[2];
  unsigned long long* spre[2];
  uint32_t* rl;
  unsigned long long* rpre;
  uint32_t* el[2];
  unsigned long long* epre[2];
  uint32_t* al;
  uint32_t* seed;
  unsigned long long* seedpre;
  uint32_t* fl[2];
  uint32_t* pl[2];
  unsigned long long* ppre[2];
  // the staged batch (out-direction arcs, mirrored when undirected)
  const uint32_t* fin;
  const unsigned long long* skeys;
  const uint8_t* act;
  const unsigned long long* nfin;  // count of the staged batch's final arcs (device)
  const uint32_t* roots;
  uint32_t k;
  uint32_t pad;
  unsigned long long limit;  // affected pairs past which the repair gives up
  MsrState* st;
};

struct MsRepair {
  DevBuf<uint32_t> d;
  DevBuf<unsigned long long> cand, aff, tm, sup, nm, pend, pm;
  DevBuf<uint32_t> tl[2], vl[2], sl[2], el[2], al, seed, fl[2], pl[2];
  DevBuf<unsigned long long> spre[2], epre[2], seedpre, ppre[2];
  DevBuf<uint32_t> rl;
  DevBuf<unsigned long long> rpre;
  uint64_t n = 0;          // vertices the arrays cover
  bool setup = false;      // the current batch allocated arrays or captured a program
  MsrArgs* d_args = nullptr;
  MsrArgs* h_args = nullptr;  // pinned staging of the argument block
  MsrState* d_st = nullptr;
  MsrState* h_st = nullptr;  // pinned
  hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
};

// Chunks of the rows longer than the short-row bound in one direction's row
// table: row v owns chunks [row_chunk[v], row_chunk[v+1]).
struct MsPlan {
  DevBuf<uint32_t> row_chunk;
  DevBuf<uint32_t> chunk_row;
  DevBuf<uint32_t> cnt;  // chunks per row (scratch of the rebuild)
  uint64_t nchunks = 0;
  uint64_t version = ~0ull;
  uint64_t n = 0;
  uint32_t short_row = 0, chunk = 0;
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
  double beta = 0.0;              // pull -> push also needs new pairs * beta < n * k (0: off)
  uint32_t short_row = vgx::kMsShortRow;
  uint32_t chunk = vgx::kMsChunk;
  vgx::MsPlan plan[2];            // [0] out-rows (push), [1] in-rows (pull)
  vgx::DevBuf<uint8_t> scan;
  uint32_t push_levels = 0;
  // captured launch program
  bool use_graph = false;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  uint64_t graph_key[5] = {0, 0, 0, 0, 0};
  uint64_t replays = 0;
  // incremental repair
  vgx::MsRepair rep;
  bool in_d32 = false;       // the current distances are rep.d (not newbits)
  uint64_t dist_version = ~0ull;  // graph version the distances belong to
  int repair_mode = 0;       // 0 adaptive, 1 always repair, 2 always a full pass
  double fallback = 0.1;     // affected pairs past fallback * n * k give up
  // running means of the two arms, answer-ready, graph update included: the
  // full pass, and the repair per batch-size class (floor(log2 ops))
  double pass_ms = 0.0;
  double repair_ms[48] = {};
  uint32_t passes_in_class[48] = {};  // full passes chosen in a row per class (a repair is retried every 16th)
  double repair_us_per_op = 0.0;
  uint64_t batches = 0, repaired = 0, full_passes = 0, fallbacks = 0;
  bool last_repaired = false, last_fallback = false;
  uint64_t last_affected_pairs = 0, last_lowered = 0, last_candidates = 0;
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

__global__ void k_ms_chunk_count(const uint2* __restrict__ row, uint32_t n, uint32_t short_row, uint32_t chunk,
                                 uint32_t* __restrict__ cnt) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t len = row[v].y;
    cnt[v] = len > short_row ? (len + chunk - 1) / chunk : 0u;
  }
}

__global__ void k_ms_chunk_fill(const uint32_t* __restrict__ row_chunk, uint32_t n, uint32_t* __restrict__ chunk_row) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    for (uint32_t c = row_chunk[v]; c < row_chunk[v + 1]; ++c) chunk_row[c] = v;
}

// push, short rows: lane per frontier vertex u.
__device__ __forceinline__ void ms_push_short(const MsArgs& a, uint32_t l, uint32_t blk, uint32_t nblk) {
  const unsigned long long* frontier = level_ptr(a, l);
  unsigned long long* next = level_ptr(a, l + 1);
  for (uint32_t u = blk * blockDim.x + threadIdx.x; u < a.n; u += nblk * blockDim.x) {
    const unsigned long long f = frontier[u];
    if (!f) continue;
    const uint2 rb = a.out_row[u];
    if (rb.y > a.short_row) continue;
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
__device__ __forceinline__ void ms_push_long(const MsArgs& a, uint32_t l, uint32_t blk, uint32_t nblk) {
  const unsigned long long* frontier = level_ptr(a, l);
  unsigned long long* next = level_ptr(a, l + 1);
  for (uint32_t c = blk * kMsWaves + threadIdx.x / kWave; c < a.out_nch; c += nblk * kMsWaves) {
    const uint32_t u = a.out_cr[c];
    const unsigned long long f = frontier[u];
    if (!f) continue;
    const uint2 rb = a.out_row[u];
    const uint32_t b = rb.x + (c - a.out_rc[u]) * a.chunk;
    const uint32_t e = min(b + a.chunk, rb.x + rb.y);
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
__device__ __forceinline__ void ms_pull_short(const MsArgs& a, uint32_t l, uint32_t blk, uint32_t nblk) {
  const unsigned long long* frontier = level_ptr(a, l);
  unsigned long long* next = level_ptr(a, l + 1);
  const unsigned long long all = a.all;
  for (uint32_t v = blk * blockDim.x + threadIdx.x; v < a.n; v += nblk * blockDim.x) {
    const unsigned long long s = a.seen[v];
    if (s == all) continue;
    const uint2 rb = a.in_row[v];
    if (rb.y > a.short_row) continue;
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
__device__ __forceinline__ void ms_pull_long(const MsArgs& a, uint32_t l, uint32_t blk, uint32_t nblk) {
  const unsigned long long* frontier = level_ptr(a, l);
  unsigned long long* next = level_ptr(a, l + 1);
  const unsigned long long all = a.all;
  for (uint32_t c = blk * kMsWaves + threadIdx.x / kWave; c < a.in_nch; c += nblk * kMsWaves) {
    const uint32_t v = a.in_cr[c];
    const unsigned long long s = a.seen[v];
    if (s == all) continue;
    const uint2 rb = a.in_row[v];
    const uint32_t b = rb.x + (c - a.in_rc[v]) * a.chunk;
    const uint32_t e = min(b + a.chunk, rb.x + rb.y);
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

// One level in the direction the previous level's decide kernel set: the
// first `shrt` workgroups take the rows of at most short_row arcs, the
// others the long-row chunks of that direction (one launch per level in
// place of a push and a pull pair, of which one would exit at once).
__global__ __launch_bounds__(kMsBlock) void k_ms_level(const MsArgs* __restrict__ pa, uint32_t l, uint32_t shrt,
                                                       uint32_t out_long, uint32_t in_long) {
  const MsArgs a = *pa;
  const MsCtl* c = a.ctl;
  if (c->done) return;
  const bool pull = c->dir[l] != 0;
  const uint32_t b = blockIdx.x;
  if (b < shrt) {
    if (pull)
      ms_pull_short(a, l, b, shrt);
    else
      ms_push_short(a, l, b, shrt);
    return;
  }
  const uint32_t lb = b - shrt;
  if (pull) {
    if (lb < in_long) ms_pull_long(a, l, lb, in_long);
  } else if (lb < out_long) {
    ms_push_long(a, l, lb, out_long);
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
      bool push = static_cast<double>(arcs) * a.alpha < a.arcs;
      // leaving pull also needs the frontier small next to all (vertex, root) pairs
      if (push && c->dir[l] == 1 && a.beta > 0.0)
        push = static_cast<double>(pairs) * a.beta < static_cast<double>(a.n) * a.k;
      c->dir[l + 1] = push ? 0u : 1u;
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
  if (pl.version == version && pl.n == nv && pl.row_chunk.p && pl.short_row == m->short_row &&
      pl.chunk == m->chunk)
    return pl;
  const uint32_t n = static_cast<uint32_t>(nv);
  pl.row_chunk.ensure(nv + 2);
  VGX_HIP(hipMemsetAsync(pl.row_chunk.p, 0, sizeof(uint32_t), s));
  uint32_t total = 0;
  if (n) {
    DevBuf<uint32_t>& cnt = pl.cnt;
    cnt.ensure(nv + 1);
    k_ms_chunk_count<<<ms_blocks(n), kMsBlock, 0, s>>>(adj.row.p, n, m->short_row, m->chunk, cnt.p);
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
  pl.short_row = m->short_row;
  pl.chunk = m->chunk;
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
    k_ms_level<<<gr.shrt + std::max(gr.out_long, gr.in_long), kMsBlock, 0, s>>>(pa, l, gr.shrt, gr.out_long,
                                                                                 gr.in_long);
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
  m->in_d32 = false;
  if (n == 0) {
    m->ran = true;
    m->dist_version = g->version;
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
    a.beta = m->beta;
    a.short_row = m->short_row;
    a.chunk = m->chunk;
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
  m->dist_version = g->version;
  m->reached.assign(m->k, 0);
}

// ------------------------------------------------------------------ repair
//
// Exact repair of every root after a batch, in the reference's order for one
// root (incremental/bfs.hpp:29-98, 206-354): on the graph without the
// deleted arcs, a (vertex, root) pair is affected once no live in-neighbour
// that is itself unaffected sits one level above it; affected pairs are
// reset, seeded from their unaffected in-neighbours in the post-batch graph,
// the added arcs are relaxed, and Bellman-Ford rounds propagate every
// lowered pair.  Lanes are roots: the masks say which lanes do work, and a
// 64-bit ballot turns a lane test back into a mask.

constexpr uint32_t kMsrWindow = 16;  // items a wave takes at a time from a plain list
// Edges a wave takes at least in the edge-balanced repair steps: each edge
// is a whole wave's work (one lane per root), so a small frontier is spread
// over eight-edge ranges, one group of eight pairs per wave, rather than
// 64-edge windows that leave most waves idle.
constexpr uint32_t kMsrQuantum = 8;
constexpr uint32_t kMsrHead = 64;    // in-arcs of a row the first support scan reads
constexpr unsigned kMsrBlocks = 1024;

__device__ __forceinline__ bool lane_bit(unsigned long long m) { return (m >> lane_id()) & 1ull; }

// A wave-uniform copy of x (taken from lane 0), so loads it indexes go
// through the scalar unit.
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Wave-aggregated append of v to a plain list (the lanes that want it).
__device__ __forceinline__ void push_plain(unsigned long long* ctr, uint32_t* list, bool want, uint32_t v) {
  const unsigned long long b = __ballot(want);
  if (!b) return;
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(ctr, static_cast<unsigned long long>(__popcll(b)));
  base = __shfl(base, 0, kWave);
  if (want) list[base + lanes_below(b)] = v;
}

// Grid-stride over windows of kMsrWindow items of a plain list: fn(first)
// with lane j < kMsrWindow holding item first + j (fn reads it itself).
template <class Fn>
__device__ __forceinline__ void msr_windows(uint64_t count, Fn&& fn) {
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  for (uint64_t w = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave;
       w * kMsrWindow < count; w += waves)
    fn(w * kMsrWindow);
}

// Eight pairs of a wave, as msr_groups8 loads them: the pair's lane t
// (-1: no pair), its ends a and b, the rows d[a][lane] and d[b][lane] and
// the masks ma[a] and mb[b] (0 when not asked for).
struct MsrGroup {
  int t[8];
  uint32_t a[8], b[8], da[8], db[8];
  unsigned long long ma[8], mb[8];
};

// Walks the pairs (a, b) a wave holds one per lane (b = kTomb: none) in
// groups of eight: the eight pairs' rows and masks are loaded together
// before group(g) runs, so a wave keeps sixteen row loads in flight
// instead of two.
template <class Group>
__device__ __forceinline__ void msr_groups8(const uint32_t* __restrict__ d, const unsigned long long* ma,
                                            const unsigned long long* mb, uint32_t av, uint32_t bv,
                                            Group&& group) {
  const unsigned lane = lane_id();
  unsigned long long todo = __ballot(!(bv & kTomb));
  while (todo) {
    MsrGroup g;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      g.t[i] = todo ? __ffsll(static_cast<long long>(todo)) - 1 : -1;
      todo = todo ? todo & (todo - 1) : 0ull;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int t = g.t[i] < 0 ? 0 : g.t[i];
      g.a[i] = uni(__shfl(av, t, kWave));
      g.b[i] = uni(__shfl(bv, t, kWave));
      const bool on = g.t[i] >= 0;
      g.da[i] = on ? d[g.a[i] * 64ull + lane] : kInf;
      g.db[i] = on ? d[g.b[i] * 64ull + lane] : kInf;
      g.ma[i] = on && ma ? ma[g.a[i]] : 0ull;
      g.mb[i] = on && mb ? mb[g.b[i]] : 0ull;
    }
    group(g);
  }
}

// The same with body(t, a, b, da, db, ma, mb) run for each pair in lane
// order.
template <class Body>
__device__ __forceinline__ void msr_pairs8(const uint32_t* __restrict__ d, const unsigned long long* ma,
                                           const unsigned long long* mb, uint32_t av, uint32_t bv, Body&& body) {
  msr_groups8(d, ma, mb, av, bv, [&](const MsrGroup& g) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (g.t[i] >= 0) body(g.t[i], g.a[i], g.b[i], g.da[i], g.db[i], g.ma[i], g.mb[i]);
  });
}

// Claims for the eight pairs of a group: lane 0 ORs bits[i] (wave-uniform)
// into claim[b[i]], the eight atomics issued before any result is read;
// lane t[i] learns whether its pair set the first bits of b[i] (then it
// lists b[i]: want, pv), and `count` adds every claimed bit.
__device__ __forceinline__ void msr_claim8(unsigned long long* claim, const MsrGroup& g,
                                           const unsigned long long (&bits)[8], bool& want, uint32_t& pv,
                                           unsigned long long& count) {
  const unsigned lane = lane_id();
  unsigned long long old[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    old[i] = 0ull;
    if (bits[i] && lane == 0) old[i] = atomicOr(&claim[g.b[i]], bits[i]);
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (!bits[i]) continue;
    count += __popcll(bits[i]);
    const unsigned long long o = __shfl(old[i], 0, kWave);
    if (!o && lane == static_cast<unsigned>(g.t[i])) {
      want = true;
      pv = g.b[i];
    }
  }
}

// Relaxations d_r(b) = min(d_r(b), d_r(a) + 1) of the pairs of a group on
// the roots in `roots` (per pair, lane mask) with d_r(a) + 1 < d_r(b): the
// eight atomics issued before any result is read; bits[i] = the roots
// whose distance the pair lowered.
template <class Roots>
__device__ __forceinline__ void msr_relax8(uint32_t* __restrict__ d, const MsrGroup& g, Roots&& roots,
                                           unsigned long long (&bits)[8]) {
  const unsigned lane = lane_id();
  bool c[8];
  uint32_t ret[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    c[i] = g.t[i] >= 0 && roots(i) && g.da[i] != kInf && g.da[i] + 1 < g.db[i];
    ret[i] = c[i] ? atomicMin(&d[g.b[i] * 64ull + lane], g.da[i] + 1) : 0u;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) bits[i] = g.t[i] >= 0 ? __ballot(c[i] && ret[i] > g.da[i] + 1) : 0ull;
}

// Heads of deleted arcs u->v that were shortest-parent arcs for some root:
// d_r(u) + 1 == d_r(v) (bfs.hpp:52-63).  They become candidates of round 0.
__global__ __launch_bounds__(kMsBlock) void k_msr_candidates(const MsrArgs* __restrict__ pa) {
  const MsrArgs& a = *pa;
  const unsigned lane = lane_id();
  const uint32_t myroot = lane < a.k ? a.roots[lane] : kInf;
  unsigned long long ncand = 0;
  const uint64_t nfin = *a.nfin;
  msr_windows(nfin, [&](uint64_t first) {
    const uint64_t j = first + lane;
    uint32_t u = 0, v = kTomb;
    if (lane < kMsrWindow && j < nfin) {
      const uint32_t p = a.fin[j];
      if (a.act[p] == kActDel) {
        u = static_cast<uint32_t>(a.skeys[p] >> 32);
        v = static_cast<uint32_t>(a.skeys[p]);
      }
    }
    bool want = false;
    uint32_t pv = 0;
    msr_groups8(a.d, nullptr, nullptr, u, v, [&](const MsrGroup& g) {
      unsigned long long bits[8];
#pragma unroll
      for (int i = 0; i < 8; ++i)
        bits[i] = g.t[i] >= 0 ? __ballot(g.da[i] != kInf && g.da[i] + 1 == g.db[i] && myroot != g.b[i]) : 0ull;
      msr_claim8(a.cand, g, bits, want, pv, ncand);
    });
    push_plain(&a.st->tq[0], a.tl[0], want, pv);
  });
  if (lane == 0 && ncand) atomicAdd(&a.st->candidates, ncand);
}

// Round `par`, step 1 (lane per vertex): take each touched vertex's
// candidate roots, minus those already affected and the roots it is; list
// the vertices with roots left (and, for the support scan, those with
// in-arcs).  Clears the other parity's counters.
__global__ __launch_bounds__(kMsBlock) void k_msr_take(const MsrArgs* __restrict__ pa, int par) {
  const MsrArgs& a = *pa;
  if (a.st->fallback) return;
  __shared__ uint32_t s_roots[64];
  if (threadIdx.x < 64) s_roots[threadIdx.x] = threadIdx.x < a.k ? a.roots[threadIdx.x] : kInf;
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->tq[par ^ 1] = 0;
    a.st->vq[par ^ 1] = 0;
    a.st->sq[par ^ 1] = 0;
    a.st->eq[par ^ 1] = 0;
    a.st->rq = 0;
  }
  const uint64_t nt = a.st->tq[par];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t j0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x - lane_id(); j0 < nt;
       j0 += stride) {
    const uint64_t j = j0 + lane_id();
    uint32_t v = 0, len = 0;
    bool keep = false;
    if (j < nt) {
      v = a.tl[par][j];
      unsigned long long m = atomicExch(&a.cand[v], 0ull) & ~a.aff[v];
      for (uint32_t r = 0; r < 64; ++r)
        if (s_roots[r] == v) m &= ~(1ull << r);
      if (m) {
        keep = true;
        a.tm[v] = m;
        a.sup[v] = 0;
        len = a.in_row[v].y;
      }
    }
    push_plain(&a.st->vq[par], a.vl[par], keep, v);
    wave_push(&a.st->sq[par], a.sl[par], a.spre[par], keep && len > 0, v, min(len, kMsrHead));
  }
}

// Step 1b (lane per vertex): the vertices the head scan left with roots
// unsupported scan the rest of their in-rows.
__global__ __launch_bounds__(kMsBlock) void k_msr_take_rest(const MsrArgs* __restrict__ pa, int par) {
  const MsrArgs& a = *pa;
  if (a.st->fallback) return;
  const uint64_t nv = a.st->vq[par];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t j0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x - lane_id(); j0 < nv;
       j0 += stride) {
    const uint64_t j = j0 + lane_id();
    uint32_t v = 0, len = 0;
    if (j < nv) {
      v = a.vl[par][j];
      if (a.tm[v] & ~a.sup[v]) len = a.in_row[v].y;
    }
    wave_push(&a.st->rq, a.rl, a.rpre, len > kMsrHead, v, len > kMsrHead ? len - kMsrHead : 0u);
  }
}

// Step 2 (edge-balanced over the in-arcs w->v of the listed vertices): root
// r of v is supported when some live w has d_r(w) + 1 == d_r(v) and is not
// affected for r.  Run over the first kMsrHead in-arcs of every row, then
// over the rest of the rows that still need it: shortest parents are found
// early in most rows, and a hub's row need not be read whole.
__global__ __launch_bounds__(kMsBlock) void k_msr_support(const MsrArgs* __restrict__ pa, int par, int rest) {
  const MsrArgs& a = *pa;
  if (a.st->fallback) return;
  const unsigned long long qc = rest ? a.st->rq : a.st->sq[par];
  const uint32_t* list = rest ? a.rl : a.sl[par];
  const unsigned long long* pre = rest ? a.rpre : a.spre[par];
  const uint32_t skip = rest ? kMsrHead : 0u;  // the rest of a row starts after its head
  const uint32_t W = gridDim.x * kMsWaves;
  const uint32_t gw = blockIdx.x * kMsWaves + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  lbs_edges<false, kMsrQuantum>(list, pre, static_cast<uint32_t>(qc >> kEdgeBits), qc & kEdgeMask, a.in_row, gw, W,
            [&](uint32_t, uint32_t v, uint2, uint32_t slot, bool valid) {
              // the roots of v still without support when its lane looked (a
              // vertex whose roots are all supported by then sends no loads);
              // a pair reads them from its lane instead of loading sup again
              // behind the previous pair's atomic
              const unsigned long long open = valid ? (a.tm[v] & ~a.sup[v]) : 0ull;
              const uint32_t w = open ? a.in_arena[slot + skip] : kTomb;
              msr_pairs8(a.d, nullptr, a.aff, v, w,
                         [&](int t, uint32_t vt, uint32_t, uint32_t dv, uint32_t dw, unsigned long long,
                             unsigned long long affw) {
                           const unsigned long long need = __shfl(open, t, kWave);
                           const unsigned long long bits =
                               __ballot(lane_bit(need) && dw != kInf && dw + 1 == dv && !lane_bit(affw));
                           if (bits && lane == 0) atomicOr(&a.sup[vt], bits);
                         });
            });
}

// Step 3 (lane per vertex): roots left without support become affected;
// the vertices with newly affected roots and out-arcs are listed for the
// emit step.
__global__ __launch_bounds__(kMsBlock) void k_msr_evaluate(const MsrArgs* __restrict__ pa, int par) {
  const MsrArgs& a = *pa;
  if (a.st->fallback) return;
  const uint64_t nv = a.st->vq[par];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  unsigned long long pairs = 0;
  for (uint64_t j0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x - lane_id(); j0 < nv;
       j0 += stride) {
    const uint64_t j = j0 + lane_id();
    uint32_t v = 0, len = 0;
    bool newly = false, first = false;
    if (j < nv) {
      v = a.vl[par][j];
      const unsigned long long nm = a.tm[v] & ~a.sup[v];
      if (nm) {
        const unsigned long long old = a.aff[v];
        a.aff[v] = old | nm;  // this round's only writer of aff[v]
        a.nm[v] = nm;
        first = old == 0;
        newly = true;
        pairs += __popcll(nm);
        len = a.out_row[v].y;
      }
    }
    push_plain(&a.st->aall, a.al, first, v);
    wave_push(&a.st->eq[par], a.el[par], a.epre[par], newly && len > 0, v, len);
  }
  pairs = wave_sum(pairs);
  if (lane_id() == 0 && pairs) atomicAdd(&a.st->apairs, pairs);
}

// Step 4 (edge-balanced over the out-arcs u->x of the newly affected): x
// loses a parent for root r when d_r(x) == d_r(u) + 1; it becomes a
// candidate of the next round.  Past the pair limit the repair stops.
__global__ __launch_bounds__(kMsBlock) void k_msr_emit(const MsrArgs* __restrict__ pa, int par) {
  const MsrArgs& a = *pa;
  if (a.st->fallback) return;
  if (a.st->apairs > a.limit) {
    if (blockIdx.x == 0 && threadIdx.x == 0) a.st->fallback = 1;
    return;
  }
  const unsigned long long qc = a.st->eq[par];
  const uint32_t W = gridDim.x * kMsWaves;
  const uint32_t gw = blockIdx.x * kMsWaves + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  lbs_edges<false, kMsrQuantum>(a.el[par], a.epre[par], static_cast<uint32_t>(qc >> kEdgeBits), qc & kEdgeMask, a.out_row, gw, W,
            [&](uint32_t, uint32_t u, uint2, uint32_t slot, bool valid) {
              const uint32_t x = valid ? a.out_arena[slot] : kTomb;
              bool want = false;
              uint32_t px = 0;
              unsigned long long claimed = 0;
              msr_groups8(a.d, a.nm, nullptr, u, x, [&](const MsrGroup& g) {
                unsigned long long bits[8];
#pragma unroll
                for (int i = 0; i < 8; ++i)
                  bits[i] = g.t[i] >= 0
                                ? __ballot(lane_bit(g.ma[i]) && g.da[i] != kInf && g.db[i] == g.da[i] + 1)
                                : 0ull;
                msr_claim8(a.cand, g, bits, want, px, claimed);
              });
              push_plain(&a.st->tq[par ^ 1], a.tl[par ^ 1], want, px);
            });
}

// Affected pairs become unreachable; affected vertices with in-arcs are
// listed for seeding.
__global__ __launch_bounds__(kMsBlock) void k_msr_reset(const MsrArgs* __restrict__ pa) {
  const MsrArgs& a = *pa;
  const unsigned lane = lane_id();
  msr_windows(a.st->aall, [&](uint64_t first) {
    const uint64_t na = a.st->aall;
    const uint64_t j = first + lane;
    uint32_t v = 0, len = 0;
    const bool in = lane < kMsrWindow && j < na;
    if (in) {
      v = a.al[j];
      len = a.in_row[v].y;
    }
    for (unsigned long long todo = __ballot(in); todo; todo &= todo - 1) {
      const int t = __ffsll(static_cast<long long>(todo)) - 1;
      const uint32_t vt = uni(__shfl(v, t, kWave));
      if (lane_bit(a.aff[vt])) a.d[vt * 64ull + lane] = kInf;
    }
    wave_push(&a.st->seedq, a.seed, a.seedpre, in && len > 0, v, len);
  });
}

// best_boundary_distance (bfs.hpp:261-269) for every affected pair, over
// the in-arcs w->v of the post-batch graph from pairs that are not affected.
__global__ __launch_bounds__(kMsBlock) void k_msr_seed(const MsrArgs* __restrict__ pa) {
  const MsrArgs& a = *pa;
  const unsigned long long qc = a.st->seedq;
  const uint32_t W = gridDim.x * kMsWaves;
  const uint32_t gw = blockIdx.x * kMsWaves + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  lbs_edges<false, kMsrQuantum>(a.seed, a.seedpre, static_cast<uint32_t>(qc >> kEdgeBits), qc & kEdgeMask, a.in_row, gw, W,
            [&](uint32_t, uint32_t v, uint2, uint32_t slot, bool valid) {
              const uint32_t w = valid ? a.in_arena[slot] : kTomb;
              msr_pairs8(a.d, a.aff, a.aff, v, w,
                         [&](int, uint32_t vt, uint32_t, uint32_t dv, uint32_t dw, unsigned long long affv,
                             unsigned long long affw) {
                           if (lane_bit(affv & ~affw) && dw != kInf && dw + 1 < dv)
                             atomicMin(&a.d[vt * 64ull + lane], dw + 1);
                         });
            });
}

// Affected pairs that found a distance start the propagation.
__global__ __launch_bounds__(kMsBlock) void k_msr_collect(const MsrArgs* __restrict__ pa) {
  const MsrArgs& a = *pa;
  const unsigned lane = lane_id();
  unsigned long long low = 0;
  msr_windows(a.st->aall, [&](uint64_t first) {
    const uint64_t na = a.st->aall;
    const uint64_t j = first + lane;
    uint32_t v = 0;
    const bool in = lane < kMsrWindow && j < na;
    if (in) v = a.al[j];
    bool want = false;
    uint32_t pv = 0;
    for (unsigned long long todo = __ballot(in); todo; todo &= todo - 1) {
      const int t = __ffsll(static_cast<long long>(todo)) - 1;
      const uint32_t vt = uni(__shfl(v, t, kWave));
      const unsigned long long bits = __ballot(lane_bit(a.aff[vt]) && a.d[vt * 64ull + lane] != kInf);
      if (!bits) continue;
      low += __popcll(bits);
      unsigned long long old = 0;
      if (lane == 0) old = atomicOr(&a.pend[vt], bits);
      old = __shfl(old, 0, kWave);
      if (!old && lane == static_cast<unsigned>(t)) {
        want = true;
        pv = vt;
      }
    }
    push_plain(&a.st->fq[0], a.fl[0], want, pv);
  });
  if (lane == 0 && low) atomicAdd(&a.st->lowered, low);
}

// relax_edge over the arcs the batch added (bfs.hpp:91-96, 332-339).
__global__ __launch_bounds__(kMsBlock) void k_msr_relax_additions(const MsrArgs* __restrict__ pa) {
  const MsrArgs& a = *pa;
  const unsigned lane = lane_id();
  unsigned long long low = 0;
  const uint64_t nfin = *a.nfin;
  msr_windows(nfin, [&](uint64_t first) {
    const uint64_t j = first + lane;
    uint32_t u = 0, v = kTomb;
    if (lane < kMsrWindow && j < nfin) {
      const uint32_t p = a.fin[j];
      if (a.act[p] == kActIns || a.act[p] == kActUndel) {
        u = static_cast<uint32_t>(a.skeys[p] >> 32);
        v = static_cast<uint32_t>(a.skeys[p]);
      }
    }
    bool want = false;
    uint32_t pv = 0;
    msr_groups8(a.d, nullptr, nullptr, u, v, [&](const MsrGroup& g) {
      unsigned long long bits[8];
      msr_relax8(a.d, g, [](int) { return true; }, bits);
      msr_claim8(a.pend, g, bits, want, pv, low);
    });
    push_plain(&a.st->fq[0], a.fl[0], want, pv);
  });
  if (lane == 0 && low) atomicAdd(&a.st->lowered, low);
}

// Propagation round `par`, step 1 (lane per vertex): take each listed
// vertex's lowered roots; list those with out-arcs.  Clears the other
// parity's counters.
__global__ __launch_bounds__(kMsBlock) void k_msr_ptake(const MsrArgs* __restrict__ pa, int par) {
  const MsrArgs& a = *pa;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->fq[par ^ 1] = 0;
    a.st->pq[par ^ 1] = 0;
  }
  const uint64_t nf = a.st->fq[par];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t j0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x - lane_id(); j0 < nf;
       j0 += stride) {
    const uint64_t j = j0 + lane_id();
    uint32_t v = 0, len = 0;
    bool keep = false;
    if (j < nf) {
      v = a.fl[par][j];
      const unsigned long long m = atomicExch(&a.pend[v], 0ull);
      a.pm[v] = m;
      len = a.out_row[v].y;
      keep = m != 0;
    }
    wave_push(&a.st->pq[par], a.pl[par], a.ppre[par], keep && len > 0, v, len);
  }
}

// Step 2 (edge-balanced over the out-arcs u->x): d_r(x) = min(d_r(x),
// d_r(u) + 1) for the taken roots of u (propagate_decreases, bfs.hpp
// :341-354); lowered pairs go to the next round.
__global__ __launch_bounds__(kMsBlock) void k_msr_propagate(const MsrArgs* __restrict__ pa, int par) {
  const MsrArgs& a = *pa;
  const unsigned long long qc = a.st->pq[par];
  const uint32_t W = gridDim.x * kMsWaves;
  const uint32_t gw = blockIdx.x * kMsWaves + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  unsigned long long low = 0;
  lbs_edges<false, kMsrQuantum>(a.pl[par], a.ppre[par], static_cast<uint32_t>(qc >> kEdgeBits), qc & kEdgeMask, a.out_row, gw, W,
            [&](uint32_t, uint32_t u, uint2, uint32_t slot, bool valid) {
              const uint32_t x = valid ? a.out_arena[slot] : kTomb;
              bool want = false;
              uint32_t px = 0;
              msr_groups8(a.d, a.pm, nullptr, u, x, [&](const MsrGroup& g) {
                unsigned long long bits[8];
                msr_relax8(a.d, g, [&](int i) { return lane_bit(g.ma[i]); }, bits);
                msr_claim8(a.pend, g, bits, want, px, low);
              });
              push_plain(&a.st->fq[par ^ 1], a.fl[par ^ 1], want, px);
            });
  if (lane == 0 && low) atomicAdd(&a.st->lowered, low);
}

// Affected masks back to zero for the next batch.
__global__ void k_msr_clear(const MsrArgs* __restrict__ pa) {
  const MsrArgs& a = *pa;
  const uint64_t na = a.st->aall;
  for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < na;
       j += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    a.aff[a.al[j]] = 0;
}

__global__ void k_msr_zero_state(MsrState* st) {
  if (threadIdx.x < sizeof(MsrState) / 8) reinterpret_cast<unsigned long long*>(st)[threadIdx.x] = 0;
}

// d[v][r] from the level masks of a full pass: bit r of newbits[L][v] is
// set for exactly one level L (v's distance from root r), so each (v, r) of
// a tile of 64 vertices is written once in LDS, then the tile goes out as
// 64 contiguous 256-byte rows.
__global__ __launch_bounds__(kMsBlock) void k_ms_to_d32(const unsigned long long* __restrict__ newbits,
                                                        uint32_t levels, uint64_t n, uint32_t* __restrict__ d) {
  __shared__ uint32_t tile[64][65];
  const uint32_t vv = threadIdx.x / 4, r0 = (threadIdx.x % 4) * 16;
  for (uint64_t v0 = blockIdx.x * 64ull; v0 < n; v0 += gridDim.x * 64ull) {
    for (uint32_t i = threadIdx.x; i < 64 * 64; i += blockDim.x) tile[i / 64][i % 64] = kInf;
    __syncthreads();
    if (v0 + vv < n) {
      for (uint32_t L = 0; L < levels; ++L) {
        const uint32_t m = static_cast<uint32_t>(newbits[static_cast<uint64_t>(L) * n + v0 + vv] >> r0) & 0xFFFFu;
        for (uint32_t b = m; b; b &= b - 1) tile[vv][r0 + __ffs(b) - 1] = L;
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 64 * 64; i += blockDim.x)
      if (v0 + i / 64 < n) d[(v0 + i / 64) * 64 + i % 64] = tile[i / 64][i % 64];
    __syncthreads();
  }
}

// Columns [from, to) of d set to unreachable (vertices a batch added).
__global__ void k_msr_fill_inf(uint32_t* d, uint64_t from, uint64_t to) {
  for (uint64_t i = from * 64 + blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < to * 64;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    d[i] = kInf;
}

// Root r's distances (uint32, kInf unreachable) from d.
__global__ void k_msr_root(const uint32_t* __restrict__ d, uint64_t n, uint32_t r, uint32_t* __restrict__ out) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    out[v] = d[v * 64 + r];
}

// Per-root uint8 levels, root-major (255 unreachable), from d; also the
// largest finite distance (for the level count) and reached counts.
__global__ __launch_bounds__(kMsBlock) void k_msr_u8(const uint32_t* __restrict__ d, uint64_t n, uint32_t k,
                                                     uint8_t* __restrict__ out) {
  // a workgroup stages 64 vertices x 64 roots through LDS so that both the
  // 256-byte row reads and the root-major byte writes are coalesced
  __shared__ uint8_t tile[64][65];
  for (uint64_t v0 = blockIdx.x * 64ull; v0 < n; v0 += gridDim.x * 64ull) {
    for (uint32_t i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
      const uint32_t vv = i / 64, r = i % 64;
      const uint32_t x = v0 + vv < n ? d[(v0 + vv) * 64 + r] : kInf;
      tile[r][vv] = x > 254u ? 255u : static_cast<uint8_t>(x);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
      const uint32_t r = i / 64, vv = i % 64;
      if (r < k && v0 + vv < n) out[static_cast<uint64_t>(r) * n + v0 + vv] = tile[r][vv];
    }
    __syncthreads();
  }
}

__global__ void k_msr_stats(const uint32_t* __restrict__ d, uint64_t n, uint32_t k,
                            unsigned long long* __restrict__ reached, unsigned int* __restrict__ maxd) {
  const unsigned lane = lane_id();
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  unsigned long long c = 0;
  uint32_t mx = 0;
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; v < n;
       v += waves) {
    const uint32_t x = d[v * 64 + lane];
    if (lane < k && x != kInf) {
      ++c;
      mx = max(mx, x);
    }
  }
  if (lane < k && c) atomicAdd(&reached[lane], c);
  mx = wave_max(mx);
  if (lane == 0) atomicMax(maxd, mx);
}

static unsigned msr_grid(uint64_t items, unsigned per = kMsBlock) { return ms_blocks(items, per, 4096); }

// Repair arrays sized for the graph's vertex count (new vertices start
// unreachable with empty masks; the list buffers hold at most one entry per
// vertex and round).
static void msr_ensure(vgx_msbfs* m, uint64_t n) {
  MsRepair& R = m->rep;
  hipStream_t s = m->g->stream;
  if (n > R.n) {
    R.setup = true;
    const uint64_t keep = R.n;
    const uint64_t cap = std::max<uint64_t>(n, R.n + R.n / 4);
    R.d.grow_copy(cap * 64 + 64, keep * 64, s);
    for (DevBuf<unsigned long long>* b : {&R.cand, &R.aff, &R.pend}) {
      b->grow_copy(cap + 1, keep, s);
      VGX_HIP(hipMemsetAsync(b->p + keep, 0, (cap + 1 - keep) * 8, s));
    }
    for (DevBuf<unsigned long long>* b : {&R.tm, &R.sup, &R.nm, &R.pm}) b->ensure(cap + 1);
    for (int p = 0; p < 2; ++p) {
      R.tl[p].ensure(cap + 1);
      R.vl[p].ensure(cap + 1);
      R.sl[p].ensure(cap + 1);
      R.spre[p].ensure(cap + 1);
      R.el[p].ensure(cap + 1);
      R.epre[p].ensure(cap + 1);
      R.fl[p].ensure(cap + 1);
      R.pl[p].ensure(cap + 1);
      R.ppre[p].ensure(cap + 1);
    }
    R.rl.ensure(cap + 1);
    R.rpre.ensure(cap + 1);
    R.al.ensure(cap + 1);
    R.seed.ensure(cap + 1);
    R.seedpre.ensure(cap + 1);
    R.n = cap;
  }
}

static MsrArgs msr_args(vgx_msbfs* m) {
  vgx_graph* g = m->g;
  MsRepair& R = m->rep;
  const BatchScratch& bs = g->bs[0];
  MsrArgs a{};
  a.out_row = g->adj[0].row.p;
  a.out_arena = g->adj[0].arena.p;
  a.in_row = g->in().row.p;
  a.in_arena = g->in().arena.p;
  a.d = R.d.p;
  a.cand = R.cand.p;
  a.aff = R.aff.p;
  a.tm = R.tm.p;
  a.sup = R.sup.p;
  a.nm = R.nm.p;
  a.pend = R.pend.p;
  a.pm = R.pm.p;
  for (int p = 0; p < 2; ++p) {
    a.tl[p] = R.tl[p].p;
    a.vl[p] = R.vl[p].p;
    a.sl[p] = R.sl[p].p;
    a.spre[p] = R.spre[p].p;
    a.el[p] = R.el[p].p;
    a.epre[p] = R.epre[p].p;
    a.fl[p] = R.fl[p].p;
    a.pl[p] = R.pl[p].p;
    a.ppre[p] = R.ppre[p].p;
  }
  a.rl = R.rl.p;
  a.rpre = R.rpre.p;
  a.al = R.al.p;
  a.seed = R.seed.p;
  a.seedpre = R.seedpre.p;
  a.fin = bs.fin.p;
  a.skeys = bs.skeys.p;
  a.act = bs.act.p;
  a.nfin = &g->d_ctr->nfinal[0];
  a.roots = m->droots.p;
  a.k = m->k;
  a.limit = static_cast<unsigned long long>(m->fallback * static_cast<double>(g->n) * m->k);
  if (a.limit < 1) a.limit = 1;
  a.st = R.d_st;
  return a;
}

static void msr_upload(vgx_msbfs* m) {
  // every upload is followed by a stream synchronisation before the next
  // one, so no copy from the pinned staging block is pending here
  *m->rep.h_args = msr_args(m);
  VGX_HIP(hipMemcpyAsync(m->rep.d_args, m->rep.h_args, sizeof(MsrArgs), hipMemcpyHostToDevice, m->g->stream));
}

// One synchronisation reads the repair state and, with `graph`, the graph's
// counters (the staged batch's counts, or a finished batch's) too.
static void msr_sync(vgx_msbfs* m, bool graph = false) {
  if (graph)
    VGX_HIP(hipMemcpyAsync(m->g->h_ctr, m->g->d_ctr, sizeof(GraphCounters), hipMemcpyDeviceToHost, m->g->stream));
  VGX_HIP(hipMemcpyAsync(m->rep.h_st, m->rep.d_st, sizeof(MsrState), hipMemcpyDeviceToHost, m->g->stream));
  VGX_HIP(hipStreamSynchronize(m->g->stream));
}

// The current distances in rep.d (from the last full pass when they are
// still held as level masks).
static void msr_materialize(vgx_msbfs* m) {
  msr_ensure(m, m->n);
  if (m->in_d32) return;
  hipStream_t s = m->g->stream;
  if (m->n) {
    k_ms_to_d32<<<ms_blocks((m->n + 63) / 64, 1, 4096), kMsBlock, 0, s>>>(m->newbits.p, m->levels, m->n,
                                                                           m->rep.d.p);
    VGX_LAUNCH_CHECK();
  }
  m->in_d32 = true;
}

constexpr uint32_t kMsrInvRounds = 2;   // invalidation rounds per host check
constexpr uint32_t kMsrPropRounds = 4;  // propagation rounds per host check

// The repair's four launch programs: 0 = state cleared, candidates and the
// first invalidation rounds; 1 = more invalidation rounds; 2 = reset, seeds,
// added arcs and the first propagation rounds; 3 = more propagation rounds.
// Fixed grids, every size read on the device, arguments in device memory:
// each is captured once into a hipGraph and replayed for every batch.
static void msr_program(vgx_msbfs* m, int which, hipStream_t s) {
  const MsrArgs* pa = m->rep.d_args;
  if (which == 0) {
    k_msr_zero_state<<<1, 64, 0, s>>>(m->rep.d_st);
    VGX_LAUNCH_CHECK();
    k_msr_candidates<<<kMsrBlocks, kMsBlock, 0, s>>>(pa);
    VGX_LAUNCH_CHECK();
  }
  if (which == 0 || which == 1) {
    for (uint32_t i = 0; i < kMsrInvRounds; ++i) {
      const int par = static_cast<int>(i & 1);  // kMsrInvRounds is even
      k_msr_take<<<kMsrBlocks, kMsBlock, 0, s>>>(pa, par);
      VGX_LAUNCH_CHECK();
      k_msr_support<<<kMsrBlocks, kMsBlock, 0, s>>>(pa, par, 0);
      VGX_LAUNCH_CHECK();
      k_msr_take_rest<<<kMsrBlocks, kMsBlock, 0, s>>>(pa, par);
      VGX_LAUNCH_CHECK();
      k_msr_support<<<kMsrBlocks, kMsBlock, 0, s>>>(pa, par, 1);
      VGX_LAUNCH_CHECK();
      k_msr_evaluate<<<kMsrBlocks, kMsBlock, 0, s>>>(pa, par);
      VGX_LAUNCH_CHECK();
      k_msr_emit<<<kMsrBlocks, kMsBlock, 0, s>>>(pa, par);
      VGX_LAUNCH_CHECK();
    }
    return;
  }
  if (which == 2) {
    k_msr_reset<<<kMsrBlocks, kMsBlock, 0, s>>>(pa);
    VGX_LAUNCH_CHECK();
    k_msr_seed<<<kMsrBlocks, kMsBlock, 0, s>>>(pa);
    VGX_LAUNCH_CHECK();
    k_msr_collect<<<kMsrBlocks, kMsBlock, 0, s>>>(pa);
    VGX_LAUNCH_CHECK();
    k_msr_relax_additions<<<kMsrBlocks, kMsBlock, 0, s>>>(pa);
    VGX_LAUNCH_CHECK();
  }
  for (uint32_t i = 0; i < kMsrPropRounds; ++i) {
    const int par = static_cast<int>(i & 1);  // kMsrPropRounds is even
    k_msr_ptake<<<kMsrBlocks, kMsBlock, 0, s>>>(pa, par);
    VGX_LAUNCH_CHECK();
    k_msr_propagate<<<kMsrBlocks, kMsBlock, 0, s>>>(pa, par);
    VGX_LAUNCH_CHECK();
  }
}

static void msr_run_program(vgx_msbfs* m, int which) {
  hipStream_t s = m->g->stream;
  MsRepair& R = m->rep;
  if (!m->use_graph) {
    msr_program(m, which, s);
    return;
  }
  if (!R.exec[which]) {
    R.setup = true;
    VGX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    msr_program(m, which, s);
    VGX_HIP(hipStreamEndCapture(s, &R.graph[which]));
    VGX_HIP(hipGraphInstantiate(&R.exec[which], R.graph[which], nullptr, nullptr, 0));
  }
  VGX_HIP(hipGraphLaunch(R.exec[which], s));
}

static void msr_drop_graphs(MsRepair& R) {
  for (int i = 0; i < 4; ++i) {
    if (R.exec[i]) (void)hipGraphExecDestroy(R.exec[i]);
    if (R.graph[i]) (void)hipGraphDestroy(R.graph[i]);
    R.exec[i] = nullptr;
    R.graph[i] = nullptr;
  }
}

// Predicted repair time of a batch of size class cls: its own running mean,
// else the nearest smaller class's scaled by the size ratio (pessimistic),
// else the nearest larger class's (optimistic), else 0 (try it).
static double msr_predict(const vgx_msbfs* m, int cls) {
  if (m->repair_ms[cls] > 0.0) return m->repair_ms[cls];
  for (int c = cls - 1; c >= 0; --c)
    if (m->repair_ms[c] > 0.0) return m->repair_ms[c] * static_cast<double>(1ull << (cls - c));
  for (int c = cls + 1; c < 48; ++c)
    if (m->repair_ms[c] > 0.0) return m->repair_ms[c];
  return 0.0;
}

static void msr_apply(vgx_msbfs* m, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k,
                      bool device, hipStream_t producer) {
  vgx_graph* g = m->g;
  hipStream_t s = g->stream;
  m->last_repaired = false;
  m->last_fallback = false;
  m->last_affected_pairs = m->last_lowered = m->last_candidates = 0;
  if (k == 0) return;
  const auto t0 = std::chrono::steady_clock::now();
  // the arm: a repair needs distances to repair; a root id the batch brings
  // into the graph needs a pass (its own distance becomes 0)
  bool repair = m->ran && m->repair_mode != 2 && m->dist_version == g->version;
  const int cls = std::min(47, 63 - __builtin_clzll(k));
  if (repair && m->repair_mode == 0 && m->pass_ms > 0.0) {
    const double pred = msr_predict(m, cls);
    repair = pred < m->pass_ms;
    // the means drift: a class within twice the pass looks again now and then
    if (!repair && pred < 2.0 * m->pass_ms && ++m->passes_in_class[cls] % 16 == 0) repair = true;
    if (repair) m->passes_in_class[cls] = 0;
  }
  m->rep.setup = false;
  if (repair) msr_materialize(m);
  // A repair leaves the staged batch's counts on the device until the
  // invalidation's first synchronisation (its kernels read them there).
  if (device) graph_stage_device(g, src, dst, add, k, producer, repair);
  else graph_stage(g, src, dst, add, k, repair);
  // distances cover the vertices of the last pass or repair (m->n): a root
  // whose id the graph now reaches needs a pass (its own distance becomes 0)
  const uint64_t n_cov = m->n;
  for (uint32_t r = 0; r < m->k; ++r)
    if (m->roots[r] >= n_cov && m->roots[r] < g->n) repair = false;
  m->batches += 1;
  if (!repair && !g->bs[0].counted) {  // staged for a repair, answered by a pass
    graph_sync_counters(g);
    graph_stage_collect(g);
  }
  auto elapsed_ms = [&]() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  };
  auto ema = [](double& x, double v) { x = x > 0.0 ? 0.75 * x + 0.25 * v : v; };
  if (!repair) {
    graph_apply_deletions(g);
    graph_apply_additions(g);
    graph_finish_batch(g);
    ms_run(m);
    ema(m->pass_ms, elapsed_ms());
    m->full_passes += 1;
    return;
  }
  MsRepair& R = m->rep;
  msr_ensure(m, g->n);
  if (g->n > n_cov) {
    k_msr_fill_inf<<<msr_grid((g->n - n_cov) * 64), kMsBlock, 0, s>>>(R.d.p, n_cov, g->n);
    VGX_LAUNCH_CHECK();
  }
  m->n = g->n;
  graph_apply_deletions(g);
  msr_upload(m);
  bool fallback = false;
  for (uint32_t r0 = 0;; r0 += kMsrInvRounds) {
    msr_run_program(m, r0 == 0 ? 0 : 1);
    msr_sync(m, r0 == 0);
    if (r0 == 0) graph_stage_collect(g);  // may grow the arena (rows keep their places)
    if (R.h_st->fallback || R.h_st->apairs > static_cast<unsigned long long>(m->fallback * g->n * m->k + 1)) {
      fallback = true;
      break;
    }
    if (R.h_st->tq[0] == 0) break;  // nothing touched for the next round
    if (r0 > 2 * g->n + 4) throw Error(VGX_DEVICE_ERROR, "multi-source invalidation did not terminate");
  }
  m->last_candidates = R.h_st->candidates;
  m->last_affected_pairs = R.h_st->apairs;
  graph_apply_additions(g);
  if (fallback) {
    graph_finish_batch(g);
    VGX_HIP(hipMemsetAsync(R.cand.p, 0, (g->n + 1) * 8, s));
    VGX_HIP(hipMemsetAsync(R.aff.p, 0, (g->n + 1) * 8, s));
    ms_run(m);
    if (!m->rep.setup) ema(m->repair_ms[cls], elapsed_ms());  // what trying the repair cost at this size
    m->fallbacks += 1;
    m->full_passes += 1;
    m->last_fallback = true;
    return;
  }
  // The batch's bookkeeping (which may compact and move rows) waits for the
  // propagation's last synchronisation: the rows stay where the additions
  // left them, which is what the propagation reads.
  msr_upload(m);
  const uint64_t n_aff = R.h_st->aall;
  for (uint32_t r0 = 0;; r0 += kMsrPropRounds) {
    msr_run_program(m, r0 == 0 ? 2 : 3);
    msr_sync(m, true);  // the finished batch's counters come along
    if (R.h_st->fq[0] == 0) break;
    if (r0 > 2 * g->n + 4) throw Error(VGX_DEVICE_ERROR, "multi-source propagation did not terminate");
  }
  if (n_aff) {
    k_msr_clear<<<msr_grid(n_aff), kMsBlock, 0, s>>>(R.d_args);
    VGX_LAUNCH_CHECK();
  }
  graph_finish_collect(g);  // may compact: rows move, the distances do not
  m->last_lowered = R.h_st->lowered;
  m->in_d32 = true;
  m->dist_version = g->version;
  m->levels = 0;  // recomputed on demand from d
  m->last_repaired = true;
  m->repaired += 1;
  if (!m->rep.setup) {  // one-time allocations and captures are not the arm's cost
    const double ms = elapsed_ms();
    ema(m->repair_ms[cls], ms);
    ema(m->repair_us_per_op, 1e3 * ms / static_cast<double>(k));
  }
}

// Reached count per root and the level count (deepest distance + 1) of
// the distances held in rep.d.
static uint32_t msr_stats(vgx_msbfs* m, std::vector<uint64_t>* reached) {
  hipStream_t s = m->g->stream;
  DevBuf<unsigned long long> acc;
  acc.ensure(65);
  VGX_HIP(hipMemsetAsync(acc.p, 0, 65 * 8, s));
  if (m->n) {
    k_msr_stats<<<msr_grid(m->n * kWave), kMsBlock, 0, s>>>(m->rep.d.p, m->n, m->k, acc.p,
                                                           reinterpret_cast<unsigned int*>(acc.p + 64));
    VGX_LAUNCH_CHECK();
  }
  unsigned long long h[65];
  VGX_HIP(hipMemcpyAsync(h, acc.p, sizeof(h), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  if (reached) reached->assign(h, h + m->k);
  bool any = false;
  for (uint32_t r = 0; r < m->k; ++r) any = any || h[r] > 0;
  return any ? static_cast<uint32_t>(h[64] & 0xFFFFFFFFu) + 1 : 0;
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
    m->rep.d_args = static_cast<MsrArgs*>(dev_alloc(sizeof(MsrArgs)));
    m->rep.d_st = static_cast<MsrState*>(dev_alloc(sizeof(MsrState)));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&m->rep.h_args), sizeof(MsrArgs), hipHostMallocDefault));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&m->rep.h_st), sizeof(MsrState), hipHostMallocDefault));
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
  msr_drop_graphs(m->rep);
  if (m->rep.d_args) dev_free(m->rep.d_args);
  if (m->rep.h_args) (void)hipHostFree(m->rep.h_args);
  if (m->rep.d_st) dev_free(m->rep.d_st);
  if (m->rep.h_st) (void)hipHostFree(m->rep.h_st);
  delete m;
  VGX_API_END
}

vgx_status vgx_msbfs_run(vgx_msbfs* m) {
  VGX_API_BEGIN
  if (!m) throw Error(VGX_INVALID_ARGUMENT, "null msbfs");
  VGX_HIP(hipSetDevice(m->g->device));
  const auto t0 = std::chrono::steady_clock::now();
  ms_run(m);
  if (m->pass_ms <= 0.0)  // the repair policy's first estimate of the full-pass arm
    m->pass_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  VGX_API_END
}

vgx_status vgx_msbfs_set_graph(vgx_msbfs* m, int on) {
  VGX_API_BEGIN
  if (!m) throw Error(VGX_INVALID_ARGUMENT, "null msbfs");
  m->use_graph = on != 0;
  if (!m->use_graph) {
    ms_drop_graph(m);
    msr_drop_graphs(m->rep);
  }
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
    if (m->in_d32)
      k_msr_root<<<ms_blocks(m->n), kMsBlock, 0, s>>>(m->rep.d.p, m->n, r, tmp.p);
    else
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
  VGX_HIP(hipSetDevice(m->g->device));
  const uint32_t levels = m->in_d32 ? msr_stats(m, nullptr) : m->levels;
  if (levels > 255) throw Error(VGX_UNSUPPORTED, "depth exceeds 254: use vgx_msbfs_distances");
  hipStream_t s = m->g->stream;
  DevBuf<uint8_t> tmp;
  tmp.ensure(static_cast<uint64_t>(m->k) * m->n + 1);
  if (m->n) {
    if (m->in_d32)
      k_msr_u8<<<ms_blocks((m->n + 63) / 64, 1, 4096), kMsBlock, 0, s>>>(m->rep.d.p, m->n, m->k, tmp.p);
    else
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
  if (m->in_d32) {
    std::vector<uint64_t> rc;
    msr_stats(m, &rc);
    for (uint32_t r = 0; r < m->k && r < cap; ++r) reached_per_root[r] = rc[r];
    vgx::t_last_error.clear();
    return VGX_OK;
  }
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

vgx_status vgx_msbfs_set_params(vgx_msbfs* m, const vgx_msbfs_tuning* t) {
  VGX_API_BEGIN
  if (!m || !t) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (t->blocks > 65535) throw Error(VGX_INVALID_ARGUMENT, "too many workgroups");
  if (!(t->alpha > 0.0)) throw Error(VGX_INVALID_ARGUMENT, "alpha must be positive");
  if (!(t->beta >= 0.0)) throw Error(VGX_INVALID_ARGUMENT, "beta must be zero or positive");
  if (t->short_row < 1 || t->short_row > 1024) throw Error(VGX_INVALID_ARGUMENT, "short_row must be 1..1024");
  if (t->chunk < 64 || t->chunk > (1u << 20)) throw Error(VGX_INVALID_ARGUMENT, "chunk must be 64..2^20");
  m->blocks = t->blocks;
  m->alpha = t->alpha;
  m->beta = t->beta;
  m->short_row = t->short_row;
  m->chunk = t->chunk;
  VGX_API_END
}

vgx_status vgx_msbfs_get_params(vgx_msbfs* m, vgx_msbfs_tuning* t) {
  VGX_API_BEGIN
  if (!m || !t) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  t->blocks = m->blocks;
  t->alpha = m->alpha;
  t->beta = m->beta;
  t->short_row = m->short_row;
  t->chunk = m->chunk;
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
  // one parameter at a time, each over its candidates, keeping the fastest
  auto sweep = [&](auto& field, std::initializer_list<std::decay_t<decltype(field)>> cands) {
    auto best = field;
    for (auto c : cands) {
      field = c;
      const double t = time_run();
      if (t < best_t) {
        best_t = t;
        best = c;
      }
    }
    field = best;
  };
  sweep(m->alpha, {2.0, 4.0, 8.0, 16.0, 32.0});
  sweep(m->beta, {0.0, 4.0, 16.0, 64.0});
  sweep(m->short_row, {8u, 16u, 32u, 64u, 128u});
  sweep(m->chunk, {256u, 512u, 1024u, 2048u, 4096u});
  sweep(m->blocks, {0u, 512u, 1024u, 2048u, 4096u, 8192u});
  ms_run(m);
  if (best_blocks) *best_blocks = m->blocks;
  if (best_alpha) *best_alpha = m->alpha;
  if (best_ms) *best_ms = best_t;
  VGX_API_END
}

// Number of levels of the last run (the deepest distance + 1).
vgx_status vgx_msbfs_level_count(vgx_msbfs* m, uint32_t* levels) {
  VGX_API_BEGIN
  if (!m || !levels) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (m->in_d32) {
    VGX_HIP(hipSetDevice(m->g->device));
    *levels = msr_stats(m, nullptr);
  } else {
    *levels = m->levels;
  }
  VGX_API_END
}

vgx_status vgx_msbfs_apply(vgx_msbfs* m, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k) {
  VGX_API_BEGIN
  if (!m || (k && (!src || !dst || !add))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(m->g->device));
  msr_apply(m, src, dst, add, k, false, nullptr);
  VGX_API_END
}

vgx_status vgx_msbfs_apply_device(vgx_msbfs* m, const uint32_t* src_dev, const uint32_t* dst_dev,
                                  const uint8_t* add_dev, uint64_t k, void* stream) {
  VGX_API_BEGIN
  if (!m || (k && (!src_dev || !dst_dev || !add_dev))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(m->g->device));
  msr_apply(m, src_dev, dst_dev, add_dev, k, true, static_cast<hipStream_t>(stream));
  VGX_API_END
}

vgx_status vgx_msbfs_set_repair(vgx_msbfs* m, int mode, double fallback_fraction) {
  VGX_API_BEGIN
  if (!m) throw Error(VGX_INVALID_ARGUMENT, "null msbfs");
  if (mode < 0 || mode > 2) throw Error(VGX_INVALID_ARGUMENT, "mode must be 0 (adaptive), 1 (repair) or 2 (full)");
  if (!(fallback_fraction > 0.0)) throw Error(VGX_INVALID_ARGUMENT, "fallback fraction must be positive");
  m->repair_mode = mode;
  m->fallback = fallback_fraction;
  VGX_API_END
}

vgx_status vgx_msbfs_repair_stats(vgx_msbfs* m, vgx_msbfs_repair_info* out) {
  VGX_API_BEGIN
  if (!m || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  out->batches = m->batches;
  out->repaired = m->repaired;
  out->full_passes = m->full_passes;
  out->fallbacks = m->fallbacks;
  out->last_repaired = m->last_repaired ? 1 : 0;
  out->last_fallback = m->last_fallback ? 1 : 0;
  out->last_candidates = m->last_candidates;
  out->last_affected_pairs = m->last_affected_pairs;
  out->last_lowered = m->last_lowered;
  out->pass_ms = m->pass_ms;
  out->repair_us_per_op = m->repair_us_per_op;
  VGX_API_END
}

}  // extern "C"
