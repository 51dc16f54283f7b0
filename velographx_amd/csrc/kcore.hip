// k-core decomposition (core numbers) on the live rows of an undirected
// graph, by level-synchronous parallel peeling.
//
// Reference: include/velographx/incremental/kcore.hpp.  A batch's repair
// region is the reference's (:31-56): the components of the batch's
// endpoints before and after the graph takes it, found here by a flood fill
// from the endpoints on the device, once on each graph.  The region is a
// union of whole components of the new graph, so every neighbour of a
// region vertex is in the region: its vertices are re-peeled with their full
// live degrees and all other vertices keep their core numbers.  The values
// are those of the reference's min-degree heap (:83-121): core numbers
// depend only on the graph.
//
// Level k: a scan claims every unpeeled region vertex whose remaining
// degree is at most k (core number k) into a queue; then rounds of peeling
// decrement the remaining degree of each unpeeled neighbour of the queued
// vertices, and the decrement that lands on k claims that neighbour into
// the next queue (after the scan every unpeeled vertex has degree above k,
// and concurrent decrements return distinct values, so exactly one does).
// A level whose scan claims nothing jumps straight to the smallest
// remaining degree.  Every arc is decremented at most once, so degrees
// never wrap.
//
// Scans read two lists: a near list of the unpeeled vertices of degree at
// most `bound`, read by every scan, and a far list of the others, read only
// when a level passes the bound -- then the far list is split again with
// bound = k + kKcWindow.  A far vertex whose degree drops to the bound
// during peeling is appended to the near list there (a degree falls by one
// at a time, so it crosses the bound once), and the far scans skip vertices
// whose degree is at or below the old bound (they are on the near list).
// Most vertices of a skewed graph stay unpeeled for many levels at degrees
// far above the level; they are read once per window instead of per level.
//
// The level loop is a device-steered step program: launch s reads launch
// s - 1's plan and results from a ring of four records and decides its own
// step (a scan, a peeling round, or nothing once done), so the host enqueues
// launches in chunks and reads the state back once per chunk.  Within a
// peeling step a workgroup keeps the first few items of its own claims in
// LDS and peels them itself, in rounds separated by workgroup barriers,
// instead of leaving every claim for the next launch (peeling within a level
// does not depend on the order of the decrements).
#include <algorithm>
#include <cstring>
#include <string>

#include "vgx_internal.hpp"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);
extern thread_local std::string t_last_error;

constexpr uint32_t kKcUnset = 0xFFFFFFFFu;
constexpr int kKcBlock = 256;
constexpr uint32_t kKcUnroll = 4;
constexpr uint32_t kKcPhInit = 0, kKcPhScan = 1, kKcPhPeel = 2, kKcPhDone = 3;
constexpr int32_t kKcWindow = 8;  // near-list degree window above the level
// Peeling items a workgroup keeps in LDS for itself within one step (its
// own claims, peeled without a launch between rounds); the rest of a step's
// claims go to the global queue for the next step.  Small on purpose: a
// long cascade is faster spread over the whole grid in the next step than
// walked by one workgroup's four waves (measured at R-MAT s20: 32 items
// 7.9 ms per recompute, 256 items 8.9 ms, 4,096 items 19.3 ms).
constexpr uint32_t kKcLocal = 32;
// Flood-fill rounds run several per host check, on a ring of three counters
// like the peeling rounds of the step program's queues.
constexpr uint32_t kKcFloodPerCheck = 4;

// One step of the level loop.  The plan fields are written by the step's
// first thread; the counters, zeroed by the step before, are added to by
// every workgroup of the step.
struct KcRec {
  uint32_t phase, k, round;
  uint32_t nl;                     // near list a scan writes (it reads nl ^ 1); kept by peeling rounds
  uint32_t fl;                     // far list in use (a splitting scan reads it and writes fl ^ 1)
  uint32_t split;                  // scan: the level passed the bound, the far list is read
  int32_t bound;                   // degree bound of the near list (-1: no near list yet)
  uint32_t pad;
  unsigned long long in_count;     // scan: near entries; peeling round: queued vertices
  unsigned long long level_rem;    // peeling round: entries on the near list being written
  unsigned long long far_count;    // entries of the far list in use
  unsigned long long queued;       // claims appended to this step's output queue
  unsigned long long remaining;    // scan: vertices written to the near list
  unsigned long long far_rem;      // splitting scan: vertices written to the new far list
  unsigned long long min_deg;      // scan: the smallest degree among the vertices it kept
  unsigned long long crossed;      // peeling round: far vertices appended to the near list
};

struct KcState {
  KcRec rec[4];                    // step s reads rec[(s + 3) % 4], writes rec[s % 4]
  unsigned long long scans, rounds, levels;
  unsigned long long flood[3];     // flood-fill queue lengths
  unsigned long long marked;       // region vertices
};

}  // namespace vgx

struct vgx_kcore {
  vgx_graph* g = nullptr;
  vgx::DevBuf<uint32_t> deg, core, mark, near[2], far[2], ends;
  vgx::DevBuf<unsigned long long> items[2];  // peeling items of the level loop
  vgx::DevBuf<uint32_t> fq[2];               // flood-fill queues
  vgx::KcState* d_st = nullptr;
  vgx::KcState* h_st = nullptr;  // pinned
  uint64_t n = 0;
  uint64_t last_repaired = 0;
  uint64_t mark_valid = 0;       // entries of mark initialised
  uint32_t epoch = 0;            // mark[v] >= the current pre-batch epoch: in the region
  uint32_t levels = 0, rounds = 0;
  uint64_t steps_hint = 0;       // steps the last level loop took
};

namespace vgx {

static unsigned kc_blocks(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t b = (items + per_block - 1) / per_block;
  return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(b, cap)));
}

// Appends the flagged lanes' vertices to q at *count (one atomic per wave).
__device__ __forceinline__ void kc_push(bool flag, uint32_t v, uint32_t* __restrict__ q, unsigned long long* count) {
  const unsigned long long m = __ballot(flag);
  if (!m) return;
  unsigned long long base = 0;
  const unsigned leader = __ffsll(static_cast<long long>(m)) - 1;
  if (lane_id() == leader) base = atomicAdd(count, static_cast<unsigned long long>(__popcll(m)));
  base = __shfl(base, static_cast<int>(leader), kWave);
  if (flag) q[base + lanes_below(m)] = v;
}

// Per-wave staging of queue pushes in LDS: a push lands in the wave's
// buffer, and one atomic per flush (or, at the end of a kernel, one per
// workgroup) reserves the queue slots -- instead of one atomic on a hot
// counter per wave step, which serialises thousands of waves at the L2.
constexpr uint32_t kKcStage = 512;
struct KcStage {
  uint32_t v[kKcStage];
};

__device__ __forceinline__ void kc_flush(KcStage& sb, uint32_t& nbuf, uint32_t* __restrict__ q,
                                         unsigned long long* count) {
  if (nbuf == 0) return;
  __builtin_amdgcn_wave_barrier();
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(count, static_cast<unsigned long long>(nbuf));
  base = __shfl(base, 0, kWave);
  for (uint32_t c = lane_id(); c < nbuf; c += kWave) q[base + c] = sb.v[c];
  __builtin_amdgcn_wave_barrier();
  nbuf = 0;
}

// Stages v for the flagged lanes; called by every lane of the wave.
__device__ __forceinline__ void kc_stage(KcStage& sb, uint32_t& nbuf, bool flag, uint32_t v,
                                         uint32_t* __restrict__ q, unsigned long long* count) {
  const unsigned long long m = __ballot(flag);
  if (!m) return;
  if (flag) sb.v[nbuf + lanes_below(m)] = v;
  nbuf += __popcll(m);
  if (nbuf > kKcStage - kWave) kc_flush(sb, nbuf, q, count);
}

// The last flush of every wave of a workgroup, with one atomic for the
// workgroup.  Every thread of the workgroup calls it; `sm` holds
// waves + 1 words.
__device__ __forceinline__ void kc_flush_block(KcStage& sb, uint32_t nbuf, uint32_t* __restrict__ q,
                                               unsigned long long* count, unsigned long long* sm) {
  const uint32_t wv = threadIdx.x / kWave, waves = blockDim.x / kWave;
  if (lane_id() == 0) sm[wv] = nbuf;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t i = 0; i < waves; ++i) t += sm[i];
    sm[waves] = t ? atomicAdd(count, t) : 0ull;
  }
  __syncthreads();
  unsigned long long base = sm[waves];
  for (uint32_t i = 0; i < wv; ++i) base += sm[i];
  for (uint32_t c = lane_id(); c < nbuf; c += kWave) q[base + c] = sb.v[c];
  __syncthreads();  // sm is reused by the next flush
}

// Peeling work items: (chunk << 32) | v names entries [chunk * kKcChunk,
// (chunk + 1) * kKcChunk) of v's row, so a hub row is spread over many
// waves instead of being walked by one.  A claimed vertex with an empty row
// has nothing to peel and queues no item.
constexpr uint32_t kKcChunk = kKcUnroll * kWave;
struct KcItems {
  unsigned long long v[kKcStage];
};

__device__ __forceinline__ void kc_flush_items(KcItems& sb, uint32_t& nbuf, unsigned long long* __restrict__ q,
                                               unsigned long long* count) {
  if (nbuf == 0) return;
  __builtin_amdgcn_wave_barrier();
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(count, static_cast<unsigned long long>(nbuf));
  base = __shfl(base, 0, kWave);
  for (uint32_t c = lane_id(); c < nbuf; c += kWave) q[base + c] = sb.v[c];
  __builtin_amdgcn_wave_barrier();
  nbuf = 0;
}

// Queues the nch items of v for every lane (nch = 0: none); called by every
// lane of the wave.  A wave step with more items than the buffer holds
// (long rows) writes them to the queue directly.
__device__ __forceinline__ void kc_stage_items(KcItems& sb, uint32_t& nbuf, uint32_t nch, uint32_t v,
                                               unsigned long long* __restrict__ q, unsigned long long* count) {
  const unsigned long long incl = wave_inclusive_scan<unsigned long long>(nch);
  const uint32_t total = static_cast<uint32_t>(__shfl(incl, kWave - 1, kWave));
  if (total == 0) return;
  const uint32_t first = static_cast<uint32_t>(incl) - nch;
  if (total > kKcStage - kWave) {
    unsigned long long base = 0;
    if (lane_id() == 0) base = atomicAdd(count, static_cast<unsigned long long>(total));
    base = __shfl(base, 0, kWave);
    for (uint32_t c = 0; c < nch; ++c) q[base + first + c] = (static_cast<unsigned long long>(c) << 32) | v;
    return;
  }
  if (nbuf + total > kKcStage) kc_flush_items(sb, nbuf, q, count);
  for (uint32_t c = 0; c < nch; ++c) sb.v[nbuf + first + c] = (static_cast<unsigned long long>(c) << 32) | v;
  nbuf += total;
  if (nbuf > kKcStage - kWave) kc_flush_items(sb, nbuf, q, count);
}

__device__ __forceinline__ void kc_flush_items_block(KcItems& sb, uint32_t nbuf, unsigned long long* __restrict__ q,
                                                     unsigned long long* count, unsigned long long* sm) {
  const uint32_t wv = threadIdx.x / kWave, waves = blockDim.x / kWave;
  if (lane_id() == 0) sm[wv] = nbuf;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t i = 0; i < waves; ++i) t += sm[i];
    sm[waves] = t ? atomicAdd(count, t) : 0ull;
  }
  __syncthreads();
  unsigned long long base = sm[waves];
  for (uint32_t i = 0; i < wv; ++i) base += sm[i];
  for (uint32_t c = lane_id(); c < nbuf; c += kWave) q[base + c] = sb.v[c];
  __syncthreads();
}

__device__ __forceinline__ uint32_t kc_chunks(uint32_t len) { return (len + kKcChunk - 1) / kKcChunk; }

// ---------------------------------------------------------------- region

// Seeds the flood fill with the batch's endpoints below n: a vertex joins
// when its mark rises to `epoch` (one flood per epoch, so each vertex is
// queued once per flood).
__global__ void k_kc_flood_seed(const uint32_t* __restrict__ ends, uint64_t nends, uint64_t n, uint32_t epoch,
                                uint32_t* __restrict__ mark, uint32_t* __restrict__ q, KcState* __restrict__ st) {
  for (uint64_t i0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); i0 < nends;
       i0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = i0 + threadIdx.x;
    const uint32_t v = i < nends ? ends[i] : 0u;
    const bool join = i < nends && v < n && atomicMax(&mark[v], epoch) < epoch;
    kc_push(join, v, q, &st->flood[0]);
  }
}

// One flood round: every neighbour of a queued vertex not yet marked in
// this flood joins the next queue.  Round r reads its length from
// flood[r % 3], appends to flood[(r + 1) % 3] and clears flood[(r + 2) % 3].
__global__ __launch_bounds__(kKcBlock) void k_kc_flood(uint32_t round, uint32_t epoch, const uint2* __restrict__ row,
                                                       const uint32_t* __restrict__ arena,
                                                       const uint32_t* __restrict__ q, uint32_t* __restrict__ mark,
                                                       uint32_t* __restrict__ qnext, KcState* __restrict__ st) {
  __shared__ KcStage s_stage[kKcBlock / kWave];
  __shared__ unsigned long long s_flush[kKcBlock / kWave + 1];
  KcStage& sb = s_stage[threadIdx.x / kWave];
  uint32_t nbuf = 0;
  const uint64_t count = st->flood[round % 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) st->flood[(round + 2) % 3] = 0;
  unsigned long long* next = &st->flood[(round + 1) % 3];
  const uint64_t W = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  const unsigned lane = lane_id();
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; i < count;
       i += W) {
    const uint2 rb = row[q[i]];
    for (uint32_t j0 = 0; j0 < rb.y; j0 += kWave) {
      const uint32_t j = j0 + lane;
      const uint32_t x = j < rb.y ? arena[rb.x + j] : kTomb;
      const uint32_t v = x & ~kTomb;
      const bool join = !(x & kTomb) && mark[v] < epoch && atomicMax(&mark[v], epoch) < epoch;
      kc_stage(sb, nbuf, join, v, qnext, next);
    }
  }
  kc_flush_block(sb, nbuf, qnext, next, s_flush);
}

// The level loop's start: every vertex of the region (mark >= lo, or all
// vertices when lo is 0) takes its live degree, loses its core number and
// joins the first scan's list (counted in rec[3].remaining, the "previous
// step" of launch 0); a vertex added since the last computation that is not
// in the region has no arcs and core number 0.
__global__ void k_kc_init(const uint32_t* __restrict__ live_deg, uint64_t n, uint64_t n_old, uint32_t lo,
                          const uint32_t* __restrict__ mark, uint32_t* __restrict__ deg, uint32_t* __restrict__ core,
                          uint32_t* __restrict__ list, KcState* __restrict__ st) {
  __shared__ KcStage s_stage[kKcBlock / kWave];
  __shared__ unsigned long long s_flush[kKcBlock / kWave + 1];
  KcStage& sb = s_stage[threadIdx.x / kWave];
  uint32_t nbuf = 0;
  for (uint64_t v0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); v0 < n;
       v0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t v = v0 + threadIdx.x;
    bool in = false;
    if (v < n) {
      in = lo == 0 || mark[v] >= lo;
      if (in) {
        deg[v] = live_deg[v];
        core[v] = kKcUnset;
      } else if (v >= n_old) {
        core[v] = 0;
      }
    }
    kc_stage(sb, nbuf, in, static_cast<uint32_t>(v), list, &st->rec[3].remaining);
  }
  kc_flush_block(sb, nbuf, list, &st->rec[3].remaining, s_flush);
}

// ------------------------------------------------------------ step program

struct KcPlan {
  uint32_t phase, k, round, nl, fl, split;
  int32_t bound;
  unsigned long long in_count, level_rem, far_count;
};

// The step after `p` (its plan and results); shared by every workgroup.
__device__ __forceinline__ KcPlan kc_next(const KcRec& p) {
  KcPlan x{kKcPhDone, p.k, 0u, p.nl, p.fl, 0u, p.bound, 0ull, p.level_rem, p.far_count};
  if (p.phase == kKcPhInit) {
    // the init put the region on far list 0; the first scan splits it
    if (p.remaining != 0) x = {kKcPhScan, 0u, 0u, 0u, 0u, 1u, -1, 0ull, 0ull, p.remaining};
  } else if (p.phase == kKcPhScan) {
    const unsigned long long far = p.split ? p.far_rem : p.far_count;
    const int32_t bound = p.split ? static_cast<int32_t>(p.k) + kKcWindow : p.bound;
    const uint32_t fl = p.split ? (p.fl ^ 1u) : p.fl;
    if (p.queued == 0) {
      // nothing at this level: the next level with vertices (past the bound
      // when the near list is empty and the far list was not read)
      if (p.remaining + far != 0) {
        const uint32_t k = (p.split || p.remaining != 0) ? static_cast<uint32_t>(p.min_deg)
                                                          : static_cast<uint32_t>(bound + 1);
        x = {kKcPhScan, k, 0u, p.nl ^ 1u, fl, static_cast<int32_t>(k) > bound ? 1u : 0u, bound, p.remaining, 0ull,
             far};
      }
    } else {
      x = {kKcPhPeel, p.k, 0u, p.nl, fl, 0u, bound, p.queued, p.remaining, far};
    }
  } else if (p.phase == kKcPhPeel) {
    const unsigned long long near = p.level_rem + p.crossed;
    if (p.queued != 0) {
      x = {kKcPhPeel, p.k, p.round + 1u, p.nl, p.fl, 0u, p.bound, p.queued, near, p.far_count};
    } else if (near + p.far_count != 0) {
      const uint32_t k = p.k + 1u;
      x = {kKcPhScan, k, 0u, p.nl ^ 1u, p.fl, static_cast<int32_t>(k) > p.bound ? 1u : 0u, p.bound, near, 0ull,
           p.far_count};
    }
  }
  return x;
}

struct KcArgs {
  const uint2* row;
  const uint32_t* arena;
  uint32_t* deg;
  uint32_t* core;
  unsigned long long* q[2];  // peeling items
  uint32_t* near[2];
  uint32_t* far[2];
  KcState* st;
};

// A workgroup's LDS worklist of peeling items: slots [0, kKcLocal) of work,
// filled through *tail (which may run past kKcLocal: those items go to the
// global queue instead).
struct KcLocal {
  unsigned long long* work;
  uint32_t* tail;
};

// Queues the nch items of v for every lane (nch = 0: none), into the
// workgroup's LDS worklist while it has room and to the global queue (q,
// counted at count) after that; called by every lane of the wave.
__device__ __forceinline__ void kc_push_items(const KcLocal& loc, uint32_t nch, uint32_t v,
                                              unsigned long long* __restrict__ q, unsigned long long* count) {
  const unsigned long long incl = wave_inclusive_scan<unsigned long long>(nch);
  const uint32_t total = static_cast<uint32_t>(__shfl(incl, kWave - 1, kWave));
  if (total == 0) return;
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(loc.tail, total);
  base = __shfl(base, 0, kWave);
  const uint32_t first = base + static_cast<uint32_t>(incl) - nch;
  uint32_t spill = 0;
  for (uint32_t c = 0; c < nch; ++c) {
    if (first + c < kKcLocal)
      loc.work[first + c] = (static_cast<unsigned long long>(c) << 32) | v;
    else
      ++spill;
  }
  if (base + total <= kKcLocal) return;  // all of the wave's items stayed local (uniform)
  // the items past the worklist's room, this lane's last `spill` chunks
  const uint32_t c0 = nch - spill;
  if (spill) {
    const unsigned long long g = atomicAdd(count, static_cast<unsigned long long>(spill));
    for (uint32_t c = 0; c < spill; ++c) q[g + c] = (static_cast<unsigned long long>(c0 + c) << 32) | v;
  }
}

// One peeling pass over the items [first, count) of q, `stride` waves
// apart: every unpeeled neighbour loses one degree; the decrement that
// lands on k claims it (its items go to the workgroup's worklist, or to
// qnext when that is full), and a far vertex that drops to the bound is
// appended at cross (counted at ccount).
__device__ __forceinline__ void kc_peel(const KcArgs& a, uint32_t k, int32_t bound,
                                        const unsigned long long* __restrict__ q, uint64_t count, uint64_t first,
                                        uint64_t stride, const KcLocal& loc, unsigned long long* qnext,
                                        unsigned long long* qcount, uint32_t* cross, unsigned long long* ccount,
                                        KcStage& sb1, uint32_t& nb1) {
  const unsigned lane = lane_id();
  for (uint64_t i = first; i < count; i += stride) {
    // one item: kKcChunk entries of a row, four arena loads in flight per
    // lane
    const unsigned long long item = q[i];
    const uint2 rb = a.row[static_cast<uint32_t>(item)];
    const uint32_t j0 = static_cast<uint32_t>(item >> 32) * kKcChunk;
    uint32_t x[kKcUnroll];
#pragma unroll
    for (uint32_t t = 0; t < kKcUnroll; ++t) {
      const uint32_t j = j0 + t * kWave + lane;
      x[t] = j < rb.y ? a.arena[rb.x + j] : kTomb;  // past the row: skipped like a tombstone
    }
#pragma unroll
    for (uint32_t t = 0; t < kKcUnroll; ++t) {
      uint32_t nch = 0;
      bool crossed = false;
      const uint32_t v = x[t] & ~kTomb;
      if (!(x[t] & kTomb) && a.core[v] == kKcUnset) {
        // unpeeled vertices hold degree > k during level k, so exactly
        // one decrement -- the one that lands on k -- claims v
        const uint32_t after = atomicSub(&a.deg[v], 1u) - 1u;
        if (after == k) {
          a.core[v] = k;
          nch = kc_chunks(a.row[v].y);
        } else {
          crossed = static_cast<long long>(after) == bound;
        }
      }
      kc_push_items(loc, nch, v, qnext, qcount);
      kc_stage(sb1, nb1, crossed, v, cross, ccount);
    }
  }
}

__global__ __launch_bounds__(kKcBlock) void k_kc_step(uint32_t step, KcArgs a) {
  const KcRec& prev = a.st->rec[(step + 3) & 3];
  const KcPlan p = kc_next(prev);
  KcRec& cur = a.st->rec[step & 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cur.phase = p.phase;
    cur.k = p.k;
    cur.round = p.round;
    cur.nl = p.nl;
    cur.fl = p.fl;
    cur.split = p.split;
    cur.bound = p.bound;
    cur.in_count = p.in_count;
    cur.level_rem = p.level_rem;
    cur.far_count = p.far_count;
    KcRec& nxt = a.st->rec[(step + 1) & 3];  // last read by step - 2
    nxt.queued = 0;
    nxt.remaining = 0;
    nxt.far_rem = 0;
    nxt.min_deg = ~0ull;
    nxt.crossed = 0;
    if (prev.phase == kKcPhInit) a.st->marked = prev.remaining;  // the region, counted by the init
    if (prev.phase == kKcPhScan && prev.queued != 0) a.st->levels += 1;
    if (p.phase == kKcPhScan) a.st->scans += 1;
    if (p.phase == kKcPhPeel) a.st->rounds += 1;
  }
  // workgroups past the step's input have nothing to do (a uniform exit:
  // no barrier below is left waiting)
  if (p.phase == kKcPhDone) return;
  if (p.phase == kKcPhScan &&
      blockIdx.x * static_cast<uint64_t>(blockDim.x) >= p.in_count + (p.split ? p.far_count : 0ull))
    return;
  if (p.phase == kKcPhPeel && blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) >= p.in_count) return;
  __shared__ KcItems s_items[kKcBlock / kWave];
  __shared__ KcStage s_stage[2][kKcBlock / kWave];
  __shared__ unsigned long long s_flush[kKcBlock / kWave + 1];
  KcItems& sb0 = s_items[threadIdx.x / kWave];
  KcStage& sb1 = s_stage[0][threadIdx.x / kWave];
  KcStage& sb2 = s_stage[1][threadIdx.x / kWave];
  uint32_t nb0 = 0, nb1 = 0, nb2 = 0;
  if (p.phase == kKcPhScan) {
    // the near list the previous scan (and peeling) wrote, then, when the
    // level passed the bound, the far list
    const uint32_t* near_in = a.near[p.nl ^ 1u];
    uint32_t* near_out = a.near[p.nl];
    const uint32_t* far_in = a.far[p.fl];
    uint32_t* far_out = a.far[p.fl ^ 1u];
    const long long nb = static_cast<long long>(p.k) + kKcWindow;
    const uint64_t total = p.in_count + (p.split ? p.far_count : 0ull);
    unsigned long long mn = ~0ull;
    for (uint64_t i0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); i0 < total;
         i0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
      const uint64_t i = i0 + threadIdx.x;
      bool keep_near = false, keep_far = false;
      uint32_t v = 0, nch = 0;
      if (i < total) {
        const bool from_far = i >= p.in_count;
        v = from_far ? far_in[i - p.in_count] : near_in[i];
        if (a.core[v] == kKcUnset) {
          const uint32_t d = a.deg[v];
          if (from_far && static_cast<long long>(d) <= p.bound) {
            // crossed the bound while peeling: it is on a near list
          } else if (d <= p.k) {
            a.core[v] = p.k;
            nch = kc_chunks(a.row[v].y);
          } else {
            keep_near = !p.split || static_cast<long long>(d) <= nb;
            keep_far = !keep_near;
            mn = min(mn, static_cast<unsigned long long>(d));
          }
        }
      }
      kc_stage_items(sb0, nb0, nch, v, a.q[0], &cur.queued);
      kc_stage(sb1, nb1, keep_near, v, near_out, &cur.remaining);
      kc_stage(sb2, nb2, keep_far, v, far_out, &cur.far_rem);
    }
    kc_flush_items_block(sb0, nb0, a.q[0], &cur.queued, s_flush);
    kc_flush_block(sb1, nb1, near_out, &cur.remaining, s_flush);
    kc_flush_block(sb2, nb2, far_out, &cur.far_rem, s_flush);
    mn = ~wave_max(~mn);  // the wave's minimum
    __shared__ unsigned long long s_mn[kKcBlock / kWave];
    if (lane_id() == 0) s_mn[threadIdx.x / kWave] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t w = 1; w < kKcBlock / kWave; ++w) mn = min(mn, s_mn[w]);
      if (mn != ~0ull) atomicMin(&cur.min_deg, mn);
    }
  } else if (p.phase == kKcPhPeel) {
    // The step's items (the scan's claims, or the claims the last step
    // could not keep), grid-stride; then every workgroup peels its own
    // claims from LDS in rounds separated by workgroup barriers, until none
    // are left.  Peeling within a level is order-free (each vertex is
    // claimed by the one decrement that lands on k, whichever order the
    // decrements come in), so a level usually ends within one step.
    __shared__ unsigned long long s_work[kKcLocal];
    __shared__ uint32_t s_tail, s_snap;
    if (threadIdx.x == 0) s_tail = 0;
    __syncthreads();
    const unsigned long long* q = a.q[p.round & 1u];
    unsigned long long* qnext = a.q[(p.round + 1u) & 1u];
    // far vertices that drop to the bound join the near list after its
    // current entries
    uint32_t* cross = a.near[p.nl] + p.level_rem;
    const uint32_t wv = threadIdx.x / kWave, waves = kKcBlock / kWave;
    const uint64_t W = static_cast<uint64_t>(gridDim.x) * waves;
    KcLocal loc{s_work, &s_tail};
    kc_peel(a, p.k, p.bound, q, p.in_count, blockIdx.x * static_cast<uint64_t>(waves) + wv, W, loc, qnext,
            &cur.queued, cross, &cur.crossed, sb1, nb1);
    uint32_t h = 0;
    for (;;) {
      __syncthreads();  // the claims appended so far are in LDS
      if (threadIdx.x == 0) s_snap = min(s_tail, kKcLocal);
      __syncthreads();
      const uint32_t t = s_snap;  // one snapshot for the whole workgroup
      if (h >= t) break;
      kc_peel(a, p.k, p.bound, s_work + h, t - h, wv, waves, loc, qnext, &cur.queued, cross, &cur.crossed, sb1,
              nb1);
      h = t;
    }
    kc_flush_block(sb1, nb1, cross, &cur.crossed, s_flush);
  }
}

// Buffers for n vertices: the core numbers of the last computation and the
// region marks are kept (new marks start at 0, below every epoch).
static void kc_ensure(vgx_kcore* kc, uint64_t n) {
  hipStream_t s = kc->g->stream;
  kc->deg.ensure(n + 1);
  kc->core.grow_copy(n + 1, kc->n, s);
  kc->mark.grow_copy(n + 1, kc->mark_valid, s);
  if (kc->mark_valid < kc->mark.n) {
    VGX_HIP(hipMemsetAsync(kc->mark.p + kc->mark_valid, 0, (kc->mark.n - kc->mark_valid) * sizeof(uint32_t), s));
    kc->mark_valid = kc->mark.n;
  }
  for (auto* b : {&kc->fq[0], &kc->fq[1], &kc->near[0], &kc->near[1], &kc->far[0], &kc->far[1]}) b->ensure(n + 1);
  // a claimed vertex queues one item per kKcChunk stored row entries
  const uint64_t items = n + kc->g->adj[0].stored / kKcChunk + 64;
  kc->items[0].ensure(items);
  kc->items[1].ensure(items);
}

// The level loop over the region: all vertices (lo = 0) or those marked at
// or above lo.  n_old: the vertex count the core numbers were computed for.
static void kc_run(vgx_kcore* kc, uint32_t lo, uint64_t n_old) {
  vgx_graph* g = kc->g;
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  KcState init{};
  init.rec[3].phase = kKcPhInit;
  init.rec[0].min_deg = ~0ull;
  std::memcpy(kc->h_st, &init, sizeof(init));
  VGX_HIP(hipMemcpyAsync(kc->d_st, kc->h_st, sizeof(KcState), hipMemcpyHostToDevice, s));
  const Adjacency& adj = g->adj[0];
  // the init writes far list 0, which the first scan splits
  k_kc_init<<<kc_blocks(n, kKcBlock, 4096), kKcBlock, 0, s>>>(adj.deg.p, n, n_old, lo, kc->mark.p, kc->deg.p,
                                                              kc->core.p, kc->far[0].p, kc->d_st);
  VGX_LAUNCH_CHECK();
  KcArgs a{adj.row.p,
           adj.arena.p,
           kc->deg.p,
           kc->core.p,
           {kc->items[0].p, kc->items[1].p},
           {kc->near[0].p, kc->near[1].p},
           {kc->far[0].p, kc->far[1].p},
           kc->d_st};
  // one grid for every step: a wave per queued vertex, grid-stride
  const unsigned blocks = kc_blocks(n, kKcBlock / kWave, 1024);
  uint64_t step = 0, chunk = std::max<uint64_t>(64, kc->steps_hint);
  // each step claims a vertex, ends a level or jumps to a higher one
  const uint64_t limit = 3 * n + 64;
  for (;;) {
    for (uint64_t i = 0; i < chunk; ++i, ++step) {
      k_kc_step<<<blocks, kKcBlock, 0, s>>>(static_cast<uint32_t>(step), a);
      VGX_LAUNCH_CHECK();
    }
    VGX_HIP(hipMemcpyAsync(kc->h_st, kc->d_st, sizeof(KcState), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    if (kc->h_st->rec[(step - 1) & 3].phase == kKcPhDone) break;
    if (step > limit) throw Error(VGX_DEVICE_ERROR, "k-core level loop did not finish");
    chunk = std::min<uint64_t>(2 * chunk, 4096);
  }
  const KcState& h = *kc->h_st;
  kc->levels = static_cast<uint32_t>(h.levels);
  kc->rounds = static_cast<uint32_t>(h.rounds);
  kc->steps_hint = h.scans + h.rounds + 2;
  kc->last_repaired = h.marked;
  kc->n = n;
}

static void kc_recompute(vgx_kcore* kc) {
  const uint64_t n = kc->g->n;
  kc_ensure(kc, n);
  kc->levels = kc->rounds = 0;
  kc->last_repaired = 0;
  if (n == 0) {
    kc->n = 0;
    return;
  }
  kc_run(kc, 0, 0);
}

// Flood fill from the endpoints (device copy in kc->ends) over the graph's
// current rows, marking with `epoch`.
static void kc_flood(vgx_kcore* kc, uint64_t nends, uint32_t epoch) {
  vgx_graph* g = kc->g;
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  if (n == 0 || nends == 0) return;
  VGX_HIP(hipMemsetAsync(kc->d_st->flood, 0, sizeof(kc->d_st->flood), s));
  k_kc_flood_seed<<<kc_blocks(nends, kKcBlock, 1024), kKcBlock, 0, s>>>(kc->ends.p, nends, n, epoch, kc->mark.p,
                                                                        kc->fq[0].p, kc->d_st);
  VGX_LAUNCH_CHECK();
  const Adjacency& adj = g->adj[0];
  const unsigned blocks = kc_blocks(n, kKcBlock / kWave, 1024);
  // each vertex joins once per flood, so at most n rounds have a queue
  for (uint32_t r = 0;;) {
    for (uint32_t i = 0; i < kKcFloodPerCheck; ++i, ++r) {
      k_kc_flood<<<blocks, kKcBlock, 0, s>>>(r, epoch, adj.row.p, adj.arena.p, kc->fq[r & 1].p, kc->mark.p,
                                             kc->fq[(r + 1) & 1].p, kc->d_st);
      VGX_LAUNCH_CHECK();
    }
    unsigned long long ring[3] = {0, 0, 0};
    VGX_HIP(hipMemcpyAsync(ring, kc->d_st->flood, sizeof(ring), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    const uint64_t c = ring[r % 3];
    if (c > n) throw Error(VGX_DEVICE_ERROR, "k-core flood queue exceeded the vertex count");
    if (c == 0) break;
    if (r > n + kKcFloodPerCheck) throw Error(VGX_DEVICE_ERROR, "k-core flood did not finish");
  }
}

// kcore.hpp:31-56: the region is marked on the graph before and after the
// batch, then re-peeled.
static void kc_apply(vgx_kcore* kc, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k) {
  vgx_graph* g = kc->g;
  hipStream_t s = g->stream;
  if (kc->epoch > 0xFFFFFFF0u) {  // marks restart below every future epoch
    if (kc->mark_valid) VGX_HIP(hipMemsetAsync(kc->mark.p, 0, kc->mark_valid * sizeof(uint32_t), s));
    kc->epoch = 0;
  }
  const uint32_t e_pre = kc->epoch + 1, e_post = kc->epoch + 2;
  kc->epoch += 2;
  kc->ends.ensure(2 * k + 1);
  VGX_HIP(hipMemcpyAsync(kc->ends.p, src, k * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  VGX_HIP(hipMemcpyAsync(kc->ends.p + k, dst, k * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  const uint64_t n_old = kc->n;
  kc_ensure(kc, g->n);
  kc_flood(kc, 2 * k, e_pre);
  const vgx_status st = vgx_graph_apply(g, src, dst, add, k);
  if (st != VGX_OK) throw Error(st, std::string(vgx_last_error()));
  kc_ensure(kc, g->n);
  kc_flood(kc, 2 * k, e_post);
  kc_run(kc, e_pre, n_old);
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

vgx_status vgx_kcore_destroy(vgx_kcore* kc) {
  VGX_API_BEGIN
  if (!kc) return VGX_OK;
  if (kc->g) (void)hipStreamSynchronize(kc->g->stream);
  if (kc->d_st) dev_free(kc->d_st);
  if (kc->h_st) (void)hipHostFree(kc->h_st);
  delete kc;
  VGX_API_END
}

vgx_status vgx_kcore_create(vgx_graph* g, vgx_kcore** out) {
  VGX_API_BEGIN
  if (!g || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (g->directed)
    throw Error(VGX_INVALID_ARGUMENT,
                "IncrementalKCore requires an undirected graph; directed k-core semantics must be selected explicitly");
  VGX_HIP(hipSetDevice(g->device));
  auto* kc = new vgx_kcore();
  try {
    kc->g = g;
    kc->d_st = static_cast<KcState*>(dev_alloc(sizeof(KcState)));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&kc->h_st), sizeof(KcState), hipHostMallocDefault));
    kc_recompute(kc);
  } catch (...) {
    vgx_kcore_destroy(kc);
    throw;
  }
  *out = kc;
  VGX_API_END
}

vgx_status vgx_kcore_recompute(vgx_kcore* kc) {
  VGX_API_BEGIN
  if (!kc) throw Error(VGX_INVALID_ARGUMENT, "null k-core");
  VGX_HIP(hipSetDevice(kc->g->device));
  kc_recompute(kc);
  VGX_API_END
}

vgx_status vgx_kcore_apply(vgx_kcore* kc, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k) {
  VGX_API_BEGIN
  if (!kc || (k && (!src || !dst || !add))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(kc->g->device));
  if (k == 0) {
    kc->last_repaired = 0;
    return VGX_OK;
  }
  kc_apply(kc, src, dst, add, k);
  VGX_API_END
}

vgx_status vgx_kcore_core(vgx_kcore* kc, uint32_t* out, uint64_t cap, uint64_t* count) {
  VGX_API_BEGIN
  if (!kc || !count) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(kc->g->device));
  *count = kc->n;
  const uint64_t c = std::min(cap, kc->n);
  if (out && c) {
    VGX_HIP(hipMemcpyAsync(out, kc->core.p, c * sizeof(uint32_t), hipMemcpyDeviceToHost, kc->g->stream));
    VGX_HIP(hipStreamSynchronize(kc->g->stream));
  }
  VGX_API_END
}

vgx_status vgx_kcore_stats(vgx_kcore* kc, uint64_t* last_repaired_vertices, uint32_t* levels, uint32_t* rounds) {
  VGX_API_BEGIN
  if (!kc) throw Error(VGX_INVALID_ARGUMENT, "null k-core");
  if (last_repaired_vertices) *last_repaired_vertices = kc->last_repaired;
  if (levels) *levels = kc->levels;
  if (rounds) *rounds = kc->rounds;
  VGX_API_END
}

}  // extern "C"
