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
// degree is at most k (core number k) into a queue, and copies the others to
// the next scan's list with the smallest of their degrees; then rounds of
// peeling, one wave per queued vertex, decrement the remaining degree of
// each unpeeled neighbour and claim (compare-and-swap on its core slot) the
// ones that drop to k or below into the next queue.  A level whose scan
// claims nothing jumps straight to the smallest remaining degree.  Every
// arc is decremented at most once, so degrees never wrap.
//
// The level loop is a device-steered step program: launch s reads launch
// s - 1's plan and results from a ring of four records and decides its own
// step (a scan, a peeling round, or nothing once done), so the host enqueues
// launches in chunks and reads the state back once per chunk.
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
// Flood-fill rounds run several per host check, on a ring of three counters
// like the peeling rounds of the step program's queues.
constexpr uint32_t kKcFloodPerCheck = 4;

// One step of the level loop.  The plan fields are written by the step's
// first thread; the counters, zeroed by the step before, are added to by
// every workgroup of the step.
struct KcRec {
  uint32_t phase, k, round;
  uint32_t list;                   // scan: the list it writes (it reads the other one);
                                   // peeling round: the list the next scan writes
  unsigned long long in_count;     // entries of this step's input (queue or list)
  unsigned long long level_rem;    // vertices the level's scan left on its list
  unsigned long long queued;       // claims appended to this step's output queue
  unsigned long long remaining;    // scan: vertices copied to the next list
  unsigned long long min_deg;      // scan: the smallest remaining degree among them
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
  vgx::DevBuf<uint32_t> deg, core, mark, q[2], list[2], ends;
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
      kc_push(join, v, qnext, next);
    }
  }
}

// The level loop's start: every vertex of the region (mark >= lo, or all
// vertices when lo is 0) takes its live degree, loses its core number and
// joins the first scan's list (counted in rec[3].remaining, the "previous
// step" of launch 0); a vertex added since the last computation that is not
// in the region has no arcs and core number 0.
__global__ void k_kc_init(const uint32_t* __restrict__ live_deg, uint64_t n, uint64_t n_old, uint32_t lo,
                          const uint32_t* __restrict__ mark, uint32_t* __restrict__ deg, uint32_t* __restrict__ core,
                          uint32_t* __restrict__ list, KcState* __restrict__ st) {
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
    kc_push(in, static_cast<uint32_t>(v), list, &st->rec[3].remaining);
    const unsigned long long m = __ballot(in);
    if (lane_id() == 0 && m) atomicAdd(&st->marked, static_cast<unsigned long long>(__popcll(m)));
  }
}

// ------------------------------------------------------------ step program

struct KcPlan {
  uint32_t phase, k, round, list;
  unsigned long long in_count, level_rem;
};

// The step after `p` (its plan and results); shared by every workgroup.
__device__ __forceinline__ KcPlan kc_next(const KcRec& p) {
  KcPlan x{kKcPhDone, p.k, 0u, p.list, 0ull, 0ull};
  if (p.phase == kKcPhInit) {
    x = {kKcPhScan, 0u, 0u, 0u, p.remaining, 0ull};
    if (p.remaining == 0) x.phase = kKcPhDone;
  } else if (p.phase == kKcPhScan) {
    if (p.queued == 0) {
      // nothing at this level: the next level with vertices
      if (p.remaining != 0)
        x = {kKcPhScan, static_cast<uint32_t>(p.min_deg), 0u, p.list ^ 1u, p.remaining, 0ull};
    } else {
      x = {kKcPhPeel, p.k, 0u, p.list ^ 1u, p.queued, p.remaining};
    }
  } else if (p.phase == kKcPhPeel) {
    if (p.queued != 0) x = {kKcPhPeel, p.k, p.round + 1u, p.list, p.queued, p.level_rem};
    else if (p.level_rem != 0) x = {kKcPhScan, p.k + 1u, 0u, p.list, p.level_rem, 0ull};
  }
  return x;
}

struct KcArgs {
  const uint2* row;
  const uint32_t* arena;
  uint32_t* deg;
  uint32_t* core;
  uint32_t* q[2];
  uint32_t* list[2];
  KcState* st;
};

__global__ __launch_bounds__(kKcBlock) void k_kc_step(uint32_t step, KcArgs a) {
  const KcRec& prev = a.st->rec[(step + 3) & 3];
  const KcPlan p = kc_next(prev);
  KcRec& cur = a.st->rec[step & 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cur.phase = p.phase;
    cur.k = p.k;
    cur.round = p.round;
    cur.list = p.list;
    cur.in_count = p.in_count;
    cur.level_rem = p.level_rem;
    KcRec& nxt = a.st->rec[(step + 1) & 3];  // last read by step - 2
    nxt.queued = 0;
    nxt.remaining = 0;
    nxt.min_deg = ~0ull;
    if (prev.phase == kKcPhScan && prev.queued != 0) a.st->levels += 1;
    if (p.phase == kKcPhScan) a.st->scans += 1;
    if (p.phase == kKcPhPeel) a.st->rounds += 1;
  }
  if (p.phase == kKcPhScan) {
    // the scan reads the list the previous scan (or the init) wrote
    const uint32_t* in = a.list[p.list ^ 1u];
    uint32_t* out = a.list[p.list];
    unsigned long long mn = ~0ull;
    for (uint64_t i0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); i0 < p.in_count;
         i0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
      const uint64_t i = i0 + threadIdx.x;
      bool claim = false, keep = false;
      uint32_t v = 0;
      if (i < p.in_count) {
        v = in[i];
        if (a.core[v] == kKcUnset) {
          const uint32_t d = a.deg[v];
          if (d <= p.k) {
            a.core[v] = p.k;
            claim = true;
          } else {
            keep = true;
            mn = min(mn, static_cast<unsigned long long>(d));
          }
        }
      }
      kc_push(claim, v, a.q[0], &cur.queued);
      kc_push(keep, v, out, &cur.remaining);
    }
    mn = ~wave_max(~mn);  // the wave's minimum
    if (lane_id() == 0 && mn != ~0ull) atomicMin(&cur.min_deg, mn);
  } else if (p.phase == kKcPhPeel) {
    const uint32_t* q = a.q[p.round & 1u];
    uint32_t* qnext = a.q[(p.round + 1u) & 1u];
    const uint64_t W = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
    const unsigned lane = lane_id();
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; i < p.in_count;
         i += W) {
      const uint2 rb = a.row[q[i]];
      // uniform across the wave; four arena loads in flight per lane (hub
      // rows are bound by load latency)
      for (uint32_t j0 = 0; j0 < rb.y; j0 += kKcUnroll * kWave) {
        uint32_t x[kKcUnroll];
#pragma unroll
        for (uint32_t t = 0; t < kKcUnroll; ++t) {
          const uint32_t j = j0 + t * kWave + lane;
          x[t] = j < rb.y ? a.arena[rb.x + j] : kTomb;  // past the row: skipped like a tombstone
        }
#pragma unroll
        for (uint32_t t = 0; t < kKcUnroll; ++t) {
          bool claim = false;
          const uint32_t v = x[t] & ~kTomb;
          if (!(x[t] & kTomb) && a.core[v] == kKcUnset) {
            const uint32_t before = atomicSub(&a.deg[v], 1u);
            if (before - 1u <= p.k) claim = atomicCAS(&a.core[v], kKcUnset, p.k) == kKcUnset;
          }
          kc_push(claim, v, qnext, &cur.queued);
        }
      }
    }
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
  for (auto* b : {&kc->q[0], &kc->q[1], &kc->list[0], &kc->list[1]}) b->ensure(n + 1);
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
  // the init writes list[1]: the first scan has list 0 and reads list 0 ^ 1
  k_kc_init<<<kc_blocks(n, kKcBlock, 4096), kKcBlock, 0, s>>>(adj.deg.p, n, n_old, lo, kc->mark.p, kc->deg.p,
                                                              kc->core.p, kc->list[1].p, kc->d_st);
  VGX_LAUNCH_CHECK();
  KcArgs a{adj.row.p, adj.arena.p, kc->deg.p, kc->core.p, {kc->q[0].p, kc->q[1].p}, {kc->list[0].p, kc->list[1].p},
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
                                                                        kc->q[0].p, kc->d_st);
  VGX_LAUNCH_CHECK();
  const Adjacency& adj = g->adj[0];
  const unsigned blocks = kc_blocks(n, kKcBlock / kWave, 1024);
  // each vertex joins once per flood, so at most n rounds have a queue
  for (uint32_t r = 0;;) {
    for (uint32_t i = 0; i < kKcFloodPerCheck; ++i, ++r) {
      k_kc_flood<<<blocks, kKcBlock, 0, s>>>(r, epoch, adj.row.p, adj.arena.p, kc->q[r & 1].p, kc->mark.p,
                                             kc->q[(r + 1) & 1].p, kc->d_st);
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
