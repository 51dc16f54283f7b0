// Direction-optimizing BFS and exact incremental BFS repair for gfx950.
//
// Reference: include/velographx/incremental/bfs.hpp (BasicIncrementalBFS:
// recompute :100-120, apply :29-98, compute_affected_prebatch :222-259,
// repair_affected :271-319, propagate_decreases :341-354) and the static
// bfs_distances in src/algorithms.cpp:11-27.  BFS distances are unique, so
// every traversal order below yields the reference's vector bit for bit.
//
// Full BFS: one launch per step on a fixed grid, where a step is a level,
// the tail (phase B) of a bottom-up level or the pack after a top-down
// level that claimed by distance.  Each workgroup derives what its step is
// and the level's direction from what the previous step recorded (Beamer's
// rule: top-down -> bottom-up when m_f > m_u / alpha, bottom-up -> top-down
// when n_f < n_u / beta, with n_u the vertices with in-arcs still
// unvisited, which is what a bottom-up pass checks: vertices without
// in-arcs start visited); no workgroup waits for another, no decision is
// taken from an earlier traversal, and the host only checks for completion
// after the number of steps this graph's traversals have needed.
//   top-down : the frontier is a packed queue (row begin, edge prefix); every
//              wave takes an equal slice of the frontier's edges (load-
//              balanced search), claims new vertices with one atomicOr on
//              the visited bitmap and pushes them wave-aggregated.  Before an
//              expected large bottom-up level no queue is built: the claim
//              is the distance alone and a pack step derives the frontier.
//   bottom-up: every wave owns 64 consecutive vertices (one visited word);
//              unvisited lanes check the first parents from the head array,
//              then the rest of their in-row (inline or in a tail step);
//              one ballot gives the next frontier word.
//
// Incremental repair (apply):
//   1. deletions are applied first (tombstones), then vertices that lost
//      every live shortest parent are found level-asynchronously with
//      counters: support[v] = live in-edges from d(v)-1, lost[v] = losses
//      from affected parents; v is affected once lost[v] >= support[v].
//      This is the reference's support-count fixpoint (bfs.hpp:222-259)
//      with the deleted edges removed from both sides of the comparison.
//   2. additions are applied; past the fallback fraction -> full BFS.
//   3. affected vertices are reset, seeded from their unaffected
//      in-neighbours, the final additions are relaxed, and decreases are
//      propagated in Bellman-Ford rounds (each round one edge-balanced
//      launch).  The fixpoint satisfies d(w) <= d(u)+1 on every arc with
//      realisable upper bounds, hence equals the BFS distances.
#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "policy.hpp"
#include "vgx_internal.hpp"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);
extern thread_local std::string t_last_error;

constexpr int kBfsBlock = 256;
constexpr uint32_t kWavesPerBlock = kBfsBlock / kWave;
// Bottom-up phase A checks the first kHead parents of a vertex from the
// dense head array; the rest of longer rows is scanned by phase B.
constexpr uint32_t kHead = 4;
// Head entry past the end of a shorter row (tombstone bit set, so it is
// never in a frontier; distinct from any tombstoned arc, since the graph
// accepts arc endpoints below 2^31 - 1 only).
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;
// How a top-down level claims new vertices: with an atomic on the visited
// bitmap and a queue push (kClaimQueue), or by storing the distance alone,
// racing stores of the same value, after which a pack step derives the
// frontier (kClaimDistance).
enum Claims : uint32_t { kClaimQueue = 0, kClaimDistance = 2 };
// The count of zero-in-degree bits is added by every workgroup of
// k_build_zin into one of kZinStripes partial sums, kStripeWords apart (one
// 128-byte line each): additions to one address serialise.
constexpr uint32_t kZinStripes = 16;
constexpr uint32_t kStripeWords = 16;
// Default direction switch (Beamer's alpha 15, beta 18), automatic grid and
// chunking, tail and claim ratios of the device rules.
constexpr vgx_bfs_tuning kDefaultTuning{15.0, 18.0, 0, 0, 32.0, 5.0};
// Words of the frontier bitmap one wave of a pack step derives per loop.
constexpr uint32_t kPackWords = 16;

// Per-level record, kept in a ring of 4 indexed by level & 3.  Level L's
// kernel accumulates the frontier of level L+1 into rec[(L+1)&3] (q, found)
// and workgroup 0 hands over the direction and visited-edge sum it used;
// rec[(L+2)&3] is cleared for level L+1.  Every workgroup of level L
// derives the same direction from rec[L&3], so no workgroup waits for
// another and no grid-wide fence is needed within a level.
// The atomically updated counters and the leader-written hand-over live in
// separate 128-byte lines.
constexpr uint32_t kLogLevels = 64;

// The frontier size is added by every workgroup; it is striped over
// kFoundStripes cache lines (workgroup b adds to stripe b % kFoundStripes)
// so that a level's final additions do not serialise on one address.
constexpr uint32_t kFoundStripes = 16;
struct alignas(128) FoundStripe {
  unsigned long long v;
};
struct alignas(128) LevelRec {
  unsigned long long q;                   // packed (count << 38 | edges) of the queued frontier
  FoundStripe found[kFoundStripes];       // frontier size, zero-length rows included
};

__device__ __forceinline__ unsigned long long found_total(const LevelRec& r) {
  unsigned long long t = 0;
#pragma unroll
  for (uint32_t i = 0; i < kFoundStripes; ++i) t += r.found[i].v;
  return t;
}
struct alignas(128) LevelCtl {
  unsigned long long visited_edges;  // edges of vertices visited before this level
  // vertices with live in-arcs still unvisited when the previous level
  // started (this level's frontier is still counted in it)
  unsigned long long unvisited;
  uint32_t prev_mode;                // direction used by the previous level
  uint32_t no_queue;                 // the previous level built no frontier queue
};

// Rows a bottom-up level leaves for phase B (same ring).
struct alignas(128) PendCtr {
  unsigned long long q;
};

// A traversal is a sequence of steps, one launch each: a level, the tail
// (phase B) of a bottom-up level that left rows pending, or the pack after
// a top-down level that claimed by distance.  Step i reads what step i - 1
// did from srec[i & 3] (written by its leader) and the counters it left,
// and its own leader records step i in srec[(i + 1) & 3], a slot no
// workgroup of step i reads.
enum Phase : uint32_t { kPhStart = 0, kPhLevel = 1, kPhTail = 2, kPhPack = 3, kPhDone = 4 };
struct alignas(128) StepRec {
  uint32_t phase;
  uint32_t level;
  uint32_t mode;         // direction of a level step (0, 1, 2 as in bfs_level)
  uint32_t claims;       // claim mode of a level step (enum Claims)
  uint32_t inline_tail;  // a bottom-up level step finished its rows itself
};
struct StepPlan {
  uint32_t phase;
  uint32_t level;
};
// The step after `prev`; `pend_q` is the pending-row counter of prev's level.
__host__ __device__ inline StepPlan next_step(const StepRec& prev, unsigned long long pend_q) {
  switch (prev.phase) {
    case kPhStart:
      return {kPhLevel, 0u};
    case kPhLevel:
      if (prev.mode == 1 && !prev.inline_tail && pend_q != 0) return {kPhTail, prev.level};
      if (prev.mode != 1 && prev.claims == kClaimDistance) return {kPhPack, prev.level};
      return {kPhLevel, prev.level + 1};
    case kPhTail:
    case kPhPack:
      return {kPhLevel, prev.level + 1};
    default:
      return {kPhDone, prev.level};
  }
}
constexpr uint32_t kLogSteps = 96;

struct BfsState {
  LevelRec rec[4];
  LevelCtl ctl[4];
  PendCtr pend[4];
  StepRec srec[4];
  unsigned long long reached;
  unsigned long long total_edges;
  uint32_t done_level;  // first level with an empty frontier (~0u while running)
  uint32_t bu_levels;
  uint32_t n;
  uint32_t pad;
  float alpha;
  float beta;
  float tail_ratio;   // bottom-up rows go to a tail step when n_f * tail_ratio < n_u
  float claim_ratio;  // top-down claims by distance when m_f * claim_ratio >= n_u
  // phase << 24 | level of every step (steps < kLogSteps), for diagnostics
  uint32_t log_step[kLogSteps];
  // per-level record for diagnostics (levels < kLogLevels)
  uint32_t log_mode[kLogLevels];
  unsigned long long log_nf[kLogLevels];
  unsigned long long log_mf[kLogLevels];
  unsigned long long log_pend[kLogLevels];  // packed count << 38 | edges left for phase B
};

__device__ __forceinline__ void write_step(BfsState* st, uint32_t step, uint32_t phase, uint32_t level, uint32_t mode,
                                           uint32_t claims, uint32_t inline_tail) {
  StepRec& r = st->srec[(step + 1) & 3];
  r.phase = phase;
  r.level = level;
  r.mode = mode;
  r.claims = claims;
  r.inline_tail = inline_tail;
  if (step < kLogSteps) st->log_step[step] = (phase << 24) | (level & 0xFFFFFFu);
}

// Counters of the incremental repair.  The lists of the invalidation
// rounds come in two parities (round r uses r & 1) and the repair frontier
// in a ring of three, so that a kernel can clear the counters the next
// round appends to while this round's are still being read: the host
// launches several rounds between synchronisations and every kernel takes
// its list sizes from here.
struct IncState {
  unsigned long long tq[2];     // touched list (plain count)
  unsigned long long tfq[2];    // first-touched list, packed (in-rows)
  unsigned long long aq[2];     // newly affected, packed (out-rows)
  unsigned long long aall;      // all affected (plain count)
  unsigned long long tall;      // all touched (plain count)
  unsigned long long candidates;
  unsigned long long fq[3];     // repair frontier, packed (out-rows)
  unsigned long long sq;        // seed list, packed (in-rows)
  unsigned long long became_reachable;
  unsigned long long relaxed;
  uint32_t fallback;            // set once more than `limit` vertices are affected
  uint32_t pad;
};

enum : uint32_t { kFAffected = 1u, kFSupport = 2u, kFCand = 4u };

struct BfsWork {
  DevBuf<uint32_t> dist;
  DevBuf<unsigned long long> visited;
  DevBuf<unsigned long long> bm;  // 3 bitmaps
  DevBuf<uint32_t> queue[2];       // top-down frontier: row begins of its vertices
  DevBuf<unsigned long long> pre[2];
  DevBuf<uint32_t> pend;           // bottom-up rows left for phase B
  DevBuf<unsigned long long> ppre; // their edge prefixes
  DevBuf<uint4> head;
  DevBuf<unsigned long long> zin;  // vertices without live in-arcs (and ids >= n)
  DevBuf<unsigned long long> zcount;  // bits set in zin, in kZinStripes striped partial sums
  uint64_t head_version = ~0ull, head_n = 0;
  // incremental
  DevBuf<uint32_t> lost, support, flags, stamp;
  DevBuf<uint32_t> tlist, tflist, alist, aall, tall, slist;
  DevBuf<unsigned long long> tfpre, apre, spre;
  uint64_t cap = 0;
};

}  // namespace vgx

struct vgx_bfs {
  vgx_graph* g = nullptr;
  uint32_t source = 0;
  double fallback = 0.35;
  uint64_t dist_len = 0;
  uint64_t reachable = 0;
  uint64_t last_candidates = 0, last_affected = 0, last_relaxed = 0;
  bool last_full = false;
  uint32_t levels = 0, bu_levels = 0;
  vgx_bfs_tuning tune = vgx::kDefaultTuning;
  vgx::BfsWork w;
  vgx::BfsState* d_st = nullptr;
  vgx::BfsState* h_st = nullptr;
  vgx::IncState* d_inc = nullptr;
  vgx::IncState* h_inc = nullptr;
  uint32_t epoch = 0;
  int profiling = 0;  // 1: events around every step, 2: around every traversal
  std::vector<hipEvent_t> events;
  uint64_t level_launches = 0;
  double level_ms = 0.0;
  float last_level_ms[64] = {};  // per level of the last traversal, its steps summed (profiling)
  uint32_t steps = 0;            // steps the last traversal needed
  // whole-traversal GPU time (profiling mode 2): events around each traversal
  hipEvent_t tev[2] = {nullptr, nullptr};
  double traversal_ms = 0.0;
  uint64_t traversals = 0;
  vgx::ViewGuard views;  // zero-copy views of w.dist
};

namespace vgx {

// ---------------------------------------------------------------- full BFS

struct LevelArgs {
  const uint2* row;       // out-rows (top-down)
  const uint32_t* arena;
  const uint2* in_row;    // in-rows (bottom-up scans parents)
  const uint32_t* in_arena;
  uint32_t* dist;
  unsigned long long* visited;
  unsigned long long* bm;
  uint32_t* queue0;
  uint32_t* queue1;
  unsigned long long* pre0;
  unsigned long long* pre1;
  uint32_t* pend;            // bottom-up rows left for phase B
  unsigned long long* ppre;
  BfsState* st;
  uint32_t n;
  uint32_t words;
  const uint4* head;         // first kHead in-neighbours of every vertex
};

// head[v] = the first kHead entries of v's in-row (tombstone bits kept;
// missing entries are kNoEntry, so a head whose last entry is kNoEntry
// holds the whole row).  Also clears the count that k_build_zin,
// launched next on the same stream, adds to.
__global__ void k_build_head(const uint2* __restrict__ in_row, const uint32_t* __restrict__ in_arena, uint32_t n,
                             uint4* __restrict__ head, unsigned long long* __restrict__ zcount) {
  if (blockIdx.x == 0 && threadIdx.x < kZinStripes) zcount[threadIdx.x * kStripeWords] = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint2 rb = in_row[v];
    uint32_t h[kHead];
#pragma unroll
    for (uint32_t t = 0; t < kHead; ++t) h[t] = t < rb.y ? in_arena[rb.x + t] : kNoEntry;
    head[v] = make_uint4(h[0], h[1], h[2], h[3]);
  }
}

// Bit v set when v has no live in-arc (or v >= n).  Such a vertex can only
// be reached as the root, so every traversal starts with these bits already
// in the visited bitmap and the bottom-up sweeps skip them.
__global__ __launch_bounds__(kBfsBlock) void k_build_zin(const uint32_t* __restrict__ in_deg, uint32_t n,
                                                         uint32_t words, unsigned long long* __restrict__ zin,
                                                         unsigned long long* __restrict__ zcount) {
  const uint32_t W = gridDim.x * kWavesPerBlock;
  unsigned long long c = 0;
  for (uint32_t wi = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; wi < words; wi += W) {
    const uint32_t v = wi * kWave + lane_id();
    const unsigned long long m = __ballot(v >= n || in_deg[v] == 0);
    if (lane_id() == 0) zin[wi] = m;
    c += __popcll(m);
  }
  __shared__ unsigned long long s_c[kWavesPerBlock];
  if (lane_id() == 0) s_c[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t i = 0; i < kWavesPerBlock; ++i) t += s_c[i];
    if (t) atomicAdd(&zcount[(blockIdx.x % kZinStripes) * kStripeWords], t);
  }
}

__device__ __forceinline__ bool in_frontier(const unsigned long long* bm, uint32_t x) {
  return !(x & kTomb) && ((bm[x >> 6] >> (x & 63)) & 1ull);
}

// Whether a parent of v beyond the head is in the frontier: rows with up to
// 64 more entries on the lane itself, longer rows by the whole wave, first
// hit ends the scan.  Must be called by every lane of the wave (all lanes
// take part in the scan of a long row); `want` marks the lanes asking.
__device__ __forceinline__ bool rest_in_frontier(const LevelArgs& a, const unsigned long long* bm_cur, uint2 rb,
                                                 bool want) {
  bool found = false;
  const bool mine = want && rb.y - kHead <= kWave;
  if (mine)
    for (uint32_t j = kHead; j < rb.y; ++j)
      if (in_frontier(bm_cur, a.in_arena[rb.x + j])) {
        found = true;
        break;
      }
  unsigned long long lm = __ballot(want && !mine);
  while (lm) {
    const int l = __ffsll(static_cast<long long>(lm)) - 1;
    lm &= lm - 1;
    const uint32_t bx = __shfl(rb.x, l, kWave), by = __shfl(rb.y, l, kWave);
    for (uint32_t j = kHead; j < by; j += kWave) {
      const bool hit = j + lane_id() < by && in_frontier(bm_cur, a.in_arena[bx + j + lane_id()]);
      if (__ballot(hit)) {
        if (static_cast<int>(lane_id()) == l) found = true;
        break;
      }
    }
  }
  return found;
}

// One level step.  Its direction follows the rule above from the previous
// level's counters.  Two launch-shape choices are made here too, by every
// workgroup alike from the same counters, so that no traversal depends on
// an earlier one: a top-down level whose frontier arcs are large next to
// the unvisited vertices (m_f * claim_ratio >= n_u: the next level will
// most likely run bottom-up over a large frontier) claims by storing the
// distance alone, with no atomic and no queue, and a pack step derives the
// next frontier from the distances; a bottom-up level with a frontier small
// next to the unvisited vertices (n_f * tail_ratio < n_u: many rows will
// find no parent in their head) leaves the rows beyond the head to a tail
// step, edge-balanced, and otherwise finishes them itself.  A wrong guess
// costs time, never results.  Without a queue the next level has no
// frontier arc count: it runs bottom-up when the frontier is large by the
// bottom-up rule (n_f * beta >= n_u), else top-down from the frontier
// bitmap, and counts the frontier's arcs as its size times the mean degree.
__device__ __forceinline__ void bfs_level(const LevelArgs& a, uint32_t level, uint32_t step) {
  BfsState* st = a.st;
  const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
  const unsigned long long qc = st->rec[level & 3].q;
  const unsigned long long nf = found_total(st->rec[level & 3]);
  if (nf == 0) {
    if (leader) {
      st->done_level = level;
      write_step(st, step, kPhDone, level, 0u, kClaimQueue, 1u);
    }
    return;
  }
  // Direction of this level, derived identically by every workgroup.
  const bool no_queue = level > 0 && st->ctl[level & 3].no_queue;
  const unsigned long long mf =
      no_queue ? static_cast<unsigned long long>(static_cast<double>(nf) * static_cast<double>(st->total_edges) /
                                                 static_cast<double>(st->n > 0 ? st->n : 1))
               : (qc & kEdgeMask);
  const unsigned long long seen = st->ctl[level & 3].visited_edges + mf;
  // vertices a bottom-up pass of this level would check
  const unsigned long long left = st->ctl[level & 3].unvisited;
  const unsigned long long unvisited = left > nf ? left - nf : 0;
  // 0: top-down from the queue, 1: bottom-up, 2: top-down from the frontier
  // bitmap (after a bottom-up level, or a top-down level that built no queue)
  uint32_t mode = 0;
  if (level > 0) {
    const uint32_t prev = st->ctl[level & 3].prev_mode;
    if (prev != 1) {
      const unsigned long long mu = st->total_edges > seen ? st->total_edges - seen : 0;
      if (no_queue)
        mode = static_cast<double>(nf) * st->beta >= static_cast<double>(unvisited) ? 1u : 2u;
      else
        mode = static_cast<double>(mf) * st->alpha > static_cast<double>(mu) ? 1u : 0u;
    } else {
      mode = static_cast<double>(nf) * st->beta < static_cast<double>(unvisited) ? 2u : 1u;
    }
  }
  uint32_t claims = kClaimQueue;
  int inline_tail = 1;
  if (mode != 1) {
    if (static_cast<double>(mf) * st->claim_ratio >= static_cast<double>(unvisited)) claims = kClaimDistance;
  } else {
    inline_tail = static_cast<double>(nf) * st->tail_ratio >= static_cast<double>(unvisited) ? 1 : 0;
  }
  LevelRec* rn = &st->rec[(level + 1) & 3];
  if (leader) {
    write_step(st, step, kPhLevel, level, mode, claims, static_cast<uint32_t>(inline_tail));
    st->ctl[(level + 1) & 3].prev_mode = mode;
    st->ctl[(level + 1) & 3].visited_edges = seen;
    st->ctl[(level + 1) & 3].unvisited = unvisited;
    // only top-down levels fill a queue (bottom-up ones never do)
    st->ctl[(level + 1) & 3].no_queue = (mode != 1 && claims != kClaimQueue) ? 1u : 0u;
    st->rec[(level + 2) & 3].q = 0;
    for (uint32_t j = 0; j < kFoundStripes; ++j) st->rec[(level + 2) & 3].found[j].v = 0;
    st->pend[(level + 2) & 3].q = 0;
    atomicAdd(&st->reached, nf);
    if (mode == 1) atomicAdd(&st->bu_levels, 1u);
    if (level < kLogLevels) {
      st->log_mode[level] = mode;
      st->log_nf[level] = nf;
      st->log_mf[level] = mf;
    }
    if (level + 1 < kLogLevels) st->log_pend[level + 1] = 0;
  }
  const int cur = level & 1, nxt = cur ^ 1;
  const uint32_t words = a.words;
  unsigned long long* bm_cur = a.bm + static_cast<uint64_t>(level % 3) * words;
  unsigned long long* bm_next = a.bm + static_cast<uint64_t>((level + 1) % 3) * words;
  unsigned long long* bm_clear = a.bm + static_cast<uint64_t>((level + 2) % 3) * words;
  const uint32_t gthreads = gridDim.x * blockDim.x;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  // level 0 starts with the bitmaps cleared by the init kernel
  if (level > 0)
    for (uint32_t i = gtid; i < words; i += gthreads) bm_clear[i] = 0;
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const uint32_t gw = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  uint32_t* qn = nxt ? a.queue1 : a.queue0;
  unsigned long long* pn = nxt ? a.pre1 : a.pre0;
  const uint32_t next_level = level + 1;
  __shared__ WaveStage s_stage[kWavesPerBlock];
  __shared__ WaveStage s_stage2[kWavesPerBlock];
  __shared__ unsigned long long s_found[kWavesPerBlock];
  WaveStage& sb = s_stage[threadIdx.x / kWave];
  WaveStage& sb2 = s_stage2[threadIdx.x / kWave];
  uint32_t nbuf = 0, nbuf2 = 0;
  unsigned long long found_local = 0, pend_local = 0;

  // Claim w for the next level if nobody has; queue its row (begin, length).
  // A level that builds no queue claims with the distance alone: every lane
  // that finds w unvisited stores the same value, with no atomic (a
  // device-scope atomic is a round trip past the L2), and the pack step then
  // derives the frontier bitmap, the visited bits and the count of new
  // vertices from the distances.
  auto claim = [&](uint32_t w, bool consider) {
    if (claims == kClaimDistance) {
      if (consider && !(w & kTomb) && !((a.visited[w >> 6] >> (w & 63)) & 1ull)) a.dist[w] = next_level;
      return;
    }
    bool newly = false;
    uint2 wr = make_uint2(0u, 0u);
    if (consider && !(w & kTomb)) {
      const unsigned long long bit = 1ull << (w & 63);
      unsigned long long* wp = a.visited + (w >> 6);
      if (!(*wp & bit)) newly = !(atomicOr(wp, bit) & bit);
    }
    if (newly) {
      a.dist[w] = next_level;
      atomicOr(bm_next + (w >> 6), 1ull << (w & 63));
      if (claims == kClaimQueue) wr = a.row[w];
    }
    found_local += __popcll(__ballot(newly));
    stage_push(sb, nbuf, &rn->q, qn, pn, newly && wr.y > 0, wr.x, wr.y);
  };

  if (mode == 2) {
    // Frontier vertices from the bitmap: rows of up to 32 entries are
    // expanded by their own lane, longer rows by the whole wave.  The wave's
    // words (stride W) are loaded 64 at a time, one per lane, and only the
    // non-empty ones are visited, so a sparse frontier costs one load.
    for (uint32_t base = gw; base < words; base += kWave * W) {
      const uint32_t wk = base + lane * W;
      const unsigned long long fk = wk < words ? bm_cur[wk] : 0ull;
      unsigned long long todo = __ballot(fk != 0);
      while (todo) {
        const int lw = __ffsll(static_cast<long long>(todo)) - 1;
        todo &= todo - 1;
        const uint32_t wi = base + static_cast<uint32_t>(lw) * W;
        const unsigned long long fw = __shfl(fk, lw, kWave);
        const uint32_t u = wi * 64 + lane;
        const bool mine = (fw >> lane) & 1ull;
        const uint2 rb = mine ? a.row[u] : make_uint2(0u, 0u);
        const bool longr = rb.y > 32;
        const uint32_t cnt = longr ? 0u : rb.y;
        uint32_t most = cnt;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) most = max(most, static_cast<uint32_t>(__shfl_xor(static_cast<int>(most), d, kWave)));
        for (uint32_t j = 0; j < most; ++j) {
          const bool act = j < cnt;
          claim(act ? a.arena[rb.x + j] : kTomb, act);
        }
        unsigned long long lm = __ballot(longr);
        while (lm) {
          const int l = __ffsll(static_cast<long long>(lm)) - 1;
          lm &= lm - 1;
          const uint32_t bx = __shfl(rb.x, l, kWave), by = __shfl(rb.y, l, kWave);
          for (uint32_t j = 0; j < by; j += kWave) {
            const bool act = j + lane < by;
            claim(act ? a.arena[bx + j + lane] : kTomb, act);
          }
        }
      }
    }
  } else if (mode == 0) {
    const uint32_t F = static_cast<uint32_t>(qc >> kEdgeBits);
    const uint32_t* qcur = cur ? a.queue1 : a.queue0;
    const unsigned long long* pcur = cur ? a.pre1 : a.pre0;
    lbs_edges<true>(qcur, pcur, F, mf, a.row, gw, W,
              [&](uint32_t, uint32_t, uint2, uint32_t slot, bool valid) {
                claim(valid ? a.arena[slot] : kTomb, valid);
              });
  } else {
    // Phase A: one lane per vertex checks the first kHead parents, read from
    // the dense head array (16 bytes per vertex, coalesced across the wave)
    // instead of the row itself.  Rows are sorted by id, so these are the
    // vertex's lowest-id, on R-MAT highest-degree, parents.  The wave's
    // visited words (stride W) are loaded 64 at a time, one per lane; only
    // words with unvisited vertices are processed, two per step with both
    // words' loads issued before use (each word belongs to one wave, so the
    // loaded values stay current).
    const uint4 none = make_uint4(kTomb, kTomb, kTomb, kTomb);
    for (uint32_t base = gw; base < words; base += kWave * W) {
      const uint32_t wk = base + lane * W;
      const unsigned long long vk = wk < words ? a.visited[wk] : ~0ull;
      unsigned long long todo = __ballot(vk != ~0ull);
      while (todo) {
        const int l0 = __ffsll(static_cast<long long>(todo)) - 1;
        todo &= todo - 1;
        const bool has_j = todo != 0;
        const int l1 = has_j ? __ffsll(static_cast<long long>(todo)) - 1 : l0;
        if (has_j) todo &= todo - 1;
        const uint32_t wi = base + static_cast<uint32_t>(l0) * W, wj = base + static_cast<uint32_t>(l1) * W;
        const unsigned long long vw0 = __shfl(vk, l0, kWave);
        const unsigned long long vw1 = has_j ? __shfl(vk, l1, kWave) : ~0ull;
        const uint32_t v0 = wi * 64 + lane, v1 = wj * 64 + lane;
        const bool unv0 = v0 < a.n && !((vw0 >> lane) & 1ull);
        const bool unv1 = has_j && v1 < a.n && !((vw1 >> lane) & 1ull);
        const uint4 h0 = unv0 ? a.head[v0] : none;
        const uint4 h1 = unv1 ? a.head[v1] : none;
        // the first parent, then the second, then the last two, each only
        // for the lanes still searching (inactive lanes send no address)
        bool f0 = in_frontier(bm_cur, h0.x);
        bool f1 = in_frontier(bm_cur, h1.x);
        if (!f0) f0 = in_frontier(bm_cur, h0.y);
        if (!f1) f1 = in_frontier(bm_cur, h1.y);
        if (!f0) f0 = in_frontier(bm_cur, h0.z) | in_frontier(bm_cur, h0.w);
        if (!f1) f1 = in_frontier(bm_cur, h1.z) | in_frontier(bm_cur, h1.w);
        // the rest of longer rows goes to phase B (a tail step);
        // a head that ends in kNoEntry is the whole row
        const uint2 r0 = (unv0 && !f0 && h0.w != kNoEntry) ? a.in_row[v0] : make_uint2(0u, 0u);
        const uint2 r1 = (unv1 && !f1 && h1.w != kNoEntry) ? a.in_row[v1] : make_uint2(0u, 0u);
        if (inline_tail) {
          // counted like phase B's list, for the level log
          if (r0.y > kHead) pend_local += (1ull << kEdgeBits) + (r0.y - kHead);
          if (r1.y > kHead) pend_local += (1ull << kEdgeBits) + (r1.y - kHead);
          // called by every lane: a long row is scanned by the whole wave
          // (a short-circuit `f0 || ...` would leave the lanes that already
          // found a parent out of that scan, and skip the entries they own)
          const bool g0 = rest_in_frontier(a, bm_cur, r0, r0.y > kHead);
          const bool g1 = rest_in_frontier(a, bm_cur, r1, r1.y > kHead);
          f0 = f0 || g0;
          f1 = f1 || g1;
        } else {
          stage_push(sb2, nbuf2, &st->pend[level & 3].q, a.pend, a.ppre, r0.y > kHead, v0, r0.y - kHead);
          stage_push(sb2, nbuf2, &st->pend[level & 3].q, a.pend, a.ppre, r1.y > kHead, v1, r1.y - kHead);
        }
        // Bottom-up levels build no queue: the next level reads the bitmap.
        const unsigned long long m0 = __ballot(f0), m1 = __ballot(f1);
        if (lane == 0) {
          if (m0) {
            bm_next[wi] = m0;
            a.visited[wi] = vw0 | m0;
          }
          if (m1) {
            bm_next[wj] = m1;
            a.visited[wj] = vw1 | m1;
          }
        }
        if (f0) a.dist[v0] = next_level;
        if (f1) a.dist[v1] = next_level;
        found_local += __popcll(m0) + __popcll(m1);
      }
    }
  }
  // one atomic per workgroup and list for the last staged pushes
  __shared__ unsigned long long s_flush[2 * kWavesPerBlock + 1];
  stage_flush_block(sb, nbuf, &rn->q, qn, pn, s_flush);
  stage_flush_block(sb2, nbuf2, &st->pend[level & 3].q, a.pend, a.ppre, s_flush);
  if (lane == 0) s_found[threadIdx.x / kWave] = found_local;
  const bool count_pend = mode == 1 && inline_tail && level < kLogLevels;
  __shared__ unsigned long long s_pend[kWavesPerBlock];
  if (count_pend) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) pend_local += __shfl_xor(pend_local, d, kWave);
    if (lane == 0) s_pend[threadIdx.x / kWave] = pend_local;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long f = 0;
    for (uint32_t i = 0; i < kWavesPerBlock; ++i) f += s_found[i];
    if (f) atomicAdd(&rn->found[blockIdx.x % kFoundStripes].v, f);
    if (count_pend) {
      unsigned long long p = 0;
      for (uint32_t i = 0; i < kWavesPerBlock; ++i) p += s_pend[i];
      if (p) atomicAdd(&st->log_pend[level], p);
    }
  }
}

// Pack step, after a top-down level that claimed by distance: the vertices
// it claimed are those at distance level + 1.  Waves read the distances
// (coalesced), set the next frontier bitmap and the visited bits from them
// and count them for the next level.
__device__ __forceinline__ void bfs_pack(const LevelArgs& a, uint32_t level) {
  BfsState* st = a.st;
  const uint32_t words = a.words, next_level = level + 1;
  unsigned long long* bm_next = a.bm + static_cast<uint64_t>((level + 1) % 3) * words;
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const uint32_t gw = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  unsigned long long found_local = 0;
  // kPackWords words per step: their distance loads and visited words are
  // issued together, and lane k writes word k's results.
  for (uint32_t w0 = gw * kPackWords; w0 < words; w0 += W * kPackWords) {
    uint32_t d[kPackWords];
#pragma unroll
    for (uint32_t k = 0; k < kPackWords; ++k) {
      const uint32_t v = (w0 + k) * 64 + lane;
      d[k] = v < a.n ? a.dist[v] : kInf;
    }
    const bool mine = lane < kPackWords && w0 + lane < words;
    const unsigned long long vis = mine ? a.visited[w0 + lane] : 0ull;
    unsigned long long mk = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPackWords; ++k) {
      const unsigned long long m = __ballot(d[k] == next_level);
      if (lane == k) mk = m;
      found_local += __popcll(m);
    }
    if (mine && mk) {
      bm_next[w0 + lane] = mk;
      a.visited[w0 + lane] = vis | mk;
    }
  }
  __shared__ unsigned long long s_found[kWavesPerBlock];
  if (lane == 0) s_found[threadIdx.x / kWave] = found_local;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long f = 0;
    for (uint32_t i = 0; i < kWavesPerBlock; ++i) f += s_found[i];
    if (f) atomicAdd(&st->rec[(level + 1) & 3].found[blockIdx.x % kFoundStripes].v, f);
  }
}

// Tail step of a bottom-up level that left rows pending (phase B): the
// parents of pending rows beyond the first kHead are scanned edge-balanced
// and coalesced; a parent in the frontier claims the row's vertex through
// the visited bitmap.
__device__ __forceinline__ void bfs_tail(const LevelArgs& a, uint32_t level) {
  BfsState* st = a.st;
  const unsigned long long pq = st->pend[level & 3].q;
  if (blockIdx.x == 0 && threadIdx.x == 0 && level < kLogLevels) st->log_pend[level] = pq;
  if (pq == 0) return;
  __shared__ unsigned long long s_found[kWavesPerBlock];
  unsigned long long found_local = 0;
  const uint32_t words = a.words;
  const unsigned long long* bm_cur = a.bm + static_cast<uint64_t>(level % 3) * words;
  unsigned long long* bm_next = a.bm + static_cast<uint64_t>((level + 1) % 3) * words;
  LevelRec* rn = &st->rec[(level + 1) & 3];
  const uint32_t next_level = level + 1;
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const uint32_t gw = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  lbs_edges(a.pend, a.ppre, static_cast<uint32_t>(pq >> kEdgeBits), pq & kEdgeMask, a.in_row, gw, W,
            [&](uint32_t, uint32_t v, uint2, uint32_t slot, bool valid) {
              const uint32_t x = valid ? a.in_arena[slot + kHead] : kTomb;
              // A window holds consecutive edges of a few rows; of the lanes
              // that find a parent of the same vertex only the lowest claims
              // it, so the wave sends one atomic per vertex, not one per hit.
              const bool hit = in_frontier(bm_cur, x);
              const unsigned long long hits = __ballot(hit);
              const uint32_t vprev = __shfl_up(v, 1, kWave);
              const unsigned long long starts = __ballot(lane == 0 || vprev != v);
              const unsigned long long below = (1ull << lane) - 1;
              const unsigned long long seg = starts & (below | (1ull << lane));
              const int first = 63 - __clzll(static_cast<long long>(seg));
              const unsigned long long seg_below = below & ~((1ull << first) - 1);
              bool newly = false;
              if (hit && !(hits & seg_below)) {
                const unsigned long long bit = 1ull << (v & 63);
                unsigned long long* wp = a.visited + (v >> 6);
                if (!(*wp & bit)) newly = !(atomicOr(wp, bit) & bit);
              }
              if (newly) {
                a.dist[v] = next_level;
                atomicOr(bm_next + (v >> 6), 1ull << (v & 63));
              }
              found_local += __popcll(__ballot(newly));
            });
  if (lane == 0) s_found[threadIdx.x / kWave] = found_local;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long f = 0;
    for (uint32_t i = 0; i < kWavesPerBlock; ++i) f += s_found[i];
    if (f) atomicAdd(&rn->found[blockIdx.x % kFoundStripes].v, f);
  }
}

// One step of a traversal: a level, the tail of a bottom-up level or the
// pack after a top-down level that claimed by distance.  Which one is
// decided on the device from what the previous step recorded (next_step),
// so every launch does work until the traversal is over and the same launch
// program serves every root; steps launched past the end exit at once.
__global__ __launch_bounds__(kBfsBlock) void k_bfs_step(LevelArgs a, uint32_t step) {
  BfsState* st = a.st;
  const StepRec prev = st->srec[step & 3];
  const StepPlan p = next_step(prev, st->pend[prev.level & 3].q);
  if (p.phase == kPhLevel) {
    bfs_level(a, p.level, step);
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) write_step(st, step, p.phase, p.level, prev.mode, prev.claims, prev.inline_tail);
  if (p.phase == kPhTail)
    bfs_tail(a, p.level);
  else if (p.phase == kPhPack)
    bfs_pack(a, p.level);
}

// -------------------------------------------------------- incremental kernels

struct IncArgs {
  const uint2* out_row;
  const uint32_t* out_arena;
  const uint2* in_row;
  const uint32_t* in_arena;
  uint32_t* dist;
  uint32_t* lost;
  uint32_t* support;
  uint32_t* flags;
  uint32_t* stamp;
  IncState* st;
  uint32_t epoch;
  uint32_t source;
  uint64_t lcap;  // slots of one parity of the round lists
};

// Workgroups of the repair kernels whose list sizes live on the device.
constexpr unsigned kRepairBlocks = 512;

// touch(): record that v received a loss (or is a candidate) in the round
// whose lists have parity `par`.
__device__ __forceinline__ void touch(const IncArgs& a, int par, uint32_t v, uint32_t* tlist, uint32_t* tall,
                                      uint32_t* tflist, unsigned long long* tfpre, bool want) {
  bool t_new = false, first = false;
  uint32_t inlen = 0;
  if (want) {
    t_new = atomicExch(&a.stamp[v], a.epoch) != a.epoch;
    uint32_t old = atomicOr(&a.flags[v], kFSupport);
    first = !(old & kFSupport);
    if (first) {
      a.support[v] = 0;
      inlen = a.in_row[v].y;
    }
  }
  const uint64_t off = par * a.lcap;
  unsigned long long bt = __ballot(t_new);
  if (bt) {
    unsigned long long base = 0;
    if (lane_id() == 0) base = atomicAdd(&a.st->tq[par], static_cast<unsigned long long>(__popcll(bt)));
    base = __shfl(base, 0, kWave);
    if (t_new) tlist[off + base + lanes_below(bt)] = v;
  }
  unsigned long long bf = __ballot(first);
  if (bf) {
    unsigned long long base = 0;
    if (lane_id() == 0) base = atomicAdd(&a.st->tall, static_cast<unsigned long long>(__popcll(bf)));
    base = __shfl(base, 0, kWave);
    if (first) tall[base + lanes_below(bf)] = v;
  }
  wave_push(&a.st->tfq[par], tflist + off, tfpre + off, first && inlen > 0, v, inlen);
}

// Heads of deleted arcs that were shortest-parent arcs (bfs.hpp:52-63, 242-245),
// touched for round 0.
__global__ void k_inc_candidates(IncArgs a, const uint32_t* fin, const unsigned long long* nfin_p,
                                 const unsigned long long* skeys, const uint8_t* act, uint32_t* tlist,
                                 uint32_t* tall, uint32_t* tflist, unsigned long long* tfpre) {
  const uint64_t nfin = *nfin_p;  // the staged batch's count, read on the device
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (uint64_t j0 = base - lane_id(); j0 < nfin; j0 += stride) {
    const uint64_t j = j0 + lane_id();
    bool want = false;
    uint32_t v = 0;
    if (j < nfin) {
      uint32_t p = fin[j];
      if (act[p] == kActDel) {
        uint32_t u = static_cast<uint32_t>(skeys[p] >> 32);
        v = static_cast<uint32_t>(skeys[p]);
        uint32_t du = a.dist[u], dv = a.dist[v];
        if (du != kInf && dv != kInf && du + 1 == dv) {
          uint32_t old = atomicOr(&a.flags[v], kFCand);
          if (!(old & kFCand)) atomicAdd(&a.st->candidates, 1ull);
          want = v != a.source;
        }
      }
    }
    touch(a, 0, v, tlist, tall, tflist, tfpre, want);
  }
}

// Round of parity par, step 1: support[v] += live in-arcs p->v with
// d(p) + 1 == d(v), over the vertices touched for the first time.  Clears
// the affected counter this round's evaluation appends to.
__global__ void k_inc_support(IncArgs a, int par, const uint32_t* tflist, const unsigned long long* tfpre) {
  if (a.st->fallback) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.st->aq[par] = 0;
  const unsigned long long qc = a.st->tfq[par];
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const uint32_t gw = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  const uint64_t off = par * a.lcap;
  lbs_edges(tflist + off, tfpre + off, static_cast<uint32_t>(qc >> kEdgeBits), qc & kEdgeMask, a.in_row, gw, W,
            [&](uint32_t, uint32_t v, uint2, uint32_t slot, bool valid) {
              if (!valid) return;
              uint32_t p = a.in_arena[slot];
              if (p & kTomb) return;
              uint32_t dp = a.dist[p];
              if (dp != kInf && dp + 1 == a.dist[v]) atomicAdd(&a.support[v], 1u);
            });
}

// Step 2: v is affected once every live shortest parent is lost.  Clears the
// touched counters of the next parity, which step 3 appends to.
__global__ void k_inc_evaluate(IncArgs a, int par, const uint32_t* tlist, uint32_t* alist,
                               unsigned long long* apre, uint32_t* aall) {
  if (a.st->fallback) return;
  const uint64_t nt = a.st->tq[par];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->tq[par ^ 1] = 0;
    a.st->tfq[par ^ 1] = 0;
  }
  const uint64_t off = par * a.lcap;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (uint64_t j0 = base - lane_id(); j0 < nt; j0 += stride) {
    const uint64_t j = j0 + lane_id();
    bool aff = false;
    uint32_t v = 0, len = 0;
    if (j < nt) {
      v = tlist[off + j];
      uint32_t f = a.flags[v];
      if (!(f & kFAffected) && a.lost[v] >= a.support[v]) {
        a.flags[v] = f | kFAffected;
        aff = true;
        len = a.out_row[v].y;
      }
    }
    unsigned long long b = __ballot(aff);
    if (b) {
      unsigned long long base2 = 0;
      if (lane_id() == 0) base2 = atomicAdd(&a.st->aall, static_cast<unsigned long long>(__popcll(b)));
      base2 = __shfl(base2, 0, kWave);
      if (aff) aall[base2 + lanes_below(b)] = v;
    }
    wave_push(&a.st->aq[par], alist + off, apre + off, aff && len > 0, v, len);
  }
}

// Step 3: affected u makes its live children at d(u)+1 lose a parent
// (bfs.hpp:247-257); they are touched for the next round (a.epoch is that
// round's).  Past `limit` affected vertices the repair stops here and falls
// back to a full traversal.
__global__ void k_inc_emit(IncArgs a, int par, uint64_t limit, const uint32_t* alist, const unsigned long long* apre,
                           uint32_t* tlist, uint32_t* tall, uint32_t* tflist, unsigned long long* tfpre) {
  if (a.st->fallback) return;
  if (a.st->aall > limit) {
    if (blockIdx.x == 0 && threadIdx.x == 0) a.st->fallback = 1;
    return;
  }
  const unsigned long long qc = a.st->aq[par];
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const uint32_t gw = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  const uint64_t off = par * a.lcap;
  lbs_edges(alist + off, apre + off, static_cast<uint32_t>(qc >> kEdgeBits), qc & kEdgeMask, a.out_row, gw, W,
            [&](uint32_t, uint32_t u, uint2, uint32_t slot, bool valid) {
              bool want = false;
              uint32_t w = 0;
              if (valid) {
                w = a.out_arena[slot];
                if (!(w & kTomb)) {
                  uint32_t du = a.dist[u];
                  if (du != kInf && a.dist[w] == du + 1) {
                    atomicAdd(&a.lost[w], 1u);
                    want = true;
                  }
                } else {
                  w = 0;
                }
              }
              touch(a, par ^ 1, w, tlist, tall, tflist, tfpre, want);
            });
}

__global__ void k_inc_reset_affected(IncArgs a, const uint32_t* aall, uint32_t* slist, unsigned long long* spre) {
  const uint64_t na = a.st->aall;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (uint64_t j0 = base - lane_id(); j0 < na; j0 += stride) {
    const uint64_t j = j0 + lane_id();
    uint32_t v = 0, len = 0;
    bool want = false;
    if (j < na) {
      v = aall[j];
      a.dist[v] = kInf;
      len = a.in_row[v].y;
      want = len > 0;
    }
    wave_push(&a.st->sq, slist, spre, want, v, len);
  }
}

// best_boundary_distance (bfs.hpp:261-269) as an atomicMin over in-arcs.
__global__ void k_inc_seed(IncArgs a, const uint32_t* slist, const unsigned long long* spre) {
  const unsigned long long qc = a.st->sq;
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const uint32_t gw = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  lbs_edges(slist, spre, static_cast<uint32_t>(qc >> kEdgeBits), qc & kEdgeMask, a.in_row, gw, W,
            [&](uint32_t, uint32_t v, uint2, uint32_t slot, bool valid) {
              if (!valid) return;
              uint32_t p = a.in_arena[slot];
              if (p & kTomb) return;
              if (a.flags[p] & kFAffected) return;
              uint32_t dp = a.dist[p];
              if (dp != kInf) atomicMin(&a.dist[v], dp + 1);
            });
}

__device__ __forceinline__ void frontier_push(const IncArgs& a, int slot, uint32_t* fq, unsigned long long* fp,
                                              uint32_t v, bool want) {
  bool push = false;
  uint32_t len = 0;
  if (want) {
    push = atomicExch(&a.stamp[v], a.epoch) != a.epoch;
    if (push) len = a.out_row[v].y;
  }
  wave_push(&a.st->fq[slot], fq, fp, push && len > 0, v, len);
}

__global__ void k_inc_collect_seeds(IncArgs a, const uint32_t* aall, uint32_t* fq, unsigned long long* fp) {
  const uint64_t na = a.st->aall;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long reach = 0;
  for (uint64_t j0 = base - lane_id(); j0 < na; j0 += stride) {
    const uint64_t j = j0 + lane_id();
    uint32_t v = 0;
    bool want = false;
    if (j < na) {
      v = aall[j];
      want = a.dist[v] != kInf;
      reach += want;
    }
    frontier_push(a, 0, fq, fp, v, want);
  }
  reach = wave_sum(reach);
  if (lane_id() == 0 && reach) {
    atomicAdd(&a.st->became_reachable, reach);
    atomicAdd(&a.st->relaxed, reach);
  }
}

// relax_edge over the final additions (bfs.hpp:91-96, 332-339).
__global__ void k_inc_relax_additions(IncArgs a, const uint32_t* fin, uint64_t nfin,
                                      const unsigned long long* skeys, const uint8_t* act, uint32_t* fq,
                                      unsigned long long* fp) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long reach = 0, relaxed = 0;
  for (uint64_t j0 = base - lane_id(); j0 < nfin; j0 += stride) {
    const uint64_t j = j0 + lane_id();
    bool want = false;
    uint32_t v = 0;
    if (j < nfin) {
      uint32_t p = fin[j];
      if (act[p] >= kActNopAdd) {
        uint32_t u = static_cast<uint32_t>(skeys[p] >> 32);
        v = static_cast<uint32_t>(skeys[p]);
        uint32_t du = a.dist[u];
        if (du != kInf && du + 1 < a.dist[v]) {
          uint32_t old = atomicMin(&a.dist[v], du + 1);
          if (old > du + 1) {
            want = true;
            reach += old == kInf;
            ++relaxed;
          }
        }
      }
    }
    frontier_push(a, 0, fq, fp, v, want);
  }
  reach = wave_sum(reach);
  relaxed = wave_sum(relaxed);
  if (lane_id() == 0 && (reach || relaxed)) {
    atomicAdd(&a.st->became_reachable, reach);
    atomicAdd(&a.st->relaxed, relaxed);
  }
}

// propagate_decreases (bfs.hpp:341-354) as one Bellman-Ford round: round r
// expands frontier r % 3 into (r + 1) % 3 and clears (r + 2) % 3, which
// round r + 1 appends to (a.epoch is this round's).
__global__ void k_inc_propagate(IncArgs a, uint32_t r, const uint32_t* fq_cur, const unsigned long long* fp_cur,
                                uint32_t* fq_next, unsigned long long* fp_next) {
  const uint32_t cur = r % 3, nxt = (r + 1) % 3;
  const unsigned long long qc = a.st->fq[cur];
  if (blockIdx.x == 0 && threadIdx.x == 0) a.st->fq[(r + 2) % 3] = 0;
  const uint32_t W = gridDim.x * kWavesPerBlock;
  const uint32_t gw = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  unsigned long long reach = 0, relaxed = 0;
  lbs_edges(fq_cur, fp_cur, static_cast<uint32_t>(qc >> kEdgeBits), qc & kEdgeMask, a.out_row, gw, W,
            [&](uint32_t, uint32_t u, uint2, uint32_t slot, bool valid) {
              bool want = false;
              uint32_t w = 0;
              if (valid) {
                w = a.out_arena[slot];
                if (!(w & kTomb)) {
                  uint32_t du = a.dist[u];
                  if (du != kInf && du + 1 < a.dist[w]) {
                    uint32_t old = atomicMin(&a.dist[w], du + 1);
                    if (old > du + 1) {
                      want = true;
                      reach += old == kInf;
                      ++relaxed;
                    }
                  }
                } else {
                  w = 0;
                }
              }
              frontier_push(a, nxt, fq_next, fp_next, w, want);
            });
  reach = wave_sum(reach);
  relaxed = wave_sum(relaxed);
  if (lane_id() == 0 && (reach || relaxed)) {
    atomicAdd(&a.st->became_reachable, reach);
    atomicAdd(&a.st->relaxed, relaxed);
  }
}

__global__ void k_inc_cleanup(IncArgs a, const uint32_t* tall) {
  const uint64_t nt = a.st->tall;
  for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < nt;
       j += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t v = tall[j];
    a.lost[v] = 0;
    a.support[v] = 0;
    a.flags[v] = 0;
  }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

// ------------------------------------------------------------- host side

static unsigned blocks_for(uint64_t items, unsigned per = kBfsBlock, unsigned cap = 4096) {
  uint64_t b = (items + per - 1) / per;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

static unsigned lbs_blocks(unsigned long long edges) {
  // about 256 edges per wave, at most 2048 workgroups
  uint64_t waves = (edges + 255) / 256;
  uint64_t b = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return static_cast<unsigned>(b);
}

static void ensure_work(vgx_bfs* b) {
  vgx_graph* g = b->g;
  BfsWork& w = b->w;
  if (w.cap >= g->ncap && w.cap != 0) return;
  uint64_t c = g->ncap;
  w.dist.grow_copy(c, b->dist_len, g->stream, b->views.keep());
  uint64_t words = (c + 63) / 64;
  w.visited.ensure(words);
  w.pend.ensure(c);
  w.ppre.ensure(c);
  w.bm.ensure(3 * words);
  for (int i = 0; i < 2; ++i) {
    w.queue[i].ensure(c);
    w.pre[i].ensure(c);
  }
  {
  }
  // incremental workspace: zero-initialised counters and flags
  uint64_t old = w.lost.n ? w.cap : 0;
  w.lost.grow_copy(c, old, g->stream);
  w.support.grow_copy(c, old, g->stream);
  w.flags.grow_copy(c, old, g->stream);
  w.stamp.grow_copy(c, old, g->stream);
  VGX_HIP(hipMemsetAsync(w.lost.p + old, 0, (c - old) * 4, g->stream));
  VGX_HIP(hipMemsetAsync(w.support.p + old, 0, (c - old) * 4, g->stream));
  VGX_HIP(hipMemsetAsync(w.flags.p + old, 0, (c - old) * 4, g->stream));
  VGX_HIP(hipMemsetAsync(w.stamp.p + old, 0, (c - old) * 4, g->stream));
  // the lists of the invalidation rounds have two parities of c slots
  for (auto* l : {&w.tlist, &w.tflist, &w.alist}) l->ensure(2 * c);
  for (auto* l : {&w.tfpre, &w.apre}) l->ensure(2 * c);
  for (auto* l : {&w.aall, &w.tall, &w.slist}) l->ensure(c);
  w.spre.ensure(c);
  w.cap = c;
}

// Entries of dist beyond dist_len belong to vertices added after the last
// recompute; the reference resizes with `unreachable` (bfs.hpp:73).
static void extend_dist(vgx_bfs* b) {
  vgx_graph* g = b->g;
  ensure_work(b);
  if (b->dist_len < g->n) {
    k_fill_u32<<<blocks_for(g->n - b->dist_len), kBfsBlock, 0, g->stream>>>(b->w.dist.p + b->dist_len,
                                                                          g->n - b->dist_len, kInf);
    VGX_LAUNCH_CHECK();
    b->dist_len = g->n;
  }
}

static LevelArgs level_args(vgx_bfs* b) {
  vgx_graph* g = b->g;
  LevelArgs a;
  a.row = g->adj[0].row.p;
  a.arena = g->adj[0].arena.p;
  a.in_row = g->in().row.p;
  a.in_arena = g->in().arena.p;
  a.dist = b->w.dist.p;
  a.visited = b->w.visited.p;
  a.bm = b->w.bm.p;
  a.queue0 = b->w.queue[0].p;
  a.queue1 = b->w.queue[1].p;
  a.pre0 = b->w.pre[0].p;
  a.pre1 = b->w.pre[1].p;
  a.st = b->d_st;
  a.n = static_cast<uint32_t>(g->n);
  a.words = static_cast<uint32_t>((g->n + 63) / 64);
  a.head = b->w.head.p;
  a.pend = b->w.pend.p;
  a.ppre = b->w.ppre.p;
  return a;
}

// Start of a traversal in one launch: distances to infinity, the visited
// bitmap from the zero-in-degree bitmap, the frontier bitmaps cleared, and
// the root placed (distance 0, visited, level-0 frontier and queue).
// Workgroup 0 resets the level counters.  A root outside the graph leaves
// the level-0 frontier empty, which ends the traversal at once.
__global__ void k_bfs_init(LevelArgs a, const unsigned long long* __restrict__ zin,
                           const unsigned long long* __restrict__ zcount, uint32_t root,
                           unsigned long long total_edges, float alpha, float beta, float tail_ratio,
                           float claim_ratio) {
  const uint64_t n = a.n, words = a.words;
  const bool has_root = root < n;
  const uint64_t rw = has_root ? root >> 6 : ~0ull;
  const unsigned long long rbit = 1ull << (root & 63);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t t0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  for (uint64_t i = t0; i < n; i += stride) a.dist[i] = i == root ? 0u : kInf;
  for (uint64_t i = t0; i < words; i += stride) {
    const unsigned long long r = i == rw ? rbit : 0ull;
    a.visited[i] = zin[i] | r;
    a.bm[i] = r;
    a.bm[words + i] = 0;
    a.bm[2 * words + i] = 0;
  }
  if (blockIdx.x != 0) return;
  BfsState* st = a.st;
  for (uint32_t t = threadIdx.x; t < 4 * kFoundStripes; t += blockDim.x) st->rec[t / kFoundStripes].found[t % kFoundStripes].v = 0;
  if (threadIdx.x < 4) {
    st->rec[threadIdx.x].q = 0;
    st->ctl[threadIdx.x].visited_edges = 0;
    st->ctl[threadIdx.x].prev_mode = 0;
    st->ctl[threadIdx.x].no_queue = 0;
    st->ctl[threadIdx.x].unvisited = 0;
    st->pend[threadIdx.x].q = 0;
    st->srec[threadIdx.x].phase = kPhStart;
    st->srec[threadIdx.x].level = 0;
    st->srec[threadIdx.x].mode = 0;
    st->srec[threadIdx.x].claims = kClaimQueue;
    st->srec[threadIdx.x].inline_tail = 1;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  st->total_edges = total_edges;
  st->n = a.n;
  st->bu_levels = 0;
  st->alpha = alpha;
  st->beta = beta;
  st->tail_ratio = tail_ratio;
  st->claim_ratio = claim_ratio;
  st->done_level = ~0u;
  st->reached = 0;
  st->log_pend[0] = 0;
  // every vertex with in-arcs, and the root (level 0's frontier) either way
  const unsigned long long cells = words * 64;
  unsigned long long z = 0;
  for (uint32_t j = 0; j < kZinStripes; ++j) z += zcount[j * kStripeWords];
  st->ctl[0].unvisited = (cells > z ? cells - z : 0) + ((has_root && (zin[rw] & rbit)) ? 1 : 0);
  if (!has_root) return;
  st->rec[0].found[0].v = 1;
  const uint2 rr = a.row[root];
  if (rr.y) {
    a.queue0[0] = rr.x;
    a.pre0[0] = 0;
    st->rec[0].q = (1ull << kEdgeBits) | rr.y;
  }
}

// Whether a traversal whose last launched step left `h` is over, and at
// which level (the first one with an empty frontier).
static bool traversal_over(const BfsState& h, uint32_t launched, uint32_t* done_level) {
  const StepRec& last = h.srec[launched & 3];
  if (last.phase == kPhDone) {
    *done_level = h.done_level;
    return true;
  }
  const StepPlan p = next_step(last, h.pend[last.level & 3].q);
  if (p.phase == kPhLevel) {
    unsigned long long f = 0;
    for (uint32_t j = 0; j < kFoundStripes; ++j) f += h.rec[p.level & 3].found[j].v;
    if (f == 0) {
      *done_level = p.level;
      return true;
    }
  }
  return false;
}

// The dense head array and the zero-in-degree bitmap of the graph's
// current in-rows (zcount is cleared by k_build_head, then summed by
// k_build_zin on the same stream).
static void build_head_zin(vgx_graph* g, uint4* head, unsigned long long* zin, unsigned long long* zcount,
                           hipStream_t s) {
  const uint64_t n = g->n, words = (n + 63) / 64;
  if (!n) return;
  // one vertex per thread: each is a chain of two dependent loads
  k_build_head<<<blocks_for(n, kBfsBlock, 65535), kBfsBlock, 0, s>>>(g->in().row.p, g->in().arena.p,
                                                                     static_cast<uint32_t>(n), head, zcount);
  VGX_LAUNCH_CHECK();
  k_build_zin<<<blocks_for(words, kWavesPerBlock, 1024), kBfsBlock, 0, s>>>(
      g->in().deg.p, static_cast<uint32_t>(n), static_cast<uint32_t>(words), zin, zcount);
  VGX_LAUNCH_CHECK();
}

struct TraversalResult {
  uint32_t done_level = 0;
  uint32_t steps = 0;
  uint32_t bu_levels = 0;
  uint64_t reached = 0;
};

// One traversal from `source` over the buffers of `a` (distances, bitmaps,
// queues, state), with head/zin/zcount already built for the graph's
// current version: the init kernel and the step program, in chunks between
// host checks of the state (copied to h_st).  `prof`, when not null, is the
// traversal object whose profiling mode asks for HIP events.
static TraversalResult run_traversal(vgx_graph* g, const LevelArgs& a, const unsigned long long* zin,
                                     const unsigned long long* zcount, uint32_t source,
                                     const vgx_bfs_tuning& tune, BfsState* h_st, vgx_bfs* prof) {
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  const uint64_t words = (n + 63) / 64;
  unsigned blocks = tune.blocks;
  if (blocks == 0) {
    // about one resident round of workgroups: more waves only add flushes
    // on the level counters (measured at R-MAT scale 22)
    uint64_t waves = std::max<uint64_t>(words, g->adj[0].stored / 512);
    blocks = static_cast<unsigned>(std::min<uint64_t>(1536, std::max<uint64_t>(1, (waves + 3) / 4)));
  }
  // Steps go out in chunks between host checks.  The first chunk is the
  // number of steps the traversals of this graph version have needed so far
  // (any root; the graph keeps the count), so a traversal usually ends
  // within it: steps launched past the end exit at once, and the host sees
  // completion in the record of the last step without a launch spent on it.
  uint32_t chunk = tune.levels_per_sync;
  if (chunk == 0) chunk = g->bfs_steps ? g->bfs_steps : 8u;
  const unsigned long long total_edges = g->adj[0].stored;
  const float alpha = static_cast<float>(tune.alpha), beta = static_cast<float>(tune.beta);
  const float tail_ratio = static_cast<float>(tune.tail_ratio);
  const float claim_ratio = static_cast<float>(tune.claim_ratio);
  const unsigned ginit = blocks_for(std::max<uint64_t>(n, 1), kBfsBlock, 2048);
  const bool per_step = prof && prof->profiling == 1;
  const bool whole = prof && prof->profiling == 2;
  if (whole && !prof->tev[0]) {
    VGX_HIP(hipEventCreate(&prof->tev[0]));
    VGX_HIP(hipEventCreate(&prof->tev[1]));
  }
  if (whole) VGX_HIP(hipEventRecord(prof->tev[0], s));
  k_bfs_init<<<ginit, kBfsBlock, 0, s>>>(a, zin, zcount, source, total_edges, alpha, beta, tail_ratio, claim_ratio);
  VGX_LAUNCH_CHECK();
  uint32_t launched = 0, done_level = 0;
  for (;;) {
    if (per_step && prof->events.size() < 2 * static_cast<size_t>(chunk)) {
      const size_t old = prof->events.size();
      prof->events.resize(2 * static_cast<size_t>(chunk));
      for (size_t i = old; i < prof->events.size(); ++i) VGX_HIP(hipEventCreate(&prof->events[i]));
    }
    for (uint32_t i = 0; i < chunk; ++i) {
      if (per_step) VGX_HIP(hipEventRecord(prof->events[2 * i], s));
      k_bfs_step<<<blocks, kBfsBlock, 0, s>>>(a, launched + i);
      VGX_LAUNCH_CHECK();
      if (per_step) VGX_HIP(hipEventRecord(prof->events[2 * i + 1], s));
    }
    launched += chunk;
    if (whole) VGX_HIP(hipEventRecord(prof->tev[1], s));
    VGX_HIP(hipMemcpyAsync(h_st, a.st, sizeof(BfsState), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    if (whole) {
      float ms = 0.f;
      VGX_HIP(hipEventElapsedTime(&ms, prof->tev[0], prof->tev[1]));
      prof->traversal_ms += ms;  // over all chunks of this traversal
      VGX_HIP(hipEventRecord(prof->tev[0], s));
    }
    if (per_step) {
      for (uint32_t i = 0; i < chunk; ++i) {
        float ms = 0.f;
        VGX_HIP(hipEventElapsedTime(&ms, prof->events[2 * i], prof->events[2 * i + 1]));
        const uint32_t step = launched - chunk + i;
        const uint32_t rec = step < kLogSteps ? h_st->log_step[step] : 0u;
        // steps past the end (phase done) count as launches of the last level
        const uint32_t lv = rec & 0xFFFFFFu;
        if (step < kLogSteps && lv < kLogLevels) prof->last_level_ms[lv] += ms;
        prof->level_ms += ms;
      }
      prof->level_launches += chunk;
    }
    if (traversal_over(*h_st, launched, &done_level)) break;
    if (launched > 3 * n + 8) throw Error(VGX_DEVICE_ERROR, "BFS did not terminate");
    chunk = std::max<uint32_t>(2, chunk / 2);
  }
  if (whole) prof->traversals += 1;
  // steps this traversal needed: the first one that found the traversal over
  uint32_t needed = launched;
  for (uint32_t i = 0; i < launched && i < kLogSteps; ++i)
    if ((h_st->log_step[i] >> 24) == kPhDone) {
      needed = i;
      break;
    }
  if (g->bfs_steps_version != g->version) {
    g->bfs_steps_version = g->version;
    g->bfs_steps = needed;
  } else {
    g->bfs_steps = std::max(g->bfs_steps, needed);
  }
  TraversalResult r;
  r.done_level = done_level;
  r.steps = needed;
  r.bu_levels = h_st->bu_levels;
  r.reached = h_st->reached;
  return r;
}

static void bfs_full(vgx_bfs* b) {
  b->views.wait(b->g->stream);
  vgx_graph* g = b->g;
  ensure_work(b);
  std::fill(std::begin(b->last_level_ms), std::end(b->last_level_ms), 0.f);
  const uint64_t n = g->n;
  b->dist_len = n;
  const uint64_t words = (n + 63) / 64;
  // the dense head array and the zero-in-degree bitmap follow the rows
  if (b->w.head_version != g->rows_epoch || b->w.head_n != n) {
    b->w.head.ensure(g->ncap + 1);
    b->w.zin.ensure(words + 1);
    b->w.zcount.ensure(kZinStripes * kStripeWords);
    build_head_zin(g, b->w.head.p, b->w.zin.p, b->w.zcount.p, g->stream);
    b->w.head_version = g->rows_epoch;
    b->w.head_n = n;
  }
  const TraversalResult r =
      run_traversal(g, level_args(b), b->w.zin.p, b->w.zcount.p, b->source, b->tune, b->h_st, b);
  b->steps = r.steps;
  b->reachable = r.reached;
  b->levels = r.done_level;
  b->bu_levels = r.bu_levels;
}

// ------------------------------------------------ allocation-free traversal
// For callers that cannot allocate device memory: the traversal's scratch
// (state, visited and frontier bitmaps, queues, pending rows, head array,
// zero-in-degree bitmap) is one file-scope device array sized for graphs of
// at most kNoAllocVertices vertices, and the distances go to the caller's
// buffer.  Every call on a GPU uses that GPU's copy of the array and leaves
// its contents behind; each call rebuilds or re-initialises every part it
// reads (head and zin from the graph, the rest in k_bfs_init and the steps)
// before reading it, so nothing has to start at zero.

constexpr uint64_t kNoAllocVertices = 1ull << 20;

struct NoAllocLayout {
  size_t st, visited, bm, queue0, queue1, pre0, pre1, pend, ppre, head, zin, zcount, total;
};
constexpr size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }
constexpr NoAllocLayout no_alloc_layout(uint64_t n) {
  NoAllocLayout l{};
  const size_t words = (n + 63) / 64;
  size_t o = 0;
  l.st = o;
  o = align256(o + sizeof(BfsState));
  l.visited = o;
  o = align256(o + words * 8);
  l.bm = o;
  o = align256(o + 3 * words * 8);
  l.queue0 = o;
  o = align256(o + n * 4);
  l.queue1 = o;
  o = align256(o + n * 4);
  l.pre0 = o;
  o = align256(o + n * 8);
  l.pre1 = o;
  o = align256(o + n * 8);
  l.pend = o;
  o = align256(o + n * 4);
  l.ppre = o;
  o = align256(o + n * 8);
  l.head = o;
  o = align256(o + (n + 1) * 16);
  l.zin = o;
  o = align256(o + (words + 1) * 8);
  l.zcount = o;
  o = align256(o + kZinStripes * kStripeWords * 8);
  l.total = o;
  return l;
}
constexpr NoAllocLayout kNoAllocLayout = no_alloc_layout(kNoAllocVertices);

}  // namespace vgx

// The scratch array (about 53 MB).  In host code this name is a host-side
// stand-in, not a device address: only the address from
// hipGetSymbolAddress (kept per device below) may reach a kernel or a copy.
static __device__ __align__(256) unsigned char scratch[vgx::kNoAllocLayout.total];

namespace vgx {

constexpr int kMaxNoAllocDevices = 64;
static std::mutex g_noalloc_mu;
static unsigned char* g_noalloc_ptr[kMaxNoAllocDevices] = {};
// host copies of the traversal state, one per device (pageable memory)
static BfsState g_noalloc_state[kMaxNoAllocDevices];

// Device address of `scratch` on `device` (current), looked up on the first
// call on that device.
static unsigned char* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxNoAllocDevices) throw Error(VGX_INVALID_ARGUMENT, "device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    VGX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(VGX_DEVICE_ERROR, "no device address for the scratch array");
    g_noalloc_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_noalloc_ptr[device];
}

static uint64_t bfs_noalloc(vgx_graph* g, uint32_t source, uint32_t* dist_dev) {
  unsigned char* base = noalloc_scratch(g->device);
  const NoAllocLayout& l = kNoAllocLayout;
  LevelArgs a;
  a.row = g->adj[0].row.p;
  a.arena = g->adj[0].arena.p;
  a.in_row = g->in().row.p;
  a.in_arena = g->in().arena.p;
  a.dist = dist_dev;
  a.visited = reinterpret_cast<unsigned long long*>(base + l.visited);
  a.bm = reinterpret_cast<unsigned long long*>(base + l.bm);
  a.queue0 = reinterpret_cast<uint32_t*>(base + l.queue0);
  a.queue1 = reinterpret_cast<uint32_t*>(base + l.queue1);
  a.pre0 = reinterpret_cast<unsigned long long*>(base + l.pre0);
  a.pre1 = reinterpret_cast<unsigned long long*>(base + l.pre1);
  a.pend = reinterpret_cast<uint32_t*>(base + l.pend);
  a.ppre = reinterpret_cast<unsigned long long*>(base + l.ppre);
  a.st = reinterpret_cast<BfsState*>(base + l.st);
  a.n = static_cast<uint32_t>(g->n);
  a.words = static_cast<uint32_t>((g->n + 63) / 64);
  auto* head = reinterpret_cast<uint4*>(base + l.head);
  auto* zin = reinterpret_cast<unsigned long long*>(base + l.zin);
  auto* zcount = reinterpret_cast<unsigned long long*>(base + l.zcount);
  a.head = head;
  build_head_zin(g, head, zin, zcount, g->stream);
  return run_traversal(g, a, zin, zcount, source, kDefaultTuning, &g_noalloc_state[g->device], nullptr).reached;
}

static IncArgs inc_args(vgx_bfs* b) {
  vgx_graph* g = b->g;
  IncArgs a;
  a.out_row = g->adj[0].row.p;
  a.out_arena = g->adj[0].arena.p;
  a.in_row = g->in().row.p;
  a.in_arena = g->in().arena.p;
  a.dist = b->w.dist.p;
  a.lost = b->w.lost.p;
  a.support = b->w.support.p;
  a.flags = b->w.flags.p;
  a.stamp = b->w.stamp.p;
  a.st = b->d_inc;
  a.epoch = b->epoch;
  a.source = b->source;
  a.lcap = b->w.cap;
  return a;
}

// Reserve `count` fresh stamp epochs and return the first; the stamps are
// cleared when the 32-bit epoch would wrap.
static uint32_t take_epochs(vgx_bfs* b, uint32_t count) {
  if (b->epoch > 0xFFFFFFFFu - count - 1) {
    VGX_HIP(hipMemsetAsync(b->w.stamp.p, 0, b->w.cap * 4, b->g->stream));
    b->epoch = 0;
  }
  const uint32_t first = b->epoch + 1;
  b->epoch += count;
  return first;
}

// One synchronisation reads the repair state and, with `graph`, the
// graph's counters too (staging or a finished batch left them for the host).
static void inc_sync(vgx_bfs* b, bool graph = false) {
  if (graph)
    VGX_HIP(hipMemcpyAsync(b->g->h_ctr, b->g->d_ctr, sizeof(GraphCounters), hipMemcpyDeviceToHost, b->g->stream));
  VGX_HIP(hipMemcpyAsync(b->h_inc, b->d_inc, sizeof(IncState), hipMemcpyDeviceToHost, b->g->stream));
  VGX_HIP(hipStreamSynchronize(b->g->stream));
}

// Invalidation rounds and propagation rounds launched between two host
// synchronisations; a round with nothing to do exits at once.
constexpr uint32_t kInvalidationRounds = 2;
constexpr uint32_t kPropagationRounds = 4;

static void bfs_apply(vgx_bfs* b, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k,
                      bool device = false, hipStream_t producer = nullptr) {
  b->views.wait(b->g->stream);
  b->last_candidates = 0;
  b->last_affected = 0;
  b->last_relaxed = 0;
  b->last_full = false;
  if (k == 0) return;
  vgx_graph* g = b->g;
  hipStream_t s = g->stream;
  const uint64_t n_pre = g->n;
  uint64_t limit = static_cast<uint64_t>(static_cast<double>(n_pre) * b->fallback);
  if (limit < 1) limit = 1;

  // Staging leaves its counts on the device: the deletions and the first
  // invalidation rounds read them there, and the host collects them at the
  // invalidation's first synchronisation.
  if (device) graph_stage_device(g, src, dst, add, k, producer, true);  // may grow vertices
  else graph_stage(g, src, dst, add, k, true);
  extend_dist(b);
  BfsWork& w = b->w;
  VGX_HIP(hipMemsetAsync(b->d_inc, 0, sizeof(IncState), s));
  graph_apply_deletions(g);
  BatchScratch& bs = g->bs[0];

  // ---- invalidation cascade on the graph without the deleted arcs
  IncArgs a = inc_args(b);
  a.epoch = take_epochs(b, 1);
  if (bs.m) {
    k_inc_candidates<<<blocks_for(bs.m), kBfsBlock, 0, s>>>(a, bs.fin.p, &g->d_ctr->nfinal[0], bs.skeys.p, bs.act.p,
                                                            w.tlist.p, w.tall.p, w.tflist.p, w.tfpre.p);
    VGX_LAUNCH_CHECK();
  }
  bool fallback = false;
  bool collected = false;
  for (uint32_t r0 = 0;; r0 += kInvalidationRounds) {
    const uint32_t e0 = take_epochs(b, kInvalidationRounds);
    for (uint32_t i = 0; i < kInvalidationRounds; ++i) {
      const int par = static_cast<int>((r0 + i) & 1);
      k_inc_support<<<kRepairBlocks, kBfsBlock, 0, s>>>(a, par, w.tflist.p, w.tfpre.p);
      VGX_LAUNCH_CHECK();
      k_inc_evaluate<<<kRepairBlocks, kBfsBlock, 0, s>>>(a, par, w.tlist.p, w.alist.p, w.apre.p, w.aall.p);
      VGX_LAUNCH_CHECK();
      IncArgs an = a;
      an.epoch = e0 + i;  // the next round's touches
      k_inc_emit<<<kRepairBlocks, kBfsBlock, 0, s>>>(an, par, limit, w.alist.p, w.apre.p, w.tlist.p, w.tall.p,
                                                     w.tflist.p, w.tfpre.p);
      VGX_LAUNCH_CHECK();
    }
    inc_sync(b, !collected);
    if (!collected) {
      graph_stage_collect(g);  // may grow the arena (rows keep their places)
      collected = true;
    }
    b->last_candidates = b->h_inc->candidates;
    if (b->h_inc->fallback || b->h_inc->aall > limit) {
      fallback = true;
      break;
    }
    if (b->h_inc->tq[(r0 + kInvalidationRounds) & 1] == 0) break;  // nothing touched for the next round
    if (r0 > n_pre + 2) throw Error(VGX_DEVICE_ERROR, "invalidation did not terminate");
  }
  const uint64_t n_aff = b->h_inc->aall;
  b->last_affected = n_aff;

  // ---- apply the additions; the graph is now the post-batch graph
  graph_apply_additions(g);
  if (fallback) {
    graph_finish_batch(g);  // may compact: row pointers change, refresh args
    extend_dist(b);
    a = inc_args(b);
    k_inc_cleanup<<<kRepairBlocks, kBfsBlock, 0, s>>>(a, w.tall.p);
    VGX_LAUNCH_CHECK();
    bfs_full(b);
    b->last_full = true;
    return;
  }
  // The batch's bookkeeping (which may compact and move rows) waits for the
  // propagation's last synchronisation: until then the rows stay where the
  // additions left them, which is what the propagation reads.
  a = inc_args(b);

  // ---- re-relaxation: seeds from unaffected in-neighbours and the added
  // arcs, then Bellman-Ford rounds, several per synchronisation
  a.epoch = take_epochs(b, 1);
  if (n_aff) {
    k_inc_reset_affected<<<blocks_for(n_aff), kBfsBlock, 0, s>>>(a, w.aall.p, w.slist.p, w.spre.p);
    VGX_LAUNCH_CHECK();
    k_inc_seed<<<kRepairBlocks, kBfsBlock, 0, s>>>(a, w.slist.p, w.spre.p);
    VGX_LAUNCH_CHECK();
    k_inc_collect_seeds<<<blocks_for(n_aff), kBfsBlock, 0, s>>>(a, w.aall.p, w.queue[0].p, w.pre[0].p);
    VGX_LAUNCH_CHECK();
  }
  // The staged batch of the out-direction still lists every final arc.
  if (bs.nfinal) {
    k_inc_relax_additions<<<blocks_for(bs.nfinal), kBfsBlock, 0, s>>>(a, bs.fin.p, bs.nfinal, bs.skeys.p,
                                                                      bs.act.p, w.queue[0].p, w.pre[0].p);
    VGX_LAUNCH_CHECK();
  }
  for (uint32_t r0 = 0;; r0 += kPropagationRounds) {
    const uint32_t e0 = take_epochs(b, kPropagationRounds);
    for (uint32_t i = 0; i < kPropagationRounds; ++i) {
      const uint32_t r = r0 + i;
      IncArgs ar = a;
      ar.epoch = e0 + i;
      k_inc_propagate<<<kRepairBlocks, kBfsBlock, 0, s>>>(ar, r, w.queue[r & 1].p, w.pre[r & 1].p,
                                                          w.queue[(r + 1) & 1].p, w.pre[(r + 1) & 1].p);
      VGX_LAUNCH_CHECK();
    }
    inc_sync(b, true);  // the finished batch's counters come along
    if (b->h_inc->fq[(r0 + kPropagationRounds) % 3] == 0) break;
    if (r0 > n_pre + 2) throw Error(VGX_DEVICE_ERROR, "propagation did not terminate");
  }
  // scratch flags only: the distances are final at the synchronisation above
  k_inc_cleanup<<<kRepairBlocks, kBfsBlock, 0, s>>>(a, w.tall.p);
  VGX_LAUNCH_CHECK();
  graph_finish_collect(g);  // may compact: rows move, the distances do not
  b->reachable = b->reachable - n_aff + b->h_inc->became_reachable;
  b->last_relaxed = b->h_inc->relaxed;
}

// ------------------------------------------------------ adaptive selector

// Shallow shortest-parent deletions of a batch on the pre-batch distances
// (benchmarks/adaptive_policy_bfs.cpp:126-148): a deletion (u, v) counts
// when d(u) + 1 = d(v) <= 3, and for undirected graphs also when
// d(v) + 1 = d(u) <= 3.
constexpr uint32_t kShallowDeletionDepth = 3;

__global__ void k_shallow_deletions(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                    const uint8_t* __restrict__ add, uint64_t k, const uint32_t* __restrict__ dist,
                                    uint64_t n, int undirected, unsigned long long* count) {
  unsigned long long c = 0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < k;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (add[i]) continue;
    const uint32_t u = src[i], v = dst[i];
    if (u >= n || v >= n) continue;
    const uint32_t du = dist[u], dv = dist[v];
    if (du == kInf || dv == kInf) continue;
    if (du + 1 == dv && dv <= kShallowDeletionDepth) ++c;
    if (undirected && dv + 1 == du && du <= kShallowDeletionDepth) ++c;
  }
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(count, c);
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

struct vgx_policy {
  vgx::AdaptivePolicy model;
};

struct vgx_selector {
  vgx_bfs* b = nullptr;
  vgx::AdaptivePolicy model{0, 0.0, 0.0};
  bool first_batch = true;
  vgx::DevBuf<uint8_t> ops;  // staged first batch: src | dst | add
  unsigned long long* d_cnt = nullptr;
  unsigned long long* h_cnt = nullptr;
};

namespace {

using SelClock = std::chrono::steady_clock;

double us_since(SelClock::time_point t0) {
  return std::chrono::duration<double, std::micro>(SelClock::now() - t0).count();
}

vgx::PolicyFeatures features_of(const vgx_selector_trace& t) {
  vgx::PolicyFeatures f;
  f.update_fraction = t.update_fraction;
  f.shallow_parent_deletion_fraction = t.shallow_parent_deletion_fraction;
  f.first_batch = t.first_batch != 0;
  return f;
}

void decide_into(const vgx::AdaptivePolicy& m, vgx_selector_trace* t) {
  const vgx::PolicyDecision d = m.decide(features_of(*t));
  t->chose_full = d.full ? 1 : 0;
  t->reason = d.reason;
  t->predicted_incremental_us = d.predicted_incremental_us;
  t->predicted_full_us = d.predicted_full_us;
}

void observe_from(vgx::AdaptivePolicy& m, const vgx_selector_trace& t) {
  vgx::PolicyDecision d;
  d.full = t.chose_full != 0;
  d.reason = t.reason;
  d.predicted_incremental_us = t.predicted_incremental_us;
  d.predicted_full_us = t.predicted_full_us;
  m.observe(features_of(t), d, t.internal_fallback != 0, t.execution_us, t.affected_vertices);
}

// Fraction of the batch's ops that delete a shallow shortest-parent arc,
// evaluated on the device distances before the batch.
double shallow_fraction(vgx_selector* sel, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                        uint64_t k) {
  vgx_bfs* b = sel->b;
  hipStream_t s = b->g->stream;
  sel->ops.ensure(k * 9 + 16);
  uint32_t* d_src = reinterpret_cast<uint32_t*>(sel->ops.p);
  uint32_t* d_dst = d_src + k;
  uint8_t* d_add = reinterpret_cast<uint8_t*>(d_dst + k);
  VGX_HIP(hipMemcpyAsync(d_src, src, k * 4, hipMemcpyHostToDevice, s));
  VGX_HIP(hipMemcpyAsync(d_dst, dst, k * 4, hipMemcpyHostToDevice, s));
  VGX_HIP(hipMemcpyAsync(d_add, add, k, hipMemcpyHostToDevice, s));
  VGX_HIP(hipMemsetAsync(sel->d_cnt, 0, 8, s));
  k_shallow_deletions<<<blocks_for(k), kBfsBlock, 0, s>>>(d_src, d_dst, d_add, k, b->w.dist.p, b->dist_len,
                                                          b->g->directed ? 0 : 1, sel->d_cnt);
  VGX_LAUNCH_CHECK();
  VGX_HIP(hipMemcpyAsync(sel->h_cnt, sel->d_cnt, 8, hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  return static_cast<double>(*sel->h_cnt) / static_cast<double>(std::max<uint64_t>(1, k));
}

}  // namespace

extern "C" {

vgx_status vgx_bfs_create(vgx_graph* g, uint32_t source, double deletion_fallback_fraction, vgx_bfs** out) {
  VGX_API_BEGIN
  if (!g || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(g->device));
  auto* b = new vgx_bfs();
  try {
    b->g = g;
    b->source = source;
    b->fallback = deletion_fallback_fraction;
    b->d_st = static_cast<BfsState*>(dev_alloc(sizeof(BfsState)));
    b->d_inc = static_cast<IncState*>(dev_alloc(sizeof(IncState)));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_st), sizeof(BfsState), hipHostMallocDefault));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_inc), sizeof(IncState), hipHostMallocDefault));
    std::memset(b->h_inc, 0, sizeof(IncState));
    bfs_full(b);
  } catch (...) {
    vgx_bfs_destroy(b);
    throw;
  }
  *out = b;
  VGX_API_END
}

vgx_status vgx_bfs_destroy(vgx_bfs* b) {
  VGX_API_BEGIN
  if (!b) return VGX_OK;
  if (b->g) (void)hipStreamSynchronize(b->g->stream);
  if (b->d_st) dev_free(b->d_st);
  if (b->d_inc) dev_free(b->d_inc);
  if (b->h_st) (void)hipHostFree(b->h_st);
  if (b->h_inc) (void)hipHostFree(b->h_inc);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->tev)
    if (e) (void)hipEventDestroy(e);
  b->views.destroy();
  delete b;
  VGX_API_END
}

vgx_status vgx_bfs_recompute(vgx_bfs* b) {
  VGX_API_BEGIN
  if (!b) throw Error(VGX_INVALID_ARGUMENT, "null bfs");
  VGX_HIP(hipSetDevice(b->g->device));
  bfs_full(b);
  VGX_API_END
}

vgx_status vgx_bfs_set_source(vgx_bfs* b, uint32_t source) {
  VGX_API_BEGIN
  if (!b) throw Error(VGX_INVALID_ARGUMENT, "null bfs");
  VGX_HIP(hipSetDevice(b->g->device));
  b->source = source;
  bfs_full(b);
  VGX_API_END
}

vgx_status vgx_bfs_distances_noalloc(vgx_graph* g, uint32_t source, uint32_t* dist_dev, uint64_t cap,
                                     uint64_t* reachable) {
  VGX_API_BEGIN
  if (!g || !dist_dev) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (g->n > kNoAllocVertices)
    throw Error(VGX_INVALID_ARGUMENT, "the allocation-free BFS takes graphs of at most 1048576 vertices; this one has " +
                                          std::to_string(g->n));
  if (cap < g->n) throw Error(VGX_INVALID_ARGUMENT, "dist_dev holds fewer entries than the graph has vertices");
  VGX_HIP(hipSetDevice(g->device));
  const uint64_t r = g->n ? bfs_noalloc(g, source, dist_dev) : 0;
  if (reachable) *reachable = r;
  VGX_API_END
}

vgx_status vgx_bfs_apply(vgx_bfs* b, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k) {
  VGX_API_BEGIN
  if (!b || (k && (!src || !dst || !add))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(b->g->device));
  bfs_apply(b, src, dst, add, k);
  VGX_API_END
}

vgx_status vgx_bfs_apply_device(vgx_bfs* b, const uint32_t* src_dev, const uint32_t* dst_dev, const uint8_t* add_dev,
                                uint64_t k, void* stream) {
  VGX_API_BEGIN
  if (!b || (k && (!src_dev || !dst_dev || !add_dev))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(b->g->device));
  bfs_apply(b, src_dev, dst_dev, add_dev, k, true, static_cast<hipStream_t>(stream));
  VGX_API_END
}

vgx_status vgx_bfs_distances(vgx_bfs* b, uint32_t* out, uint64_t cap, uint64_t* count) {
  VGX_API_BEGIN
  if (!b || !count) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(b->g->device));
  *count = b->dist_len;
  uint64_t c = std::min<uint64_t>(cap, b->dist_len);
  if (out && c) {
    VGX_HIP(hipMemcpyAsync(out, b->w.dist.p, c * sizeof(uint32_t), hipMemcpyDeviceToHost, b->g->stream));
    VGX_HIP(hipStreamSynchronize(b->g->stream));
  }
  VGX_API_END
}

vgx_status vgx_bfs_graph(const vgx_bfs* b, vgx_graph** g) {
  VGX_API_BEGIN
  if (!b || !g) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  *g = b->g;
  VGX_API_END
}

vgx_status vgx_bfs_hold_view(vgx_bfs* b, int delta, void* stream) {
  VGX_API_BEGIN
  if (!b || (delta != 1 && delta != -1)) throw Error(VGX_INVALID_ARGUMENT, "delta must be +1 or -1");
  VGX_HIP(hipSetDevice(b->g->device));
  if (delta > 0) b->views.hold(static_cast<hipStream_t>(stream));
  else b->views.release();
  VGX_API_END
}

vgx_status vgx_bfs_distances_device(vgx_bfs* b, const uint32_t** dist_dev, uint64_t* count) {
  VGX_API_BEGIN
  if (!b || !dist_dev) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  *dist_dev = b->w.dist.p;
  if (count) *count = b->dist_len;
  VGX_API_END
}

vgx_status vgx_bfs_stats_get(vgx_bfs* b, vgx_bfs_stats* out) {
  VGX_API_BEGIN
  if (!b || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  out->reachable_count = b->reachable;
  out->last_deletion_candidates = b->last_candidates;
  out->last_affected_vertices = b->last_affected;
  out->last_used_full_recompute = b->last_full ? 1 : 0;
  out->levels = b->levels;
  out->bottom_up_levels = b->bu_levels;
  out->last_relaxed_vertices = b->last_relaxed;
  out->level_launches = b->level_launches;
  out->level_kernel_ms = b->level_ms;
  out->steps = b->steps;
  out->traversal_ms = b->traversal_ms;
  out->traversals = b->traversals;
  VGX_API_END
}

vgx_status vgx_bfs_set_profiling(vgx_bfs* b, int on) {
  VGX_API_BEGIN
  if (!b) throw Error(VGX_INVALID_ARGUMENT, "null bfs");
  if (on < 0 || on > 2) throw Error(VGX_INVALID_ARGUMENT, "profiling mode is 0, 1 or 2");
  b->profiling = on;
  b->level_launches = 0;
  b->level_ms = 0.0;
  b->traversal_ms = 0.0;
  b->traversals = 0;
  VGX_API_END
}

vgx_status vgx_bfs_set_tuning(vgx_bfs* b, const vgx_bfs_tuning* t) {
  VGX_API_BEGIN
  if (!b || !t) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (!(t->alpha > 0) || !(t->beta > 0)) throw Error(VGX_INVALID_ARGUMENT, "alpha and beta must be positive");
  if (t->blocks > 65535) throw Error(VGX_INVALID_ARGUMENT, "too many workgroups");
  if (!(t->tail_ratio >= 0) || !(t->claim_ratio >= 0))
    throw Error(VGX_INVALID_ARGUMENT, "tail_ratio and claim_ratio must be non-negative");
  b->tune = *t;
  VGX_API_END
}

vgx_status vgx_bfs_get_tuning(vgx_bfs* b, vgx_bfs_tuning* t) {
  VGX_API_BEGIN
  if (!b || !t) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  *t = b->tune;
  VGX_API_END
}

vgx_status vgx_bfs_autotune(vgx_bfs* b, uint32_t reps, vgx_bfs_tuning* best, double* best_ms) {
  VGX_API_BEGIN
  if (!b) throw Error(VGX_INVALID_ARGUMENT, "null bfs");
  if (reps == 0) reps = 1;
  VGX_HIP(hipSetDevice(b->g->device));
  const int prof = b->profiling;
  b->profiling = 0;
  auto time_of = [&](const vgx_bfs_tuning& t) {
    b->tune = t;
    bfs_full(b);  // settles the graph's step count for this setting
    const auto t0 = SelClock::now();
    for (uint32_t r = 0; r < reps; ++r) bfs_full(b);
    return us_since(t0) / 1e3 / reps;
  };
  vgx_bfs_tuning cur = b->tune;
  double cur_ms = time_of(cur);
  auto sweep = [&](auto setter, const auto& values) {
    for (auto v : values) {
      vgx_bfs_tuning t = cur;
      setter(t, v);
      const double ms = time_of(t);
      if (ms < cur_ms) {
        cur_ms = ms;
        cur = t;
      }
    }
  };
  const double alphas[] = {5.0, 10.0, 15.0, 25.0, 40.0};
  const double betas[] = {12.0, 18.0, 24.0, 48.0};
  const uint32_t blocks[] = {0u, 512u, 1024u, 2048u};
  const double tails[] = {8.0, 16.0, 32.0, 64.0, 128.0};
  const double claims[] = {2.0, 3.0, 5.0, 8.0, 12.0};
  sweep([](vgx_bfs_tuning& t, double v) { t.alpha = v; }, alphas);
  sweep([](vgx_bfs_tuning& t, double v) { t.beta = v; }, betas);
  sweep([](vgx_bfs_tuning& t, uint32_t v) { t.blocks = v; }, blocks);
  sweep([](vgx_bfs_tuning& t, double v) { t.tail_ratio = v; }, tails);
  sweep([](vgx_bfs_tuning& t, double v) { t.claim_ratio = v; }, claims);
  b->tune = cur;
  b->profiling = prof;
  if (best) *best = cur;
  if (best_ms) *best_ms = cur_ms;
  VGX_API_END
}

vgx_status vgx_bfs_level_log(vgx_bfs* b, vgx_bfs_level* out, uint32_t cap, uint32_t* count) {
  VGX_API_BEGIN
  if (!b || !count) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(b->g->device));
  VGX_HIP(hipMemcpyAsync(b->h_st, b->d_st, sizeof(BfsState), hipMemcpyDeviceToHost, b->g->stream));
  VGX_HIP(hipStreamSynchronize(b->g->stream));
  const uint32_t levels = std::min<uint32_t>(b->levels, kLogLevels);
  *count = levels;
  for (uint32_t i = 0; i < levels && i < cap && out; ++i) {
    out[i].bottom_up = b->h_st->log_mode[i];
    out[i].frontier_vertices = b->h_st->log_nf[i];
    out[i].frontier_edges = b->h_st->log_mf[i];
    out[i].pending_vertices = b->h_st->log_pend[i] >> kEdgeBits;
    out[i].pending_edges = b->h_st->log_pend[i] & kEdgeMask;
    out[i].kernel_ms = b->last_level_ms[i];
  }
  VGX_API_END
}

const char* vgx_policy_reason_name(int reason) { return vgx::policy_reason_name(reason); }

vgx_status vgx_policy_create(uint64_t vertices, double initial_full_us, double initial_reachable_fraction,
                             vgx_policy** out) {
  VGX_API_BEGIN
  if (!out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (!(initial_full_us >= 0.0)) throw Error(VGX_INVALID_ARGUMENT, "initial_full_us must be non-negative");
  *out = new vgx_policy{vgx::AdaptivePolicy(vertices, initial_full_us, initial_reachable_fraction)};
  VGX_API_END
}

vgx_status vgx_policy_destroy(vgx_policy* p) {
  VGX_API_BEGIN
  delete p;
  VGX_API_END
}

vgx_status vgx_policy_decide(vgx_policy* p, vgx_selector_trace* t) {
  VGX_API_BEGIN
  if (!p || !t) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  decide_into(p->model, t);
  VGX_API_END
}

vgx_status vgx_policy_observe(vgx_policy* p, const vgx_selector_trace* t) {
  VGX_API_BEGIN
  if (!p || !t) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  observe_from(p->model, *t);
  VGX_API_END
}

vgx_status vgx_selector_create(vgx_bfs* b, vgx_selector** out) {
  VGX_API_BEGIN
  if (!b || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(b->g->device));
  auto* sel = new vgx_selector();
  try {
    sel->b = b;
    sel->d_cnt = static_cast<unsigned long long*>(dev_alloc(8));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&sel->h_cnt), 8, hipHostMallocDefault));
    // the initial full BFS is the full arm's first sample (harness :59-65)
    const auto t0 = SelClock::now();
    bfs_full(b);
    const double full_us = us_since(t0);
    const uint64_t n = b->g->n;
    const double reach = static_cast<double>(b->reachable) / static_cast<double>(std::max<uint64_t>(1, n));
    sel->model = vgx::AdaptivePolicy(n, full_us, reach);
  } catch (...) {
    vgx_selector_destroy(sel);
    throw;
  }
  *out = sel;
  VGX_API_END
}

vgx_status vgx_selector_destroy(vgx_selector* sel) {
  VGX_API_BEGIN
  if (!sel) return VGX_OK;
  if (sel->b) (void)hipStreamSynchronize(sel->b->g->stream);
  if (sel->d_cnt) dev_free(sel->d_cnt);
  if (sel->h_cnt) (void)hipHostFree(sel->h_cnt);
  delete sel;
  VGX_API_END
}

vgx_status vgx_selector_apply(vgx_selector* sel, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                              uint64_t k, vgx_selector_trace* trace) {
  VGX_API_BEGIN
  if (!sel || (k && (!src || !dst || !add))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  vgx_bfs* b = sel->b;
  vgx_graph* g = b->g;
  VGX_HIP(hipSetDevice(g->device));
  vgx_selector_trace t{};
  const auto t_dec = SelClock::now();
  t.first_batch = sel->first_batch ? 1 : 0;
  t.update_fraction =
      static_cast<double>(k) / static_cast<double>(std::max<uint64_t>(1, g->adj[0].live));
  t.shallow_parent_deletion_fraction = (sel->first_batch && k) ? shallow_fraction(sel, src, dst, add, k) : 0.0;
  decide_into(sel->model, &t);
  t.decision_us = us_since(t_dec);
  const auto t_exec = SelClock::now();
  if (t.chose_full) {
    // graph.apply(batch); bfs.recompute()  (harness :245-248)
    if (k) {
      graph_stage(g, src, dst, add, k);
      graph_apply_deletions(g);
      graph_apply_additions(g);
      graph_finish_batch(g);
    }
    bfs_full(b);
    b->last_candidates = 0;
    b->last_affected = 0;
    b->last_relaxed = 0;
    b->last_full = true;
  } else {
    bfs_apply(b, src, dst, add, k);
    t.affected_vertices = b->last_affected;
    t.internal_fallback = b->last_full ? 1 : 0;
  }
  VGX_HIP(hipStreamSynchronize(g->stream));
  t.execution_us = us_since(t_exec);
  observe_from(sel->model, t);
  sel->first_batch = false;
  if (trace) *trace = t;
  VGX_API_END
}

}  // extern "C"
