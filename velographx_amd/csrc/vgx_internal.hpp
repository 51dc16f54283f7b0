// Internal definitions shared by the gfx950 kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "vgx.h"

namespace vgx {

// Row entries are vertex ids below 2^31; bit 31 marks a deleted (tombstoned)
// entry that stays in place so rows keep their sorted order.
constexpr uint32_t kTomb = 0x80000000u;
constexpr uint32_t kIdMask = 0x7FFFFFFFu;
constexpr uint32_t kInf = 0xFFFFFFFFu;
constexpr uint64_t kInvalidKey = ~0ull;
// Packed frontier counter: vertex count in the high 26 bits, edge prefix in
// the low 38 bits, so one 64-bit atomic reserves a queue slot and its edge
// offset consistently.
constexpr int kEdgeBits = 38;
constexpr unsigned long long kEdgeMask = (1ull << kEdgeBits) - 1;

struct Error : std::runtime_error {
  vgx_status code;
  Error(vgx_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define VGX_HIP(x)                                                  \
  do {                                                              \
    hipError_t e_ = (x);                                            \
    if (e_ != hipSuccess) ::vgx::throw_hip(e_, #x, __FILE__, __LINE__); \
  } while (0)

#define VGX_LAUNCH_CHECK() VGX_HIP(hipGetLastError())

void* dev_alloc(size_t bytes);
void dev_free(void* p);

// Owning device buffer that grows on demand (contents are not preserved by
// ensure(); use grow_copy() when they must be).
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) dev_free(p);
    p = nullptr;
    n = 0;
  }
  // Capacity for m elements (contents not kept).  A buffer that has to grow
  // takes a quarter more than asked, so sizes that creep up from call to
  // call (batches, snapshots) do not reallocate every time.
  void ensure(size_t m) {
    if (m <= n) return;
    const size_t grown = n ? m + m / 4 : m;
    release();
    size_t c = grown < 1024 ? 1024 : grown;
    p = static_cast<T*>(dev_alloc(c * sizeof(T)));
    n = c;
  }
  // The same, keeping the old allocation in *keep (not freed) when given.
  void ensure(size_t m, std::vector<void*>* keep) {
    if (m <= n) return;
    if (keep && p) {
      keep->push_back(p);
      p = nullptr;
      n = 0;
    }
    ensure(m);
  }
  // Capacity for m elements keeping the first `used`; the old allocation
  // goes to *keep instead of being freed when keep is given.
  void grow_copy(size_t m, size_t used, hipStream_t s, std::vector<void*>* keep = nullptr) {
    if (m <= n) return;
    size_t c = m < 1024 ? 1024 : m;
    T* q = static_cast<T*>(dev_alloc(c * sizeof(T)));
    if (used) VGX_HIP(hipMemcpyAsync(q, p, used * sizeof(T), hipMemcpyDeviceToDevice, s));
    VGX_HIP(hipStreamSynchronize(s));
    if (keep && p) keep->push_back(p);
    else release();
    p = q;
    n = c;
  }
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
};

// An array that zero-copy views alias (a traversal's distances, a PageRank
// object's ranks).  While views are held, the engine's stream waits (on the
// device) for the work queued so far on the streams the views were taken
// for before each write to the array, and an allocation the array outgrows
// is kept, not freed, until the last view is released (then, once those
// streams have finished, it is freed).
struct ViewGuard {
  int views = 0;
  std::vector<hipStream_t> streams;
  std::vector<void*> retired;
  hipEvent_t ev = nullptr;

  std::vector<void*>* keep() { return views > 0 ? &retired : nullptr; }
  void hold(hipStream_t s) {
    ++views;
    for (hipStream_t x : streams)
      if (x == s) return;
    streams.push_back(s);
  }
  void wait(hipStream_t engine) {
    if (views <= 0) return;
    if (!ev) VGX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    for (hipStream_t x : streams) {
      VGX_HIP(hipEventRecord(ev, x));
      VGX_HIP(hipStreamWaitEvent(engine, ev, 0));
    }
  }
  void drop_retired() {
    for (hipStream_t x : streams) VGX_HIP(hipStreamSynchronize(x));
    for (void* q : retired) dev_free(q);
    retired.clear();
  }
  void release() {
    if (views <= 0) throw Error(VGX_INVALID_ARGUMENT, "no view is held");
    if (--views > 0) return;
    if (!retired.empty()) drop_retired();
    streams.clear();
  }
  // At the owner's destruction (no view can be out: a view keeps its owner).
  void destroy() {
    for (void* q : retired) dev_free(q);
    retired.clear();
    if (ev) (void)hipEventDestroy(ev);
    ev = nullptr;
  }
};

// Pinned host staging buffer.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~HostBuf();
  void ensure(size_t b);
};

// One direction of adjacency: a row table {beg, len} into an id arena.
// len counts stored entries including tombstones; cap >= len is the slot
// count reserved for the row.  Rows are sorted by id (tombstones included).
struct Adjacency {
  DevBuf<uint2> row;      // [ncap]
  DevBuf<uint32_t> cap;   // [ncap]
  DevBuf<uint32_t> deg;   // [ncap] live degree
  DevBuf<uint32_t> base;  // [ncap] len at the row's last bulk load or compaction
  DevBuf<uint32_t> arena; // [acap]
  uint64_t acap = 0;       // arena capacity (entries)
  uint64_t used = 0;       // host mirror of the bump pointer
  uint64_t stored = 0;     // sum of len (live + tombstones)
  uint64_t live = 0;       // live arcs
  uint64_t relocated = 0;  // rows moved since the last compaction
};

// Device-side counters shared with the host through one small copy.
struct GraphCounters {
  unsigned long long bump[2];        // arena bump pointer per direction
  unsigned long long need[2];        // arena slots requested by relocations
  unsigned long long live_delta[2];  // signed changes, as two's complement
  unsigned long long tomb_delta[2];
  unsigned long long stored_delta[2];
  unsigned long long relocated[2];
  unsigned long long nfinal[2];
  unsigned long long nins[2];
  unsigned long long nbig[2];        // rows relocated (outgrow their slots, or very long)
  unsigned long long nmid[2];        // rows of more than 32 entries merged in place
  unsigned long long reloc_chunks[2];  // output chunks of the relocated rows
};

// Staged update batch (device side).
struct BatchScratch {
  DevBuf<uint32_t> src, dst;
  DevBuf<uint8_t> add;
  DevBuf<unsigned long long> keys, skeys;
  DevBuf<uint32_t> vals, svals;
  DevBuf<uint8_t> act;      // per sorted arc: action code
  DevBuf<uint32_t> pos;     // per sorted arc: arena slot of the match
  DevBuf<uint32_t> fin;     // sorted positions of final arcs
  DevBuf<uint32_t> ins;     // sorted positions of inserts
  DevBuf<uint32_t> segk;    // per insert index: inserts in its row (segment start) or 0
  DevBuf<uint32_t> big;     // insert indices of rows relocated
  DevBuf<uint32_t> mid;     // insert indices of rows merged in place by a workgroup
  DevBuf<uint32_t> reloc;   // relocation plan: 8 words per relocated row + chunk prefix
  DevBuf<uint8_t> flags;
  DevBuf<uint8_t> tmp;      // library temp storage
  DevBuf<unsigned long long> extent;  // device batches: largest non-loop endpoint + 1
  DevBuf<unsigned long long> bcount;  // per-workgroup counts of final arcs and inserts (compaction)
  uint64_t k = 0, m = 0;    // ops, arcs
  bool counted = true;      // nfinal .. need below hold the device counts
  uint64_t nfinal = 0, nins = 0, nbig = 0, nmid = 0, need = 0;
};

enum ActCode : uint8_t { kActNopDel = 0, kActDel = 1, kActNopAdd = 2, kActUndel = 3, kActIns = 4 };

struct CsrSnapshot {
  DevBuf<uint32_t> off;  // [n+1]
  DevBuf<uint32_t> adj;  // [arcs]
  uint64_t version = ~0ull;
  uint64_t n = 0, arcs = 0;
};

// Chunk decomposition of one CSR direction: row v owns chunks
// [row_chunk[v], row_chunk[v+1]) of at most kPprChunk arcs each (an empty
// row owns one empty chunk), so hub rows spread over many waves.
struct PprPlan {
  DevBuf<uint32_t> row_chunk;  // [n+1]
  DevBuf<uint32_t> chunk_row;  // [nchunks]
  DevBuf<uint32_t> multi;      // rows of more than one chunk
  DevBuf<uint32_t> lchunk;     // chunks of rows longer than the short-row bound
  uint64_t nchunks = 0;
  uint64_t nmulti = 0;
  uint64_t nlchunks = 0;
  uint64_t version = ~0ull;
  uint64_t n = 0;
};

// Scratch of the weighted personalized PageRank kernels.
struct PprWork {
  DevBuf<float> pn;     // normalised weights, out-arc order
  DevBuf<float> pin;    // the same, in-arc order
  DevBuf<float> wsum;   // out-weight sums
  DevBuf<float> tmp;    // ping-pong n x K
  DevBuf<float> part;   // per-chunk partial rows, nchunks x K
  DevBuf<float> cpart;  // per-chunk scalar partials
  DevBuf<uint8_t> scan; // hipcub temporary storage
  PprPlan plan[2];      // [0] out-CSR, [1] in-CSR
};

}  // namespace vgx

namespace vgx {
// Scratch of snapshot packing and compaction, kept between calls so that a
// consolidation allocates nothing once the sizes have been seen: the spare
// arena and row table receive the compacted copy and are swapped with the
// live ones (the old ones become the spares).
struct PackScratch {
  DevBuf<uint32_t> off;          // row offsets of the packed copy (n + 2)
  DevBuf<uint32_t> spare_arena;
  DevBuf<uint2> spare_row;
  DevBuf<uint32_t> row_chunks, chunk_first;  // chunks per row and their exclusive scan
  DevBuf<uint32_t> chunk_u, chunk_j, chunk_live, chunk_scan;
  DevBuf<unsigned long long> counts;  // [0] chunks
  DevBuf<unsigned long long> rowsum;  // per-row state sums (delta queries, row compaction)
};
}  // namespace vgx

struct vgx_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  bool directed = false;
  uint64_t n = 0;
  uint64_t ncap = 0;
  uint64_t version = 0;
  // Advances whenever the arc set changes (with `version`) and whenever
  // compaction moves or shortens rows without changing it: the key of
  // every cache of row positions or stored lengths.
  uint64_t rows_epoch = 0;
  vgx::Adjacency adj[2];  // [0] out, [1] in (unused when undirected)
  vgx::GraphCounters* d_ctr = nullptr;
  vgx::GraphCounters* h_ctr = nullptr;  // pinned
  vgx::BatchScratch bs[2];
  vgx::HostBuf stage;
  vgx::CsrSnapshot snap[2];
  vgx::PprWork ppr;
  vgx::PackScratch pk;
  // launches (steps) the single-source traversals of this version have
  // needed at most, any root: the first chunk a new traversal enqueues
  uint32_t bfs_steps = 0;
  uint64_t bfs_steps_version = ~0ull;
  int ndir() const { return directed ? 2 : 1; }
  vgx::Adjacency& in() { return directed ? adj[1] : adj[0]; }
};

namespace vgx {

// Graph internals used by the algorithm files.
void graph_grow_vertices(vgx_graph* g, uint64_t n_new);
// Stages a batch: uploads, sorts, resolves actions; grows vertices; sizes
// the arena for the inserts.  Does not change adjacency.  With `defer` it
// returns without synchronising: the host counts of the staged batch (and
// the arena sizing) wait for graph_stage_collect, which the caller runs
// after its own next synchronisation of the counters (graph_sync_counters or
// a copy of g->d_ctr to g->h_ctr); graph_apply_deletions may run before.
void graph_stage(vgx_graph* g, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k,
                 bool defer = false);
// The same for a batch already in device memory (src, dst, add device
// pointers), written on `producer` (waited for first when it is not the
// graph's stream; null = the default stream).
void graph_stage_device(vgx_graph* g, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k,
                        hipStream_t producer, bool defer = false);
void graph_stage_collect(vgx_graph* g);
void graph_apply_deletions(vgx_graph* g);
void graph_apply_additions(vgx_graph* g);
// Host bookkeeping of a finished batch (live and stored counts, version,
// storage maintenance, which may compact and move rows): graph_finish_batch
// synchronises the counters first, graph_finish_collect expects them synced.
void graph_finish_batch(vgx_graph* g);
void graph_finish_collect(vgx_graph* g);
const CsrSnapshot& graph_snapshot(vgx_graph* g, int dir);
void graph_sync_counters(vgx_graph* g);

// -------------------------------------------------------------- wave helpers

constexpr int kWave = 64;

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
  const unsigned lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T y = __shfl_up(x, d, kWave);
    if (lane >= static_cast<unsigned>(d)) x += y;
  }
  return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

template <class T>
__device__ __forceinline__ T wave_max(T x) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    T y = __shfl_xor(x, d, kWave);
    x = x > y ? x : y;
  }
  return x;
}

__device__ __forceinline__ unsigned long long atomic_load_u64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_load_u32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-aggregated push of (v, len) into a packed frontier: one 64-bit atomic
// per wave reserves consecutive queue slots and their edge prefixes.  Must be
// called by all lanes of the wave.
__device__ __forceinline__ void wave_push(unsigned long long* ctr, uint32_t* queue, unsigned long long* pre,
                                          bool want, uint32_t v, uint32_t len) {
  unsigned long long ball = __ballot(want);
  if (ball == 0) return;
  uint32_t e = want ? len : 0u;
  unsigned long long einc = wave_inclusive_scan<unsigned long long>(e);
  unsigned long long etot = __shfl(einc, kWave - 1, kWave);
  uint32_t cnt = __popcll(ball);
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(ctr, (static_cast<unsigned long long>(cnt) << kEdgeBits) + etot);
  base = __shfl(base, 0, kWave);
  if (want) {
    uint32_t slot = static_cast<uint32_t>(base >> kEdgeBits) + lanes_below(ball);
    queue[slot] = v;
    pre[slot] = (base & kEdgeMask) + einc - e;
  }
}

// Load-balanced edge loop over a packed frontier (queue, pre, count F, edge
// total E): each wave takes a contiguous edge range and, for every edge,
// calls fn(item_index, vertex, row, arena_slot, lane_valid).  fn is called
// by all lanes (valid=false for padding) so it may use wave intrinsics.
// With Begins set the queue holds each item's row begin instead of its
// vertex (row is not read; fn gets the begin as both vertex and row.x), so
// no dependent row load stands between an item and its arena entries.
// A wave's range is a multiple of Quantum edges (a full window by default;
// less spreads a small frontier over more waves when fn does a wave's work
// per edge).
template <bool Begins = false, uint32_t Quantum = kWave, class Fn>
__device__ __forceinline__ void lbs_edges(const uint32_t* queue, const unsigned long long* pre, uint32_t F,
                                          unsigned long long E, const uint2* row, uint32_t gw, uint32_t W,
                                          Fn&& fn) {
  static_assert(Quantum > 0 && Quantum <= kWave && (Quantum & (Quantum - 1)) == 0, "quantum: a power of two <= 64");
  if (E == 0) return;
  unsigned long long S = (E + W - 1) / W;
  S = (S + Quantum - 1) & ~static_cast<unsigned long long>(Quantum - 1);
  unsigned long long a = static_cast<unsigned long long>(gw) * S;
  if (a >= E) return;
  unsigned long long b = a + S < E ? a + S : E;
  const unsigned lane = lane_id();
  // Last item whose prefix <= a, by a 64-way search: the wave samples 64
  // evenly spaced prefixes of [lo, hi) and keeps the stretch after the last
  // one <= a, so F items take ceil(log64 F) dependent loads instead of
  // log2 F.  Prefixes strictly increase, so the lanes that pass are a prefix
  // of the wave; lane 0 always passes (pre[lo] <= a).
  uint32_t lo = 0, hi = F;
  while (hi - lo > 1) {
    const uint32_t step = (hi - lo + kWave - 1) / kWave;
    const uint32_t idx = lo + lane * step;
    const bool ok = idx < hi && pre[idx] <= a;
    const unsigned long long m = __ballot(ok);
    const uint32_t last = 63 - __clzll(static_cast<long long>(m));
    lo += last * step;
    hi = lo + step < hi ? lo + step : hi;
  }
  // Invariant: item i0 contains edge e (pre[i0] <= e < pre[i0 + 1]).  The
  // items of the next window are loaded before fn runs on this one, so
  // their latency hides behind fn's own loads.
  uint32_t i0 = lo;
  unsigned long long pk = ~0ull;
  uint32_t qk = 0u;
  if (i0 + lane < F) {
    pk = pre[i0 + lane];
    qk = queue[i0 + lane];
  }
  for (unsigned long long e = a; e < b; e += kWave) {
    // Lane k holds item i0 + k of the window.  Every item has at least one
    // edge and item i0 + 1 starts after e, so the items starting inside this
    // 64-edge chunk are among the next 63; their start offsets, OR-ed into
    // one mask, give each lane its item with a popcount instead of a search.
    unsigned long long bit = (lane > 0 && pk < e + kWave) ? (1ull << (pk - e)) : 0ull;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) bit |= __shfl_xor(bit, d, kWave);
    const unsigned long long g = e + lane;
    const bool valid = g < b;
    const uint32_t k = __popcll(bit & ((2ull << lane) - 1ull));
    const uint32_t v = __shfl(qk, static_cast<int>(k), kWave);
    const unsigned long long pv = __shfl(pk, static_cast<int>(k), kWave);
    uint2 rb;
    if constexpr (Begins)
      rb = make_uint2(v, 0u);
    else
      rb = valid ? row[v] : make_uint2(0u, 0u);
    const uint32_t slot = valid ? rb.x + static_cast<uint32_t>(g - pv) : 0u;
    // the item containing edge e + 64: items of the window that start at or
    // before it, plus item i0 + 64 when all 63 others do
    const bool more = e + kWave < b;
    uint32_t adv = __popcll(__ballot(lane > 0 && pk <= e + kWave));
    if (more && adv == kWave - 1 && i0 + kWave < F && pre[i0 + kWave] <= e + kWave) adv = kWave;
    const uint32_t i1 = i0 + adv;
    unsigned long long pk1 = ~0ull;
    uint32_t qk1 = 0u;
    if (more && i1 + lane < F) {
      pk1 = pre[i1 + lane];
      qk1 = queue[i1 + lane];
    }
    fn(i0 + k, v, rb, slot, valid);
    if (!more) break;
    i0 = i1;
    pk = pk1;
    qk = qk1;
  }
}

// Per-wave staging of frontier pushes in LDS: pushes land in the wave's
// buffer and one packed atomic per flush reserves their queue slots and
// edge prefixes, instead of one atomic on a hot counter per push.
constexpr uint32_t kStage = 256;
struct WaveStage {
  uint32_t v[kStage];
  uint32_t len[kStage];
};

__device__ __forceinline__ void stage_flush(WaveStage& sb, uint32_t& nbuf, unsigned long long* ctr,
                                            uint32_t* queue, unsigned long long* pre) {
  if (nbuf == 0) return;
  __builtin_amdgcn_wave_barrier();
  const unsigned lane = lane_id();
  unsigned long long total = 0;
  for (uint32_t c = 0; c < nbuf; c += kWave) {
    const uint32_t l = c + lane < nbuf ? sb.len[c + lane] : 0u;
    total += wave_sum<unsigned long long>(l);
  }
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(ctr, (static_cast<unsigned long long>(nbuf) << kEdgeBits) + total);
  base = __shfl(base, 0, kWave);
  unsigned long long slot = base >> kEdgeBits;
  unsigned long long eoff = base & kEdgeMask;
  for (uint32_t c = 0; c < nbuf; c += kWave) {
    const bool in = c + lane < nbuf;
    const uint32_t l = in ? sb.len[c + lane] : 0u;
    const unsigned long long inc = wave_inclusive_scan<unsigned long long>(l);
    if (in) {
      queue[slot + c + lane] = sb.v[c + lane];
      pre[slot + c + lane] = eoff + inc - l;
    }
    eoff += __shfl(inc, kWave - 1, kWave);
  }
  __builtin_amdgcn_wave_barrier();
  nbuf = 0;
}

// Stage (v, len) for the lanes that want it; flushes when nearly full.
// Called by all lanes of the wave.
// Final flush of all waves of a workgroup with one atomic for the whole
// workgroup (instead of one per wave).  Every thread of the workgroup must
// call it; `sm` is a __shared__ array of 2 * waves + 1 words.
__device__ __forceinline__ void stage_flush_block(WaveStage& sb, uint32_t& nbuf, unsigned long long* ctr,
                                                  uint32_t* queue, unsigned long long* pre, unsigned long long* sm) {
  const unsigned lane = lane_id();
  const uint32_t wv = threadIdx.x / kWave, waves = blockDim.x / kWave;
  unsigned long long total = 0;
  for (uint32_t c = 0; c < nbuf; c += kWave) {
    const uint32_t l = c + lane < nbuf ? sb.len[c + lane] : 0u;
    total += wave_sum<unsigned long long>(l);
  }
  if (lane == 0) {
    sm[2 * wv] = nbuf;
    sm[2 * wv + 1] = total;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long cnt = 0, edges = 0;
    for (uint32_t i = 0; i < waves; ++i) {
      cnt += sm[2 * i];
      edges += sm[2 * i + 1];
    }
    sm[2 * waves] = cnt ? atomicAdd(ctr, (cnt << kEdgeBits) + edges) : 0ull;
  }
  __syncthreads();
  const unsigned long long base = sm[2 * waves];
  unsigned long long slot = base >> kEdgeBits;
  unsigned long long eoff = base & kEdgeMask;
  for (uint32_t i = 0; i < wv; ++i) {
    slot += sm[2 * i];
    eoff += sm[2 * i + 1];
  }
  for (uint32_t c = 0; c < nbuf; c += kWave) {
    const bool in = c + lane < nbuf;
    const uint32_t l = in ? sb.len[c + lane] : 0u;
    const unsigned long long inc = wave_inclusive_scan<unsigned long long>(l);
    if (in) {
      queue[slot + c + lane] = sb.v[c + lane];
      pre[slot + c + lane] = eoff + inc - l;
    }
    eoff += __shfl(inc, kWave - 1, kWave);
  }
  __syncthreads();
  nbuf = 0;
}

__device__ __forceinline__ void stage_push(WaveStage& sb, uint32_t& nbuf, unsigned long long* ctr,
                                           uint32_t* queue, unsigned long long* pre, bool want, uint32_t v,
                                           uint32_t len) {
  const unsigned long long ball = __ballot(want);
  if (!ball) return;
  if (want) {
    const uint32_t p = nbuf + lanes_below(ball);
    sb.v[p] = v;
    sb.len[p] = len;
  }
  nbuf += __popcll(ball);
  if (nbuf > kStage - kWave) stage_flush(sb, nbuf, ctr, queue, pre);
}

}  // namespace vgx
