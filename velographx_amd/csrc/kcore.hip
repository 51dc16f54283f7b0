// k-core decomposition (core numbers) on the live rows of an undirected
// graph, by level-synchronous parallel peeling.
//
// Reference: include/velographx/incremental/kcore.hpp.  The reference
// recomputes the components a batch touches by min-degree heap peeling
// (:83-121); core numbers depend only on the graph, so a full device
// recompute after every batch gives the same values.
//
// Level k: one scan claims every unpeeled vertex whose remaining degree is
// at most k (core number k) into a queue and finds the smallest remaining
// degree; then rounds of peeling, one wave per queued vertex, decrement the
// remaining degree of each unpeeled neighbour and claim (compare-and-swap on
// its core slot) the ones that drop to k or below into the next queue.  A
// level with nothing to claim jumps straight to the smallest remaining
// degree.  Every arc is decremented at most once, so degrees never wrap.
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

// Peel rounds run several per host check: round r reads its queue length
// from ring[r % 3], appends to ring[(r + 1) % 3] and clears ring[(r + 2) % 3]
// (the next round's output; its last reader, round r - 1, has finished).
// The scan appends the level's first queue to ring[0].
constexpr uint32_t kKcRoundsPerCheck = 4;

struct KcCounters {
  unsigned long long ring[3];
  unsigned long long queued;     // the scan's claims (copied by the level's first round)
  unsigned long long remaining;  // unpeeled vertices left after the scan's claims
  unsigned long long min_deg;    // smallest remaining degree among them
};

}  // namespace vgx

struct vgx_kcore {
  vgx_graph* g = nullptr;
  vgx::DevBuf<uint32_t> deg, core, q[2];
  vgx::KcCounters* d_ctr = nullptr;
  vgx::KcCounters* h_ctr = nullptr;  // pinned
  uint64_t n = 0;
  uint64_t last_repaired = 0;
  uint32_t levels = 0, rounds = 0;
};

namespace vgx {

static unsigned kc_blocks(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t b = (items + per_block - 1) / per_block;
  return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(b, cap)));
}

__global__ void k_kc_init(const uint32_t* __restrict__ live_deg, uint64_t n, uint32_t* __restrict__ deg,
                          uint32_t* __restrict__ core) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    deg[v] = live_deg[v];
    core[v] = kKcUnset;
  }
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

__global__ __launch_bounds__(kKcBlock) void k_kc_scan(uint32_t k, uint64_t n, const uint32_t* __restrict__ deg,
                                                      uint32_t* __restrict__ core, uint32_t* __restrict__ q,
                                                      KcCounters* __restrict__ ctr) {
  unsigned long long rem = 0, mn = ~0ull;
  // the loop bound is uniform across the block, so every lane takes part in
  // each ballot
  for (uint64_t v0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); v0 < n;
       v0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t v = v0 + threadIdx.x;
    bool claim = false;
    if (v < n && core[v] == kKcUnset) {
      const uint32_t d = deg[v];
      if (d <= k) {
        core[v] = k;
        claim = true;
      } else {
        rem += 1;
        mn = min(mn, static_cast<unsigned long long>(d));
      }
    }
    kc_push(claim, static_cast<uint32_t>(v), q, &ctr->ring[0]);
  }
  // one atomic pair per workgroup: thousands of waves on one counter line
  // serialise at the L2
  __shared__ unsigned long long s_rem[kKcBlock / kWave], s_mn[kKcBlock / kWave];
  rem = wave_sum(rem);
  mn = ~wave_max(~mn);  // the wave's minimum
  if (lane_id() == 0) {
    s_rem[threadIdx.x / kWave] = rem;
    s_mn[threadIdx.x / kWave] = mn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kKcBlock / kWave; ++w) {
      rem += s_rem[w];
      mn = min(mn, s_mn[w]);
    }
    if (rem) atomicAdd(&ctr->remaining, rem);
    if (mn != ~0ull) atomicMin(&ctr->min_deg, mn);
  }
}

__global__ __launch_bounds__(kKcBlock) void k_kc_peel(uint32_t k, uint32_t round, const uint2* __restrict__ row,
                                                      const uint32_t* __restrict__ arena,
                                                      const uint32_t* __restrict__ q, uint32_t* __restrict__ deg,
                                                      uint32_t* __restrict__ core, uint32_t* __restrict__ qnext,
                                                      KcCounters* __restrict__ ctr) {
  const uint64_t count = ctr->ring[round % 3];  // written by the previous launch
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctr->ring[(round + 2) % 3] = 0;
    if (round == 0) ctr->queued = count;
  }
  unsigned long long* next = &ctr->ring[(round + 1) % 3];
  const uint64_t W = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  const unsigned lane = lane_id();
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; i < count;
       i += W) {
    const uint2 rb = row[q[i]];
    // uniform across the wave; four arena loads in flight per lane (hub rows
    // are bound by load latency)
    for (uint32_t j0 = 0; j0 < rb.y; j0 += kKcUnroll * kWave) {
      uint32_t x[kKcUnroll];
#pragma unroll
      for (uint32_t t = 0; t < kKcUnroll; ++t) {
        const uint32_t j = j0 + t * kWave + lane;
        x[t] = j < rb.y ? arena[rb.x + j] : kTomb;  // past the row: skipped like a tombstone
      }
#pragma unroll
      for (uint32_t t = 0; t < kKcUnroll; ++t) {
        bool claim = false;
        const uint32_t v = x[t] & ~kTomb;
        if (!(x[t] & kTomb) && core[v] == kKcUnset) {
          const uint32_t before = atomicSub(&deg[v], 1u);
          if (before - 1u <= k) claim = atomicCAS(&core[v], kKcUnset, k) == kKcUnset;
        }
        kc_push(claim, v, qnext, next);
      }
    }
  }
}

static void kc_recompute(vgx_kcore* kc) {
  vgx_graph* g = kc->g;
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  kc->n = n;
  kc->levels = 0;
  kc->rounds = 0;
  kc->last_repaired = n;
  kc->deg.ensure(n + 1);
  kc->core.ensure(n + 1);
  kc->q[0].ensure(n + 1);
  kc->q[1].ensure(n + 1);
  if (n == 0) return;
  const Adjacency& a = g->adj[0];
  k_kc_init<<<kc_blocks(n, kKcBlock, 4096), kKcBlock, 0, s>>>(a.deg.p, n, kc->deg.p, kc->core.p);
  VGX_LAUNCH_CHECK();
  // the peel grid is fixed (grid-stride over the device-side queue length)
  const unsigned peel_blocks = kc_blocks(n, kKcBlock / kWave, 1024);
  uint32_t k = 0;
  // Per level: the scan and the first rounds go out together and one
  // readback tells whether the level claimed anything, what remains and the
  // smallest remaining degree; longer levels continue a few rounds per check.
  auto rounds = [&](uint32_t& r) {
    for (uint32_t i = 0; i < kKcRoundsPerCheck; ++i, ++r) {
      k_kc_peel<<<peel_blocks, kKcBlock, 0, s>>>(k, r, a.row.p, a.arena.p, kc->q[r & 1].p, kc->deg.p, kc->core.p,
                                                 kc->q[(r + 1) & 1].p, kc->d_ctr);
      VGX_LAUNCH_CHECK();
    }
    kc->rounds += kKcRoundsPerCheck;
    VGX_HIP(hipMemcpyAsync(kc->h_ctr, kc->d_ctr, sizeof(KcCounters), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    const uint64_t c = kc->h_ctr->ring[r % 3];
    if (c > n || kc->h_ctr->queued > n) throw Error(VGX_DEVICE_ERROR, "k-core queue exceeded the vertex count");
    return c;
  };
  for (;;) {
    KcCounters zero{{0, 0, 0}, 0, 0, ~0ull};
    VGX_HIP(hipMemcpyAsync(kc->d_ctr, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
    k_kc_scan<<<kc_blocks(n, kKcBlock, 512), kKcBlock, 0, s>>>(k, n, kc->deg.p, kc->core.p, kc->q[0].p, kc->d_ctr);
    VGX_LAUNCH_CHECK();
    uint32_t r = 0;
    uint64_t c = rounds(r);
    const uint64_t rem = kc->h_ctr->remaining;
    const unsigned long long mn = kc->h_ctr->min_deg;
    if (kc->h_ctr->queued == 0) {
      if (rem == 0) break;
      k = static_cast<uint32_t>(mn);  // nothing at this level: the next one with vertices
      continue;
    }
    kc->levels += 1;
    while (c) c = rounds(r);
    if (rem == 0) break;  // the scan left nothing unpeeled beyond this level's queue
    k += 1;
  }
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
  if (kc->d_ctr) dev_free(kc->d_ctr);
  if (kc->h_ctr) (void)hipHostFree(kc->h_ctr);
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
    kc->d_ctr = static_cast<KcCounters*>(dev_alloc(sizeof(KcCounters)));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&kc->h_ctr), sizeof(KcCounters), hipHostMallocDefault));
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
  const vgx_status st = vgx_graph_apply(kc->g, src, dst, add, k);
  if (st != VGX_OK) throw Error(st, std::string(vgx_last_error()));
  kc_recompute(kc);
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
