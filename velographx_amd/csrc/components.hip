// Connected components of an undirected graph on the device: min-label
// hooking with pointer jumping over the live rows.
//
// Reference: include/velographx/incremental/connected_components.hpp.  The
// reference labels a component by its union-find root, which depends on the
// order of its sequential unions; here a component's label is its smallest
// vertex id, so the partition is the reference's and the label is a fixed,
// order-free representative.
//
// Labels start at the smallest of each vertex and its live neighbours,
// jumped to their roots.  Every round: each live arc (u, v)
// whose other end has the smaller label lowers u's root's parent slot to it
// (atomicMin), then every vertex jumps to its root.  Parent values only ever
// decrease and start at the vertex itself, so parent[v] <= v and the
// pointers form a forest; a round in which no arc sees a smaller label ends
// the loop (every arc is stored in both rows, so its two labels are then
// equal), and then every vertex of a component points at the component's
// smallest id (its own parent is itself).
//
// The arcs are walked edge-balanced, a lane per stored entry, over the list
// of vertices with stored entries (built once per labelling by two scans);
// the lanes of one row in a window take their minimum first, so a row sends
// one atomic per window, and no lane waits for an atomic's result.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <string>

#include "vgx_internal.hpp"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);
extern thread_local std::string t_last_error;

constexpr int kCcBlock = 256;

}  // namespace vgx

struct vgx_components {
  vgx_graph* g = nullptr;
  vgx::DevBuf<uint32_t> parent;
  vgx::DevBuf<uint32_t> aff;  // pre-batch labels whose component lost an arc
  vgx::DevBuf<unsigned long long> hooked;
  // the edge-balanced loop's items: vertices with stored entries (queue)
  // and their edge prefix (pre), from scans of flag and len
  vgx::DevBuf<uint32_t> flag, fpos, queue;
  vgx::DevBuf<unsigned long long> len, lpre, pre;
  vgx::DevBuf<uint8_t> scan_tmp;
  uint64_t n = 0;
  uint64_t last_repaired = 0;
  uint32_t rounds = 0;
};

namespace vgx {

static unsigned cc_blocks(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t b = (items + per_block - 1) / per_block;
  return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(b, cap)));
}

// flag[v] = 1 and len[v] = stored entries of v's row for v < n (0 past it,
// which closes the scans), parent[v] = v.
__global__ void k_cc_items(const uint2* __restrict__ row, uint64_t n, uint32_t* __restrict__ flag,
                           unsigned long long* __restrict__ len, uint32_t* __restrict__ parent) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v <= n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t l = v < n ? row[v].y : 0u;
    flag[v] = l ? 1u : 0u;
    len[v] = l;
    if (v < n) parent[v] = static_cast<uint32_t>(v);
  }
}

__global__ void k_cc_list(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ fpos,
                          const unsigned long long* __restrict__ lpre, uint64_t n, uint32_t* __restrict__ queue,
                          unsigned long long* __restrict__ pre) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (!flag[v]) continue;
    queue[fpos[v]] = static_cast<uint32_t>(v);
    pre[fpos[v]] = lpre[v];
  }
}

// One round of hooks, a lane per stored entry (u, x): when x's label is
// below u's, u's root slot is lowered to it.  The lanes of a row in one
// window (a run of equal u) take their minimum and the run's last lane
// sends the one atomic.  hooked counts the waves that saw a smaller label.
// With kStart (the first round, on identity labels) each row's run takes
// the smallest of its live neighbours instead, and a run that holds the
// whole row stores min(u, that) without an atomic.
template <bool kStart>
__global__ __launch_bounds__(kCcBlock) void k_cc_hook(const uint32_t* __restrict__ queue,
                                                      const unsigned long long* __restrict__ pre, uint32_t F,
                                                      unsigned long long E, const uint2* __restrict__ row,
                                                      const uint32_t* __restrict__ arena, uint32_t* __restrict__ parent,
                                                      unsigned long long* __restrict__ hooked) {
  const uint32_t W = gridDim.x * (kCcBlock / kWave);
  const uint32_t gw = blockIdx.x * (kCcBlock / kWave) + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  bool any = false;
  lbs_edges(queue, pre, F, E, row, gw, W, [&](uint32_t, uint32_t u, uint2 rb, uint32_t slot, bool valid) {
    const uint32_t x = valid ? arena[slot] : kTomb;
    const bool live = valid && !(x & kTomb);
    uint32_t pu, m;
    if constexpr (kStart) {
      pu = u;
      m = live ? x : kInf;
    } else {
      pu = valid ? parent[u] : 0u;
      const uint32_t px = live ? parent[x] : kInf;
      m = px < pu ? px : kInf;
    }
    // runs: consecutive valid lanes of one row (the valid lanes are a
    // prefix of the wave; each invalid lane is a run of its own)
    const uint32_t uprev = __shfl_up(u, 1, kWave);
    const bool vprev = __shfl_up(static_cast<int>(valid), 1, kWave) != 0;
    const unsigned long long starts = __ballot(lane == 0 || !valid || !vprev || uprev != u);
    const unsigned long long upto = starts & (((1ull << lane) - 1ull) | (1ull << lane));
    const unsigned first = 63u - static_cast<unsigned>(__clzll(static_cast<long long>(upto)));
#pragma unroll
    for (unsigned d = 1; d < kWave; d <<= 1) {
      const uint32_t y = __shfl_up(m, d, kWave);
      if (lane >= first + d) m = min(m, y);
    }
    const bool last = lane == kWave - 1 || ((starts >> (lane + 1)) & 1ull);
    if constexpr (kStart) {
      // the run holds the whole row when its first entry is the row's first
      // and its last the row's last
      const uint32_t at = slot - rb.x;
      const uint32_t at_first = __shfl(at, static_cast<int>(first), kWave);
      if (valid && last) {
        if (at_first == 0 && at + 1 == rb.y)
          parent[u] = min(u, m);
        else if (m < u)
          atomicMin(&parent[u], m);
      }
    } else if (valid && last && m != kInf) {
      atomicMin(&parent[pu], m);
      any = true;
    }
  });
  if (__ballot(any) && lane == 0) atomicAdd(hooked, 1ull);
}

// Pointer jumping to the root (parent[v] <= v, so every chain ends).
__global__ void k_cc_compress(uint64_t n, uint32_t* __restrict__ parent) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t p = parent[v];
    uint32_t gp = parent[p];
    while (gp != p) {
      p = gp;
      gp = parent[p];
    }
    parent[v] = p;
  }
}

// Marks the pre-batch label of every arc the staged batch deletes for good
// (its last op removes an arc that exists: connected_components.hpp:45-52).
__global__ void k_cc_mark_deleted(const unsigned long long* __restrict__ skeys, const uint8_t* __restrict__ fflag,
                                  const uint8_t* __restrict__ act, uint64_t M, uint64_t n_old,
                                  const uint32_t* __restrict__ parent, uint32_t* __restrict__ aff) {
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < M;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (!fflag[p] || act[p] != kActDel) continue;
    const uint32_t u = static_cast<uint32_t>(skeys[p] >> 32);
    if (u < n_old) aff[parent[u]] = 1u;
  }
}

// Vertices of the marked components (the reference's repair count, :54-62).
__global__ void k_cc_count_marked(uint64_t n_old, const uint32_t* __restrict__ parent,
                                  const uint32_t* __restrict__ aff, unsigned long long* __restrict__ count) {
  unsigned long long c = 0;
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n_old;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    c += aff[parent[v]] ? 1ull : 0ull;
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(count, c);
}

template <class F>
static void cc_scan(vgx_components* cc, F&& f) {
  size_t bytes = 0;
  VGX_HIP(f(static_cast<void*>(nullptr), bytes));
  cc->scan_tmp.ensure(bytes + 16);
  VGX_HIP(f(static_cast<void*>(cc->scan_tmp.p), bytes));
}

static void cc_recompute(vgx_components* cc) {
  vgx_graph* g = cc->g;
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  cc->n = n;
  cc->rounds = 0;
  cc->last_repaired = n;
  cc->parent.ensure(n + 1);
  cc->hooked.ensure(1);
  if (n == 0) return;
  const Adjacency& a = g->adj[0];
  cc->flag.ensure(n + 1);
  cc->fpos.ensure(n + 1);
  cc->len.ensure(n + 1);
  cc->lpre.ensure(n + 1);
  cc->queue.ensure(n + 1);
  cc->pre.ensure(n + 1);
  k_cc_items<<<cc_blocks(n + 1, kCcBlock, 2048), kCcBlock, 0, s>>>(a.row.p, n, cc->flag.p, cc->len.p, cc->parent.p);
  VGX_LAUNCH_CHECK();
  const int items = static_cast<int>(n + 1);
  cc_scan(cc, [&](void* t, size_t& b) {
    return hipcub::DeviceScan::ExclusiveSum(t, b, cc->flag.p, cc->fpos.p, items, s);
  });
  cc_scan(cc, [&](void* t, size_t& b) {
    return hipcub::DeviceScan::ExclusiveSum(t, b, cc->len.p, cc->lpre.p, items, s);
  });
  k_cc_list<<<cc_blocks(n, kCcBlock, 2048), kCcBlock, 0, s>>>(cc->flag.p, cc->fpos.p, cc->lpre.p, n, cc->queue.p,
                                                              cc->pre.p);
  VGX_LAUNCH_CHECK();
  uint32_t F = 0;
  unsigned long long E = 0;
  VGX_HIP(hipMemcpyAsync(&F, cc->fpos.p + n, sizeof(F), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipMemcpyAsync(&E, cc->lpre.p + n, sizeof(E), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  if (E == 0) return;
  // about one wave per 64 entries, up to 2048 workgroups
  const unsigned grid = cc_blocks(E, kCcBlock, 2048);
  // start: every vertex at the smallest of itself and its live neighbours
  k_cc_hook<true><<<grid, kCcBlock, 0, s>>>(cc->queue.p, cc->pre.p, F, E, a.row.p, a.arena.p, cc->parent.p,
                                            cc->hooked.p);
  VGX_LAUNCH_CHECK();
  k_cc_compress<<<cc_blocks(n, kCcBlock, 2048), kCcBlock, 0, s>>>(n, cc->parent.p);
  VGX_LAUNCH_CHECK();
  // every round either hooks (a label strictly decreases) or ends the loop;
  // labels are bounded below, so the loop ends
  for (;;) {
    VGX_HIP(hipMemsetAsync(cc->hooked.p, 0, sizeof(unsigned long long), s));
    k_cc_hook<false><<<grid, kCcBlock, 0, s>>>(cc->queue.p, cc->pre.p, F, E, a.row.p, a.arena.p, cc->parent.p,
                                               cc->hooked.p);
    VGX_LAUNCH_CHECK();
    k_cc_compress<<<cc_blocks(n, kCcBlock, 2048), kCcBlock, 0, s>>>(n, cc->parent.p);
    VGX_LAUNCH_CHECK();
    unsigned long long h = 0;
    VGX_HIP(hipMemcpyAsync(&h, cc->hooked.p, sizeof(h), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    cc->rounds += 1;
    if (h == 0) break;
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

vgx_status vgx_components_destroy(vgx_components* cc) {
  VGX_API_BEGIN
  if (!cc) return VGX_OK;
  if (cc->g) (void)hipStreamSynchronize(cc->g->stream);
  delete cc;
  VGX_API_END
}

vgx_status vgx_components_create(vgx_graph* g, vgx_components** out) {
  VGX_API_BEGIN
  if (!g || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (g->directed)
    throw Error(VGX_INVALID_ARGUMENT,
                "IncrementalComponents requires an undirected graph; use an explicitly defined weak/strong "
                "connectivity algorithm for directed graphs");
  VGX_HIP(hipSetDevice(g->device));
  auto* cc = new vgx_components();
  try {
    cc->g = g;
    cc_recompute(cc);
  } catch (...) {
    vgx_components_destroy(cc);
    throw;
  }
  *out = cc;
  VGX_API_END
}

vgx_status vgx_components_recompute(vgx_components* cc) {
  VGX_API_BEGIN
  if (!cc) throw Error(VGX_INVALID_ARGUMENT, "null components");
  VGX_HIP(hipSetDevice(cc->g->device));
  cc_recompute(cc);
  VGX_API_END
}

vgx_status vgx_components_apply(vgx_components* cc, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                                uint64_t k) {
  VGX_API_BEGIN
  if (!cc || (k && (!src || !dst || !add))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(cc->g->device));
  if (k == 0) {
    cc->last_repaired = 0;
    return VGX_OK;
  }
  vgx_graph* g = cc->g;
  hipStream_t s = g->stream;
  const uint64_t n_old = cc->n;
  graph_stage(g, src, dst, add, k);
  // the staged, sorted batch says which arcs go for good; their pre-batch
  // components are the ones the reference rebuilds
  const BatchScratch& b = g->bs[0];  // flags.p starts with the final-op flags
  cc->aff.ensure(n_old + 1);
  VGX_HIP(hipMemsetAsync(cc->aff.p, 0, (n_old + 1) * sizeof(uint32_t), s));
  VGX_HIP(hipMemsetAsync(cc->hooked.p, 0, sizeof(unsigned long long), s));
  if (b.m && n_old) {
    k_cc_mark_deleted<<<cc_blocks(b.m, kCcBlock, 2048), kCcBlock, 0, s>>>(b.skeys.p, b.flags.p, b.act.p, b.m, n_old,
                                                                         cc->parent.p, cc->aff.p);
    VGX_LAUNCH_CHECK();
    k_cc_count_marked<<<cc_blocks(n_old, kCcBlock, 1024), kCcBlock, 0, s>>>(n_old, cc->parent.p, cc->aff.p,
                                                                           cc->hooked.p);
    VGX_LAUNCH_CHECK();
  }
  unsigned long long repaired = 0;
  VGX_HIP(hipMemcpyAsync(&repaired, cc->hooked.p, sizeof(repaired), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  graph_apply_deletions(g);
  graph_apply_additions(g);
  graph_finish_batch(g);
  cc_recompute(cc);
  cc->last_repaired = repaired;
  VGX_API_END
}

vgx_status vgx_components_labels(vgx_components* cc, uint32_t* out, uint64_t cap, uint64_t* count) {
  VGX_API_BEGIN
  if (!cc || !count) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(cc->g->device));
  *count = cc->n;
  const uint64_t c = std::min(cap, cc->n);
  if (out && c) {
    VGX_HIP(hipMemcpyAsync(out, cc->parent.p, c * sizeof(uint32_t), hipMemcpyDeviceToHost, cc->g->stream));
    VGX_HIP(hipStreamSynchronize(cc->g->stream));
  }
  VGX_API_END
}

vgx_status vgx_components_component(vgx_components* cc, uint32_t v, uint32_t* label) {
  VGX_API_BEGIN
  if (!cc || !label) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (v >= cc->n) throw Error(VGX_OUT_OF_RANGE, "vertex id outside graph");
  VGX_HIP(hipSetDevice(cc->g->device));
  VGX_HIP(hipMemcpyAsync(label, cc->parent.p + v, sizeof(uint32_t), hipMemcpyDeviceToHost, cc->g->stream));
  VGX_HIP(hipStreamSynchronize(cc->g->stream));
  VGX_API_END
}

vgx_status vgx_components_stats(vgx_components* cc, uint64_t* last_repaired_vertices, uint32_t* rounds) {
  VGX_API_BEGIN
  if (!cc) throw Error(VGX_INVALID_ARGUMENT, "null components");
  if (last_repaired_vertices) *last_repaired_vertices = cc->last_repaired;
  if (rounds) *rounds = cc->rounds;
  VGX_API_END
}

}  // extern "C"
