// Connected components of an undirected graph on the device: min-label
// hooking with pointer jumping over the live rows.
//
// Reference: include/velographx/incremental/connected_components.hpp.  The
// reference labels a component by its union-find root, which depends on the
// order of its sequential unions; here a component's label is its smallest
// vertex id, so the partition is the reference's and the label is a fixed,
// order-free representative.
//
// Labels start at the smallest of each vertex and its neighbours (no
// atomics), jumped to their roots.  Every round: each live arc (u, v) lowers the larger of the two current
// labels' parent slot to the smaller label (atomicMin, from the row of the
// endpoint holding the larger label), then every vertex jumps to its root.  Parent values only ever decrease and start at the
// vertex itself, so parent[v] <= v and the pointers form a forest; a round
// with no hook ends the loop, and then every vertex of a component points at
// the component's smallest id (its own parent is itself).
#include <algorithm>
#include <cstring>
#include <string>

#include "vgx_internal.hpp"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);
extern thread_local std::string t_last_error;

constexpr int kCcBlock = 256;
constexpr uint32_t kCcUnroll = 4;

}  // namespace vgx

struct vgx_components {
  vgx_graph* g = nullptr;
  vgx::DevBuf<uint32_t> parent;
  vgx::DevBuf<uint32_t> aff;  // pre-batch labels whose component lost an arc
  vgx::DevBuf<unsigned long long> hooked;
  uint64_t n = 0;
  uint64_t last_repaired = 0;
  uint32_t rounds = 0;
};

namespace vgx {

static unsigned cc_blocks(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t b = (items + per_block - 1) / per_block;
  return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(b, cap)));
}

// Start labels without atomics: each vertex points at the smallest of itself
// and its live neighbours (a vertex of its own component, never above it),
// one wave per row.
__global__ __launch_bounds__(256) void k_cc_init(const uint2* __restrict__ row, const uint32_t* __restrict__ arena,
                                                 uint64_t n, uint32_t* __restrict__ parent) {
  const uint64_t W = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  const unsigned lane = lane_id();
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += W) {
    const uint2 rb = row[u];
    uint32_t m = static_cast<uint32_t>(u);
    // four loads in flight per lane: hub rows are bound by load latency
    for (uint32_t j0 = 0; j0 < rb.y; j0 += kCcUnroll * kWave) {
      uint32_t x[kCcUnroll];
#pragma unroll
      for (uint32_t q = 0; q < kCcUnroll; ++q) {
        const uint32_t j = j0 + q * kWave + lane;
        x[q] = j < rb.y ? arena[rb.x + j] : kTomb;  // past the row: skipped like a tombstone
      }
#pragma unroll
      for (uint32_t q = 0; q < kCcUnroll; ++q)
        if (!(x[q] & kTomb)) m = min(m, x[q]);
    }
    m = ~wave_max(~m);  // the wave's minimum
    if (lane == 0) parent[u] = m;
  }
}

// One wave per vertex u: a live arc (u, v) whose label is below u's lowers
// u's label slot to it.  Each undirected arc is stored in both rows, so the
// row of the endpoint with the larger label does the hook; a plain read
// first skips the atomic where the slot is already low enough.
__global__ __launch_bounds__(kCcBlock) void k_cc_hook(const uint2* __restrict__ row, const uint32_t* __restrict__ arena,
                                                      uint64_t n, uint32_t* __restrict__ parent,
                                                      unsigned long long* __restrict__ hooked) {
  const uint64_t W = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  const unsigned lane = lane_id();
  unsigned long long any = 0;
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += W) {
    const uint2 rb = row[u];
    const uint32_t pu = parent[u];
    for (uint32_t j0 = 0; j0 < rb.y; j0 += kCcUnroll * kWave) {
      uint32_t pv[kCcUnroll];
#pragma unroll
      for (uint32_t q = 0; q < kCcUnroll; ++q) {
        const uint32_t j = j0 + q * kWave + lane;
        const uint32_t x = j < rb.y ? arena[rb.x + j] : kTomb;
        pv[q] = (x & kTomb) ? pu : parent[x];  // past the row or deleted: no hook
      }
#pragma unroll
      for (uint32_t q = 0; q < kCcUnroll; ++q) {
        if (pv[q] >= pu || parent[pu] <= pv[q]) continue;
        if (atomicMin(&parent[pu], pv[q]) > pv[q]) any = 1;
      }
    }
  }
  if (__ballot(any != 0) && lane == 0) atomicAdd(hooked, 1ull);
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
  k_cc_init<<<cc_blocks(n, kCcBlock / kWave, 2048), kCcBlock, 0, s>>>(a.row.p, a.arena.p, n, cc->parent.p);
  VGX_LAUNCH_CHECK();
  k_cc_compress<<<cc_blocks(n, kCcBlock, 2048), kCcBlock, 0, s>>>(n, cc->parent.p);
  VGX_LAUNCH_CHECK();
  // every round either hooks (a label strictly decreases) or ends the loop;
  // labels are bounded below, so the loop ends
  for (;;) {
    VGX_HIP(hipMemsetAsync(cc->hooked.p, 0, sizeof(unsigned long long), s));
    k_cc_hook<<<cc_blocks(n, kCcBlock / kWave, 2048), kCcBlock, 0, s>>>(a.row.p, a.arena.p, n, cc->parent.p,
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
