// Weighted personalized PageRank, forward and adjoint, for gfx950.
//
// The reference has no weighted or personalized PageRank (its PageRank is
// pagerank.hpp:277-321 with uniform teleport); this is the differentiable
// extension the benchmark configuration asks for.  Definitions:
//   P_uv = w_uv / W_u,  W_u = sum of u's out-weights (rows with W_u = 0 send
//                       no mass: P is substochastic there)
//   forward  X = (1-d) S + d P^T X          X, S: n x K, row-major float32
//   adjoint  L = C + d P L                  C = dLoss/dX
//   grad_S   = (1-d) L
//   grad_P_uv = d * <X[u,:], L[v,:]>
//   grad_w_uv = (grad_P_uv - sum_v' P_uv' grad_P_uv') / W_u
// Both fixed points are reached by `iterations` Jacobi sweeps.  Arc weights
// are indexed by the arcs of the graph's out-CSR snapshot
// (vgx_graph_csr_device / vgx_graph_export_csr order).
//
// One wave per vertex; the two half-waves take one arc each per step and
// their 32 lanes the K <= 32 columns, so every arc's row of X is one
// coalesced 128-byte load.
#include <algorithm>

#include "vgx_internal.hpp"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);
extern thread_local std::string t_last_error;

constexpr int kPprBlock = 256;

// W_u and normalised weights in out-arc order.
__global__ void k_ppr_normalise(const uint32_t* __restrict__ off, const float* __restrict__ w, uint32_t n,
                                float* __restrict__ pn, float* __restrict__ wsum) {
  const uint32_t W = gridDim.x * (blockDim.x / kWave);
  for (uint32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += W) {
    const uint32_t b = off[u], e = off[u + 1];
    float s = 0.f;
    for (uint32_t i = b + lane_id(); i < e; i += kWave) s += w[i];
    s = wave_sum(s);
    for (uint32_t i = b + lane_id(); i < e; i += kWave) pn[i] = s > 0.f ? w[i] / s : 0.f;
    if (lane_id() == 0) wsum[u] = s;
  }
}

// Position of each in-arc (v <- u) in the out-CSR: binary search of v in
// u's sorted out-row.
__global__ void k_ppr_in_to_out(const uint32_t* __restrict__ in_off, const uint32_t* __restrict__ in_adj,
                                const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj, uint32_t n,
                                const float* __restrict__ pn, float* __restrict__ pin) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    for (uint32_t e = in_off[v]; e < in_off[v + 1]; ++e) {
      const uint32_t u = in_adj[e];
      uint32_t lo = off[u], hi = off[u + 1];
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (adj[mid] < v) lo = mid + 1;
        else hi = mid;
      }
      pin[e] = pn[lo];
    }
  }
}

// y[v,:] = a * b[v,:] + d * sum_e p[e] * x[nbr_e,:] over the row of v.
__global__ __launch_bounds__(kPprBlock) void k_ppr_sweep(const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ nbr,
                                                         const float* __restrict__ p, uint32_t n, uint32_t K,
                                                         float a, const float* __restrict__ b, float d,
                                                         const float* __restrict__ x, float* __restrict__ y) {
  const uint32_t W = gridDim.x * (blockDim.x / kWave);
  const unsigned lane = lane_id();
  const unsigned half = lane >> 5, col = lane & 31;
  for (uint32_t v = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; v < n; v += W) {
    const uint32_t s = off[v], e = off[v + 1];
    float acc = 0.f;
    for (uint32_t i = s + half; i < e; i += 2) {
      const uint32_t u = nbr[i];
      const float pw = p[i];
      if (col < K) acc += pw * x[static_cast<uint64_t>(u) * K + col];
    }
    acc += __shfl_xor(acc, 32, kWave);
    if (half == 0 && col < K) {
      const uint64_t o = static_cast<uint64_t>(v) * K + col;
      y[o] = a * b[o] + d * acc;
    }
  }
}

// grad_w over the out-CSR: first grad_P per arc and the row's weighted sum,
// then the normalisation's chain rule.
__global__ __launch_bounds__(kPprBlock) void k_ppr_edge_grad(const uint32_t* __restrict__ off,
                                                             const uint32_t* __restrict__ adj,
                                                             const float* __restrict__ pn,
                                                             const float* __restrict__ wsum, uint32_t n, uint32_t K,
                                                             float d, const float* __restrict__ x,
                                                             const float* __restrict__ lam,
                                                             float* __restrict__ gw) {
  const uint32_t W = gridDim.x * (blockDim.x / kWave);
  const unsigned lane = lane_id();
  const unsigned half = lane >> 5, col = lane & 31;
  for (uint32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += W) {
    const uint32_t s = off[u], e = off[u + 1];
    if (s == e) continue;
    const float xu = col < K ? x[static_cast<uint64_t>(u) * K + col] : 0.f;
    float t = 0.f;
    for (uint32_t i0 = s; i0 < e; i0 += 2) {
      const uint32_t i = i0 + half;
      float prod = 0.f;
      if (i < e && col < K) prod = xu * lam[static_cast<uint64_t>(adj[i]) * K + col];
#pragma unroll
      for (int m = 16; m > 0; m >>= 1) prod += __shfl_xor(prod, m, kWave);
      const float gp = d * prod;
      if (i < e && col == 0) gw[i] = gp;
      if (i < e) t += pn[i] * gp;
    }
    // t is the same on the 32 lanes of a half; add the two halves
    t = __shfl(t, 0, kWave) + __shfl(t, 32, kWave);
    // the lanes below read grad_P values other lanes of this wave stored
    __threadfence_block();
    const float ws = wsum[u];
    for (uint32_t i = s + lane; i < e; i += kWave) gw[i] = ws > 0.f ? (gw[i] - t) / ws : 0.f;
  }
}

__global__ void k_scale(const float* __restrict__ x, float a, uint64_t n, float* __restrict__ y) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    y[i] = a * x[i];
}

static unsigned ppr_blocks(uint64_t items, unsigned per, unsigned cap = 8192) {
  uint64_t b = (items + per - 1) / per;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

static PprWork& ppr_work(vgx_graph* g) { return g->ppr; }

// Normalised weights (and their in-arc copy) on the caller's stream, after
// the graph's CSR snapshots are complete.
static void ppr_prepare(vgx_graph* g, const float* w, hipStream_t s, bool need_in) {
  const CsrSnapshot& out = graph_snapshot(g, 0);
  const CsrSnapshot* in = need_in ? &graph_snapshot(g, 1) : nullptr;
  VGX_HIP(hipStreamSynchronize(g->stream));
  PprWork& pw = ppr_work(g);
  const uint32_t n = static_cast<uint32_t>(out.n);
  pw.pn.ensure(out.arcs + 1);
  pw.wsum.ensure(out.n + 1);
  if (n) {
    k_ppr_normalise<<<ppr_blocks(n, kPprBlock / kWave), kPprBlock, 0, s>>>(out.off.p, w, n, pw.pn.p, pw.wsum.p);
    VGX_LAUNCH_CHECK();
  }
  if (in) {
    pw.pin.ensure(in->arcs + 1);
    if (n) {
      k_ppr_in_to_out<<<ppr_blocks(n, kPprBlock), kPprBlock, 0, s>>>(in->off.p, in->adj.p, out.off.p, out.adj.p, n,
                                                                      pw.pn.p, pw.pin.p);
      VGX_LAUNCH_CHECK();
    }
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

vgx_status vgx_ppr_forward(vgx_graph* g, const float* w_dev, const float* s_dev, float* x_dev, uint32_t K,
                           double damping, uint32_t iterations, void* stream) {
  VGX_API_BEGIN
  if (!g || !w_dev || !s_dev || !x_dev) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (K == 0 || K > 32) throw Error(VGX_INVALID_ARGUMENT, "K must be in [1, 32]");
  if (!(damping >= 0.0 && damping <= 1.0)) throw Error(VGX_INVALID_ARGUMENT, "damping must be in [0, 1]");
  VGX_HIP(hipSetDevice(g->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->stream;
  if (s_dev == x_dev) throw Error(VGX_INVALID_ARGUMENT, "S and X must not alias");
  ppr_prepare(g, w_dev, s, true);
  const CsrSnapshot& in = graph_snapshot(g, 1);
  PprWork& pw = ppr_work(g);
  const uint32_t n = static_cast<uint32_t>(in.n);
  const uint64_t nk = static_cast<uint64_t>(n) * K;
  pw.tmp.ensure(nk + 1);
  const float a = static_cast<float>(1.0 - damping), d = static_cast<float>(damping);
  if (nk) {
    // X_0 = (1-d) S, then X <- (1-d) S + d P^T X
    k_scale<<<ppr_blocks(nk, 256), 256, 0, s>>>(s_dev, a, nk, x_dev);
    VGX_LAUNCH_CHECK();
    float* cur = x_dev;
    float* nxt = pw.tmp.p;
    for (uint32_t it = 0; it < iterations; ++it) {
      k_ppr_sweep<<<ppr_blocks(n, kPprBlock / kWave), kPprBlock, 0, s>>>(in.off.p, in.adj.p, pw.pin.p, n, K, a,
                                                                         s_dev, d, cur, nxt);
      VGX_LAUNCH_CHECK();
      std::swap(cur, nxt);
    }
    if (cur != x_dev) VGX_HIP(hipMemcpyAsync(x_dev, cur, nk * 4, hipMemcpyDeviceToDevice, s));
  }
  VGX_HIP(hipStreamSynchronize(s));
  VGX_API_END
}

vgx_status vgx_ppr_backward(vgx_graph* g, const float* w_dev, const float* x_dev, const float* c_dev,
                            float* lambda_dev, float* grad_s_dev, float* grad_w_dev, uint32_t K, double damping,
                            uint32_t iterations, void* stream) {
  VGX_API_BEGIN
  if (!g || !w_dev || !x_dev || !c_dev || !lambda_dev) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (K == 0 || K > 32) throw Error(VGX_INVALID_ARGUMENT, "K must be in [1, 32]");
  VGX_HIP(hipSetDevice(g->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->stream;
  if (c_dev == lambda_dev) throw Error(VGX_INVALID_ARGUMENT, "C and Lambda must not alias");
  ppr_prepare(g, w_dev, s, false);
  const CsrSnapshot& out = graph_snapshot(g, 0);
  PprWork& pw = ppr_work(g);
  const uint32_t n = static_cast<uint32_t>(out.n);
  const uint64_t nk = static_cast<uint64_t>(n) * K;
  pw.tmp.ensure(nk + 1);
  const float d = static_cast<float>(damping);
  if (nk) {
    // L_0 = C, then L <- C + d P L over the out-CSR
    VGX_HIP(hipMemcpyAsync(lambda_dev, c_dev, nk * 4, hipMemcpyDeviceToDevice, s));
    float* cur = lambda_dev;
    float* nxt = pw.tmp.p;
    for (uint32_t it = 0; it < iterations; ++it) {
      k_ppr_sweep<<<ppr_blocks(n, kPprBlock / kWave), kPprBlock, 0, s>>>(out.off.p, out.adj.p, pw.pn.p, n, K, 1.f,
                                                                         c_dev, d, cur, nxt);
      VGX_LAUNCH_CHECK();
      std::swap(cur, nxt);
    }
    if (cur != lambda_dev) VGX_HIP(hipMemcpyAsync(lambda_dev, cur, nk * 4, hipMemcpyDeviceToDevice, s));
    if (grad_s_dev) {
      k_scale<<<ppr_blocks(nk, 256), 256, 0, s>>>(lambda_dev, static_cast<float>(1.0 - damping), nk, grad_s_dev);
      VGX_LAUNCH_CHECK();
    }
    if (grad_w_dev && out.arcs) {
      k_ppr_edge_grad<<<ppr_blocks(n, kPprBlock / kWave), kPprBlock, 0, s>>>(out.off.p, out.adj.p, pw.pn.p,
                                                                             pw.wsum.p, n, K, d, x_dev, lambda_dev,
                                                                             grad_w_dev);
      VGX_LAUNCH_CHECK();
    }
  }
  VGX_HIP(hipStreamSynchronize(s));
  VGX_API_END
}

}  // extern "C"
