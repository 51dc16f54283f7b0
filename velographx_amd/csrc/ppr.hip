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
// Work is split into chunks of at most 512 arcs of one row (PprPlan), one
// wave per chunk, so a hub row of R-MAT (tens of thousands of arcs) spreads
// over many waves instead of serialising one.  Rows of several chunks get
// per-chunk partials that a second pass adds in a fixed order: every sum
// has a fixed order and the results do not depend on scheduling.
// In a sweep with K = 32 and 16-byte aligned buffers, eight lanes hold one
// 128-byte row of X as float4s and a wave gathers eight arcs per load;
// otherwise the two half-waves take one arc each per step and their 32
// lanes the K <= 32 columns.
// A sweep gives rows of at most 8 arcs (empty ones included: most rows of
// a skewed graph) to eight-lane groups, a row per group and eight
// consecutive rows per wave, in place of a wave each: a wave's pass over such
// a row is a chain of dependent loads with most lanes idle.  The two kinds of
// work share one launch, in separate workgroups.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "vgx_internal.hpp"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);
extern thread_local std::string t_last_error;

constexpr int kPprBlock = 256;
constexpr uint32_t kPprChunk = 512;
constexpr uint32_t kPprWaves = kPprBlock / kWave;
constexpr uint32_t kPprShort = 8;  // rows of at most this many arcs: one eight-lane group each

// ------------------------------------------------------------ chunk plans

__global__ void k_ppr_chunk_count(const uint32_t* __restrict__ off, uint32_t n, uint32_t* __restrict__ cnt) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t len = off[v + 1] - off[v];
    cnt[v] = len > kPprChunk ? (len + kPprChunk - 1) / kPprChunk : 1u;
  }
}

__global__ void k_ppr_chunk_fill(const uint32_t* __restrict__ row_chunk, uint32_t n, uint32_t* __restrict__ chunk_row) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    for (uint32_t c = row_chunk[v]; c < row_chunk[v + 1]; ++c) chunk_row[c] = v;
}

// Arc range of chunk c of row v.
__device__ __forceinline__ void chunk_range(const uint32_t* off, const uint32_t* row_chunk, uint32_t v, uint32_t c,
                                            uint32_t& b, uint32_t& e) {
  b = off[v] + (c - row_chunk[v]) * kPprChunk;
  e = min(b + kPprChunk, off[v + 1]);
}

// --------------------------------------------------------- normalisation

// Per-chunk sums of the out-weights.
__global__ void k_ppr_wsum_part(const uint32_t* __restrict__ off, const uint32_t* __restrict__ row_chunk,
                                const uint32_t* __restrict__ chunk_row, uint32_t nchunks,
                                const float* __restrict__ w, float* __restrict__ cpart) {
  for (uint32_t c = blockIdx.x * kPprWaves + threadIdx.x / kWave; c < nchunks; c += gridDim.x * kPprWaves) {
    const uint32_t u = chunk_row[c];
    uint32_t b, e;
    chunk_range(off, row_chunk, u, c, b, e);
    float s = 0.f;
    for (uint32_t i = b + lane_id(); i < e; i += kWave) s += w[i];
    s = wave_sum(s);
    if (lane_id() == 0) cpart[c] = s;
  }
}

// W_u = chunk sums of row u in chunk order.
__global__ void k_ppr_wsum_row(const uint32_t* __restrict__ row_chunk, uint32_t n, const float* __restrict__ cpart,
                               float* __restrict__ wsum) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (uint32_t c = row_chunk[u]; c < row_chunk[u + 1]; ++c) s += cpart[c];
    wsum[u] = s;
  }
}

__global__ void k_ppr_normalise(const uint32_t* __restrict__ off, const uint32_t* __restrict__ row_chunk,
                                const uint32_t* __restrict__ chunk_row, uint32_t nchunks,
                                const float* __restrict__ w, const float* __restrict__ wsum,
                                float* __restrict__ pn) {
  for (uint32_t c = blockIdx.x * kPprWaves + threadIdx.x / kWave; c < nchunks; c += gridDim.x * kPprWaves) {
    const uint32_t u = chunk_row[c];
    uint32_t b, e;
    chunk_range(off, row_chunk, u, c, b, e);
    const float s = wsum[u];
    for (uint32_t i = b + lane_id(); i < e; i += kWave) pn[i] = s > 0.f ? w[i] / s : 0.f;
  }
}

// Position of each in-arc (v <- u) in the out-CSR: binary search of v in
// u's sorted out-row, one lane per in-arc.
__global__ void k_ppr_in_to_out(const uint32_t* __restrict__ in_off, const uint32_t* __restrict__ in_adj,
                                const uint32_t* __restrict__ row_chunk, const uint32_t* __restrict__ chunk_row,
                                uint32_t nchunks, const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                const float* __restrict__ pn, float* __restrict__ pin) {
  for (uint32_t c = blockIdx.x * kPprWaves + threadIdx.x / kWave; c < nchunks; c += gridDim.x * kPprWaves) {
    const uint32_t v = chunk_row[c];
    uint32_t b, e;
    chunk_range(in_off, row_chunk, v, c, b, e);
    for (uint32_t i = b + lane_id(); i < e; i += kWave) {
      const uint32_t u = in_adj[i];
      uint32_t lo = off[u], hi = off[u + 1];
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (adj[mid] < v) lo = mid + 1;
        else hi = mid;
      }
      pin[i] = pn[lo];
    }
  }
}

// ------------------------------------------------------------------ sweeps

// Over one chunk of row v: sum_e p[e] * x[nbr_e, col].  Half-wave h takes
// arcs b+h, b+h+2, ...; four of them are loaded before any is used.
__device__ __forceinline__ float chunk_gather(const uint32_t* __restrict__ nbr, const float* __restrict__ p,
                                              const float* __restrict__ x, uint32_t b, uint32_t e, uint32_t K,
                                              unsigned half, unsigned col) {
  float acc = 0.f;
  const bool on = col < K;
  uint32_t i = b + half;
  for (; i + 6 < e; i += 8) {
    const uint32_t u0 = nbr[i], u1 = nbr[i + 2], u2 = nbr[i + 4], u3 = nbr[i + 6];
    const float p0 = p[i], p1 = p[i + 2], p2 = p[i + 4], p3 = p[i + 6];
    if (on) {
      const float x0 = x[static_cast<uint64_t>(u0) * K + col];
      const float x1 = x[static_cast<uint64_t>(u1) * K + col];
      const float x2 = x[static_cast<uint64_t>(u2) * K + col];
      const float x3 = x[static_cast<uint64_t>(u3) * K + col];
      acc += p0 * x0;
      acc += p1 * x1;
      acc += p2 * x2;
      acc += p3 * x3;
    }
  }
  for (; i < e; i += 2) {
    const uint32_t u = nbr[i];
    const float pw = p[i];
    if (on) acc += pw * x[static_cast<uint64_t>(u) * K + col];
  }
  return acc + __shfl_xor(acc, 32, kWave);
}

// The same for K = 32 with 16-byte loads: lane group g = lane / 8 takes arcs
// b+g, b+g+8, ...; lane s = lane % 8 of the group holds columns 4s..4s+3, so
// one load instruction fetches eight 128-byte rows.  Four steps are in flight
// (two in the tail).
// Returns the chunk's sums on lanes 0..7 (columns 4*lane..4*lane+3).
__device__ __forceinline__ float4 chunk_gather32(const uint32_t* __restrict__ nbr, const float* __restrict__ p,
                                                 const float* __restrict__ x, uint32_t b, uint32_t e,
                                                 unsigned lane) {
  const unsigned grp = lane >> 3, sub = lane & 7;
  const float4* x4 = reinterpret_cast<const float4*>(x);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t i = b + grp;
  for (; i + 24 < e; i += 32) {
    const uint32_t u0 = nbr[i], u1 = nbr[i + 8], u2 = nbr[i + 16], u3 = nbr[i + 24];
    const float p0 = p[i], p1 = p[i + 8], p2 = p[i + 16], p3 = p[i + 24];
    const float4 a0 = x4[static_cast<uint64_t>(u0) * 8 + sub];
    const float4 a1 = x4[static_cast<uint64_t>(u1) * 8 + sub];
    const float4 a2 = x4[static_cast<uint64_t>(u2) * 8 + sub];
    const float4 a3 = x4[static_cast<uint64_t>(u3) * 8 + sub];
    acc.x += p0 * a0.x;
    acc.y += p0 * a0.y;
    acc.z += p0 * a0.z;
    acc.w += p0 * a0.w;
    acc.x += p1 * a1.x;
    acc.y += p1 * a1.y;
    acc.z += p1 * a1.z;
    acc.w += p1 * a1.w;
    acc.x += p2 * a2.x;
    acc.y += p2 * a2.y;
    acc.z += p2 * a2.z;
    acc.w += p2 * a2.w;
    acc.x += p3 * a3.x;
    acc.y += p3 * a3.y;
    acc.z += p3 * a3.z;
    acc.w += p3 * a3.w;
  }
  for (; i + 8 < e; i += 16) {
    const uint32_t u0 = nbr[i], u1 = nbr[i + 8];
    const float p0 = p[i], p1 = p[i + 8];
    const float4 a0 = x4[static_cast<uint64_t>(u0) * 8 + sub];
    const float4 a1 = x4[static_cast<uint64_t>(u1) * 8 + sub];
    acc.x += p0 * a0.x;
    acc.y += p0 * a0.y;
    acc.z += p0 * a0.z;
    acc.w += p0 * a0.w;
    acc.x += p1 * a1.x;
    acc.y += p1 * a1.y;
    acc.z += p1 * a1.z;
    acc.w += p1 * a1.w;
  }
  if (i < e) {
    const uint32_t u0 = nbr[i];
    const float p0 = p[i];
    const float4 a0 = x4[static_cast<uint64_t>(u0) * 8 + sub];
    acc.x += p0 * a0.x;
    acc.y += p0 * a0.y;
    acc.z += p0 * a0.z;
    acc.w += p0 * a0.w;
  }
#pragma unroll
  for (int d = 8; d < kWave; d <<= 1) {
    acc.x += __shfl_xor(acc.x, d, kWave);
    acc.y += __shfl_xor(acc.y, d, kWave);
    acc.z += __shfl_xor(acc.z, d, kWave);
    acc.w += __shfl_xor(acc.w, d, kWave);
  }
  return acc;
}

// y[v,:] = a * b[v,:] + d * sum_e p[e] * x[nbr_e,:] over the chunks of rows
// longer than kPprShort; rows of several chunks leave per-chunk partial rows
// for k_ppr_finish.
template <bool kVec32>
__device__ __forceinline__ void ppr_long_chunks(uint32_t blk, uint32_t nblk, const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ nbr,
                                                         const float* __restrict__ p,
                                                         const uint32_t* __restrict__ row_chunk,
                                                         const uint32_t* __restrict__ chunk_row,
                                                         const uint32_t* __restrict__ lchunk, uint32_t nlchunks,
                                                         uint32_t K, float a, const float* __restrict__ b, float d,
                                                         const float* __restrict__ x, float* __restrict__ y,
                                                         float* __restrict__ part) {
  const unsigned lane = lane_id();
  const unsigned half = lane >> 5, col = lane & 31;
  for (uint32_t j = blk * kPprWaves + threadIdx.x / kWave; j < nlchunks; j += nblk * kPprWaves) {
    const uint32_t c = lchunk[j];
    const uint32_t v = chunk_row[c];
    const bool single = row_chunk[v + 1] - row_chunk[v] == 1;
    uint32_t s, e;
    chunk_range(off, row_chunk, v, c, s, e);
    if (kVec32) {
      const float4 acc = chunk_gather32(nbr, p, x, s, e, lane);
      if (lane < 8) {
        if (single) {
          const uint64_t o = static_cast<uint64_t>(v) * 8 + lane;
          const float4 bb = reinterpret_cast<const float4*>(b)[o];
          reinterpret_cast<float4*>(y)[o] =
              make_float4(a * bb.x + d * acc.x, a * bb.y + d * acc.y, a * bb.z + d * acc.z, a * bb.w + d * acc.w);
        } else {
          reinterpret_cast<float4*>(part)[static_cast<uint64_t>(c) * 8 + lane] = acc;
        }
      }
    } else {
      const float acc = chunk_gather(nbr, p, x, s, e, K, half, col);
      if (half == 0 && col < K) {
        if (single) {
          const uint64_t o = static_cast<uint64_t>(v) * K + col;
          y[o] = a * b[o] + d * acc;
        } else {
          part[static_cast<uint64_t>(c) * K + col] = acc;
        }
      }
    }
  }
}

// Rows of at most kPprShort arcs, eight lanes each, rows in vertex order:
// y[v,:] = a * b[v,:] + d * sum_e p[e] * x[nbr_e,:], the arcs added in row
// order.  Lane s of a group holds columns 4s..4s+3 (K = 32, float4) or
// s, s+8, s+16, s+24 (below K).  Longer rows are ppr_long_chunks'.
template <bool kVec32>
__device__ __forceinline__ void ppr_short_rows(uint32_t blk, uint32_t nblk, const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ nbr,
                                                         const float* __restrict__ p, uint32_t n, uint32_t K,
                                                         float a, const float* __restrict__ b, float d,
                                                         const float* __restrict__ x, float* __restrict__ y) {
  const unsigned sub = lane_id() & 7;
  const uint32_t stride = nblk * (kPprBlock / 8);
  for (uint32_t v = (blk * kPprBlock + threadIdx.x) >> 3; v < n; v += stride) {
    const uint32_t s = off[v], e = off[v + 1];
    if (e - s > kPprShort) continue;
    if (kVec32) {
      const float4* x4 = reinterpret_cast<const float4*>(x);
      const uint64_t o = static_cast<uint64_t>(v) * 8 + sub;
      const float4 bb = reinterpret_cast<const float4*>(b)[o];
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
      uint32_t i = s;
      for (; i + 3 < e; i += 4) {
        const uint32_t u0 = nbr[i], u1 = nbr[i + 1], u2 = nbr[i + 2], u3 = nbr[i + 3];
        const float p0 = p[i], p1 = p[i + 1], p2 = p[i + 2], p3 = p[i + 3];
        const float4 a0 = x4[static_cast<uint64_t>(u0) * 8 + sub];
        const float4 a1 = x4[static_cast<uint64_t>(u1) * 8 + sub];
        const float4 a2 = x4[static_cast<uint64_t>(u2) * 8 + sub];
        const float4 a3 = x4[static_cast<uint64_t>(u3) * 8 + sub];
        acc.x += p0 * a0.x;
        acc.y += p0 * a0.y;
        acc.z += p0 * a0.z;
        acc.w += p0 * a0.w;
        acc.x += p1 * a1.x;
        acc.y += p1 * a1.y;
        acc.z += p1 * a1.z;
        acc.w += p1 * a1.w;
        acc.x += p2 * a2.x;
        acc.y += p2 * a2.y;
        acc.z += p2 * a2.z;
        acc.w += p2 * a2.w;
        acc.x += p3 * a3.x;
        acc.y += p3 * a3.y;
        acc.z += p3 * a3.z;
        acc.w += p3 * a3.w;
      }
      for (; i < e; ++i) {
        const float p0 = p[i];
        const float4 a0 = x4[static_cast<uint64_t>(nbr[i]) * 8 + sub];
        acc.x += p0 * a0.x;
        acc.y += p0 * a0.y;
        acc.z += p0 * a0.z;
        acc.w += p0 * a0.w;
      }
      reinterpret_cast<float4*>(y)[o] =
          make_float4(a * bb.x + d * acc.x, a * bb.y + d * acc.y, a * bb.z + d * acc.z, a * bb.w + d * acc.w);
    } else {
      float acc[4] = {0.f, 0.f, 0.f, 0.f};
      uint32_t i = s;
      for (; i + 1 < e; i += 2) {
        const uint32_t u0 = nbr[i], u1 = nbr[i + 1];
        const float p0 = p[i], p1 = p[i + 1];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t col = sub + 8 * j;
          if (col < K) {
            const float x0 = x[static_cast<uint64_t>(u0) * K + col];
            const float x1 = x[static_cast<uint64_t>(u1) * K + col];
            acc[j] += p0 * x0;
            acc[j] += p1 * x1;
          }
        }
      }
      if (i < e) {
        const uint32_t u0 = nbr[i];
        const float p0 = p[i];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t col = sub + 8 * j;
          if (col < K) acc[j] += p0 * x[static_cast<uint64_t>(u0) * K + col];
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t col = sub + 8 * j;
        if (col < K) {
          const uint64_t o = static_cast<uint64_t>(v) * K + col;
          y[o] = a * b[o] + d * acc[j];
        }
      }
    }
  }
}

// One sweep: the first `long_blocks` workgroups take the chunks of long rows
// (the hub chunks start first), the rest the short rows.
template <bool kVec32>
__global__ __launch_bounds__(kPprBlock) void k_ppr_sweep(const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ nbr,
                                                         const float* __restrict__ p,
                                                         const uint32_t* __restrict__ row_chunk,
                                                         const uint32_t* __restrict__ chunk_row,
                                                         const uint32_t* __restrict__ lchunk, uint32_t nlchunks,
                                                         uint32_t long_blocks, uint32_t n, uint32_t K, float a,
                                                         const float* __restrict__ b, float d,
                                                         const float* __restrict__ x, float* __restrict__ y,
                                                         float* __restrict__ part) {
  if (blockIdx.x < long_blocks)
    ppr_long_chunks<kVec32>(blockIdx.x, long_blocks, off, nbr, p, row_chunk, chunk_row, lchunk, nlchunks, K, a, b,
                            d, x, y, part);
  else
    ppr_short_rows<kVec32>(blockIdx.x - long_blocks, gridDim.x - long_blocks, off, nbr, p, n, K, a, b, d, x, y);
}

// Chunks of rows longer than kPprShort, one append per wave.
__global__ void k_ppr_long_collect(const uint32_t* __restrict__ off, const uint32_t* __restrict__ chunk_row,
                                   uint32_t nchunks, uint32_t* __restrict__ lchunk,
                                   unsigned long long* __restrict__ count) {
  const unsigned lane = lane_id();
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t c0 = blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1)); c0 < nchunks; c0 += stride) {
    const uint32_t c = c0 + lane;
    bool take = false;
    if (c < nchunks) {
      const uint32_t v = chunk_row[c];
      take = off[v + 1] - off[v] > kPprShort;
    }
    const unsigned long long m = __ballot(take);
    if (!m) continue;
    const unsigned first = __ffsll(static_cast<long long>(m)) - 1;
    unsigned long long base = 0;
    if (lane == first) base = atomicAdd(count, static_cast<unsigned long long>(__popcll(m)));
    base = __shfl(base, first, kWave);
    if (take) lchunk[base + lanes_below(m)] = c;
  }
}

__global__ void k_ppr_multi_collect(const uint32_t* __restrict__ row_chunk, uint32_t n, uint32_t* __restrict__ multi,
                                    unsigned long long* count) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    if (row_chunk[v + 1] - row_chunk[v] > 1) multi[atomicAdd(count, 1ull)] = v;
}

// Rows of several chunks, one wave each: half-wave h adds the partial rows
// of chunks h, h+2, ... (four in flight), then the halves are combined.
__global__ __launch_bounds__(kPprBlock) void k_ppr_finish(const uint32_t* __restrict__ row_chunk,
                                                          const uint32_t* __restrict__ multi, uint32_t nmulti,
                                                          uint32_t K, float a, const float* __restrict__ b, float d,
                                                          const float* __restrict__ part, float* __restrict__ y) {
  const unsigned lane = lane_id();
  const unsigned half = lane >> 5, col = lane & 31;
  for (uint32_t j = blockIdx.x * kPprWaves + threadIdx.x / kWave; j < nmulti; j += gridDim.x * kPprWaves) {
    const uint32_t r = multi[j];
    const uint32_t c0 = row_chunk[r], c1 = row_chunk[r + 1];
    float acc = 0.f;
    if (col < K) {
      uint32_t c = c0 + half;
      for (; c + 6 < c1; c += 8)
        acc += (part[static_cast<uint64_t>(c) * K + col] + part[static_cast<uint64_t>(c + 2) * K + col]) +
               (part[static_cast<uint64_t>(c + 4) * K + col] + part[static_cast<uint64_t>(c + 6) * K + col]);
      for (; c < c1; c += 2) acc += part[static_cast<uint64_t>(c) * K + col];
    }
    acc += __shfl_xor(acc, 32, kWave);
    if (half == 0 && col < K) {
      const uint64_t o = static_cast<uint64_t>(r) * K + col;
      y[o] = a * b[o] + d * acc;
    }
  }
}

// ---------------------------------------------------------- edge gradient

// Per out-arc i = (u -> v): G = d <X[u,:], L[v,:]>, stored in gw; the
// chunk's sum of P_uv * G goes to cpart.  One lane per arc; X[u,:] is the
// same for the whole chunk and sits in registers.
__global__ __launch_bounds__(kPprBlock) void k_ppr_edge_dot(const uint32_t* __restrict__ off,
                                                            const uint32_t* __restrict__ adj,
                                                            const float* __restrict__ pn,
                                                            const uint32_t* __restrict__ row_chunk,
                                                            const uint32_t* __restrict__ chunk_row, uint32_t nchunks,
                                                            uint32_t K, float d, const float* __restrict__ x,
                                                            const float* __restrict__ lam, float* __restrict__ gw,
                                                            float* __restrict__ cpart, int vec32) {
  for (uint32_t c = blockIdx.x * kPprWaves + threadIdx.x / kWave; c < nchunks; c += gridDim.x * kPprWaves) {
    const uint32_t u = chunk_row[c];
    uint32_t b, e;
    chunk_range(off, row_chunk, u, c, b, e);
    float xu[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) xu[k] = k < static_cast<int>(K) ? x[static_cast<uint64_t>(u) * K + k] : 0.f;
    float t = 0.f;
    for (uint32_t i = b + lane_id(); i < e; i += kWave) {
      const float* lv = lam + static_cast<uint64_t>(adj[i]) * K;
      float dot = 0.f;
      if (vec32) {
        const float4* l4 = reinterpret_cast<const float4*>(lv);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const float4 z = l4[q];
          dot += xu[4 * q] * z.x + xu[4 * q + 1] * z.y + xu[4 * q + 2] * z.z + xu[4 * q + 3] * z.w;
        }
      } else {
#pragma unroll
        for (int k = 0; k < 32; ++k)
          if (k < static_cast<int>(K)) dot += xu[k] * lv[k];
      }
      const float g = d * dot;
      gw[i] = g;
      t += pn[i] * g;
    }
    t = wave_sum(t);
    if (lane_id() == 0) cpart[c] = t;
  }
}

// grad_w_uv = (G_uv - sum over u's chunks of cpart, in chunk order) / W_u.
__global__ __launch_bounds__(kPprBlock) void k_ppr_edge_norm(const uint32_t* __restrict__ off,
                                                             const uint32_t* __restrict__ row_chunk,
                                                             const uint32_t* __restrict__ chunk_row,
                                                             uint32_t nchunks, const float* __restrict__ wsum,
                                                             const float* __restrict__ cpart,
                                                             float* __restrict__ gw) {
  for (uint32_t c = blockIdx.x * kPprWaves + threadIdx.x / kWave; c < nchunks; c += gridDim.x * kPprWaves) {
    const uint32_t u = chunk_row[c];
    uint32_t b, e;
    chunk_range(off, row_chunk, u, c, b, e);
    if (b == e) continue;
    float t = 0.f;
    for (uint32_t j = row_chunk[u]; j < row_chunk[u + 1]; ++j) t += cpart[j];
    const float ws = wsum[u];
    for (uint32_t i = b + lane_id(); i < e; i += kWave) gw[i] = ws > 0.f ? (gw[i] - t) / ws : 0.f;
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

// Chunk plan of one snapshot, rebuilt when the snapshot changes.
static const PprPlan& ppr_plan(vgx_graph* g, int dir, const CsrSnapshot& sn, hipStream_t s) {
  PprWork& pw = ppr_work(g);
  PprPlan& pl = pw.plan[dir];
  if (pl.version == sn.version && pl.n == sn.n && pl.row_chunk.p) return pl;
  const uint32_t n = static_cast<uint32_t>(sn.n);
  pl.row_chunk.ensure(sn.n + 2);
  VGX_HIP(hipMemsetAsync(pl.row_chunk.p, 0, sizeof(uint32_t), s));
  uint32_t total = 0;
  if (n) {
    DevBuf<uint32_t> cnt;
    cnt.ensure(sn.n + 1);
    k_ppr_chunk_count<<<ppr_blocks(n, kPprBlock), kPprBlock, 0, s>>>(sn.off.p, n, cnt.p);
    VGX_LAUNCH_CHECK();
    size_t bytes = 0;
    VGX_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, cnt.p, pl.row_chunk.p + 1, static_cast<int>(n), s));
    pw.scan.ensure(bytes + 16);
    VGX_HIP(hipcub::DeviceScan::InclusiveSum(pw.scan.p, bytes, cnt.p, pl.row_chunk.p + 1, static_cast<int>(n), s));
    VGX_HIP(hipMemcpyAsync(&total, pl.row_chunk.p + n, sizeof(total), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    pl.chunk_row.ensure(static_cast<uint64_t>(total) + 1);
    k_ppr_chunk_fill<<<ppr_blocks(n, kPprBlock), kPprBlock, 0, s>>>(pl.row_chunk.p, n, pl.chunk_row.p);
    VGX_LAUNCH_CHECK();
    // rows of several chunks (their order in the list does not matter: each
    // row's sum runs over its own chunks in a fixed order)
    const uint64_t most = std::min<uint64_t>(n, total / 2 + 1);
    pl.multi.ensure(most + 1);
    pl.lchunk.ensure(static_cast<uint64_t>(total) + 1);
    DevBuf<unsigned long long> cnt64;
    cnt64.ensure(2);
    VGX_HIP(hipMemsetAsync(cnt64.p, 0, 16, s));
    k_ppr_multi_collect<<<ppr_blocks(n, kPprBlock), kPprBlock, 0, s>>>(pl.row_chunk.p, n, pl.multi.p, cnt64.p);
    VGX_LAUNCH_CHECK();
    k_ppr_long_collect<<<ppr_blocks(total, kPprBlock), kPprBlock, 0, s>>>(sn.off.p, pl.chunk_row.p, total,
                                                                         pl.lchunk.p, cnt64.p + 1);
    VGX_LAUNCH_CHECK();
    unsigned long long nm[2] = {0, 0};
    VGX_HIP(hipMemcpyAsync(nm, cnt64.p, 16, hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    pl.nmulti = nm[0];
    pl.nlchunks = nm[1];
  } else {
    pl.nmulti = 0;
    pl.nlchunks = 0;
  }
  pl.nchunks = total;
  pl.version = sn.version;
  pl.n = sn.n;
  return pl;
}

// Normalised weights (and their in-arc copy) on the caller's stream, after
// the graph's CSR snapshots are complete.
static void ppr_prepare(vgx_graph* g, const float* w, hipStream_t s, bool need_in) {
  const CsrSnapshot& out = graph_snapshot(g, 0);
  const CsrSnapshot* in = need_in ? &graph_snapshot(g, 1) : nullptr;
  VGX_HIP(hipStreamSynchronize(g->stream));
  PprWork& pw = ppr_work(g);
  const PprPlan& po = ppr_plan(g, 0, out, s);
  const uint32_t n = static_cast<uint32_t>(out.n);
  pw.pn.ensure(out.arcs + 1);
  pw.wsum.ensure(out.n + 1);
  pw.cpart.ensure(po.nchunks + 1);
  if (n) {
    const unsigned gc = ppr_blocks(po.nchunks, kPprWaves);
    k_ppr_wsum_part<<<gc, kPprBlock, 0, s>>>(out.off.p, po.row_chunk.p, po.chunk_row.p,
                                             static_cast<uint32_t>(po.nchunks), w, pw.cpart.p);
    VGX_LAUNCH_CHECK();
    k_ppr_wsum_row<<<ppr_blocks(n, kPprBlock), kPprBlock, 0, s>>>(po.row_chunk.p, n, pw.cpart.p, pw.wsum.p);
    VGX_LAUNCH_CHECK();
    k_ppr_normalise<<<gc, kPprBlock, 0, s>>>(out.off.p, po.row_chunk.p, po.chunk_row.p,
                                             static_cast<uint32_t>(po.nchunks), w, pw.wsum.p, pw.pn.p);
    VGX_LAUNCH_CHECK();
  }
  if (in) {
    const PprPlan& pi = ppr_plan(g, 1, *in, s);
    pw.pin.ensure(in->arcs + 1);
    if (n) {
      k_ppr_in_to_out<<<ppr_blocks(pi.nchunks, kPprWaves), kPprBlock, 0, s>>>(
          in->off.p, in->adj.p, pi.row_chunk.p, pi.chunk_row.p, static_cast<uint32_t>(pi.nchunks), out.off.p,
          out.adj.p, pw.pn.p, pw.pin.p);
      VGX_LAUNCH_CHECK();
    }
  }
}

// `iterations` sweeps y <- a*b + d * (P-gather of cur) over one direction,
// ping-ponging between `io` and the scratch buffer; the result ends in io.
static void ppr_sweeps(vgx_graph* g, const CsrSnapshot& sn, const PprPlan& pl, const float* p, uint32_t K, float a,
                       const float* b, float d, uint32_t iterations, float* io, hipStream_t s) {
  PprWork& pw = ppr_work(g);
  const uint32_t n = static_cast<uint32_t>(sn.n);
  const uint64_t nk = static_cast<uint64_t>(n) * K;
  pw.tmp.ensure(nk + 1);
  pw.part.ensure(pl.nchunks * K + 1);
  const unsigned gs = ppr_blocks(pl.nlchunks, kPprWaves);
  const unsigned gr = ppr_blocks(n, kPprBlock / 8);
  const unsigned gf = ppr_blocks(pl.nmulti, kPprWaves);
  // 16-byte rows need K = 32 and 16-byte aligned X, S and output buffers
  const bool vec = K == 32 && ((reinterpret_cast<uintptr_t>(io) | reinterpret_cast<uintptr_t>(b) |
                                reinterpret_cast<uintptr_t>(pw.tmp.p) | reinterpret_cast<uintptr_t>(pw.part.p)) &
                               15) == 0;
  float* cur = io;
  float* nxt = pw.tmp.p;
  for (uint32_t it = 0; it < iterations; ++it) {
    const uint32_t lb = pl.nlchunks ? gs : 0;
    if (vec)
      k_ppr_sweep<true><<<lb + gr, kPprBlock, 0, s>>>(sn.off.p, sn.adj.p, p, pl.row_chunk.p, pl.chunk_row.p,
                                                      pl.lchunk.p, static_cast<uint32_t>(pl.nlchunks), lb, n, K, a,
                                                      b, d, cur, nxt, pw.part.p);
    else
      k_ppr_sweep<false><<<lb + gr, kPprBlock, 0, s>>>(sn.off.p, sn.adj.p, p, pl.row_chunk.p, pl.chunk_row.p,
                                                       pl.lchunk.p, static_cast<uint32_t>(pl.nlchunks), lb, n, K, a,
                                                       b, d, cur, nxt, pw.part.p);
    VGX_LAUNCH_CHECK();
    if (pl.nmulti) {
      k_ppr_finish<<<gf, kPprBlock, 0, s>>>(pl.row_chunk.p, pl.multi.p, static_cast<uint32_t>(pl.nmulti), K, a, b,
                                            d, pw.part.p, nxt);
      VGX_LAUNCH_CHECK();
    }
    std::swap(cur, nxt);
  }
  if (cur != io) VGX_HIP(hipMemcpyAsync(io, cur, nk * 4, hipMemcpyDeviceToDevice, s));
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
  hipStream_t s = static_cast<hipStream_t>(stream);  // null: the default stream
  if (s_dev == x_dev) throw Error(VGX_INVALID_ARGUMENT, "S and X must not alias");
  ppr_prepare(g, w_dev, s, true);
  const CsrSnapshot& in = graph_snapshot(g, 1);
  PprWork& pw = ppr_work(g);
  const uint64_t nk = in.n * K;
  const float a = static_cast<float>(1.0 - damping), d = static_cast<float>(damping);
  if (nk) {
    // X_0 = (1-d) S, then X <- (1-d) S + d P^T X over the in-CSR
    k_scale<<<ppr_blocks(nk, 256), 256, 0, s>>>(s_dev, a, nk, x_dev);
    VGX_LAUNCH_CHECK();
    ppr_sweeps(g, in, pw.plan[1], pw.pin.p, K, a, s_dev, d, iterations, x_dev, s);
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
  hipStream_t s = static_cast<hipStream_t>(stream);  // null: the default stream
  if (c_dev == lambda_dev) throw Error(VGX_INVALID_ARGUMENT, "C and Lambda must not alias");
  ppr_prepare(g, w_dev, s, false);
  const CsrSnapshot& out = graph_snapshot(g, 0);
  PprWork& pw = ppr_work(g);
  const PprPlan& po = pw.plan[0];
  const uint64_t nk = out.n * K;
  const float d = static_cast<float>(damping);
  if (nk) {
    // L_0 = C, then L <- C + d P L over the out-CSR
    VGX_HIP(hipMemcpyAsync(lambda_dev, c_dev, nk * 4, hipMemcpyDeviceToDevice, s));
    ppr_sweeps(g, out, po, pw.pn.p, K, 1.f, c_dev, d, iterations, lambda_dev, s);
    if (grad_s_dev) {
      k_scale<<<ppr_blocks(nk, 256), 256, 0, s>>>(lambda_dev, static_cast<float>(1.0 - damping), nk, grad_s_dev);
      VGX_LAUNCH_CHECK();
    }
    if (grad_w_dev && out.arcs) {
      const unsigned gc = ppr_blocks(po.nchunks, kPprWaves);
      k_ppr_edge_dot<<<gc, kPprBlock, 0, s>>>(out.off.p, out.adj.p, pw.pn.p, po.row_chunk.p, po.chunk_row.p,
                                              static_cast<uint32_t>(po.nchunks), K, d, x_dev, lambda_dev,
                                              grad_w_dev, pw.cpart.p,
                                              K == 32 && (reinterpret_cast<uintptr_t>(lambda_dev) & 15) == 0);
      VGX_LAUNCH_CHECK();
      k_ppr_edge_norm<<<gc, kPprBlock, 0, s>>>(out.off.p, po.row_chunk.p, po.chunk_row.p,
                                               static_cast<uint32_t>(po.nchunks), pw.wsum.p, pw.cpart.p,
                                               grad_w_dev);
      VGX_LAUNCH_CHECK();
    }
  }
  VGX_HIP(hipStreamSynchronize(s));
  VGX_API_END
}

}  // extern "C"
