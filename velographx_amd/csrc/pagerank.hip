// Pull PageRank and localized repair for gfx950.
//
// Reference: include/velographx/incremental/pagerank.hpp
//   full_solve :277-321   next[v] = (1-d)/n + d*dangling/n + d*sum rank[u]/outdeg[u]
//                         (dangling mass spread uniformly, L-inf stop)
//   apply      :47-191    full recompute when n changed, when dangling vertices
//                         existed, or when an endpoint became dangling;
//                         otherwise Jacobi sweeps over the active set
//   validate_against_full / apply_validated :204-258
// plus the static push pagerank of src/algorithms.cpp:56-81 (L1 stop).
//
// One iteration is two launches over the graph's live in-rows (deleted
// entries tombstoned and skipped).  k_pr_iter splits the vertices by stored
// row length: rows without entries take the base rank, rows of <= 32
// entries are summed by groups of 8 lanes, rows up to 1024 by one wave, and
// larger rows in 1024-entry chunks by several waves that leave partial
// sums; every workgroup leaves its residual and dangling partials.
// k_pr_finish then sums each large row's chunk partials in chunk order and
// its last workgroup reduces all partials in a fixed order, sets the next
// iteration's dangling share and the convergence flag, so the host only
// polls every few iterations and the sweep kernel needs no fence.  FP64
// reproduces the reference's arithmetic; FP32 keeps ranks and sums in float
// with Kahan-compensated lane sums.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "vgx_internal.hpp"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);
extern thread_local std::string t_last_error;

constexpr int kPrBlock = 256;
constexpr uint32_t kSmallRow = 32;
// Lanes per short row in a sweep.
constexpr uint32_t kSmallGroup = 8;
constexpr uint32_t kChunk = 1024;
// Workgroups per in-degree class and launch; each class loops grid-stride,
// so the last workgroup reduces at most 3 x 1024 partials.
constexpr uint64_t kMaxClassBlocks = 1024;
// Every kPushProbe-th batch marked to skip the residual pushes tries them.
constexpr uint32_t kPushProbe = 8;

struct PrState {
  double dangling_share;  // d * dangling / n for the next iteration
  double l1, linf;
  unsigned long long iterations;
  uint32_t done;
  uint32_t blocks_done;
  uint32_t stop_l1;  // 1: stop when l1 < tol, 0: when linf <= tol
  uint32_t pad;
  double tol;
  double teleport;
  double damping;
  unsigned long long max_iterations;
};

struct PrPartial {
  double l1, linf, dangling;
};

// Counters of the residual push.  Round k's claim reads the active list of
// acnt[k % 3] and appends the rows it pushes to q[k & 1] (packed count << 38
// | arcs); its distribute appends the vertices whose residual crosses the
// threshold to acnt[(k + 1) % 3].  Claim k clears acnt[(k + 1) % 3] and
// q[(k + 1) & 1], slots no kernel of round k reads, so rounds go out back to
// back with no host step between them.
struct PushState {
  unsigned long long acnt[3];
  unsigned long long q[2];
  unsigned long long pushed_vertices;
  double c;   // uniform residual: every entry of the residual is r_s[v] + c
  double l1;  // sum |r_s| (residual reduction)
};

}  // namespace vgx

struct vgx_pagerank {
  vgx_graph* g = nullptr;
  vgx::ViewGuard views;  // zero-copy views of rank[cur]
  double damping = 0.85;
  int precision = 64;
  uint64_t n = 0;
  bool have_values = false;
  // per graph version
  uint64_t lists_version = ~0ull;
  vgx::DevBuf<uint32_t> zlist, slist, mlist, hlist;
  vgx::DevBuf<uint32_t> hchunk_v, hchunk_off;  // chunk -> H index, chunk -> arc start
  vgx::DevBuf<uint32_t> hchunk_first;          // H index -> first chunk (size nh+1)
  vgx::DevBuf<uint32_t> class_counts, class_bases, class_tail;  // list building scratch
  uint64_t nz = 0, ns = 0, nm = 0, nh = 0, nchunks = 0;
  vgx::DevBuf<uint8_t> rank[2], contrib[2], hpart, ref_rank[2];
  vgx::DevBuf<vgx::PrPartial> partials;
  vgx::PrState* d_st = nullptr;
  vgx::PrState* h_st = nullptr;
  int cur = 0;  // buffer holding the current ranks
  vgx_pagerank_stats stats{};
  // localized repair workspace
  vgx::DevBuf<uint32_t> act_flag, ever_flag, alist[2], counters;
  vgx::DevBuf<uint8_t> upd;
  vgx::DevBuf<uint8_t> scan_tmp;
  // residual push: the residual of the current ranks (exact while
  // res_valid), its entries' bound, and the push lists
  vgx::DevBuf<uint8_t> res, push_amt, push_save;
  vgx::DevBuf<uint32_t> push_stamp, push_list[3], push_q;
  vgx::DevBuf<unsigned long long> push_pre;
  vgx::PushState* d_push = nullptr;
  vgx::PushState* h_push = nullptr;
  bool res_valid = false;
  uint64_t res_version = ~0ull;
  double res_bound = INFINITY;  // every |residual entry| is at most this
  uint32_t push_epoch = 0;
  // batches of at least half this many ops skip the pushes (the last push
  // phase of such a batch exceeded its budget in the first round), except
  // every kPushProbe-th one, which tries again
  uint64_t push_skip_k = 0;
  uint32_t push_skipped = 0;
};

namespace vgx {

template <class T>
struct Acc;
template <>
struct Acc<double> {
  double s = 0.0;
  __device__ void add(double x) { s += x; }
  __device__ double get() const { return s; }
};
template <>
struct Acc<float> {
  float s = 0.f, c = 0.f;
  __device__ void add(float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ float get() const { return s; }
};

// The sweeps read the graph's live in-rows (sorted, deleted entries
// tombstoned), so an update batch needs no repacked copy; a row's sum runs
// over its live entries in row order whatever tombstones it holds.
struct PrArgs {
  const uint2* row;
  const uint32_t* arena;
  const uint32_t* outdeg;
  const uint32_t* zlist;
  const uint32_t* slist;
  const uint32_t* mlist;
  const uint32_t* hlist;
  const uint32_t* hchunk_v;
  const uint32_t* hchunk_off;
  const uint32_t* hchunk_first;
  uint32_t nz, ns, nm, nh, nchunks;
  uint32_t bz, bs, bm;  // workgroups for Z, S and M (H uses the rest)
  uint32_t n;
  void* rank0;
  void* rank1;
  void* contrib0;
  void* contrib1;
  void* hpart;
  PrPartial* partials;
  PrState* st;
};

template <class T>
__device__ __forceinline__ T finish_vertex(const PrArgs& a, uint32_t v, T incoming, T base, const T* rank, T* next,
                                           T* cnext, double& l1, double& linf, double& dang) {
  const T updated = base + static_cast<T>(a.st->damping) * incoming;
  const T delta = updated > rank[v] ? updated - rank[v] : rank[v] - updated;
  next[v] = updated;
  const uint32_t od = a.outdeg[v];
  cnext[v] = od ? updated / static_cast<T>(od) : T(0);
  if (!od) dang += static_cast<double>(updated);
  l1 += static_cast<double>(delta);
  linf = linf > static_cast<double>(delta) ? linf : static_cast<double>(delta);
  return updated;
}

template <class T>
__device__ void block_reduce3(double& a, double& b, double& c, double* sm) {
  a = wave_sum(a);
  b = wave_max(b);
  c = wave_sum(c);
  const unsigned w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    sm[w * 3 + 0] = a;
    sm[w * 3 + 1] = b;
    sm[w * 3 + 2] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (unsigned i = 1; i < blockDim.x / kWave; ++i) {
      a += sm[i * 3 + 0];
      b = b > sm[i * 3 + 1] ? b : sm[i * 3 + 1];
      c += sm[i * 3 + 2];
    }
  }
}

// Lane's sum of contrib over the live entries e, e + 64, ... below end,
// two entries per step with both loads issued before either is used; the
// additions run in entry order, as a one-entry loop would add them.
template <class T>
__device__ __forceinline__ void sum_range(const uint32_t* __restrict__ arena, const T* __restrict__ contrib,
                                          uint32_t e, uint32_t end, Acc<T>& acc) {
  for (; e < end; e += 2 * kWave) {
    const uint32_t x0 = arena[e];
    const uint32_t x1 = e + kWave < end ? arena[e + kWave] : kTomb;
    const T c0 = (x0 & kTomb) ? T(0) : contrib[x0];
    const T c1 = (x1 & kTomb) ? T(0) : contrib[x1];
    if (!(x0 & kTomb)) acc.add(c0);
    if (!(x1 & kTomb)) acc.add(c1);
  }
}

template <class T>
__global__ __launch_bounds__(kPrBlock) void k_pr_iter(PrArgs a) {
  __shared__ double sm[3 * (kPrBlock / kWave)];
  PrState* st = a.st;
  if (st->done) return;
  const int par = static_cast<int>(st->iterations & 1ull);
  const T* rank = static_cast<const T*>(par ? a.rank1 : a.rank0);
  T* next = static_cast<T*>(par ? a.rank0 : a.rank1);
  const T* contrib = static_cast<const T*>(par ? a.contrib1 : a.contrib0);
  T* cnext = static_cast<T*>(par ? a.contrib0 : a.contrib1);
  T* hpart = static_cast<T*>(a.hpart);
  const T base = static_cast<T>(st->teleport + st->dangling_share);
  double l1 = 0.0, linf = 0.0, dang = 0.0;
  const uint32_t b0 = blockIdx.x;
  // Each class is covered grid-stride by its own range of workgroups.
  if (b0 < a.bz) {
    // no in-arcs: the next rank is the base, one vertex per thread
    for (uint32_t i = b0 * blockDim.x + threadIdx.x; i < a.nz; i += a.bz * blockDim.x)
      finish_vertex<T>(a, a.zlist[i], T(0), base, rank, next, cnext, l1, linf, dang);
  }
  const uint32_t b = b0 - a.bz;  // wraps for the Z workgroups, which match nothing below
  if (b0 < a.bz) {
  } else if (b < a.bs) {
    // Short rows by groups of kSmallGroup lanes: a group's loads of one row
    // are consecutive (coalesced), and a wave's lanes stay busy however the
    // row lengths vary, where a lane per row would send 64 scattered
    // addresses per arena load and idle on the wave's shortest rows.
    constexpr uint32_t G = kSmallGroup;
    const uint32_t lane = lane_id(), sub = lane % G;
    const uint32_t gwave = b * (kPrBlock / kWave) + threadIdx.x / kWave;
    const uint32_t nwaves = a.bs * (kPrBlock / kWave);
    for (uint32_t i0 = gwave * (kWave / G); i0 < a.ns; i0 += nwaves * (kWave / G)) {
      const uint32_t i = i0 + lane / G;
      const bool act = i < a.ns;
      const uint32_t v = act ? a.slist[i] : 0u;
      const uint2 rb = act ? a.row[v] : make_uint2(0u, 0u);
      Acc<T> acc;
      for (uint32_t e = rb.x + sub; e < rb.x + rb.y; e += G) {
        const uint32_t x = a.arena[e];
        if (!(x & kTomb)) acc.add(contrib[x]);
      }
      T sum = acc.get();
#pragma unroll
      for (uint32_t d = 1; d < G; d <<= 1) sum += __shfl_xor(sum, static_cast<int>(d), kWave);
      if (act && sub == 0) finish_vertex<T>(a, v, sum, base, rank, next, cnext, l1, linf, dang);
    }
  } else if (b < a.bs + a.bm) {
    for (uint32_t i = (b - a.bs) * (kPrBlock / kWave) + threadIdx.x / kWave; i < a.nm;
         i += a.bm * (kPrBlock / kWave)) {
      const uint32_t v = a.mlist[i];
      Acc<T> acc;
      const uint2 rb = a.row[v];
      sum_range<T>(a.arena, contrib, rb.x + lane_id(), rb.x + rb.y, acc);
      T s = wave_sum(acc.get());
      if (lane_id() == 0) finish_vertex<T>(a, v, s, base, rank, next, cnext, l1, linf, dang);
    }
  } else {
    const uint32_t bh = gridDim.x - a.bz - a.bs - a.bm;
    for (uint32_t c = (b - a.bs - a.bm) * (kPrBlock / kWave) + threadIdx.x / kWave; c < a.nchunks;
         c += bh * (kPrBlock / kWave)) {
      const uint32_t hv = a.hlist[a.hchunk_v[c]];
      const uint2 rb = a.row[hv];
      const uint32_t e0 = a.hchunk_off[c];
      const uint32_t e1 = min(e0 + kChunk, rb.x + rb.y);
      Acc<T> acc;
      sum_range<T>(a.arena, contrib, e0 + lane_id(), e1, acc);
      const T s = wave_sum(acc.get());
      if (lane_id() == 0) hpart[c] = s;
    }
  }
  // this workgroup's partials; k_pr_finish reduces them after this launch
  block_reduce3<T>(l1, linf, dang, sm);
  if (threadIdx.x == 0) a.partials[b0] = PrPartial{l1, linf, dang};
}

// After k_pr_iter (the kernel boundary makes its writes visible, so no
// fence is needed): one thread per large row sums its chunk partials in
// chunk order and finishes it; the last workgroup reduces the sweep's
// partials (k_pr_iter's nprev, then this launch's) in a fixed order, sets
// the next dangling share and the convergence flag.
template <class T>
__global__ __launch_bounds__(kPrBlock) void k_pr_finish(PrArgs a, uint32_t nprev) {
  __shared__ double sm[3 * (kPrBlock / kWave)];
  __shared__ int s_last;
  PrState* st = a.st;
  if (st->done) return;
  const int par = static_cast<int>(st->iterations & 1ull);
  const T* rank = static_cast<const T*>(par ? a.rank1 : a.rank0);
  T* next = static_cast<T*>(par ? a.rank0 : a.rank1);
  T* cnext = static_cast<T*>(par ? a.contrib0 : a.contrib1);
  const T* hpart = static_cast<const T*>(a.hpart);
  const T base = static_cast<T>(st->teleport + st->dangling_share);
  double l1 = 0.0, linf = 0.0, dang = 0.0;
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < a.nh; h += gridDim.x * blockDim.x) {
    Acc<T> acc;
    for (uint32_t c = a.hchunk_first[h]; c < a.hchunk_first[h + 1]; ++c) acc.add(hpart[c]);
    finish_vertex<T>(a, a.hlist[h], acc.get(), base, rank, next, cnext, l1, linf, dang);
  }
  block_reduce3<T>(l1, linf, dang, sm);
  if (threadIdx.x == 0) {
    a.partials[nprev + blockIdx.x] = PrPartial{l1, linf, dang};
    __threadfence();
    s_last = atomicAdd(&st->blocks_done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  // every partial, thread t taking t, t + 256, ... in order, then the
  // fixed-order block reduction (the result does not depend on timing)
  double hl1 = 0.0, hlinf = 0.0, hdang = 0.0;
  const uint32_t total = nprev + gridDim.x;
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
    hl1 += __hip_atomic_load(&a.partials[i].l1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double pl = __hip_atomic_load(&a.partials[i].linf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    hlinf = hlinf > pl ? hlinf : pl;
    hdang += __hip_atomic_load(&a.partials[i].dangling, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  block_reduce3<T>(hl1, hlinf, hdang, sm);
  if (threadIdx.x == 0) {
    const double L1 = hl1, LI = hlinf, D = hdang;
    st->l1 = L1;
    st->linf = LI;
    st->dangling_share = st->damping * D / static_cast<double>(a.n);
    st->iterations += 1;
    const bool conv = st->stop_l1 ? (L1 < st->tol) : (LI <= st->tol);
    if (conv || st->iterations >= st->max_iterations) st->done = conv ? 1u : 2u;
    st->blocks_done = 0;
    __threadfence();
  }
}

// rank = 1/n (or keep), contrib = rank/outdeg, dangling partials.
template <class T>
__global__ __launch_bounds__(kPrBlock) void k_pr_init(PrArgs a, int warm) {
  __shared__ double sm[3 * (kPrBlock / kWave)];
  __shared__ int s_last;
  PrState* st = a.st;
  T* rank = static_cast<T*>(a.rank0);
  T* contrib = static_cast<T*>(a.contrib0);
  double dang = 0.0, z1 = 0.0, z2 = 0.0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += gridDim.x * blockDim.x) {
    T r = warm ? rank[v] : static_cast<T>(1.0 / static_cast<double>(a.n));
    rank[v] = r;
    const uint32_t od = a.outdeg[v];
    contrib[v] = od ? r / static_cast<T>(od) : T(0);
    if (!od) dang += static_cast<double>(r);
  }
  block_reduce3<T>(z1, z2, dang, sm);
  if (threadIdx.x == 0) {
    a.partials[blockIdx.x] = PrPartial{0.0, 0.0, dang};
    __threadfence();
    s_last = atomicAdd(&st->blocks_done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  double D = 0.0, y1 = 0.0, y2 = 0.0;
  for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x)
    D += __hip_atomic_load(&a.partials[i].dangling, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  block_reduce3<T>(y1, y2, D, sm);
  if (threadIdx.x == 0) {
    st->dangling_share = st->damping * D / static_cast<double>(a.n);
    st->blocks_done = 0;
    __threadfence();
  }
}

// Split vertices by in-degree into four lists kept in vertex order: Z (no
// in-arcs: their next rank is the base alone), S (at most kSmallRow stored
// entries), M (at most kChunk), H (longer, chunked).  Wave w owns 64-vertex
// groups [w*gpw, (w+1)*gpw): k_pr_class_count counts its members per
// class, an exclusive scan over [Z counts | S | M | H] gives every wave its
// write positions, k_pr_class_write fills the lists.  No atomics, and the
// lists come out sorted.
constexpr uint32_t kClassWaves = 4096;
constexpr uint32_t kClasses = 4;

__device__ __forceinline__ void class_ballots(const uint2* row, uint32_t n, uint32_t g, unsigned long long* b) {
  const uint32_t v = g * kWave + lane_id();
  const uint32_t d = v < n ? row[v].y : 0u;  // stored entries, tombstones included
  b[0] = __ballot(v < n && d == 0);
  b[1] = __ballot(v < n && d > 0 && d <= kSmallRow);
  b[2] = __ballot(v < n && d > kSmallRow && d <= kChunk);
  b[3] = __ballot(v < n && d > kChunk);
}

__global__ void k_pr_class_count(const uint2* __restrict__ row, uint32_t n, uint32_t gpw,
                                 uint32_t* __restrict__ counts) {
  const uint32_t w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const uint32_t groups = (n + kWave - 1) / kWave;
  uint32_t c[kClasses] = {0, 0, 0, 0};
  for (uint32_t g = w * gpw; g < min((w + 1) * gpw, groups); ++g) {
    unsigned long long b[kClasses];
    class_ballots(row, n, g, b);
#pragma unroll
    for (uint32_t k = 0; k < kClasses; ++k) c[k] += __popcll(b[k]);
  }
  if (lane_id() == 0)
#pragma unroll
    for (uint32_t k = 0; k < kClasses; ++k) counts[k * kClassWaves + w] = c[k];
}

__global__ void k_pr_class_write(const uint2* __restrict__ row, uint32_t n, uint32_t gpw,
                                 const uint32_t* __restrict__ bases, uint32_t* __restrict__ zlist,
                                 uint32_t* __restrict__ slist, uint32_t* __restrict__ mlist,
                                 uint32_t* __restrict__ hlist) {
  const uint32_t w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const uint32_t groups = (n + kWave - 1) / kWave;
  uint32_t* lists[kClasses] = {zlist, slist, mlist, hlist};
  uint32_t o[kClasses];
#pragma unroll
  for (uint32_t k = 0; k < kClasses; ++k) o[k] = bases[k * kClassWaves + w] - bases[k * kClassWaves];
  for (uint32_t g = w * gpw; g < min((w + 1) * gpw, groups); ++g) {
    unsigned long long b[kClasses];
    class_ballots(row, n, g, b);
    const uint32_t v = g * kWave + lane_id();
#pragma unroll
    for (uint32_t k = 0; k < kClasses; ++k) {
      if ((b[k] >> lane_id()) & 1ull) lists[k][o[k] + lanes_below(b[k])] = v;
      o[k] += __popcll(b[k]);
    }
  }
}

// Chunk table of the large rows (one workgroup of kHubThreads, each a
// contiguous run of H rows): hub h owns chunks [first[h], first[h+1]), chunk
// c starts at arena entry off[c] of row hlist[cv[c]].  tail[0..4] receive
// the class list sizes (Z, S, M, H) and the chunk count for the host.
constexpr uint32_t kHubThreads = 1024;
__global__ __launch_bounds__(kHubThreads) void k_pr_hub_chunks(const uint2* __restrict__ row,
                                                               const uint32_t* __restrict__ hlist,
                                                               const uint32_t* __restrict__ counts,
                                                               const uint32_t* __restrict__ bases,
                                                               uint32_t* __restrict__ cv, uint32_t* __restrict__ co,
                                                               uint32_t* __restrict__ first, uint32_t* __restrict__ tail) {
  constexpr uint32_t kW = kHubThreads / kWave;
  __shared__ unsigned long long s_tot[kW];
  const uint32_t W = kClassWaves;
  const uint32_t nz = bases[W], ns = bases[2 * W] - bases[W], nm = bases[3 * W] - bases[2 * W];
  const uint32_t nh = bases[4 * W - 1] + counts[4 * W - 1] - bases[3 * W];
  const uint32_t per = (nh + kHubThreads - 1) / kHubThreads;
  const uint32_t h0 = min(nh, threadIdx.x * per), h1 = min(nh, h0 + per);
  unsigned long long mine = 0;
  for (uint32_t h = h0; h < h1; ++h) mine += (row[hlist[h]].y + kChunk - 1) / kChunk;
  const unsigned long long inc = wave_inclusive_scan<unsigned long long>(mine);
  const uint32_t wv = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) s_tot[wv] = inc;
  __syncthreads();
  unsigned long long base = inc - mine, total = 0;
  for (uint32_t w = 0; w < kW; ++w) {
    if (w < wv) base += s_tot[w];
    total += s_tot[w];
  }
  for (uint32_t h = h0; h < h1; ++h) {
    const uint2 rb = row[hlist[h]];
    const uint32_t nc = (rb.y + kChunk - 1) / kChunk;
    first[h] = static_cast<uint32_t>(base);
    for (uint32_t q = 0; q < nc; ++q) {
      cv[base + q] = h;
      co[base + q] = rb.x + q * kChunk;
    }
    base += nc;
  }
  if (threadIdx.x == 0) {
    first[nh] = static_cast<uint32_t>(total);
    tail[0] = nz;
    tail[1] = ns;
    tail[2] = nm;
    tail[3] = nh;
    tail[4] = static_cast<uint32_t>(total);
  }
}

template <class T>
__global__ void k_to_double(const T* src, double* dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[i] = static_cast<double>(src[i]);
}

template <class T>
__global__ void k_from_double(const double* src, T* dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[i] = static_cast<T>(src[i]);
}

// ------------------------------------------------- localized Jacobi repair

struct LocalArgs {
  const uint2* in_row;
  const uint32_t* in_arena;
  const uint2* out_row;
  const uint32_t* out_arena;
  const uint32_t* outdeg;
  uint32_t* act_flag;   // epoch-stamped "in next active set"
  uint32_t* ever_flag;  // 1 once ever active
  uint32_t* alist_cur;
  uint32_t* alist_next;
  unsigned long long* cnt;  // [0] next active count, [1] ever count, [2] l1 bits, [3] linf bits
  uint32_t epoch;
  uint32_t n;
};

__device__ __forceinline__ void activate(const LocalArgs& a, uint32_t w, bool want) {
  bool push = false;
  if (want && w < a.n) {
    push = atomicExch(&a.act_flag[w], a.epoch) != a.epoch;
    if (push && atomicExch(&a.ever_flag[w], 1u) == 0u) atomicAdd(&a.cnt[1], 1ull);
  }
  unsigned long long bb = __ballot(push);
  if (bb) {
    unsigned long long base = 0;
    if (lane_id() == 0) base = atomicAdd(&a.cnt[0], static_cast<unsigned long long>(__popcll(bb)));
    base = __shfl(base, 0, kWave);
    if (push) a.alist_next[base + lanes_below(bb)] = w;
  }
}

// Initial active set: endpoints and the out-neighbours of src (and of dst
// when undirected), pagerank.hpp:112-117.  One wave per update.
__global__ void k_local_seed(LocalArgs a, const uint32_t* src, const uint32_t* dst, uint64_t k, int undirected) {
  const uint64_t W = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; i < k; i += W) {
    const uint32_t s = src[i], d = dst[i];
    activate(a, lane_id() == 0 ? s : d, lane_id() < 2);
    for (int side = 0; side < (undirected ? 2 : 1); ++side) {
      const uint32_t u = side ? d : s;
      if (u >= a.n) continue;
      const uint2 rb = a.out_row[u];
      for (uint32_t j = 0; j < rb.y; j += kWave) {
        uint32_t w = j + lane_id() < rb.y ? a.out_arena[rb.x + j + lane_id()] : kTomb;
        activate(a, w & kIdMask, !(w & kTomb));
      }
    }
  }
}

// One Jacobi sweep over the active list (pagerank.hpp:157-173): wave per
// vertex; updated values go to upd[i] and are written back afterwards.
template <class T>
__global__ void k_local_sweep(LocalArgs a, const T* rank, T* upd, uint64_t na, double base, double damping,
                              double tol) {
  const uint64_t W = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  double l1 = 0.0, linf = 0.0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; i < na; i += W) {
    const uint32_t v = a.alist_cur[i];
    const uint2 rb = a.in_row[v];
    Acc<T> acc;
    for (uint32_t j = lane_id(); j < rb.y; j += kWave) {
      const uint32_t u = a.in_arena[rb.x + j];
      if (u & kTomb) continue;
      const uint32_t od = a.outdeg[u];
      if (od) acc.add(rank[u] / static_cast<T>(od));
    }
    const T incoming = wave_sum(acc.get());
    const T updated = static_cast<T>(base + damping * static_cast<double>(incoming));
    const double delta = fabs(static_cast<double>(updated) - static_cast<double>(rank[v]));
    if (lane_id() == 0) {
      upd[i] = updated;
      l1 += delta;
      linf = linf > delta ? linf : delta;
    }
    const bool spread = delta > tol;
    if (spread) {
      const uint2 ob = a.out_row[v];
      for (uint32_t j = 0; j < ob.y; j += kWave) {
        uint32_t w = j + lane_id() < ob.y ? a.out_arena[ob.x + j + lane_id()] : kTomb;
        activate(a, w & kIdMask, !(w & kTomb));
      }
    }
  }
  l1 = wave_sum(l1);
  linf = wave_max(linf);
  if (lane_id() == 0) {
    atomicAdd(reinterpret_cast<double*>(&a.cnt[2]), l1);
    // non-negative doubles order like their bit patterns
    atomicMax(&a.cnt[3], static_cast<unsigned long long>(__double_as_longlong(linf)));
  }
}

template <class T>
__global__ void k_local_writeback(const uint32_t* alist, const T* upd, uint64_t na, T* rank) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < na;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    rank[alist[i]] = upd[i];
}

__global__ void k_any_dangling(const uint32_t* src, const uint32_t* dst, uint64_t k, int undirected,
                               const uint32_t* outdeg, uint32_t n, unsigned long long* flag) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < k;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t s = src[i], d = dst[i];
    bool dang = s < n && outdeg[s] == 0;
    if (undirected && d != s && d < n && outdeg[d] == 0) dang = true;
    if (dang) atomicOr(flag, 1ull);
  }
}

__global__ void k_count_zero(const uint32_t* deg, uint64_t n, unsigned long long* out) {
  unsigned long long c = 0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    c += deg[i] == 0;
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(out, c);
}

template <class T>
__global__ void k_diff(const T* a, const double* ref, uint64_t n, double* out2) {
  double l1 = 0.0, linf = 0.0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    double e = fabs(static_cast<double>(a[i]) - ref[i]);
    l1 += e;
    linf = linf > e ? linf : e;
  }
  l1 = wave_sum(l1);
  linf = wave_max(linf);
  if (lane_id() == 0) {
    atomicAdd(&out2[0], l1);
    atomicMax(reinterpret_cast<unsigned long long*>(&out2[1]),
              static_cast<unsigned long long>(__double_as_longlong(linf)));
  }
}

// ------------------------------------------------------ residual push
//
// The residual of ranks x is r = b + d M x - x (b = (1-d)/n, M the pull
// operator with the dangling mass spread uniformly); the exact ranks are
// x* = x + A^-1 r with A = I - d M.  A batch changes the rows of its
// sources: their old contributions d x_u / deg_old(u) leave the residual of
// their old out-neighbours and the new ones enter at the new out-neighbours
// (a dangling row contributes d x_u / n to every vertex: that part goes to
// the uniform scalar c).  A push moves r_u into x_u and d r_u / deg(u) to
// each out-neighbour's residual, or d r_u / n into c when u is dangling; it
// keeps x* = x + A^-1 (r_s + c 1) exact.  Since A x* = b with b uniform,
// A^-1 1 = n / (1 - d) x*, so once r_s is small the uniform part is folded
// exactly: x* = (x + A^-1 r_s) / (1 - k), k = c n / (1 - d), and the residual
// of x / (1 - k) is r_s / (1 - k).

struct PushArgs {
  const uint2* out_row;
  const uint32_t* out_arena;
  const uint32_t* outdeg;
  void* x;
  void* r;
  void* amt;                  // per pushed row: d r_u / deg(u)
  uint32_t* stamp;            // epoch-stamped list membership
  uint32_t* list[3];          // active lists, ring by round
  uint32_t* q;                // rows pushed this round
  unsigned long long* pre;    // their arc prefixes
  PushState* st;
  double damping;
  double theta;               // push entries with |r| > theta
  uint32_t n;
  uint32_t epoch;
};

// Append w to the active list `li` once per epoch (wave-aggregated).  Must
// be called by every lane of the wave.
__device__ __forceinline__ void push_activate(const PushArgs& a, int li, uint32_t epoch, uint32_t w, bool want) {
  const bool add = want && atomicExch(&a.stamp[w], epoch) != epoch;
  const unsigned long long bb = __ballot(add);
  if (!bb) return;
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(&a.st->acnt[li], static_cast<unsigned long long>(__popcll(bb)));
  base = __shfl(base, 0, kWave);
  if (add) a.list[li][base + lanes_below(bb)] = w;
}

template <class T>
__device__ __forceinline__ void add_residual(const PushArgs& a, int li, uint32_t epoch, uint32_t w, T amount,
                                             bool want) {
  bool cross = false;
  if (want) {
    T* r = static_cast<T*>(a.r);
    const T old = atomicAdd(&r[w], amount);
    const T nw = old + amount;
    cross = fabs(static_cast<double>(old)) <= a.theta && fabs(static_cast<double>(nw)) > a.theta;
  }
  push_activate(a, li, epoch, w, cross);
}

// Rows whose out-arcs the batch changes: the sources of its ops (both
// endpoints when undirected), each once (stamp epoch), into list 2.
__global__ void k_push_mark_rows(PushArgs a, const uint32_t* src, const uint32_t* dst, uint64_t k, int undirected) {
  for (uint64_t i0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); i0 < k;
       i0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = i0 + threadIdx.x;
    const bool in = i < k;
    const uint32_t u = in ? src[i] : 0u, v = in ? dst[i] : 0u;
    const bool op = in && u != v && u < a.n && v < a.n;
    push_activate(a, 2, a.epoch, u, op);
    push_activate(a, 2, a.epoch, v, op && undirected);
  }
}

// Wave-aggregated reservation of queue slot and arc prefix for (u, len,
// part) in the packed counter *qc.  Called by every lane of the wave.
template <class T>
__device__ __forceinline__ void push_queue(const PushArgs& a, unsigned long long* qc, bool row, uint32_t u,
                                           uint32_t len, T part) {
  const unsigned long long ball = __ballot(row);
  if (!ball) return;
  const uint32_t e = row ? len : 0u;
  const unsigned long long einc = wave_inclusive_scan<unsigned long long>(e);
  const unsigned long long etot = __shfl(einc, kWave - 1, kWave);
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(qc, (static_cast<unsigned long long>(__popcll(ball)) << kEdgeBits) + etot);
  base = __shfl(base, 0, kWave);
  if (row) {
    const uint64_t slot = (base >> kEdgeBits) + lanes_below(ball);
    a.q[slot] = u;
    a.pre[slot] = (base & kEdgeMask) + einc - e;
    static_cast<T*>(a.amt)[slot] = part;
  }
}

// Contributions of the rows of list 2, with `sign` -1 (before the batch:
// the old rows) or +1 (after it: the new rows): rows with live out-arcs go
// to the pushed queue q[0] with sign d x_u / deg(u) for k_push_spread, rows
// without add sign d x_u / n to the uniform part.
template <class T>
__global__ void k_push_queue_rows(PushArgs a, double sign) {
  const T* x = static_cast<const T*>(a.x);
  const uint64_t nrows = a.st->acnt[2];
  double cl = 0.0;
  for (uint64_t i0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); i0 < nrows;
       i0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = i0 + threadIdx.x;
    bool row = false;
    uint32_t u = 0, len = 0;
    T part = T(0);
    if (i < nrows) {
      u = a.list[2][i];
      const uint32_t deg = a.outdeg[u];
      const double xu = static_cast<double>(x[u]);
      if (deg == 0) {
        cl += sign * a.damping * xu / static_cast<double>(a.n);
      } else {
        row = true;
        len = a.out_row[u].y;
        part = static_cast<T>(sign * a.damping * xu / static_cast<double>(deg));
      }
    }
    push_queue<T>(a, &a.st->q[0], row, u, len, part);
  }
  cl = wave_sum(cl);
  if (lane_id() == 0 && cl != 0.0) atomicAdd(&a.st->c, cl);
}

// Every vertex with |r| > theta into active list `li` (after the threshold
// was lowered, or when no bound on the entries is known).
template <class T>
__global__ void k_push_seed(PushArgs a, int li) {
  const T* r = static_cast<const T*>(a.r);
  for (uint32_t v0 = blockIdx.x * blockDim.x; v0 < a.n; v0 += gridDim.x * blockDim.x) {
    const uint32_t v = v0 + threadIdx.x;
    const bool want = v < a.n && fabs(static_cast<double>(r[v])) > a.theta;
    push_activate(a, li, a.epoch, v, want);
  }
}

// Claim of round k: every listed vertex whose residual is above the
// threshold moves it into its rank; rows with live out-arcs go to the
// pushed list q[k & 1] with d r_u / deg(u), dangling ones add d r_u / n to
// the uniform part.  Workgroup 0 clears the counters round k + 1 appends to.
template <class T>
__global__ void k_push_claim(PushArgs a, uint32_t k) {
  PushState* st = a.st;
  const int cur = static_cast<int>(k % 3);
  const uint64_t na = st->acnt[cur];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->acnt[(k + 1) % 3] = 0;
    st->q[(k + 1) & 1] = 0;
  }
  T* x = static_cast<T*>(a.x);
  T* r = static_cast<T*>(a.r);
  T* amt = static_cast<T*>(a.amt);
  unsigned long long* qc = &st->q[k & 1];
  double cl = 0.0;
  unsigned long long pushed = 0;
  for (uint64_t i0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); i0 < na;
       i0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = i0 + threadIdx.x;
    bool row = false;
    uint32_t u = 0, len = 0;
    T part = T(0);
    if (i < na) {
      u = a.list[cur][i];
      const T ru = r[u];
      if (fabs(static_cast<double>(ru)) > a.theta) {
        r[u] = T(0);
        x[u] += ru;
        ++pushed;
        const uint32_t deg = a.outdeg[u];
        if (deg == 0) {
          cl += a.damping * static_cast<double>(ru) / static_cast<double>(a.n);
        } else {
          row = true;
          len = a.out_row[u].y;
          part = static_cast<T>(a.damping * static_cast<double>(ru) / static_cast<double>(deg));
        }
      }
    }
    push_queue<T>(a, qc, row, u, len, part);
  }
  cl = wave_sum(cl);
  pushed = wave_sum(pushed);
  if (lane_id() == 0) {
    if (cl != 0.0) atomicAdd(&st->c, cl);
    if (pushed) atomicAdd(&st->pushed_vertices, pushed);
  }
}

// The queued rows' amounts (queue slot qs) to their live out-neighbours'
// residuals, edge-balanced; entries that cross the threshold join list li
// (epoch `epoch`).  Round k's distribute is (k & 1, (k + 1) % 3, its epoch).
template <class T>
__global__ __launch_bounds__(kPrBlock) void k_push_spread(PushArgs a, int qs, int li, uint32_t epoch) {
  const unsigned long long qc = a.st->q[qs];
  const uint32_t F = static_cast<uint32_t>(qc >> kEdgeBits);
  const unsigned long long E = qc & kEdgeMask;
  const T* amt = static_cast<const T*>(a.amt);
  const uint32_t W = gridDim.x * (blockDim.x / kWave);
  const uint32_t gw = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  lbs_edges(a.q, a.pre, F, E, a.out_row, gw, W, [&](uint32_t item, uint32_t, uint2, uint32_t slot, bool valid) {
    const uint32_t w = valid ? a.out_arena[slot] : kTomb;
    const T amount = valid ? amt[item] : T(0);
    add_residual<T>(a, li, epoch, w & kIdMask, amount, !(w & kTomb));
  });
}

template <class T>
__global__ void k_push_l1(const T* r, uint32_t n, double* out) {
  double s = 0.0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    s += fabs(static_cast<double>(r[v]));
  s = wave_sum(s);
  if (lane_id() == 0 && s != 0.0) atomicAdd(out, s);
}

// x <- x * scale, r <- r * scale with scale = 1 / (1 - k), k = c n / (1 - d):
// the uniform part of the residual folded into the ranks (see above).
template <class T>
__global__ void k_push_fold(T* x, T* r, uint32_t n, double scale) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    x[v] = static_cast<T>(static_cast<double>(x[v]) * scale);
    r[v] = static_cast<T>(static_cast<double>(r[v]) * scale);
  }
}

// r = next - x: the residual of x from one sweep's output next.
template <class T>
__global__ void k_push_resync(const T* next, const T* x, T* r, uint32_t n) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) r[v] = next[v] - x[v];
}

// ---------------------------------------------------------------- host side

static unsigned gblocks(uint64_t items, unsigned per = kPrBlock, unsigned cap = 4096) {
  uint64_t b = (items + per - 1) / per;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

static size_t elt(const vgx_pagerank* p) { return p->precision == 32 ? 4 : 8; }

static void build_lists(vgx_pagerank* p) {
  vgx_graph* g = p->g;
  if (p->lists_version == g->rows_epoch && p->n == g->n) return;
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  const Adjacency& in = g->in();
  p->slist.ensure(n + 1);
  p->mlist.ensure(n + 1);
  p->hlist.ensure(n + 1);
  // at most stored / kChunk + one chunk per large row; large rows exceed kChunk
  const uint64_t max_chunks = 2 * (in.arena.n / kChunk) + 2;
  p->hchunk_v.ensure(max_chunks);
  p->hchunk_off.ensure(max_chunks);
  p->hchunk_first.ensure(std::min<uint64_t>(n, in.arena.n / kChunk + 1) + 1);
  p->zlist.ensure(n + 1);
  p->class_tail.ensure(kClasses + 1);
  uint32_t tail[kClasses + 1] = {0, 0, 0, 0, 0};
  if (n) {
    const uint32_t W = kClassWaves;
    const uint32_t groups = static_cast<uint32_t>((n + kWave - 1) / kWave);
    const uint32_t gpw = (groups + W - 1) / W;
    p->class_counts.ensure(kClasses * W);
    p->class_bases.ensure(kClasses * W);
    const unsigned blocks = W / (kPrBlock / kWave);
    k_pr_class_count<<<blocks, kPrBlock, 0, s>>>(in.row.p, static_cast<uint32_t>(n), gpw, p->class_counts.p);
    VGX_LAUNCH_CHECK();
    size_t bytes = 0;
    VGX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, p->class_counts.p, p->class_bases.p,
                                             static_cast<int>(kClasses * W), s));
    p->scan_tmp.ensure(bytes + 16);
    VGX_HIP(hipcub::DeviceScan::ExclusiveSum(p->scan_tmp.p, bytes, p->class_counts.p, p->class_bases.p,
                                             static_cast<int>(kClasses * W), s));
    k_pr_class_write<<<blocks, kPrBlock, 0, s>>>(in.row.p, static_cast<uint32_t>(n), gpw, p->class_bases.p,
                                                 p->zlist.p, p->slist.p, p->mlist.p, p->hlist.p);
    VGX_LAUNCH_CHECK();
    k_pr_hub_chunks<<<1, kHubThreads, 0, s>>>(in.row.p, p->hlist.p, p->class_counts.p, p->class_bases.p,
                                              p->hchunk_v.p, p->hchunk_off.p, p->hchunk_first.p,
                                              p->class_tail.p);
    VGX_LAUNCH_CHECK();
    VGX_HIP(hipMemcpyAsync(tail, p->class_tail.p, sizeof(tail), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
  }
  p->nz = tail[0];
  p->ns = tail[1];
  p->nm = tail[2];
  p->nh = tail[3];
  p->nchunks = tail[4];
  p->hpart.ensure((p->nchunks + 1) * 8);
  p->lists_version = g->rows_epoch;
}

static PrArgs pr_args(vgx_pagerank* p, DevBuf<uint8_t>* rk) {
  vgx_graph* g = p->g;
  PrArgs a{};
  a.row = g->in().row.p;
  a.arena = g->in().arena.p;
  a.outdeg = g->adj[0].deg.p;
  a.zlist = p->zlist.p;
  a.slist = p->slist.p;
  a.mlist = p->mlist.p;
  a.hlist = p->hlist.p;
  a.hchunk_v = p->hchunk_v.p;
  a.hchunk_off = p->hchunk_off.p;
  a.hchunk_first = p->hchunk_first.p;
  a.nz = static_cast<uint32_t>(p->nz);
  a.ns = static_cast<uint32_t>(p->ns);
  a.nm = static_cast<uint32_t>(p->nm);
  a.nh = static_cast<uint32_t>(p->nh);
  a.nchunks = static_cast<uint32_t>(p->nchunks);
  a.bz = static_cast<uint32_t>(std::min<uint64_t>((p->nz + kPrBlock - 1) / kPrBlock, kMaxClassBlocks));
  a.bs = static_cast<uint32_t>(
      std::min<uint64_t>((p->ns + kPrBlock / kSmallGroup - 1) / (kPrBlock / kSmallGroup), kMaxClassBlocks));
  a.bm = static_cast<uint32_t>(std::min<uint64_t>((p->nm + 3) / 4, kMaxClassBlocks));
  a.n = static_cast<uint32_t>(g->n);
  a.rank0 = rk[0].p;
  a.rank1 = rk[1].p;
  a.contrib0 = p->contrib[0].p;
  a.contrib1 = p->contrib[1].p;
  a.hpart = p->hpart.p;
  a.partials = p->partials.p;
  a.st = p->d_st;
  return a;
}

// Power iteration into rk[0..1]; the result ends in rk[result index].
static int full_solve(vgx_pagerank* p, DevBuf<uint8_t>* rk, int precision, uint64_t max_iterations, double tol,
                      bool stop_l1, bool warm) {
  vgx_graph* g = p->g;
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  const size_t es = precision == 32 ? 4 : 8;
  build_lists(p);
  std::vector<void*>* keep = rk == p->rank ? p->views.keep() : nullptr;
  rk[0].ensure(n * es + 64, keep);
  rk[1].ensure(n * es + 64, keep);
  p->contrib[0].ensure(n * es + 64);
  p->contrib[1].ensure(n * es + 64);
  PrState h{};
  h.damping = p->damping;
  h.teleport = n ? (1.0 - p->damping) / static_cast<double>(n) : 0.0;
  h.tol = tol;
  h.stop_l1 = stop_l1 ? 1 : 0;
  h.max_iterations = max_iterations;
  h.done = (n == 0 || max_iterations == 0) ? 1u : 0u;
  VGX_HIP(hipMemcpyAsync(p->d_st, &h, sizeof(h), hipMemcpyHostToDevice, s));
  PrArgs a = pr_args(p, rk);
  const unsigned grid = std::max<unsigned>(
      1, a.bz + a.bs + a.bm + static_cast<uint32_t>(std::min<uint64_t>((p->nchunks + 3) / 4, kMaxClassBlocks)));
  const unsigned ginit = gblocks(n, kPrBlock, 2048);
  const unsigned gfin = gblocks(std::max<uint64_t>(p->nh, 1), kPrBlock, 64);
  p->partials.ensure(std::max<unsigned>(grid + gfin, ginit) + 1);
  a.partials = p->partials.p;
  if (n == 0) {
    VGX_HIP(hipStreamSynchronize(s));
    *p->h_st = h;
    p->h_st->done = 1;
    return 0;
  }
  if (precision == 32) k_pr_init<float><<<ginit, kPrBlock, 0, s>>>(a, warm ? 1 : 0);
  else k_pr_init<double><<<ginit, kPrBlock, 0, s>>>(a, warm ? 1 : 0);
  VGX_LAUNCH_CHECK();
  uint64_t launched = 0;
  uint32_t chunk = 4;
  while (launched < max_iterations) {
    uint64_t c = std::min<uint64_t>(chunk, max_iterations - launched);
    for (uint64_t i = 0; i < c; ++i) {
      if (precision == 32) {
        k_pr_iter<float><<<grid, kPrBlock, 0, s>>>(a);
        k_pr_finish<float><<<gfin, kPrBlock, 0, s>>>(a, grid);
      } else {
        k_pr_iter<double><<<grid, kPrBlock, 0, s>>>(a);
        k_pr_finish<double><<<gfin, kPrBlock, 0, s>>>(a, grid);
      }
      VGX_LAUNCH_CHECK();
    }
    launched += c;
    VGX_HIP(hipMemcpyAsync(p->h_st, p->d_st, sizeof(PrState), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    if (p->h_st->done) break;
    chunk = std::min<uint32_t>(chunk * 2, 16);
  }
  return static_cast<int>(p->h_st->iterations & 1ull);
}

static void count_dangling(vgx_pagerank* p) {
  vgx_graph* g = p->g;
  p->counters.ensure(16);
  auto* cnt = reinterpret_cast<unsigned long long*>(p->counters.p);
  VGX_HIP(hipMemsetAsync(cnt + 4, 0, 8, g->stream));
  if (g->n) {
    k_count_zero<<<gblocks(g->n, kPrBlock, 512), kPrBlock, 0, g->stream>>>(g->adj[0].deg.p, g->n, cnt + 4);
    VGX_LAUNCH_CHECK();
  }
  unsigned long long c = 0;
  VGX_HIP(hipMemcpyAsync(&c, cnt + 4, 8, hipMemcpyDeviceToHost, g->stream));
  VGX_HIP(hipStreamSynchronize(g->stream));
  p->stats.dangling_vertices = c;
}

// recompute(), pagerank.hpp:193-202
static void recompute(vgx_pagerank* p, uint64_t max_iterations, double tol, bool stop_l1 = false, bool warm = false) {
  p->res_valid = false;
  int r = full_solve(p, p->rank, p->precision, max_iterations, tol, stop_l1, warm);
  p->cur = r;
  p->n = p->g->n;
  p->have_values = true;
  count_dangling(p);
  p->stats.last_repaired_vertices = p->g->n;
  p->stats.last_repair_iterations = p->h_st->iterations;
  p->stats.last_residual_l1 = p->h_st->l1;
  p->stats.last_residual_linf = p->h_st->linf;
  p->stats.last_full_recompute_converged = p->h_st->done == 1 || p->g->n == 0;
  p->stats.last_used_full_recompute = 1;
}

static void reset_local_stats(vgx_pagerank* p) {
  p->stats.last_repaired_vertices = 0;
  p->stats.last_repair_iterations = 0;
  p->stats.last_residual_l1 = 0.0;
  p->stats.last_residual_linf = 0.0;
}

template <class T>
static void local_repair(vgx_pagerank* p, const uint32_t* d_src, const uint32_t* d_dst, uint64_t k,
                         uint64_t local_iterations, double tol, uint64_t fallback_limit) {
  vgx_graph* g = p->g;
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  p->act_flag.ensure(n + 1);
  p->ever_flag.ensure(n + 1);
  p->alist[0].ensure(n + 1);
  p->alist[1].ensure(n + 1);
  p->upd.ensure((n + 1) * sizeof(T));
  VGX_HIP(hipMemsetAsync(p->act_flag.p, 0, (n + 1) * 4, s));
  VGX_HIP(hipMemsetAsync(p->ever_flag.p, 0, (n + 1) * 4, s));
  auto* cnt = reinterpret_cast<unsigned long long*>(p->counters.p);
  VGX_HIP(hipMemsetAsync(cnt, 0, 4 * 8, s));
  LocalArgs a{};
  a.in_row = g->in().row.p;
  a.in_arena = g->in().arena.p;
  a.out_row = g->adj[0].row.p;
  a.out_arena = g->adj[0].arena.p;
  a.outdeg = g->adj[0].deg.p;
  a.act_flag = p->act_flag.p;
  a.ever_flag = p->ever_flag.p;
  a.cnt = cnt;
  a.n = static_cast<uint32_t>(n);
  uint32_t epoch = 1;
  int cur = 0;
  a.epoch = epoch;
  a.alist_cur = p->alist[1].p;
  a.alist_next = p->alist[0].p;
  k_local_seed<<<gblocks(k, 4, 2048), kPrBlock, 0, s>>>(a, d_src, d_dst, k, g->directed ? 0 : 1);
  VGX_LAUNCH_CHECK();
  unsigned long long h[4];
  VGX_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  uint64_t active = h[0];
  if (active == 0) {
    reset_local_stats(p);
    return;
  }
  if (active >= fallback_limit) {
    recompute(p, 200, 1e-12);
    return;
  }
  const double base = (1.0 - p->damping) / static_cast<double>(n);
  reset_local_stats(p);
  p->stats.last_full_recompute_converged = 0;
  p->stats.last_used_full_recompute = 0;
  T* rank = reinterpret_cast<T*>(p->rank[p->cur].p);
  for (uint64_t it = 0; it < local_iterations && active != 0; ++it) {
    p->stats.last_repair_iterations += 1;
    ++epoch;
    a.epoch = epoch;
    a.alist_cur = p->alist[cur].p;
    a.alist_next = p->alist[cur ^ 1].p;
    VGX_HIP(hipMemsetAsync(cnt, 0, 8, s));
    VGX_HIP(hipMemsetAsync(cnt + 2, 0, 16, s));
    k_local_sweep<T><<<gblocks(active, 4, 4096), kPrBlock, 0, s>>>(a, rank, reinterpret_cast<T*>(p->upd.p), active,
                                                                 base, p->damping, tol);
    VGX_LAUNCH_CHECK();
    k_local_writeback<T><<<gblocks(active), kPrBlock, 0, s>>>(a.alist_cur, reinterpret_cast<T*>(p->upd.p), active,
                                                              rank);
    VGX_LAUNCH_CHECK();
    VGX_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    active = h[0];
    double l1, linf;
    std::memcpy(&l1, &h[2], 8);
    std::memcpy(&linf, &h[3], 8);
    p->stats.last_residual_l1 = l1;
    p->stats.last_residual_linf = linf;
    cur ^= 1;
    if (h[1] >= fallback_limit) {
      recompute(p, 200, 1e-12);
      return;
    }
  }
  p->stats.last_repaired_vertices = h[1];
}

static void apply(vgx_pagerank* p, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k,
                  uint64_t local_iterations, double tol, double frac) {
  vgx_graph* g = p->g;
  p->res_valid = false;
  if (k == 0) {
    reset_local_stats(p);
    return;
  }
  const uint64_t prev_n = p->n;
  const bool had_dangling = p->stats.dangling_vertices != 0;
  graph_stage(g, src, dst, add, k);
  graph_apply_deletions(g);
  graph_apply_additions(g);
  graph_finish_batch(g);
  const uint64_t n = g->n;
  p->stats.last_used_full_recompute = 0;
  if (n == 0) {
    reset_local_stats(p);
    p->stats.last_full_recompute_converged = 1;
    p->stats.dangling_vertices = 0;
    p->n = 0;
    return;
  }
  if (n != prev_n || had_dangling) {
    recompute(p, 200, 1e-12);
    return;
  }
  hipStream_t s = g->stream;
  const uint32_t* d_src = g->bs[0].src.p;
  const uint32_t* d_dst = g->bs[0].src.p + k;
  p->counters.ensure(16);
  auto* cnt = reinterpret_cast<unsigned long long*>(p->counters.p);
  VGX_HIP(hipMemsetAsync(cnt + 5, 0, 8, s));
  k_any_dangling<<<gblocks(k), kPrBlock, 0, s>>>(d_src, d_dst, k, g->directed ? 0 : 1, g->adj[0].deg.p,
                                                 static_cast<uint32_t>(n), cnt + 5);
  VGX_LAUNCH_CHECK();
  unsigned long long any = 0;
  VGX_HIP(hipMemcpyAsync(&any, cnt + 5, 8, hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  if (any) {
    recompute(p, 200, 1e-12);
    return;
  }
  p->stats.dangling_vertices = 0;
  double fl = frac * static_cast<double>(n);
  uint64_t limit = static_cast<uint64_t>(fl > 1.0 ? fl : 1.0);
  if (p->precision == 32) local_repair<float>(p, d_src, d_dst, k, local_iterations, tol, limit);
  else local_repair<double>(p, d_src, d_dst, k, local_iterations, tol, limit);
}

// ------------------------------------------------------ residual push (host)

static void push_ensure(vgx_pagerank* p) {
  const uint64_t n = p->g->n;
  const size_t es = elt(p);
  p->res.ensure(n * es + 64);
  p->push_amt.ensure((n + 1) * es + 64);
  if (p->push_stamp.n < n + 1) {
    p->push_stamp.ensure(n + 1);
    VGX_HIP(hipMemsetAsync(p->push_stamp.p, 0, p->push_stamp.n * 4, p->g->stream));
    p->push_epoch = 0;
  }
  for (auto& l : p->push_list) l.ensure(n + 1);
  p->push_q.ensure(n + 1);
  p->push_pre.ensure(n + 1);
  if (!p->d_push) {
    p->d_push = static_cast<PushState*>(dev_alloc(sizeof(PushState)));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_push), sizeof(PushState), hipHostMallocDefault));
  }
}

static PushArgs push_args(vgx_pagerank* p, double theta, uint32_t epoch) {
  vgx_graph* g = p->g;
  PushArgs a{};
  a.out_row = g->adj[0].row.p;
  a.out_arena = g->adj[0].arena.p;
  a.outdeg = g->adj[0].deg.p;
  a.x = p->rank[p->cur].p;
  a.r = p->res.p;
  a.amt = p->push_amt.p;
  a.stamp = p->push_stamp.p;
  for (int i = 0; i < 3; ++i) a.list[i] = p->push_list[i].p;
  a.q = p->push_q.p;
  a.pre = p->push_pre.p;
  a.st = p->d_push;
  a.damping = p->damping;
  a.theta = theta;
  a.n = static_cast<uint32_t>(g->n);
  a.epoch = epoch;
  return a;
}

// Sweeps from the current ranks (warm) until the L1 change is below tol,
// then the residual of the ranks kept: the last sweep's output minus its
// input, with the input kept as the ranks (its residual is that L1 change).
template <class T>
static uint64_t push_sweeps(vgx_pagerank* p, uint64_t max_sweeps, double tol) {
  vgx_graph* g = p->g;
  const uint64_t n = g->n;
  push_ensure(p);  // sized for the current vertex count (it may have grown)
  if (p->cur != 0) {
    VGX_HIP(hipMemcpyAsync(p->rank[0].p, p->rank[p->cur].p, n * sizeof(T), hipMemcpyDeviceToDevice, g->stream));
    p->cur = 0;
  }
  const int r = full_solve(p, p->rank, p->precision, std::max<uint64_t>(max_sweeps, 1), tol, true, true);
  const uint64_t sweeps = p->h_st->iterations;
  const int prev = r ^ 1;
  k_push_resync<T><<<gblocks(n, kPrBlock, 2048), kPrBlock, 0, g->stream>>>(
      reinterpret_cast<const T*>(p->rank[r].p), reinterpret_cast<const T*>(p->rank[prev].p),
      reinterpret_cast<T*>(p->res.p), static_cast<uint32_t>(n));
  VGX_LAUNCH_CHECK();
  p->cur = prev;
  p->n = n;
  p->have_values = true;
  p->res_valid = true;
  p->res_version = g->version;
  p->res_bound = INFINITY;
  p->stats.last_residual_l1 = p->h_st->l1;
  p->stats.last_residual_linf = p->h_st->linf;
  return sweeps;
}

template <class T>
static double push_residual_l1(vgx_pagerank* p) {
  vgx_graph* g = p->g;
  VGX_HIP(hipMemsetAsync(&p->d_push->l1, 0, 8, g->stream));
  k_push_l1<T><<<gblocks(g->n, kPrBlock, 1024), kPrBlock, 0, g->stream>>>(reinterpret_cast<const T*>(p->res.p),
                                                                          static_cast<uint32_t>(g->n), &p->d_push->l1);
  VGX_LAUNCH_CHECK();
  double l1 = 0.0;
  VGX_HIP(hipMemcpyAsync(&l1, &p->d_push->l1, 8, hipMemcpyDeviceToHost, g->stream));
  VGX_HIP(hipStreamSynchronize(g->stream));
  return l1;
}

template <class T>
static void apply_push(vgx_pagerank* p, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k,
                       double tol, double theta, double max_push_fraction, uint64_t max_sweeps,
                       vgx_pagerank_push_stats* out) {
  vgx_graph* g = p->g;
  hipStream_t s = g->stream;
  vgx_pagerank_push_stats st{};
  if (theta <= 0.0) theta = tol / 1024.0;
  st.threshold = theta;
  if (k == 0) {
    reset_local_stats(p);
    if (out) *out = st;
    return;
  }
  push_ensure(p);
  const uint64_t n0 = g->n;
  // the residual of the current ranks, exact before the batch is injected
  if (!p->have_values || !p->res_valid || p->res_version != g->version || p->n != n0) {
    if (!p->have_values || p->n != n0) recompute(p, 200, 1e-12);
    st.fallback_sweeps += push_sweeps<T>(p, 1, 0.0);
  }
  graph_stage(g, src, dst, add, k);
  const uint32_t* d_src = g->bs[0].src.p;
  const uint32_t* d_dst = g->bs[0].src.p + k;
  const uint64_t n = g->n;
  if (n != n0) {
    // vertex growth changes the teleport of every vertex: full solve
    graph_apply_deletions(g);
    graph_apply_additions(g);
    graph_finish_batch(g);
    recompute(p, 200, 1e-12);
    st.fallback_sweeps += p->h_st->iterations;
    st.fallback_sweeps += push_sweeps<T>(p, 1, 0.0);
    st.used_fallback = 1;
    st.residual_l1 = p->h_st->l1;
    count_dangling(p);
    p->stats.last_used_full_recompute = 1;
    if (out) *out = st;
    return;
  }
  if (p->push_skip_k && 2 * k >= p->push_skip_k && ++p->push_skipped % kPushProbe != 0) {
    // the pushes of a batch this size did not fit their budget: warm sweeps
    // straight away (they keep the residual exact: the last sweep's change)
    graph_apply_deletions(g);
    graph_apply_additions(g);
    graph_finish_batch(g);
    st.used_fallback = 1;
    st.skipped_push = 1;
    st.fallback_sweeps += push_sweeps<T>(p, max_sweeps, tol);
    st.residual_l1 = p->h_st->l1;
    count_dangling(p);
    reset_local_stats(p);
    p->stats.last_used_full_recompute = 1;
    p->stats.last_repair_iterations = st.fallback_sweeps;
    p->stats.last_repaired_vertices = n;
    p->stats.last_residual_l1 = st.residual_l1;
    if (out) *out = st;
    return;
  }
  if (p->push_epoch > 0xFFFFFFFFu - 4096) {
    VGX_HIP(hipMemsetAsync(p->push_stamp.p, 0, p->push_stamp.n * 4, s));
    p->push_epoch = 0;
  }
  const uint32_t base = p->push_epoch + 1;
  VGX_HIP(hipMemsetAsync(p->d_push, 0, sizeof(PushState), s));
  const unsigned grid = 1024;
  // rows the batch changes; their old contributions leave the residual
  PushArgs a = push_args(p, theta, base);
  k_push_mark_rows<<<gblocks(k, kPrBlock, 1024), kPrBlock, 0, s>>>(a, d_src, d_dst, k, g->directed ? 0 : 1);
  VGX_LAUNCH_CHECK();
  auto inject = [&](double sign) {
    VGX_HIP(hipMemsetAsync(&p->d_push->q[0], 0, sizeof(unsigned long long), s));
    k_push_queue_rows<T><<<grid, kPrBlock, 0, s>>>(a, sign);
    VGX_LAUNCH_CHECK();
    k_push_spread<T><<<2 * grid, kPrBlock, 0, s>>>(a, 0, 0, base + 1);
    VGX_LAUNCH_CHECK();
  };
  inject(-1.0);
  graph_apply_deletions(g);
  graph_apply_additions(g);
  graph_finish_batch(g);
  // the new rows' contributions enter (the arena may have moved)
  a = push_args(p, theta, base);
  inject(1.0);
  VGX_HIP(hipMemsetAsync(&p->d_push->q[0], 0, sizeof(unsigned long long), s));
  const uint64_t budget = static_cast<uint64_t>(max_push_fraction * static_cast<double>(g->adj[0].stored));
  uint32_t round = 0;
  bool fallback = false;
  double l1 = 0.0;
  auto seed = [&](int li, uint32_t epoch) {
    PushArgs sa = push_args(p, theta, epoch);
    k_push_seed<T><<<gblocks(n, kPrBlock, 2048), kPrBlock, 0, s>>>(sa, li);
    VGX_LAUNCH_CHECK();
  };
  // entries above the threshold from before the batch
  if (p->res_bound > theta) seed(0, base + 1);
  // the ranks before any push: a fallback re-solves from them (sweeps from
  // partly pushed ranks converge more slowly: the pushes leave the error in
  // the slowly decaying directions of the iteration)
  p->push_save.ensure(n * sizeof(T) + 64);
  VGX_HIP(hipMemcpyAsync(p->push_save.p, p->rank[p->cur].p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
  // One round: the claim, a host check of its pushed arcs against the
  // budget, then the distribute.  A claim that finds nothing above the
  // threshold ends the level: the residual's L1 norm is checked and, when
  // still above l1_tol, the threshold drops 16-fold and the entries above
  // it seed the next round.
  uint64_t pushed_arcs = 0, last_pv = 0;
  bool first = true;
  for (;;) {
    PushArgs ra = push_args(p, theta, base);
    k_push_claim<T><<<grid, kPrBlock, 0, s>>>(ra, round);
    VGX_LAUNCH_CHECK();
    VGX_HIP(hipMemcpyAsync(p->h_push, p->d_push, sizeof(PushState), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    if (first) st.touched_rows = p->h_push->acnt[2];
    first = false;
    const uint64_t E = p->h_push->q[round & 1] & kEdgeMask;
    const uint64_t pv = p->h_push->pushed_vertices;
    if (E == 0 && pv == last_pv) {
      l1 = push_residual_l1<T>(p);
      if (l1 <= tol) break;
      theta /= 16.0;
      if (theta < 1e-300) {
        fallback = true;
        break;
      }
      ++round;
      seed(static_cast<int>(round % 3), base + 1 + round);
      continue;
    }
    last_pv = pv;
    if (pushed_arcs + E > budget) {
      fallback = true;
      break;
    }
    pushed_arcs += E;
    k_push_spread<T><<<2 * grid, kPrBlock, 0, s>>>(ra, static_cast<int>(round & 1), static_cast<int>((round + 1) % 3),
                                                  base + 2 + round);
    VGX_LAUNCH_CHECK();
    ++round;
  }
  st.rounds = round;
  st.pushed_vertices = p->h_push->pushed_vertices;
  st.pushed_arcs = pushed_arcs;
  st.threshold = theta;
  p->push_epoch = base + 2 + round;
  // fold the uniform part: x /= 1 - k, r /= 1 - k, k = c n / (1 - d)
  const double kc = p->h_push->c * static_cast<double>(n) / (1.0 - p->damping);
  const double scale = 1.0 / (1.0 - kc);
  if (!fallback) {
    k_push_fold<T><<<gblocks(n, kPrBlock, 2048), kPrBlock, 0, s>>>(reinterpret_cast<T*>(p->rank[p->cur].p),
                                                                  reinterpret_cast<T*>(p->res.p),
                                                                  static_cast<uint32_t>(n), scale);
    VGX_LAUNCH_CHECK();
  }
  p->res_valid = true;
  p->res_version = g->version;
  count_dangling(p);
  reset_local_stats(p);
  // the selector: a push phase over budget in its first round marks this
  // batch size as one to skip; a push phase that finishes clears the mark
  if (fallback && round == 0) {
    p->push_skip_k = p->push_skip_k ? std::min(p->push_skip_k, k) : k;
  } else if (!fallback) {
    p->push_skip_k = 0;
    p->push_skipped = 0;
  }
  if (fallback) {
    st.used_fallback = 1;
    VGX_HIP(hipMemcpyAsync(p->rank[0].p, p->push_save.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    p->cur = 0;
    st.fallback_sweeps += push_sweeps<T>(p, max_sweeps, tol);
    st.residual_l1 = p->h_st->l1;
    p->stats.last_used_full_recompute = 1;
    p->stats.last_repair_iterations = st.fallback_sweeps;
    p->stats.last_repaired_vertices = n;
    p->stats.last_full_recompute_converged = p->h_st->done == 1;
  } else {
    p->res_bound = theta * fabs(scale);
    st.residual_l1 = l1 * fabs(scale);
    p->stats.last_used_full_recompute = 0;
    p->stats.last_repair_iterations = round;
    p->stats.last_repaired_vertices = st.pushed_vertices;
    p->stats.last_residual_l1 = st.residual_l1;
  }
  if (out) *out = st;
}

static vgx_pagerank_stats validate(vgx_pagerank* p, uint64_t ref_it, double ref_tol, double l1_tol, double linf_tol,
                                   double out[6], bool& ref_conv, uint64_t& ref_iters) {
  // FP64 reference solve (the reference's full_solve) into ref_rank.
  int r = full_solve(p, p->ref_rank, 64, ref_it, ref_tol, false, false);
  ref_conv = p->h_st->done == 1 || p->g->n == 0;
  ref_iters = p->h_st->iterations;
  out[2] = p->h_st->l1;
  out[3] = p->h_st->linf;
  vgx_graph* g = p->g;
  hipStream_t s = g->stream;
  p->counters.ensure(16);
  auto* d2 = reinterpret_cast<double*>(p->counters.p) + 6;
  VGX_HIP(hipMemsetAsync(d2, 0, 16, s));
  const uint64_t n = g->n;
  if (p->n != n || !p->have_values) {
    out[0] = out[1] = INFINITY;
  } else {
    if (n) {
      const double* ref = reinterpret_cast<const double*>(p->ref_rank[r].p);
      if (p->precision == 32)
        k_diff<float><<<gblocks(n), kPrBlock, 0, s>>>(reinterpret_cast<float*>(p->rank[p->cur].p), ref, n, d2);
      else
        k_diff<double><<<gblocks(n), kPrBlock, 0, s>>>(reinterpret_cast<double*>(p->rank[p->cur].p), ref, n, d2);
      VGX_LAUNCH_CHECK();
    }
    double h2[2] = {0, 0};
    VGX_HIP(hipMemcpyAsync(h2, d2, 16, hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    out[0] = h2[0];
    out[1] = h2[1];
  }
  out[4] = r;
  (void)l1_tol;
  (void)linf_tol;
  return p->stats;
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

vgx_status vgx_pagerank_create(vgx_graph* g, double damping, int precision, vgx_pagerank** out) {
  VGX_API_BEGIN
  if (!g || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (!(damping >= 0.0 && damping <= 1.0)) throw Error(VGX_INVALID_ARGUMENT, "PageRank damping must be in [0, 1]");
  if (precision != 32 && precision != 64) throw Error(VGX_INVALID_ARGUMENT, "precision must be 32 or 64");
  VGX_HIP(hipSetDevice(g->device));
  auto* p = new vgx_pagerank();
  try {
    p->g = g;
    p->damping = damping;
    p->precision = precision;
    p->d_st = static_cast<PrState*>(dev_alloc(sizeof(PrState)));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_st), sizeof(PrState), hipHostMallocDefault));
    std::memset(p->h_st, 0, sizeof(PrState));
    p->counters.ensure(16);
    recompute(p, 200, 1e-12);
  } catch (...) {
    vgx_pagerank_destroy(p);
    throw;
  }
  *out = p;
  VGX_API_END
}

vgx_status vgx_pagerank_destroy(vgx_pagerank* p) {
  VGX_API_BEGIN
  if (!p) return VGX_OK;
  if (p->g) (void)hipStreamSynchronize(p->g->stream);
  if (p->d_st) dev_free(p->d_st);
  if (p->h_st) (void)hipHostFree(p->h_st);
  if (p->d_push) dev_free(p->d_push);
  if (p->h_push) (void)hipHostFree(p->h_push);
  p->views.destroy();
  delete p;
  VGX_API_END
}

vgx_status vgx_pagerank_recompute(vgx_pagerank* p, uint64_t max_iterations, double tol) {
  VGX_API_BEGIN
  if (!p) throw Error(VGX_INVALID_ARGUMENT, "null pagerank");
  VGX_HIP(hipSetDevice(p->g->device));
  p->views.wait(p->g->stream);
  recompute(p, max_iterations, tol);
  VGX_API_END
}

vgx_status vgx_pagerank_solve_l1(vgx_pagerank* p, uint64_t max_iterations, double l1_tol, int warm_start) {
  VGX_API_BEGIN
  if (!p) throw Error(VGX_INVALID_ARGUMENT, "null pagerank");
  VGX_HIP(hipSetDevice(p->g->device));
  p->views.wait(p->g->stream);
  const bool warm = warm_start && p->have_values && p->n == p->g->n;
  if (warm && p->cur != 0) {
    // warm start reads rank[0]
    VGX_HIP(hipMemcpyAsync(p->rank[0].p, p->rank[p->cur].p, p->n * elt(p), hipMemcpyDeviceToDevice, p->g->stream));
    p->cur = 0;
  }
  recompute(p, max_iterations, l1_tol, true, warm);
  VGX_API_END
}

vgx_status vgx_pagerank_static(vgx_graph* g, double damping, uint64_t max_iterations, double tolerance, double* out,
                               uint64_t cap, uint64_t* count, uint64_t* iterations) {
  VGX_API_BEGIN
  if (!g || !count) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (!(damping >= 0.0 && damping <= 1.0)) throw Error(VGX_INVALID_ARGUMENT, "PageRank damping must be in [0, 1]");
  *count = 0;
  if (iterations) *iterations = 0;
  if (g->n == 0) return VGX_OK;
  VGX_HIP(hipSetDevice(g->device));
  // A solver object without the constructor's solve: uniform start, FP64,
  // L1 stop, then the ranks copied out and the object released.
  struct Owned {
    vgx_pagerank* p = new vgx_pagerank();
    ~Owned() { vgx_pagerank_destroy(p); }
  } o;
  vgx_pagerank* p = o.p;
  p->g = g;
  p->damping = damping;
  p->precision = 64;
  p->d_st = static_cast<PrState*>(dev_alloc(sizeof(PrState)));
  VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_st), sizeof(PrState), hipHostMallocDefault));
  std::memset(p->h_st, 0, sizeof(PrState));
  p->counters.ensure(16);
  recompute(p, max_iterations, tolerance, true, false);
  const uint64_t n = g->n;
  *count = n;
  if (iterations) *iterations = max_iterations == 0 ? 0 : p->h_st->iterations;
  const uint64_t c = std::min(cap, n);
  if (out && c) {
    VGX_HIP(hipMemcpyAsync(out, p->rank[p->cur].p, c * 8, hipMemcpyDeviceToHost, g->stream));
    VGX_HIP(hipStreamSynchronize(g->stream));
  }
  VGX_API_END
}

vgx_status vgx_pagerank_apply(vgx_pagerank* p, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                              uint64_t k, uint64_t local_iterations, double tol, double full_fallback_fraction) {
  VGX_API_BEGIN
  if (!p || (k && (!src || !dst || !add))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(p->g->device));
  p->views.wait(p->g->stream);
  apply(p, src, dst, add, k, local_iterations, tol, full_fallback_fraction);
  VGX_API_END
}

vgx_status vgx_pagerank_apply_push(vgx_pagerank* p, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                                   uint64_t k, double l1_tol, double threshold, double max_push_fraction,
                                   uint64_t max_sweeps, vgx_pagerank_push_stats* out) {
  VGX_API_BEGIN
  if (!p || (k && (!src || !dst || !add))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (!(l1_tol > 0.0)) throw Error(VGX_INVALID_ARGUMENT, "l1_tol must be positive");
  if (!(threshold >= 0.0) || !(max_push_fraction >= 0.0))
    throw Error(VGX_INVALID_ARGUMENT, "threshold and max_push_fraction must be non-negative");
  VGX_HIP(hipSetDevice(p->g->device));
  p->views.wait(p->g->stream);
  if (p->precision == 32)
    apply_push<float>(p, src, dst, add, k, l1_tol, threshold, max_push_fraction, max_sweeps, out);
  else
    apply_push<double>(p, src, dst, add, k, l1_tol, threshold, max_push_fraction, max_sweeps, out);
  VGX_API_END
}

vgx_status vgx_pagerank_values(vgx_pagerank* p, double* out, uint64_t cap, uint64_t* count) {
  VGX_API_BEGIN
  if (!p || !count) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(p->g->device));
  const uint64_t n = p->have_values ? p->n : 0;
  *count = n;
  const uint64_t c = std::min(cap, n);
  if (out && c) {
    hipStream_t s = p->g->stream;
    if (p->precision == 64) {
      VGX_HIP(hipMemcpyAsync(out, p->rank[p->cur].p, c * 8, hipMemcpyDeviceToHost, s));
    } else {
      DevBuf<double> tmp;
      tmp.ensure(c);
      k_to_double<float><<<gblocks(c), kPrBlock, 0, s>>>(reinterpret_cast<float*>(p->rank[p->cur].p), tmp.p, c);
      VGX_LAUNCH_CHECK();
      VGX_HIP(hipMemcpyAsync(out, tmp.p, c * 8, hipMemcpyDeviceToHost, s));
      VGX_HIP(hipStreamSynchronize(s));
    }
    VGX_HIP(hipStreamSynchronize(s));
  }
  VGX_API_END
}

vgx_status vgx_pagerank_stats_get(vgx_pagerank* p, vgx_pagerank_stats* out) {
  VGX_API_BEGIN
  if (!p || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  *out = p->stats;
  VGX_API_END
}

// validate_against_full (pagerank.hpp:204-232); out_errors = {l1, linf,
// reference residual l1, reference residual linf}.
vgx_status vgx_pagerank_validate(vgx_pagerank* p, uint64_t reference_max_iterations, double reference_tol,
                                 double* out_errors, uint64_t* reference_iterations, int* reference_converged) {
  VGX_API_BEGIN
  if (!p || !out_errors) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  VGX_HIP(hipSetDevice(p->g->device));
  double o[6];
  bool conv = false;
  uint64_t it = 0;
  validate(p, reference_max_iterations, reference_tol, 0, 0, o, conv, it);
  for (int i = 0; i < 4; ++i) out_errors[i] = o[i];
  if (reference_iterations) *reference_iterations = it;
  if (reference_converged) *reference_converged = conv ? 1 : 0;
  VGX_API_END
}

// Install the FP64 reference solve as the current values (apply_validated's
// fallback, pagerank.hpp:246-256).
vgx_status vgx_pagerank_install_reference(vgx_pagerank* p, uint64_t reference_max_iterations, double reference_tol) {
  VGX_API_BEGIN
  if (!p) throw Error(VGX_INVALID_ARGUMENT, "null pagerank");
  VGX_HIP(hipSetDevice(p->g->device));
  p->views.wait(p->g->stream);
  int r = full_solve(p, p->ref_rank, 64, reference_max_iterations, reference_tol, false, false);
  vgx_graph* g = p->g;
  const uint64_t n = g->n;
  p->rank[0].ensure(n * elt(p) + 64, p->views.keep());
  if (n) {
    if (p->precision == 64) {
      VGX_HIP(hipMemcpyAsync(p->rank[0].p, p->ref_rank[r].p, n * 8, hipMemcpyDeviceToDevice, g->stream));
    } else {
      k_from_double<float><<<gblocks(n), kPrBlock, 0, g->stream>>>(
          reinterpret_cast<const double*>(p->ref_rank[r].p), reinterpret_cast<float*>(p->rank[0].p), n);
      VGX_LAUNCH_CHECK();
    }
  }
  p->cur = 0;
  p->n = n;
  p->have_values = true;
  p->res_valid = false;
  count_dangling(p);
  p->stats.last_repaired_vertices = n;
  p->stats.last_repair_iterations = p->h_st->iterations;
  p->stats.last_residual_l1 = p->h_st->l1;
  p->stats.last_residual_linf = p->h_st->linf;
  p->stats.last_full_recompute_converged = p->h_st->done == 1 || n == 0;
  VGX_HIP(hipStreamSynchronize(g->stream));
  VGX_API_END
}

vgx_status vgx_pagerank_graph(const vgx_pagerank* p, vgx_graph** g) {
  VGX_API_BEGIN
  if (!p || !g) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  *g = p->g;
  VGX_API_END
}

// Device pointer to the current ranks (float32 or float64 per precision).
vgx_status vgx_pagerank_hold_view(vgx_pagerank* p, int delta, void* stream) {
  VGX_API_BEGIN
  if (!p || (delta != 1 && delta != -1)) throw Error(VGX_INVALID_ARGUMENT, "delta must be +1 or -1");
  VGX_HIP(hipSetDevice(p->g->device));
  if (delta > 0) p->views.hold(static_cast<hipStream_t>(stream));
  else p->views.release();
  VGX_API_END
}

vgx_status vgx_pagerank_values_device(vgx_pagerank* p, const void** dev, uint64_t* count, int* precision) {
  VGX_API_BEGIN
  if (!p || !dev) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  *dev = p->rank[p->cur].p;
  if (count) *count = p->have_values ? p->n : 0;
  if (precision) *precision = p->precision;
  VGX_API_END
}

}  // extern "C"
