// HBM-resident DynamicGraph for gfx950.
//
// Reference: include/velographx/storage/dynamic_graph.hpp (SegmentedCsr
// :33-219, PackedDeltaStore :221-446, CompactRowPatches :448-473,
// DynamicGraph :479-816) and storage/consolidation.hpp:132-163.
//
// Layout (per direction; the in-direction is shared when undirected):
//   row[u] = {beg, len}  one 8-byte load gives a row's arena range
//   cap[u]               slots reserved for the row (len <= cap)
//   deg[u]               live degree (neighbor_count)
//   arena[]              vertex ids, each row sorted ascending; a deleted
//                        entry keeps its place with bit 31 set (tombstone)
// The base CSR of the reference is the initial packed arena; its delta
// overlay becomes in-place tombstones plus inserts merged into the row's
// slack; a row that outgrows its slack is rewritten at the arena's bump
// pointer (the role of a row patch).  Compaction rewrites every row packed
// without tombstones, which is the consolidation snapshot.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "vgx_internal.hpp"

namespace vgx {

thread_local std::string t_last_error;

vgx_status set_error(vgx_status s, const std::string& msg) {
  t_last_error = msg;
  return s;
}

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
  std::string m = std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" +
                  std::to_string(line) + ")";
  if (e == hipErrorOutOfMemory) throw Error(VGX_OUT_OF_MEMORY, m);
  throw Error(VGX_DEVICE_ERROR, m);
}

// Freed device buffers of at least kCacheMinBytes are kept, per device, for
// the next allocation of about their size: a snapshot or cutover builds a
// graph of the size of the one it replaces, and each of its row tables,
// degree arrays and arena would otherwise be a fresh hipMalloc (hundreds of
// MB, and a wait for the queued work) right after the old graph's hipFree.
// A buffer is cached only once the device is idle, the wait a hipFree makes,
// so no queued kernel can still use it; an allocation takes the smallest
// cached buffer of at least its size and at most half as large again.  At
// most kCacheBuffers buffers and kCacheBytes bytes per device are kept (the
// oldest go first); a failed allocation releases them all and retries.
namespace {
constexpr size_t kCacheMinBytes = size_t(1) << 20;
constexpr size_t kCacheBytes = size_t(4) << 30;
constexpr size_t kCacheBuffers = 32;
struct CacheEntry {
  int device;
  void* p;
  size_t bytes;
};
std::mutex g_cache_mu;
// every cacheable live allocation's size and device (the free path needs
// them), and the cache
struct LiveBuf {
  size_t bytes;
  int device;
};
std::unordered_map<void*, LiveBuf>& live_sizes() {
  static auto* m = new std::unordered_map<void*, LiveBuf>();  // lives to process exit
  return *m;
}
std::vector<CacheEntry>& cache() {
  static auto* v = new std::vector<CacheEntry>();
  return *v;
}
void cache_release_all() {
  std::vector<CacheEntry> drop;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    drop.swap(cache());
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (const CacheEntry& e : drop) {
    (void)hipSetDevice(e.device);
    (void)hipFree(e.p);
  }
  (void)hipSetDevice(cur);
}
}  // namespace

void* dev_alloc(size_t bytes) {
  if (bytes == 0) bytes = 256;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (bytes >= kCacheMinBytes) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto& v = cache();
    size_t best = v.size();
    for (size_t i = 0; i < v.size(); ++i) {
      const CacheEntry& e = v[i];
      if (e.device != dev || e.bytes < bytes || e.bytes > bytes + bytes / 2) continue;
      if (best == v.size() || e.bytes < v[best].bytes) best = i;
    }
    if (best != v.size()) {
      const CacheEntry e = v[best];
      v.erase(v.begin() + static_cast<std::ptrdiff_t>(best));
      live_sizes()[e.p] = LiveBuf{e.bytes, e.device};
      return e.p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    cache_release_all();
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(VGX_OUT_OF_MEMORY, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
  }
  if (bytes >= kCacheMinBytes) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    live_sizes()[p] = LiveBuf{bytes, dev};
  }
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  LiveBuf lb{0, 0};
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto& m = live_sizes();
    auto it = m.find(p);
    if (it != m.end()) {
      lb = it->second;
      m.erase(it);
    }
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (lb.bytes && lb.device != cur) (void)hipSetDevice(lb.device);
  // small buffers, and any free that cannot first see the device idle, go
  // back to the runtime
  if (lb.bytes == 0 || lb.bytes > kCacheBytes || hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(p);
    if (lb.bytes && lb.device != cur) (void)hipSetDevice(cur);
    return;
  }
  const int dev = lb.device;
  const size_t bytes = lb.bytes;
  std::vector<void*> drop;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto& v = cache();
    v.push_back({dev, p, bytes});
    for (;;) {
      size_t count = 0, total = 0;
      for (const CacheEntry& e : v)
        if (e.device == dev) {
          ++count;
          total += e.bytes;
        }
      if (count <= kCacheBuffers && total <= kCacheBytes) break;
      for (size_t i = 0; i < v.size(); ++i)
        if (v[i].device == dev) {
          drop.push_back(v[i].p);
          v.erase(v.begin() + static_cast<std::ptrdiff_t>(i));
          break;
        }
    }
  }
  for (void* q : drop) (void)hipFree(q);
  if (dev != cur) (void)hipSetDevice(cur);
}

HostBuf::~HostBuf() {
  if (p) (void)hipHostFree(p);
}

void HostBuf::ensure(size_t b) {
  if (b <= bytes) return;
  if (p) (void)hipHostFree(p);
  p = nullptr;
  size_t c = b < 4096 ? 4096 : b + b / 4;
  VGX_HIP(hipHostMalloc(&p, c, hipHostMallocDefault));
  bytes = c;
}

static inline uint32_t slack_for(uint32_t len) { return len == 0 ? 0u : (len >> 3) + 2u; }
__host__ __device__ static inline uint32_t relocated_cap(uint32_t need) { return need + (need >> 2) + 4u; }

static int bits_for(uint64_t n) {
  int b = 1;
  while (b < 32 && (1ull << b) < n) ++b;
  return b;
}

// A row takes its inserts in place on one thread while it stays this short
// and fits its slack; a longer row that fits is merged in place by one
// workgroup (right-to-left tiles) up to kInPlaceMax entries; rows that
// outgrow their slots or are longer are rewritten at the bump pointer by as
// many workgroups as they have kRelocChunk-entry chunks.  The in-place
// merge walks its tiles one after another, so rows beyond four tiles are
// faster rewritten by parallel chunks.
constexpr uint32_t kBigRow = 32;
constexpr uint32_t kInPlaceMax = 4096;
constexpr uint32_t kMidInsMax = 1024;  // most inserts one in-place merge takes
constexpr uint32_t kRelocChunk = 2048;
constexpr int kBlock = 256;

static unsigned grid_for(uint64_t items, unsigned per_block = kBlock, unsigned cap = 65535u * 4u) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// ------------------------------------------------------------------ kernels

__global__ void k_expand_keys(const uint32_t* __restrict__ s, const uint32_t* __restrict__ d, uint64_t m,
                              int undirected, int swap, unsigned long long* __restrict__ keys,
                              uint32_t* __restrict__ vals) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t u = s[i], v = d[i];
    if (swap) {
      uint32_t t = u;
      u = v;
      v = t;
    }
    bool loop = u == v;
    // Both arcs of an undirected op sit next to each other so that the
    // (stable) radix sort keeps every arc's ops in batch order.
    const uint64_t o = undirected ? 2 * i : i;
    keys[o] = loop ? kInvalidKey : ((static_cast<unsigned long long>(u) << 32) | v);
    if (vals) vals[o] = static_cast<uint32_t>(i);
    if (undirected) {
      keys[o + 1] = loop ? kInvalidKey : ((static_cast<unsigned long long>(v) << 32) | u);
      if (vals) vals[o + 1] = static_cast<uint32_t>(i);
    }
  }
}

// Row bounds of a sorted, unique key array: first index and one-past-last.
__global__ void k_row_bounds(const unsigned long long* __restrict__ keys, uint64_t M, uint32_t* __restrict__ rstart,
                             uint32_t* __restrict__ rend) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < M;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    unsigned long long k = keys[i];
    if (k == kInvalidKey) continue;
    uint32_t u = static_cast<uint32_t>(k >> 32);
    if (i == 0 || static_cast<uint32_t>(keys[i - 1] >> 32) != u) rstart[u] = static_cast<uint32_t>(i);
    if (i + 1 == M || keys[i + 1] == kInvalidKey || static_cast<uint32_t>(keys[i + 1] >> 32) != u)
      rend[u] = static_cast<uint32_t>(i + 1);
  }
}

__global__ void k_row_caps(const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ rend, uint64_t n,
                           uint32_t* __restrict__ cap, uint32_t* __restrict__ deg, uint2* __restrict__ row) {
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; u < n;
       u += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t len = rend[u] - rstart[u];
    cap[u] = len == 0 ? 0u : len + (len >> 3) + 2u;
    deg[u] = len;
    row[u].y = len;
  }
}

__global__ void k_row_begin(const uint32_t* __restrict__ beg, uint64_t n, uint2* __restrict__ row) {
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; u < n;
       u += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    row[u].x = beg[u];
}

__global__ void k_scatter(const unsigned long long* __restrict__ keys, uint64_t M, const uint32_t* __restrict__ rstart,
                          const uint2* __restrict__ row, uint32_t* __restrict__ arena) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < M;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    unsigned long long k = keys[i];
    if (k == kInvalidKey) continue;
    uint32_t u = static_cast<uint32_t>(k >> 32);
    arena[row[u].x + (static_cast<uint32_t>(i) - rstart[u])] = static_cast<uint32_t>(k);
  }
}

__device__ __forceinline__ uint32_t row_lower_bound(const uint32_t* arena, uint2 rb, uint32_t v) {
  uint32_t lo = 0, hi = rb.y;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if ((arena[rb.x + mid] & kIdMask) < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// A batch of up to kSortSmall arcs is sorted by one workgroup in LDS (one
// launch instead of the device-wide sort's histogram and digit passes),
// the workgroup sized to the batch (64, 256 or 1024 threads of four keys):
// keys and op indices in blocked order, padded with invalid keys, through a
// stable block radix sort, so the ops on an arc stay in batch order.  The
// sort runs over a compact key (u << b) | v with b = bits of the vertex
// count, and an invalid key as 1 << 2b, so 2b + 1 bits instead of 32 + b.
constexpr uint32_t kSortSmallItems = 4;
constexpr uint64_t kSortSmall = 1024 * kSortSmallItems;
template <uint32_t kThreads>
__global__ __launch_bounds__(kThreads) void k_sort_small(const unsigned long long* __restrict__ keys,
                                                         const uint32_t* __restrict__ vals, uint32_t M, int b,
                                                         unsigned long long* __restrict__ skeys,
                                                         uint32_t* __restrict__ svals) {
  using Sort = hipcub::BlockRadixSort<unsigned long long, kThreads, kSortSmallItems, uint32_t>;
  __shared__ typename Sort::TempStorage tmp;
  const unsigned long long inv = 1ull << (2 * b);
  const unsigned long long low = (1ull << b) - 1ull;
  unsigned long long k[kSortSmallItems];
  uint32_t v[kSortSmallItems];
#pragma unroll
  for (uint32_t j = 0; j < kSortSmallItems; ++j) {
    const uint32_t i = threadIdx.x * kSortSmallItems + j;
    const unsigned long long x = i < M ? keys[i] : kInvalidKey;
    k[j] = x == kInvalidKey ? inv : (((x >> 32) << b) | (x & 0xFFFFFFFFull));
    v[j] = i < M ? vals[i] : 0u;
  }
  Sort(tmp).Sort(k, v, 0, 2 * b + 1);
#pragma unroll
  for (uint32_t j = 0; j < kSortSmallItems; ++j) {
    const uint32_t i = threadIdx.x * kSortSmallItems + j;
    if (i < M) {
      skeys[i] = k[j] == inv ? kInvalidKey : (((k[j] >> b) << 32) | (k[j] & low));
      svals[i] = v[j];
    }
  }
}

// Batches of up to kSelectSmall arc ops are compacted by k_resolve's
// per-workgroup counts and k_select_write; larger ones use the library's
// device-wide select.
constexpr uint64_t kSelectSmall = 1u << 17;

// Decide the action of every final arc op (the last op on its arc wins).
// With bcount set the grid covers M exactly (one arc per thread) and each
// workgroup also leaves its counts of final arcs and inserts in
// bcount[blockIdx.x] (final << 32 | inserts) for k_select_write.
__global__ __launch_bounds__(kBlock) void k_resolve(const unsigned long long* __restrict__ skeys,
                                                    const uint32_t* __restrict__ svals, uint64_t M,
                                                    const uint8_t* __restrict__ add, const uint2* __restrict__ row,
                                                    const uint32_t* __restrict__ arena, uint64_t n,
                                                    uint8_t* __restrict__ act, uint32_t* __restrict__ pos,
                                                    uint8_t* __restrict__ fflag, uint8_t* __restrict__ iflag,
                                                    unsigned long long* __restrict__ bcount) {
  unsigned long long cnt = 0;
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < M;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    unsigned long long key = skeys[p];
    bool fin = key != kInvalidKey && (p + 1 == M || skeys[p + 1] != key);
    uint8_t f = 0, ins = 0;
    if (fin) {
      uint32_t u = static_cast<uint32_t>(key >> 32), v = static_cast<uint32_t>(key);
      bool want = add[svals[p]] != 0;
      uint2 rb = u < n ? row[u] : make_uint2(0u, 0u);
      uint32_t i = row_lower_bound(arena, rb, v);
      bool found = i < rb.y && (arena[rb.x + i] & kIdMask) == v;
      bool tomb = found && (arena[rb.x + i] & kTomb);
      uint8_t a;
      if (!want) a = (found && !tomb) ? kActDel : kActNopDel;
      else if (found) a = tomb ? kActUndel : kActNopAdd;
      else a = kActIns;
      act[p] = a;
      pos[p] = found ? rb.x + i : 0u;
      f = 1;
      ins = a == kActIns;
    }
    fflag[p] = f;
    iflag[p] = ins;
    cnt += (static_cast<unsigned long long>(f) << 32) + ins;
  }
  if (bcount) {
    __shared__ unsigned long long s_cnt[kBlock / kWave];
    cnt = wave_sum(cnt);
    if (lane_id() == 0) s_cnt[threadIdx.x / kWave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (uint32_t w = 0; w < kBlock / kWave; ++w) t += s_cnt[w];
      bcount[blockIdx.x] = t;
    }
  }
}

// Order-keeping compaction of both flag arrays over the tiling of k_resolve
// with bcount: workgroup b sums the counts of the workgroups before it, scans
// its own flags and writes the positions of final arcs to out1 and of
// inserts to out2; the last workgroup writes the totals.  Many workgroups
// share the scattered stores that one workgroup would issue alone.
__global__ __launch_bounds__(kBlock) void k_select_write(const uint8_t* __restrict__ f1, const uint8_t* __restrict__ f2,
                                                         uint64_t M, const unsigned long long* __restrict__ bcount,
                                                         uint32_t* __restrict__ out1, unsigned long long* n1,
                                                         uint32_t* __restrict__ out2, unsigned long long* n2) {
  __shared__ unsigned long long s_w[kBlock / kWave];
  __shared__ unsigned long long s_base;
  const uint32_t wv = threadIdx.x / kWave;
  if (wv == 0) {
    unsigned long long b = 0;
    for (uint32_t i = lane_id(); i < blockIdx.x; i += kWave) b += bcount[i];
    b = wave_sum(b);
    if (lane_id() == 0) s_base = b;
  }
  const uint64_t p = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
  const bool in = p < M;
  const unsigned long long c = in ? (static_cast<unsigned long long>(f1[p] != 0) << 32) + (f2[p] != 0) : 0ull;
  const unsigned long long inc = wave_inclusive_scan<unsigned long long>(c);
  if (lane_id() == kWave - 1) s_w[wv] = inc;
  __syncthreads();
  unsigned long long base = s_base + inc - c;
  for (uint32_t w = 0; w < wv; ++w) base += s_w[w];
  if (in) {
    if (c >> 32) out1[base >> 32] = static_cast<uint32_t>(p);
    if (c & 0xFFFFFFFFull) out2[base & 0xFFFFFFFFull] = static_cast<uint32_t>(p);
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) {
    const unsigned long long total = base + c;
    *n1 = total >> 32;
    *n2 = total & 0xFFFFFFFFull;
  }
}

// Per insert-run planning: segment starts record the run length and the
// arena slots a relocation will take.
__global__ void k_plan_inserts(const uint32_t* __restrict__ ins, const unsigned long long* __restrict__ nins_p,
                               const unsigned long long* __restrict__ skeys, const uint2* __restrict__ row,
                               const uint32_t* __restrict__ cap, uint32_t* __restrict__ segk,
                               uint32_t* __restrict__ big, uint32_t* __restrict__ mid, GraphCounters* ctr, int dir) {
  const uint64_t nins = *nins_p;
  const unsigned lane = lane_id();
  // Every counter below takes one atomic per wave (a hot counter takes an
  // atomic per row otherwise: thousands in a row on one address).
  unsigned long long mid_added = 0;  // inserts merged in place by k_insert_mid
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t j0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x - lane; j0 < nins; j0 += stride) {
    const uint64_t j = j0 + lane;
    bool is_big = false, is_mid = false;
    unsigned long long nc = 0;
    if (j < nins) {
      const uint32_t u = static_cast<uint32_t>(skeys[ins[j]] >> 32);
      if (j > 0 && static_cast<uint32_t>(skeys[ins[j - 1]] >> 32) == u) {
        segk[j] = 0;
      } else {
        uint64_t e = j + 1;
        while (e < nins && static_cast<uint32_t>(skeys[ins[e]] >> 32) == u) ++e;
        const uint32_t k = static_cast<uint32_t>(e - j);
        segk[j] = k;
        const uint32_t len = row[u].y, c = cap[u];
        const uint32_t need = len + k;
        // a row that outgrows its slots, or is too long to shift in place
        // by one workgroup, is rewritten at the bump pointer by many
        // workgroups; a row of more than kBigRow entries that fits is merged
        // in place
        is_big = need > c || len > kInPlaceMax || (need > kBigRow && k > kMidInsMax);
        if (is_big) {
          nc = relocated_cap(need);
        } else if (need > kBigRow) {
          is_mid = true;
          mid_added += k;
        }
      }
    }
    const unsigned long long bb = __ballot(is_big), bm = __ballot(is_mid);
    unsigned long long base_big = 0, base_mid = 0;
    nc = wave_sum(nc);
    if (lane == 0) {
      if (bb) base_big = atomicAdd(&ctr->nbig[dir], static_cast<unsigned long long>(__popcll(bb)));
      if (bm) base_mid = atomicAdd(&ctr->nmid[dir], static_cast<unsigned long long>(__popcll(bm)));
      if (nc) atomicAdd(&ctr->need[dir], nc);
    }
    base_big = __shfl(base_big, 0, kWave);
    base_mid = __shfl(base_mid, 0, kWave);
    if (is_big) big[base_big + lanes_below(bb)] = static_cast<uint32_t>(j);
    if (is_mid) mid[base_mid + lanes_below(bm)] = static_cast<uint32_t>(j);
  }
  mid_added = wave_sum(mid_added);
  if (lane == 0 && mid_added) {
    atomicAdd(&ctr->live_delta[dir], mid_added);
    atomicAdd(&ctr->stored_delta[dir], mid_added);
  }
}

__device__ __forceinline__ void block_add(unsigned long long* dst, long long v) {
  if (v != 0) atomicAdd(dst, static_cast<unsigned long long>(v));
}

__global__ void k_apply_flip(const uint32_t* __restrict__ fin, const unsigned long long* __restrict__ nfin_p,
                             const unsigned long long* __restrict__ skeys, const uint8_t* __restrict__ act,
                             const uint32_t* __restrict__ pos, uint32_t* __restrict__ arena,
                             uint32_t* __restrict__ deg, GraphCounters* ctr, int dir, int undelete) {
  const uint64_t nf = *nfin_p;
  long long live = 0, tomb = 0;
  for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < nf;
       j += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t p = fin[j];
    uint8_t a = act[p];
    uint32_t u = static_cast<uint32_t>(skeys[p] >> 32);
    if (!undelete && a == kActDel) {
      arena[pos[p]] |= kTomb;
      atomicSub(&deg[u], 1u);
      --live;
      ++tomb;
    } else if (undelete && a == kActUndel) {
      arena[pos[p]] &= kIdMask;
      atomicAdd(&deg[u], 1u);
      ++live;
      --tomb;
    }
  }
  live = wave_sum(live);
  tomb = wave_sum(tomb);
  if (lane_id() == 0) {
    block_add(&ctr->live_delta[dir], live);
    block_add(&ctr->tomb_delta[dir], tomb);
  }
}

__global__ void k_insert_small(const uint32_t* __restrict__ ins, const unsigned long long* __restrict__ nins_p,
                               const uint32_t* __restrict__ segk, const unsigned long long* __restrict__ skeys,
                               uint2* __restrict__ row, uint32_t* __restrict__ cap, uint32_t* __restrict__ arena,
                               uint32_t* __restrict__ deg, GraphCounters* ctr, int dir) {
  const uint64_t nins = *nins_p;
  long long added = 0;
  for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < nins;
       j += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t k = segk[j];
    if (k == 0) continue;
    uint32_t u = static_cast<uint32_t>(skeys[ins[j]] >> 32);
    uint2 rb = row[u];
    uint32_t len = rb.y, c = cap[u];
    if (len + k > kBigRow || len + k > c) continue;  // merged by k_insert_big
    // backward in-place merge into the row's slack
    int i = static_cast<int>(len) - 1, t = static_cast<int>(k) - 1, w = static_cast<int>(len + k) - 1;
    while (t >= 0) {
      uint32_t vt = static_cast<uint32_t>(skeys[ins[j + t]]);
      if (i >= 0 && (arena[rb.x + i] & kIdMask) > vt) {
        arena[rb.x + w] = arena[rb.x + i];
        --i;
      } else {
        arena[rb.x + w] = vt;
        --t;
      }
      --w;
    }
    row[u].y = len + k;
    atomicAdd(&deg[u], k);
    added += k;
  }
  added = wave_sum(added);
  if (lane_id() == 0) {
    block_add(&ctr->live_delta[dir], added);
    block_add(&ctr->stored_delta[dir], added);
  }
}

// One wave per mid row, four rows per workgroup: rows of kBigRow < len + k
// <= cap, len <= kInPlaceMax, k <= kMidMaxIns, merged in place.  An insert's
// id is in neither the live nor the deleted entries of its row (a deleted
// entry of the same id is flipped back instead), so the final positions of
// the old entries, i + (inserts below it), and of the inserts, t + (old
// entries below it), are disjoint.  The wave finds each insert's position by
// a binary search of the old row and keeps it in LDS with the insert; then it
// shifts the entries from the first insert's position on, 64 at a time from
// the row's tail (an entry only moves right, so a chunk's stores land on
// entries already read), and stores the inserts last.
constexpr uint32_t kMidMaxIns = kMidInsMax;
constexpr uint32_t kMidWaves = 4;
__global__ __launch_bounds__(kMidWaves * kWave) void k_insert_mid(const uint32_t* __restrict__ mid, uint32_t nmid,
                                                                  const uint32_t* __restrict__ ins,
                                                                  const uint32_t* __restrict__ segk,
                                                                  const unsigned long long* __restrict__ skeys,
                                                                  uint2* __restrict__ row, uint32_t* arena,
                                                                  uint32_t* __restrict__ deg) {
  __shared__ uint32_t s_ins[kMidWaves][kMidMaxIns];
  __shared__ uint32_t s_pos[kMidWaves][kMidMaxIns];
  const uint32_t w = threadIdx.x / kWave;
  const uint32_t r = blockIdx.x * kMidWaves + w;
  if (r >= nmid) return;  // no workgroup barrier below
  const uint32_t lane = lane_id();
  const uint32_t j = mid[r];
  const uint32_t k = segk[j];  // at least 1
  const uint32_t u = static_cast<uint32_t>(skeys[ins[j]] >> 32);
  const uint2 rb = row[u];
  const uint32_t len = rb.y;
  uint32_t* a = arena + rb.x;
  uint32_t* si = s_ins[w];
  uint32_t* sp = s_pos[w];
  for (uint32_t t = lane; t < k; t += kWave) {
    const uint32_t x = static_cast<uint32_t>(skeys[ins[j + t]]);
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
      const uint32_t m = (lo + hi) >> 1;
      if ((a[m] & kIdMask) < x) lo = m + 1;
      else hi = m;
    }
    si[t] = x;
    sp[t] = t + lo;
  }
  __builtin_amdgcn_wave_barrier();
  const uint32_t p0 = sp[0];  // entries below the first insert stay
  for (uint32_t end = len; end > p0;) {
    const uint32_t beg = end - p0 > kWave ? end - kWave : p0;
    const uint32_t i = beg + lane;
    if (i < end) {
      const uint32_t v = a[i];
      const uint32_t x = v & kIdMask;
      uint32_t lo = 0, hi = k;
      while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (si[m] < x) lo = m + 1;
        else hi = m;
      }
      a[i + lo] = v;
    }
    end = beg;
  }
  for (uint32_t t = lane; t < k; t += kWave) a[sp[t]] = si[t];
  if (lane == 0) {  // the live and stored counts were added by k_plan_inserts
    row[u].y = len + k;
    atomicAdd(&deg[u], k);
  }
}

// Relocation plan (a thread per relocated row b, insert index big[b]):
// plan[8b ..] = {u, j, k, old begin, old len, new begin, -, first output
// chunk}.  Each wave reserves its rows' new ranges and output chunks with
// one atomicAdd on the bump pointer and one on the chunk counter (rows of a
// wave are contiguous in both), and chunk_row[c] names the row that owns
// output chunk c.  The row table switches to the new range; the old range
// stays intact for k_reloc_copy.
__global__ void k_reloc_plan(const uint32_t* __restrict__ big, uint32_t nbig, const uint32_t* __restrict__ ins,
                             const uint32_t* __restrict__ segk, const unsigned long long* __restrict__ skeys,
                             uint2* __restrict__ row, uint32_t* __restrict__ cap, uint32_t* __restrict__ deg,
                             uint32_t* __restrict__ plan, uint32_t* __restrict__ chunk_row, GraphCounters* ctr,
                             int dir) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t b0 = blockIdx.x * blockDim.x + threadIdx.x - lane_id(); b0 < nbig; b0 += stride) {
    const uint32_t b = b0 + lane_id();
    const bool in = b < nbig;
    uint32_t j = 0, k = 0, u = 0, nc = 0, chunks = 0;
    uint2 rb = make_uint2(0u, 0u);
    if (in) {
      j = big[b];
      k = segk[j];
      u = static_cast<uint32_t>(skeys[ins[j]] >> 32);
      rb = row[u];
      const uint32_t need = rb.y + k;
      nc = relocated_cap(need);
      chunks = (need + kRelocChunk - 1) / kRelocChunk;
    }
    const unsigned long long inc_nc = wave_inclusive_scan<unsigned long long>(nc);
    const unsigned long long inc_ch = wave_inclusive_scan<unsigned long long>(chunks);
    const unsigned long long ksum = wave_sum<unsigned long long>(k);
    const unsigned long long rows = __popcll(__ballot(in));
    unsigned long long base_nc = 0, base_ch = 0;
    if (lane_id() == kWave - 1) {
      base_nc = atomicAdd(&ctr->bump[dir], inc_nc);
      base_ch = atomicAdd(&ctr->reloc_chunks[dir], inc_ch);
      // live and stored arcs grow by the inserts of every relocated row
      atomicAdd(&ctr->live_delta[dir], ksum);
      atomicAdd(&ctr->stored_delta[dir], ksum);
      atomicAdd(&ctr->relocated[dir], rows);
    }
    base_nc = __shfl(base_nc, kWave - 1, kWave);
    base_ch = __shfl(base_ch, kWave - 1, kWave);
    if (in) {
      const uint32_t nb = static_cast<uint32_t>(base_nc + inc_nc - nc);
      const uint32_t cf = static_cast<uint32_t>(base_ch + inc_ch - chunks);
      uint4* p = reinterpret_cast<uint4*>(plan + 8ull * b);
      p[0] = make_uint4(u, j, k, rb.x);
      p[1] = make_uint4(rb.y, nb, 0u, cf);
      row[u] = make_uint2(nb, rb.y + k);
      cap[u] = nc;
      atomicAdd(&deg[u], k);
      for (uint32_t c = 0; c < chunks; ++c) chunk_row[cf + c] = b;
    }
  }
}

// Output chunk c of a relocated row: the merge of the old row and its
// inserts over outputs [cj * kRelocChunk, ...), each thread a contiguous
// eighth found by a merge-path split.  Workgroups loop over the chunks.
__global__ __launch_bounds__(256) void k_reloc_copy(const uint32_t* __restrict__ plan,
                                                    const uint32_t* __restrict__ chunk_row,
                                                    const GraphCounters* __restrict__ ctr, int dir,
                                                    const uint32_t* __restrict__ ins,
                                                    const unsigned long long* __restrict__ skeys,
                                                    uint32_t* __restrict__ arena) {
  const uint32_t total = static_cast<uint32_t>(ctr->reloc_chunks[dir]);
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const uint32_t* p = plan + 8ull * chunk_row[c];
    const uint32_t j = p[1], k = p[2], ox = p[3], len = p[4], nb = p[5], cf = p[7];
    const uint32_t need = len + k;
    const uint32_t c0 = (c - cf) * kRelocChunk;
    const uint32_t c1 = min(c0 + kRelocChunk, need);
    constexpr uint32_t per = kRelocChunk / 256;
    const uint32_t o0 = c0 + threadIdx.x * per;
    const uint32_t o1 = min(o0 + per, c1);
    if (o0 < c1) {
      // merge path: old entries among the first o0 outputs
      uint32_t l = o0 > k ? o0 - k : 0u, h = o0 < len ? o0 : len;
      while (l < h) {
        const uint32_t m = (l + h) >> 1;
        const uint32_t av = arena[ox + m] & kIdMask;
        const uint32_t bv = static_cast<uint32_t>(skeys[ins[j + (o0 - m - 1)]]);
        if (av < bv) l = m + 1;
        else h = m;
      }
      uint32_t i = l, t = o0 - l;
      for (uint32_t w = o0; w < o1; ++w) {
        const uint32_t vt = t < k ? static_cast<uint32_t>(skeys[ins[j + t]]) : kInf;
        if (i < len && (arena[ox + i] & kIdMask) < vt) {
          arena[nb + w] = arena[ox + i];
          ++i;
        } else {
          arena[nb + w] = vt;
          ++t;
        }
      }
    }
  }
}

// Packing (snapshots and compaction).  Rows of up to kPackShort stored
// entries are copied by k_pack_rows, eight rows side by side per wave;
// every longer row is cut into kPackChunk-entry chunks, one wave per chunk,
// from a chunk table built on the device, so the work stays balanced however
// skewed the degrees are (R-MAT's heavy rows cluster at ids with few bits
// set).  A chunk of a row without deleted entries (live degree = stored
// length) is a plain copy to off[u] + its start; otherwise its live
// entries are counted first and placed after those of the row's earlier
// chunks (an exclusive scan over all chunks).
constexpr uint32_t kPackShort = 64;
constexpr uint32_t kPackGroup = 8;
constexpr uint32_t kPackChunk = 1024;

// Copy the live entries of every row of up to kPackShort entries to
// out + off[u] (order preserved), a wave per 64 consecutive rows.  Rows
// without deleted entries (all of them after a bulk load) are copied
// flattened: the wave scans their lengths, and lane l takes positions
// l, l + 64, ... of their concatenation, finding its row by a binary search
// over the 64 starts held in the lanes, four positions in flight.  Rows with
// deleted entries are compacted by groups of eight lanes, eight rows side by
// side, each row's entries loaded at once (eight per lane) and placed eight
// per step.  Optionally writes the new row table (every row) for compaction.
__global__ void k_pack_rows(const uint2* __restrict__ row, const uint32_t* __restrict__ arena, uint64_t n,
                            const uint32_t* __restrict__ off, uint32_t* __restrict__ out, uint64_t out_cap,
                            uint2* __restrict__ new_row, const uint32_t* __restrict__ deg) {
  const uint64_t W = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  const uint64_t gw = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave;
  const unsigned lane = lane_id();
  const unsigned grp = lane / kPackGroup, gl = lane % kPackGroup;
  for (uint64_t u0 = gw * kWave; u0 < n; u0 += W * kWave) {
    const uint64_t u = u0 + lane;
    uint2 rb = make_uint2(0u, 0u);
    uint32_t o = 0, dg = 0;
    if (u < n) {
      rb = row[u];
      o = off[u];
      dg = deg[u];
      if (new_row != nullptr) new_row[u] = make_uint2(o, dg);
    }
    const bool shrt = rb.y <= kPackShort;
    const bool dirty = shrt && dg != rb.y;
    // clean short rows, flattened
    const uint32_t flen = shrt && !dirty ? rb.y : 0u;
    const uint32_t fend = wave_inclusive_scan(flen);
    const uint32_t fstart = fend - flen;
    const uint32_t total = __shfl(fend, kWave - 1, kWave);
    constexpr uint32_t kFlat = 4;
    for (uint32_t k = 0; k < total; k += kFlat * kWave) {
      uint32_t x[kFlat];
      uint64_t at[kFlat];
#pragma unroll
      for (uint32_t q = 0; q < kFlat; ++q) {
        const uint32_t pos = k + q * kWave + lane;
        const bool in = pos < total;
        const uint32_t pq = in ? pos : 0u;
        // the last row whose start is <= pq has pq inside it (zero-length
        // rows share their start with the next row)
        uint32_t r = 0;
#pragma unroll
        for (uint32_t step = kWave / 2; step; step >>= 1)
          if (__shfl(fstart, static_cast<int>(r + step), kWave) <= pq) r += step;
        const uint32_t rs = __shfl(fstart, static_cast<int>(r), kWave);
        const uint32_t rx = __shfl(rb.x, static_cast<int>(r), kWave);
        const uint32_t ro = __shfl(o, static_cast<int>(r), kWave);
        at[q] = in ? static_cast<uint64_t>(ro) + (pq - rs) : out_cap;
        x[q] = in ? arena[rx + (pq - rs)] : 0u;
      }
#pragma unroll
      for (uint32_t q = 0; q < kFlat; ++q)
        if (at[q] < out_cap) out[at[q]] = x[q];  // the host checks the total
    }
    // short rows with deleted entries, by groups
    if (!__ballot(dirty)) continue;
    for (unsigned r = 0; r < kWave / kPackGroup; ++r) {
      // row handled by group grp in this round: lane grp * 8 + r
      const int src = static_cast<int>(grp * kPackGroup + r);
      const uint32_t bx = __shfl(rb.x, src, kWave);
      const uint32_t by = __shfl(dirty ? rb.y : 0u, src, kWave);
      uint64_t og = __shfl(o, src, kWave);
      const uint32_t most = wave_max(by);
      if (most == 0) continue;
      // the row's entries (at most kPackShort, kPer per lane) are all
      // loaded before the first is placed
      constexpr uint32_t kPer = kPackShort / kPackGroup;
      uint32_t x[kPer];
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t j = q * kPackGroup + gl;
        x[q] = j < by ? arena[bx + j] : kTomb;
      }
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) {
        if (q * kPackGroup >= most) break;
        const bool keep = !(x[q] & kTomb);
        const unsigned long long m = (__ballot(keep) >> (grp * kPackGroup)) & 0xFFull;
        const uint64_t at = og + __popcll(m & ((1ull << gl) - 1ull));
        if (keep && at < out_cap) out[at] = x[q];  // the host checks the total
        og += __popcll(m);
      }
    }
  }
}

// chunks[u] = kPackChunk-entry chunks of row u when it is longer than
// kPackShort (else 0); chunks[n] = 0 closes the scan.
__global__ void k_pack_chunk_count(const uint2* __restrict__ row, uint64_t n, uint32_t* __restrict__ chunks) {
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; u <= n;
       u += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t len = u < n ? row[u].y : 0u;
    chunks[u] = len > kPackShort ? (len + kPackChunk - 1) / kPackChunk : 0u;
  }
}

// Chunk table: chunk first[u] + q covers entries [q * kPackChunk, ...) of
// row u; counts[0] receives the number of chunks.
__global__ void k_pack_chunk_fill(const uint32_t* __restrict__ first, uint64_t n, uint32_t* __restrict__ chunk_u,
                                  uint32_t* __restrict__ chunk_j, unsigned long long* counts) {
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; u <= n;
       u += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (u == n) {
      counts[0] = first[n];
      continue;
    }
    const uint32_t f = first[u], e = first[u + 1];
    for (uint32_t c = f; c < e; ++c) {
      chunk_u[c] = static_cast<uint32_t>(u);
      chunk_j[c] = c - f;
    }
  }
}

// Live entries of every chunk, a wave per chunk: counted for rows with
// deleted entries, the chunk's length otherwise.  Slots past the last chunk
// are left as they are: the scan's results there are never read.
__global__ void k_pack_chunk_live(const uint2* __restrict__ row, const uint32_t* __restrict__ arena,
                                  const uint32_t* __restrict__ deg, const uint32_t* __restrict__ chunk_u,
                                  const uint32_t* __restrict__ chunk_j, const unsigned long long* counts,
                                  uint32_t* __restrict__ live) {
  const uint32_t nch = static_cast<uint32_t>(counts[0]);
  const unsigned lane = lane_id();
  const uint32_t W = gridDim.x * (blockDim.x / kWave);
  for (uint32_t c = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nch; c += W) {
    const uint32_t u = chunk_u[c];
    const uint2 rb = row[u];
    const uint32_t b = chunk_j[c] * kPackChunk;
    const uint32_t len = min(b + kPackChunk, rb.y) - b;
    uint32_t k = len;
    if (deg[u] != rb.y) {
      constexpr uint32_t kPer = kPackChunk / kWave;
      const uint32_t* src = arena + rb.x + b;
      uint32_t x[kPer];
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t i = q * kWave + lane;
        x[q] = i < len ? src[i] : kTomb;
      }
      k = 0;
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) k += (x[q] & kTomb) ? 0u : 1u;
      k = wave_sum(k);
    }
    if (lane == 0) live[c] = k;
  }
}

// Copy chunk c to off[u] + the live entries of the row's earlier chunks
// (lscan: exclusive scan of the live counts over all chunks).
__global__ void k_pack_chunk_write(const uint2* __restrict__ row, const uint32_t* __restrict__ arena,
                                   const uint32_t* __restrict__ deg, const uint32_t* __restrict__ off,
                                   const uint32_t* __restrict__ chunk_u, const uint32_t* __restrict__ chunk_j,
                                   const uint32_t* __restrict__ lscan, const unsigned long long* counts,
                                   uint32_t* __restrict__ out, uint64_t out_cap) {
  const uint32_t nch = static_cast<uint32_t>(counts[0]);
  const uint32_t W = gridDim.x * (blockDim.x / kWave);
  const unsigned lane = lane_id();
  for (uint32_t c = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nch; c += W) {
    const uint32_t u = chunk_u[c], j = chunk_j[c];
    const uint2 rb = row[u];
    uint64_t o = static_cast<uint64_t>(off[u]) + (lscan[c] - lscan[c - j]);  // the row's chunks are consecutive
    const uint32_t b = j * kPackChunk;
    const uint32_t e = min(b + kPackChunk, rb.y);
    // the chunk's entries (at most kPackChunk = 16 per lane) are all loaded
    // before any is stored
    constexpr uint32_t kPer = kPackChunk / kWave;
    const uint32_t len = e - b;
    const uint32_t* src = arena + rb.x + b;
    uint32_t x[kPer];
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
      const uint32_t i = q * kWave + lane;
      x[q] = i < len ? src[i] : kTomb;
    }
    if (deg[u] == rb.y) {
      if (o + len <= out_cap) {  // the host checks the total
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
          const uint32_t i = q * kWave + lane;
          if (i < len) out[o + i] = x[q];
        }
      }
      continue;
    }
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
      const bool keep = !(x[q] & kTomb);
      const unsigned long long m = __ballot(keep);
      const uint64_t at = o + lanes_below(m);
      if (keep && at < out_cap) out[at] = x[q];  // the host checks the total
      o += __popcll(m);
    }
  }
}

// bump pointer of direction dir := the packed total off[n]
__global__ void k_set_bump(const uint32_t* off, uint64_t n, GraphCounters* ctr, int dir) {
  if (threadIdx.x == 0) ctr->bump[dir] = off[n];
}

__global__ void k_compact_caps(const uint32_t* __restrict__ deg, uint64_t n, uint32_t* __restrict__ cap) {
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; u < n;
       u += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t d = deg[u];
    cap[u] = d == 0 ? 0u : d + (d >> 3) + 2u;
  }
}

// ------------------------------------------------------------ host helpers

struct DeviceGuard {
  explicit DeviceGuard(int dev) { VGX_HIP(hipSetDevice(dev)); }
};

template <class F>
static void with_temp(vgx_graph* g, DevBuf<uint8_t>& tmp, F&& f) {
  size_t bytes = 0;
  VGX_HIP(f(static_cast<void*>(nullptr), bytes));
  tmp.ensure(bytes + 16);
  VGX_HIP(f(static_cast<void*>(tmp.p), bytes));
  (void)g;
}

void graph_sync_counters(vgx_graph* g) {
  VGX_HIP(hipMemcpyAsync(g->h_ctr, g->d_ctr, sizeof(GraphCounters), hipMemcpyDeviceToHost, g->stream));
  VGX_HIP(hipStreamSynchronize(g->stream));
}

static void reset_deltas(vgx_graph* g) {
  // The bump pointers survive; every other counter restarts at zero.
  VGX_HIP(hipMemsetAsync(&g->d_ctr->need, 0, sizeof(GraphCounters) - offsetof(GraphCounters, need), g->stream));
}

void graph_grow_vertices(vgx_graph* g, uint64_t n_new) {
  if (n_new <= g->n) return;
  if (n_new > (1ull << 31)) throw Error(VGX_INVALID_ARGUMENT, "vertex ids must be below 2^31");
  if (n_new > g->ncap) {
    uint64_t c = std::max<uint64_t>(n_new, g->ncap * 2);
    c = std::max<uint64_t>(c, 64);
    for (int d = 0; d < g->ndir(); ++d) {
      Adjacency& a = g->adj[d];
      a.row.grow_copy(c, g->ncap, g->stream);
      a.cap.grow_copy(c, g->ncap, g->stream);
      a.deg.grow_copy(c, g->ncap, g->stream);
      a.base.grow_copy(c, g->ncap, g->stream);
      VGX_HIP(hipMemsetAsync(a.row.p + g->ncap, 0, (c - g->ncap) * sizeof(uint2), g->stream));
      VGX_HIP(hipMemsetAsync(a.cap.p + g->ncap, 0, (c - g->ncap) * sizeof(uint32_t), g->stream));
      VGX_HIP(hipMemsetAsync(a.deg.p + g->ncap, 0, (c - g->ncap) * sizeof(uint32_t), g->stream));
      VGX_HIP(hipMemsetAsync(a.base.p + g->ncap, 0, (c - g->ncap) * sizeof(uint32_t), g->stream));
    }
    g->ncap = c;
  }
  g->n = n_new;
}

// A freshly packed direction: every row's base length is its live degree.
static void rebase_rows(vgx_graph* g, int d) {
  Adjacency& a = g->adj[d];
  if (g->n)
    VGX_HIP(hipMemcpyAsync(a.base.p, a.deg.p, g->n * sizeof(uint32_t), hipMemcpyDeviceToDevice, g->stream));
}

static void ensure_arena(vgx_graph* g, int d, uint64_t extra) {
  Adjacency& a = g->adj[d];
  if (a.used + extra <= a.acap) return;
  uint64_t want = std::max<uint64_t>(a.acap + a.acap / 2, a.used + extra + (1u << 16));
  if (want >= (1ull << 32)) throw Error(VGX_OUT_OF_MEMORY, "row arena exceeds 2^32 entries");
  a.arena.grow_copy(want, a.used, g->stream);
  a.acap = want;
}

// Canonical build from an arc-key array (already mirrored/swapped).
static void build_direction(vgx_graph* g, int d, DevBuf<unsigned long long>& keys, DevBuf<unsigned long long>& skeys,
                            uint64_t M) {
  Adjacency& a = g->adj[d];
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  int end_bit = 32 + bits_for(n);
  if (M) {
    with_temp(g, g->bs[0].tmp, [&](void* t, size_t& b) {
      return hipcub::DeviceRadixSort::SortKeys(t, b, keys.p, skeys.p, static_cast<int>(M), 0, end_bit, s);
    });
  }
  // unique (invalid keys sort last and collapse to one entry)
  DevBuf<unsigned long long> nsel;
  nsel.ensure(1);
  if (M) {
    with_temp(g, g->bs[0].tmp, [&](void* t, size_t& b) {
      return hipcub::DeviceSelect::Unique(t, b, skeys.p, keys.p, nsel.p, static_cast<int>(M), s);
    });
  }
  unsigned long long U = 0;
  if (M) {
    VGX_HIP(hipMemcpyAsync(&U, nsel.p, sizeof(U), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    unsigned long long last = 0;
    if (U) {
      VGX_HIP(hipMemcpyAsync(&last, keys.p + U - 1, sizeof(last), hipMemcpyDeviceToHost, s));
      VGX_HIP(hipStreamSynchronize(s));
      if (last == kInvalidKey) --U;
    }
  }
  if (U >= (1ull << 32)) throw Error(VGX_OUT_OF_MEMORY, "more than 2^32 arcs per direction");
  DevBuf<uint32_t> rstart, rend, beg;
  rstart.ensure(n + 1);
  rend.ensure(n + 1);
  beg.ensure(n + 1);
  VGX_HIP(hipMemsetAsync(rstart.p, 0, (n + 1) * sizeof(uint32_t), s));
  VGX_HIP(hipMemsetAsync(rend.p, 0, (n + 1) * sizeof(uint32_t), s));
  if (U) {
    k_row_bounds<<<grid_for(U), kBlock, 0, s>>>(keys.p, U, rstart.p, rend.p);
    VGX_LAUNCH_CHECK();
  }
  if (n) {
    k_row_caps<<<grid_for(n), kBlock, 0, s>>>(rstart.p, rend.p, n, a.cap.p, a.deg.p, a.row.p);
    VGX_LAUNCH_CHECK();
    with_temp(g, g->bs[0].tmp, [&](void* t, size_t& b) {
      return hipcub::DeviceScan::InclusiveSum(t, b, a.cap.p, beg.p + 1, static_cast<int>(n), s);
    });
    VGX_HIP(hipMemsetAsync(beg.p, 0, sizeof(uint32_t), s));
  }
  uint32_t total = 0;
  if (n) {
    VGX_HIP(hipMemcpyAsync(&total, beg.p + n, sizeof(total), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
  }
  uint64_t acap = static_cast<uint64_t>(total) + static_cast<uint64_t>(total) / 4 + (1u << 16);
  a.arena.release();
  a.arena.ensure(acap);
  a.acap = a.arena.n;
  if (n) {
    k_row_begin<<<grid_for(n), kBlock, 0, s>>>(beg.p, n, a.row.p);
    VGX_LAUNCH_CHECK();
  }
  if (U) {
    k_scatter<<<grid_for(U), kBlock, 0, s>>>(keys.p, U, rstart.p, a.row.p, a.arena.p);
    VGX_LAUNCH_CHECK();
  }
  a.used = total;
  a.stored = U;
  a.live = U;
  a.relocated = 0;
  unsigned long long bump = total;
  VGX_HIP(hipMemcpyAsync(&g->d_ctr->bump[d], &bump, sizeof(bump), hipMemcpyHostToDevice, s));
  VGX_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------- edge lists in device memory
//
// The same builds from edge lists that already sit in HBM (a torch tensor, a
// DLPack column): nothing crosses PCIe but two 8-byte results.  Ids come as
// 32-bit (signed or unsigned: both read as uint32, so a negative int32 is a
// large id and is refused) or int64 words; a list is checked on the device
// before any state changes: `bad` counts ids outside [0, 2^31 - 1), and the
// extent is the largest endpoint + 1 over the non-loop pairs.

template <class T>
__device__ __forceinline__ bool edge_id(const T* p, uint64_t i, uint32_t& out) {
  const T x = p[i];
  if (sizeof(T) == 8) {
    const long long y = static_cast<long long>(x);
    if (y < 0 || y >= static_cast<long long>(kIdMask)) return false;
    out = static_cast<uint32_t>(y);
    return true;
  }
  out = static_cast<uint32_t>(x);
  return out < kIdMask;
}

// Pair list: src[i * stride], dst[i * stride].
template <class T>
__global__ void k_edges_check(const T* __restrict__ src, const T* __restrict__ dst, uint64_t m, uint64_t stride,
                              unsigned long long* __restrict__ res) {
  unsigned long long ext = 0, bad = 0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t u, v;
    const bool ok = edge_id(src, i * stride, u) & edge_id(dst, i * stride, v);
    if (!ok)
      ++bad;
    else if (u != v)
      ext = max(ext, static_cast<unsigned long long>(max(u, v)) + 1ull);
  }
  ext = wave_max(ext);
  bad = wave_sum(bad);
  if (lane_id() == 0) {
    if (ext) atomicMax(&res[0], ext);
    if (bad) atomicAdd(&res[1], bad);
  }
}

// Arc keys of a checked pair list (as k_expand_keys, for a bulk load: no
// batch order to keep).
template <class T>
__global__ void k_edges_keys(const T* __restrict__ src, const T* __restrict__ dst, uint64_t m, uint64_t stride,
                             int undirected, int swap, unsigned long long* __restrict__ keys) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t u, v;
    edge_id(src, i * stride, u);
    edge_id(dst, i * stride, v);
    if (swap) {
      const uint32_t t = u;
      u = v;
      v = t;
    }
    const bool loop = u == v;
    const uint64_t o = undirected ? 2 * i : i;
    keys[o] = loop ? kInvalidKey : ((static_cast<unsigned long long>(u) << 32) | v);
    if (undirected) keys[o + 1] = loop ? kInvalidKey : ((static_cast<unsigned long long>(v) << 32) | u);
  }
}

// CSR of a square matrix of n rows, offsets and column ids read as signed
// integers (int32 or int64): res[1] counts rows whose offsets start below 0,
// fall, or end past nnz; res[0] counts column ids of the rows' entries
// outside [0, n); res[2] counts the entries a build keeps (all of a row's
// entries when directed, those with v >= u when not, loops included: the
// ops of the batch the host path applies).
template <class T>
__global__ void k_csr_check(const T* __restrict__ indptr, const T* __restrict__ indices, uint64_t n, uint64_t nnz,
                            int undirected, unsigned long long* __restrict__ res) {
  unsigned long long bad_off = 0, bad_id = 0, kept = 0;
  const long long first = static_cast<long long>(indptr[0]), last = static_cast<long long>(indptr[n]);
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < nnz + n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    if (i < n) {
      const long long b = static_cast<long long>(indptr[i]), e = static_cast<long long>(indptr[i + 1]);
      if (b < 0 || e < b || static_cast<uint64_t>(e) > nnz) ++bad_off;
    } else {
      const uint64_t p = i - n;
      if (static_cast<long long>(p) < first || static_cast<long long>(p) >= last) continue;  // in no row
      const long long v = static_cast<long long>(indices[p]);
      if (v < 0 || static_cast<uint64_t>(v) >= n) {
        ++bad_id;
      } else if (undirected) {
        // the row of entry p: the last u with indptr[u] <= p
        uint64_t lo = 0, hi = n;
        while (hi - lo > 1) {
          const uint64_t mid = (lo + hi) / 2;
          if (static_cast<long long>(indptr[mid]) <= static_cast<long long>(p))
            lo = mid;
          else
            hi = mid;
        }
        if (static_cast<uint64_t>(v) >= lo) ++kept;
      } else {
        ++kept;
      }
    }
  }
  bad_off = wave_sum(bad_off);
  bad_id = wave_sum(bad_id);
  kept = wave_sum(kept);
  if (lane_id() == 0) {
    if (bad_id) atomicAdd(&res[0], bad_id);
    if (bad_off) atomicAdd(&res[1], bad_off);
    if (kept) atomicAdd(&res[2], kept);
  }
}

// Arc keys of a checked CSR, one wave per row: entry (u, v) of row u gives
// arc u->v; an undirected graph keeps only v >= u (each pair once, as the
// host path does) and mirrors it; loops and skipped entries are invalid
// keys, which sort last and are cut off.
template <class T>
__global__ void k_csr_keys(const T* __restrict__ indptr, const T* __restrict__ indices, uint64_t n, int undirected,
                           int swap, unsigned long long* __restrict__ keys) {
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; u < n;
       u += waves) {
    const uint64_t b = static_cast<uint64_t>(indptr[u]), e = static_cast<uint64_t>(indptr[u + 1]);
    for (uint64_t p = b + lane_id(); p < e; p += kWave) {
      const uint32_t v = static_cast<uint32_t>(indices[p]);
      const uint32_t uu = static_cast<uint32_t>(u);
      const bool skip = uu == v || (undirected && v < uu);
      uint32_t a = uu, c = v;
      if (swap) {
        a = v;
        c = uu;
      }
      const uint64_t o = undirected ? 2 * p : p;
      keys[o] = skip ? kInvalidKey : ((static_cast<unsigned long long>(a) << 32) | c);
      if (undirected) keys[o + 1] = skip ? kInvalidKey : ((static_cast<unsigned long long>(c) << 32) | a);
    }
  }
}

// Builds every direction from `arcs` keys written by fill(keys, d), replacing
// the adjacency (bulk_load_edges semantics: vertex count grows, version + 1).
template <class Fill>
static void bulk_load_keys(vgx_graph* g, uint64_t arcs, Fill&& fill) {
  for (int d = 0; d < g->ndir(); ++d) {
    DevBuf<unsigned long long> keys, skeys;
    keys.ensure(arcs + 1);
    skeys.ensure(arcs + 1);
    if (arcs) fill(keys.p, d);
    build_direction(g, d, keys, skeys, arcs);
    rebase_rows(g, d);
  }
  g->version += 1;
  g->rows_epoch += 1;
}

// Waits for the producer stream, runs `check` into res[0..2] on the graph's
// stream and returns them.
template <class Check>
static std::array<uint64_t, 3> device_check(vgx_graph* g, hipStream_t producer, Check&& check) {
  hipStream_t s = g->stream;
  if (producer != s) VGX_HIP(hipStreamSynchronize(producer));  // null: the default stream
  BatchScratch& b0 = g->bs[0];
  b0.extent.ensure(3);
  VGX_HIP(hipMemsetAsync(b0.extent.p, 0, 3 * sizeof(unsigned long long), s));
  check(b0.extent.p);
  VGX_LAUNCH_CHECK();
  unsigned long long res[3] = {0, 0, 0};
  VGX_HIP(hipMemcpyAsync(res, b0.extent.p, sizeof(res), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  return {res[0], res[1], res[2]};
}

template <class T>
static void bulk_load_device_t(vgx_graph* g, const T* src, const T* dst, uint64_t m, uint64_t stride,
                               hipStream_t producer) {
  hipStream_t s = g->stream;
  const auto [ext, bad, unused] = device_check(g, producer, [&](unsigned long long* res) {
    if (m) k_edges_check<T><<<grid_for(m, kBlock, 4096), kBlock, 0, s>>>(src, dst, m, stride, res);
  });
  (void)unused;
  if (bad) throw Error(VGX_INVALID_ARGUMENT, "vertex ids must lie in [0, 2^31 - 1)");
  if (ext) graph_grow_vertices(g, ext);
  const int undirected = g->directed ? 0 : 1;
  bulk_load_keys(g, undirected ? 2 * m : m, [&](unsigned long long* keys, int d) {
    k_edges_keys<T><<<grid_for(m), kBlock, 0, s>>>(src, dst, m, stride, undirected, d == 1 ? 1 : 0, keys);
    VGX_LAUNCH_CHECK();
  });
}

// A host edge list (src[i * stride], dst[i * stride]) already uploaded to
// dsrc/ddst with the same stride, checked on the device.  When the check
// finds ids at or past 2^31 - 1, which it does not size, the host form's rule
// (grow to max id + 1 over the non-loop edges; an id of 2^31 or more is
// refused by the growth) is applied by a scan of the host list.
static void bulk_load_host(vgx_graph* g, const uint32_t* src, const uint32_t* dst, const uint32_t* dsrc,
                           const uint32_t* ddst, uint64_t m, uint64_t stride) {
  hipStream_t s = g->stream;
  const auto [ext_dev, bad, unused] = device_check(g, s, [&](unsigned long long* res) {
    if (m) k_edges_check<uint32_t><<<grid_for(m, kBlock, 4096), kBlock, 0, s>>>(dsrc, ddst, m, stride, res);
  });
  (void)unused;
  uint64_t ext = ext_dev;
  if (bad) {
    uint64_t maxid = 0;
    bool any = false;
    for (uint64_t i = 0; i < m; ++i) {
      const uint32_t u = src[i * stride], v = dst[i * stride];
      if (u == v) continue;
      any = true;
      maxid = std::max<uint64_t>(maxid, std::max(u, v));
    }
    ext = any ? maxid + 1 : 0;
  }
  if (ext) graph_grow_vertices(g, ext);
  const int undirected = g->directed ? 0 : 1;
  bulk_load_keys(g, undirected ? 2 * m : m, [&](unsigned long long* keys, int d) {
    k_edges_keys<uint32_t><<<grid_for(m), kBlock, 0, s>>>(dsrc, ddst, m, stride, undirected, d == 1 ? 1 : 0, keys);
    VGX_LAUNCH_CHECK();
  });
}

template <class T>
static void bulk_load_csr_device_t(vgx_graph* g, const T* indptr, const T* indices, uint64_t n, uint64_t nnz,
                                   hipStream_t producer) {
  hipStream_t s = g->stream;
  const auto [bad_id, bad, kept] = device_check(g, producer, [&](unsigned long long* res) {
    if (n + nnz) k_csr_check<T><<<grid_for(n + nnz, kBlock, 4096), kBlock, 0, s>>>(indptr, indices, n, nnz,
                                                                            g->directed ? 0 : 1, res);
  });
  if (bad) throw Error(VGX_INVALID_ARGUMENT, "invalid CSR offsets");
  if (bad_id) throw Error(VGX_INVALID_ARGUMENT, "CSR contains out-of-range vertex id");
  graph_grow_vertices(g, n);
  if (kept == 0) return;  // an empty batch: adjacency and version unchanged
  const int undirected = g->directed ? 0 : 1;
  bulk_load_keys(g, undirected ? 2 * nnz : nnz, [&](unsigned long long* keys, int d) {
    // entries past indptr[n] (a CSR whose nnz exceeds its last offset) stay invalid
    VGX_HIP(hipMemsetAsync(keys, 0xFF, (undirected ? 2 * nnz : nnz) * sizeof(unsigned long long), s));
    if (n) k_csr_keys<T><<<grid_for(n * kWave, kBlock, 8192), kBlock, 0, s>>>(indptr, indices, n, undirected,
                                                                            d == 1 ? 1 : 0, keys);
    VGX_LAUNCH_CHECK();
  });
}

// ------------------------------------------------------------ batch apply

static void stage_sorted(vgx_graph* g, uint64_t k, bool defer);

void graph_stage(vgx_graph* g, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k, bool defer) {
  if (k >= (1ull << 31)) throw Error(VGX_INVALID_ARGUMENT, "update batch too large");
  hipStream_t s = g->stream;
  // one pinned staging copy (src | dst | add), the ids copied in the same
  // branch-free pass that finds the largest endpoint of the non-loop ops (a
  // non-loop op's larger end is at least 1, so 0 means only loops); uploaded
  // in pieces, so that each piece's upload overlaps the pass over the next
  g->stage.ensure(k * 9 + 64);
  BatchScratch& b0 = g->bs[0];
  b0.src.ensure(2 * k + 64);
  b0.add.ensure(k + 64);
  auto* hs = static_cast<uint32_t*>(g->stage.p);
  auto* hadd = reinterpret_cast<uint8_t*>(hs + 2 * k);
  constexpr uint64_t kPiece = 1ull << 18;
  uint32_t mx = 0;
  for (uint64_t a = 0; a < k; a += kPiece) {
    const uint64_t e = std::min(k, a + kPiece);
    uint32_t* __restrict__ hs_src = hs;
    uint32_t* __restrict__ hs_dst = hs + k;
    const uint32_t* __restrict__ ps = src;
    const uint32_t* __restrict__ pd = dst;
    uint32_t m_piece = 0;
    for (uint64_t i = a; i < e; ++i) {
      const uint32_t u = ps[i], v = pd[i];
      hs_src[i] = u;
      hs_dst[i] = v;
      const uint32_t m = u == v ? 0u : (u > v ? u : v);
      m_piece = m > m_piece ? m : m_piece;
    }
    mx = std::max(mx, m_piece);
    std::memcpy(hadd + a, add + a, e - a);
    VGX_HIP(hipMemcpyAsync(b0.src.p + a, hs + a, 4 * (e - a), hipMemcpyHostToDevice, s));
    VGX_HIP(hipMemcpyAsync(b0.src.p + k + a, hs + k + a, 4 * (e - a), hipMemcpyHostToDevice, s));
    VGX_HIP(hipMemcpyAsync(b0.add.p + a, hadd + a, e - a, hipMemcpyHostToDevice, s));
  }
  // an id check that fails leaves only the staging scratch written
  if (mx >= (1u << 31) - 1) throw Error(VGX_INVALID_ARGUMENT, "vertex ids must be below 2^31 - 1");
  if (mx) graph_grow_vertices(g, static_cast<uint64_t>(mx) + 1);
  stage_sorted(g, k, defer);
}

// Copies a device batch into the staging layout (src | dst, add) and
// computes the largest endpoint + 1 over its non-loop ops (0 when all are
// loops) in the same pass.
__global__ void k_batch_ingest(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                               const uint8_t* __restrict__ add, uint64_t k, uint32_t* __restrict__ sd,
                               uint8_t* __restrict__ add_out, unsigned long long* __restrict__ extent) {
  unsigned long long m = 0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < k;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t u = src[i], v = dst[i];
    sd[i] = u;
    sd[k + i] = v;
    add_out[i] = add[i];
    if (u != v) m = max(m, static_cast<unsigned long long>(max(u, v)) + 1ull);
  }
  m = wave_max(m);
  if (lane_id() == 0 && m) atomicMax(extent, m);
}

void graph_stage_device(vgx_graph* g, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k,
                        hipStream_t producer, bool defer) {
  if (k >= (1ull << 31)) throw Error(VGX_INVALID_ARGUMENT, "update batch too large");
  hipStream_t s = g->stream;
  if (producer != s) VGX_HIP(hipStreamSynchronize(producer));  // null: the default stream
  BatchScratch& b0 = g->bs[0];
  b0.src.ensure(2 * k + 64);
  b0.add.ensure(k + 64);
  b0.extent.ensure(1);
  VGX_HIP(hipMemsetAsync(b0.extent.p, 0, sizeof(unsigned long long), s));
  k_batch_ingest<<<grid_for(k, kBlock, 1024), kBlock, 0, s>>>(src, dst, add, k, b0.src.p, b0.add.p, b0.extent.p);
  VGX_LAUNCH_CHECK();
  unsigned long long ext = 0;
  VGX_HIP(hipMemcpyAsync(&ext, b0.extent.p, sizeof(ext), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  if (ext > 0 && ext - 1 >= (1ull << 31) - 1) throw Error(VGX_INVALID_ARGUMENT, "vertex ids must be below 2^31 - 1");
  if (ext > 0) graph_grow_vertices(g, ext);
  stage_sorted(g, k, defer);
}

// Sort, resolve and plan a batch whose src | dst and add already sit in
// bs[0] (from either staging path).  With `defer` the host counts are left
// for graph_stage_collect after the caller's next synchronisation.
static void stage_sorted(vgx_graph* g, uint64_t k, bool defer) {
  hipStream_t s = g->stream;
  BatchScratch& b0 = g->bs[0];
  const uint32_t* d_src = b0.src.p;
  const uint32_t* d_dst = b0.src.p + k;
  reset_deltas(g);
  const int end_bit = 32 + bits_for(g->n);
  for (int d = 0; d < g->ndir(); ++d) {
    BatchScratch& b = g->bs[d];
    Adjacency& a = g->adj[d];
    const uint64_t M = g->directed ? k : 2 * k;
    b.k = k;
    b.m = M;
    b.keys.ensure(M + 1);
    b.skeys.ensure(M + 1);
    b.vals.ensure(M + 1);
    b.svals.ensure(M + 1);
    b.act.ensure(M + 1);
    b.pos.ensure(M + 1);
    b.fin.ensure(M + 1);
    b.ins.ensure(M + 1);
    b.segk.ensure(M + 1);
    b.big.ensure(M + 1);
    b.mid.ensure(M + 1);
    // two flag arrays, each 16-byte aligned and padded to whole 16-byte words
    const uint64_t fstride = (M + 16) & ~static_cast<uint64_t>(15);
    b.flags.ensure(2 * fstride + 16);
    k_expand_keys<<<grid_for(k), kBlock, 0, s>>>(d_src, d_dst, k, g->directed ? 0 : 1, d == 1 ? 1 : 0, b.keys.p,
                                                  b.vals.p);
    VGX_LAUNCH_CHECK();
    if (M <= kSortSmall) {
      const uint32_t m32 = static_cast<uint32_t>(M);
      const int vb = bits_for(g->n);
      if (M <= 64 * kSortSmallItems)
        k_sort_small<64><<<1, 64, 0, s>>>(b.keys.p, b.vals.p, m32, vb, b.skeys.p, b.svals.p);
      else if (M <= 256 * kSortSmallItems)
        k_sort_small<256><<<1, 256, 0, s>>>(b.keys.p, b.vals.p, m32, vb, b.skeys.p, b.svals.p);
      else
        k_sort_small<1024><<<1, 1024, 0, s>>>(b.keys.p, b.vals.p, m32, vb, b.skeys.p, b.svals.p);
      VGX_LAUNCH_CHECK();
    } else {
      with_temp(g, b.tmp, [&](void* t, size_t& bytes) {
        return hipcub::DeviceRadixSort::SortPairs(t, bytes, b.keys.p, b.skeys.p, b.vals.p, b.svals.p,
                                                  static_cast<int>(M), 0, end_bit, s);
      });
    }
    uint8_t* fflag = b.flags.p;
    uint8_t* iflag = b.flags.p + fstride;
    if (M <= kSelectSmall) {
      // one arc per thread: the workgroups leave their counts for the
      // order-keeping write (a zero-arc batch still writes zero totals)
      const unsigned nb = static_cast<unsigned>(std::max<uint64_t>(1, (M + kBlock - 1) / kBlock));
      b.bcount.ensure(nb + 1);
      k_resolve<<<nb, kBlock, 0, s>>>(b.skeys.p, b.svals.p, M, b0.add.p, a.row.p, a.arena.p, g->n, b.act.p,
                                       b.pos.p, fflag, iflag, b.bcount.p);
      VGX_LAUNCH_CHECK();
      k_select_write<<<nb, kBlock, 0, s>>>(fflag, iflag, M, b.bcount.p, b.fin.p, &g->d_ctr->nfinal[d], b.ins.p,
                                           &g->d_ctr->nins[d]);
      VGX_LAUNCH_CHECK();
    } else {
      k_resolve<<<grid_for(M), kBlock, 0, s>>>(b.skeys.p, b.svals.p, M, b0.add.p, a.row.p, a.arena.p, g->n,
                                               b.act.p, b.pos.p, fflag, iflag, nullptr);
      VGX_LAUNCH_CHECK();
      hipcub::CountingInputIterator<uint32_t> it(0);
      with_temp(g, b.tmp, [&](void* t, size_t& bytes) {
        return hipcub::DeviceSelect::Flagged(t, bytes, it, fflag, b.fin.p, &g->d_ctr->nfinal[d],
                                             static_cast<int>(M), s);
      });
      with_temp(g, b.tmp, [&](void* t, size_t& bytes) {
        return hipcub::DeviceSelect::Flagged(t, bytes, it, iflag, b.ins.p, &g->d_ctr->nins[d],
                                             static_cast<int>(M), s);
      });
    }
    k_plan_inserts<<<grid_for(M), kBlock, 0, s>>>(b.ins.p, &g->d_ctr->nins[d], b.skeys.p, a.row.p, a.cap.p,
                                                  b.segk.p, b.big.p, b.mid.p, g->d_ctr, d);
    VGX_LAUNCH_CHECK();
  }
  for (int d = 0; d < g->ndir(); ++d) g->bs[d].counted = false;
  if (defer) return;
  graph_sync_counters(g);
  graph_stage_collect(g);
}

void graph_stage_collect(vgx_graph* g) {
  for (int d = 0; d < g->ndir(); ++d) {
    BatchScratch& b = g->bs[d];
    b.nfinal = g->h_ctr->nfinal[d];
    b.nins = g->h_ctr->nins[d];
    b.nbig = g->h_ctr->nbig[d];
    b.nmid = g->h_ctr->nmid[d];
    b.need = g->h_ctr->need[d];
    b.counted = true;
    ensure_arena(g, d, g->h_ctr->need[d]);
  }
}

void graph_apply_deletions(vgx_graph* g) {
  for (int d = 0; d < g->ndir(); ++d) {
    BatchScratch& b = g->bs[d];
    Adjacency& a = g->adj[d];
    // before the counts reach the host, the grid covers every staged arc
    // (the kernel reads the count of final arcs on the device)
    const uint64_t items = b.counted ? b.nfinal : b.m;
    if (!items) continue;
    k_apply_flip<<<grid_for(items), kBlock, 0, g->stream>>>(b.fin.p, &g->d_ctr->nfinal[d], b.skeys.p, b.act.p,
                                                            b.pos.p, a.arena.p, a.deg.p, g->d_ctr, d, 0);
    VGX_LAUNCH_CHECK();
  }
}

void graph_apply_additions(vgx_graph* g) {
  for (int d = 0; d < g->ndir(); ++d) {
    BatchScratch& b = g->bs[d];
    Adjacency& a = g->adj[d];
    if (b.nfinal) {
      k_apply_flip<<<grid_for(b.nfinal), kBlock, 0, g->stream>>>(b.fin.p, &g->d_ctr->nfinal[d], b.skeys.p,
                                                                 b.act.p, b.pos.p, a.arena.p, a.deg.p, g->d_ctr, d,
                                                                 1);
      VGX_LAUNCH_CHECK();
    }
    if (b.nins) {
      k_insert_small<<<grid_for(b.nins), kBlock, 0, g->stream>>>(b.ins.p, &g->d_ctr->nins[d], b.segk.p,
                                                                 b.skeys.p, a.row.p, a.cap.p, a.arena.p, a.deg.p,
                                                                 g->d_ctr, d);
      VGX_LAUNCH_CHECK();
    }
    if (b.nmid) {
      k_insert_mid<<<static_cast<unsigned>((b.nmid + kMidWaves - 1) / kMidWaves), kMidWaves * kWave, 0,
                     g->stream>>>(b.mid.p, static_cast<uint32_t>(b.nmid), b.ins.p, b.segk.p, b.skeys.p, a.row.p,
                                  a.arena.p, a.deg.p);
      VGX_LAUNCH_CHECK();
    }
    if (b.nbig) {
      // plan: 8 words per relocated row, then the owner row of every output
      // chunk (at most need / chunk + one per row)
      const uint64_t chunks = b.need / kRelocChunk + b.nbig;
      b.reloc.ensure(8 * b.nbig + chunks + 1);
      uint32_t* chunk_row = b.reloc.p + 8 * b.nbig;
      k_reloc_plan<<<grid_for(b.nbig), kBlock, 0, g->stream>>>(b.big.p, static_cast<uint32_t>(b.nbig), b.ins.p,
                                                                b.segk.p, b.skeys.p, a.row.p, a.cap.p, a.deg.p,
                                                                b.reloc.p, chunk_row, g->d_ctr, d);
      VGX_LAUNCH_CHECK();
      const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(4096, std::max<uint64_t>(1, chunks)));
      k_reloc_copy<<<grid, 256, 0, g->stream>>>(b.reloc.p, chunk_row, g->d_ctr, d, b.ins.p, b.skeys.p,
                                                a.arena.p);
      VGX_LAUNCH_CHECK();
    }
  }
}

static void compact_direction(vgx_graph* g, int d);

void graph_finish_batch(vgx_graph* g) {
  graph_sync_counters(g);
  graph_finish_collect(g);
}

void graph_finish_collect(vgx_graph* g) {
  bool need_compact = false;
  for (int d = 0; d < g->ndir(); ++d) {
    Adjacency& a = g->adj[d];
    a.live += static_cast<long long>(g->h_ctr->live_delta[d]);
    a.stored += static_cast<long long>(g->h_ctr->stored_delta[d]);
    a.relocated += g->h_ctr->relocated[d];
    a.used = g->h_ctr->bump[d];
    // Storage maintenance (the role of automatic_storage_maintenance,
    // dynamic_graph.hpp:791-802): repack once tombstones or abandoned
    // relocation space outweigh half of the live arcs.
    uint64_t tomb = a.stored - a.live;
    if (tomb > a.live / 2 + 4096 || a.used > 2 * (a.live + g->n) + (1u << 20)) need_compact = true;
  }
  g->version += 1;
  g->rows_epoch += 1;
  if (need_compact)
    for (int d = 0; d < g->ndir(); ++d) compact_direction(g, d);
}

static void graph_apply(vgx_graph* g, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k,
                        bool device = false, hipStream_t producer = nullptr) {
  if (k == 0) return;
  if (device) graph_stage_device(g, src, dst, add, k, producer);
  else graph_stage(g, src, dst, add, k);
  graph_apply_deletions(g);
  graph_apply_additions(g);
  graph_finish_batch(g);
}

// ------------------------------------------------------ snapshots, compaction

// Chunked copy of the rows longer than kPackShort; the chunk table is built
// on the device, so packing needs no host round trip.  `stored` bounds the
// entries of all rows (the arena's size does).
// Chunk table and live counts of the rows longer than kPackShort.
static void pack_chunk_prep(vgx_graph* g, Adjacency& a, uint64_t stored) {
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  PackScratch& pk = g->pk;
  // rows longer than kPackShort hold more than kPackShort entries each, so
  // chunks <= stored / kPackChunk + stored / kPackShort
  const uint64_t max_chunks = stored / kPackChunk + stored / kPackShort + 2;
  pk.row_chunks.ensure(n + 1);
  pk.chunk_first.ensure(n + 1);
  pk.chunk_u.ensure(max_chunks);
  pk.chunk_j.ensure(max_chunks);
  pk.chunk_live.ensure(max_chunks + 1);
  pk.chunk_scan.ensure(max_chunks + 1);
  pk.counts.ensure(2);
  k_pack_chunk_count<<<grid_for(n + 1, kBlock, 4096), kBlock, 0, s>>>(a.row.p, n, pk.row_chunks.p);
  VGX_LAUNCH_CHECK();
  with_temp(g, g->bs[0].tmp, [&](void* t, size_t& b) {
    return hipcub::DeviceScan::ExclusiveSum(t, b, pk.row_chunks.p, pk.chunk_first.p, static_cast<int>(n + 1), s);
  });
  k_pack_chunk_fill<<<grid_for(n + 1, kBlock, 4096), kBlock, 0, s>>>(pk.chunk_first.p, n, pk.chunk_u.p,
                                                                      pk.chunk_j.p, pk.counts.p);
  VGX_LAUNCH_CHECK();
  const unsigned grid = grid_for(max_chunks, 4, 4096);
  k_pack_chunk_live<<<grid, kBlock, 0, s>>>(a.row.p, a.arena.p, a.deg.p, pk.chunk_u.p, pk.chunk_j.p, pk.counts.p,
                                            pk.chunk_live.p);
  VGX_LAUNCH_CHECK();
  with_temp(g, g->bs[0].tmp, [&](void* t, size_t& b) {
    return hipcub::DeviceScan::ExclusiveSum(t, b, pk.chunk_live.p, pk.chunk_scan.p,
                                            static_cast<int>(max_chunks + 1), s);
  });
}

static void pack_chunk_write(vgx_graph* g, Adjacency& a, const uint32_t* off, uint32_t* out, uint64_t out_cap,
                             uint64_t stored) {
  PackScratch& pk = g->pk;
  const uint64_t max_chunks = stored / kPackChunk + stored / kPackShort + 2;
  k_pack_chunk_write<<<grid_for(max_chunks, 4, 4096), kBlock, 0, g->stream>>>(
      a.row.p, a.arena.p, a.deg.p, off, pk.chunk_u.p, pk.chunk_j.p, pk.chunk_scan.p, pk.counts.p, out, out_cap);
  VGX_LAUNCH_CHECK();
}

static void pack_direction(vgx_graph* g, int d, DevBuf<uint32_t>& off, DevBuf<uint32_t>& out, uint2* new_row,
                           const uint32_t* caps) {
  Adjacency& a = g->adj[d];
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  off.ensure(n + 2);
  VGX_HIP(hipMemsetAsync(off.p, 0, sizeof(uint32_t), s));
  if (n) {
    with_temp(g, g->bs[0].tmp, [&](void* t, size_t& b) {
      return hipcub::DeviceScan::InclusiveSum(t, b, caps ? caps : a.deg.p, off.p + 1, static_cast<int>(n), s);
    });
  }
  // capacities are deg + deg/8 + 2 (k_compact_caps), plus slack for the
  // next inserts; a snapshot holds exactly the live arcs
  const uint64_t bound = (caps ? a.live + a.live / 8 + 2 * n : a.live) + 1;
  // grown with slack, so the small growth of later batches reuses it
  if (out.n < bound) out.ensure(bound + bound / 8 + (1u << 16));
  if (n) {
    k_pack_rows<<<grid_for(n, 4 * kWave, 4096), kBlock, 0, s>>>(a.row.p, a.arena.p, n, off.p, out.p, out.n,
                                                                new_row, a.deg.p);
    VGX_LAUNCH_CHECK();
    pack_chunk_prep(g, a, a.arena.n);
    pack_chunk_write(g, a, off.p, out.p, out.n, a.arena.n);
  }
}

// Consolidation: every row is repacked without tombstones into the spare
// arena with capacity deg + deg/8 + 2, and the spare and live arena (and
// row tables) trade places.  One host synchronisation reads the new bump
// pointer.
static void compact_direction(vgx_graph* g, int d) {
  Adjacency& a = g->adj[d];
  hipStream_t s = g->stream;
  const uint64_t n = g->n;
  if (n == 0) return;
  PackScratch& pk = g->pk;
  k_compact_caps<<<grid_for(n), kBlock, 0, s>>>(a.deg.p, n, a.cap.p);
  VGX_LAUNCH_CHECK();
  pk.spare_row.ensure(g->ncap);
  VGX_HIP(hipMemsetAsync(pk.spare_row.p, 0, g->ncap * sizeof(uint2), s));
  pack_direction(g, d, pk.off, pk.spare_arena, pk.spare_row.p, a.cap.p);
  k_set_bump<<<1, 64, 0, s>>>(pk.off.p, n, g->d_ctr, d);
  VGX_LAUNCH_CHECK();
  a.arena.swap(pk.spare_arena);
  a.acap = a.arena.n;
  a.row.swap(pk.spare_row);
  rebase_rows(g, d);
  g->rows_epoch += 1;
  a.stored = a.live;
  a.relocated = 0;
  uint32_t total = 0;
  VGX_HIP(hipMemcpyAsync(&total, pk.off.p + n, sizeof(total), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  if (total > a.arena.n) throw Error(VGX_DEVICE_ERROR, "compaction exceeded its arena bound");
  a.used = total;
}

// ---------------------------------------------------- per-row delta state
// A row's changes since its last bulk load or compaction, read off its
// stored length `len`, live degree `deg` and base length `base` (its stored
// length right after that bulk load or compaction): the entries inserted
// since (len - base: an insert adds a stored entry, a deletion only marks
// one, and only compaction removes entries) plus its tombstones
// (len - deg).  An arc inserted and deleted again since counts twice; a
// base arc deleted and re-added (its tombstone revived) counts nothing.
__device__ __forceinline__ unsigned long long row_delta(uint32_t len, uint32_t deg, uint32_t base) {
  return static_cast<unsigned long long>(len - deg) + static_cast<unsigned long long>(len - base);
}

// Sums over the rows of one direction: out[0] base entries, out[1] row
// deltas, out[2] rows whose delta is not zero.
__global__ __launch_bounds__(kBlock) void k_row_delta_sums(const uint2* __restrict__ row,
                                                          const uint32_t* __restrict__ deg,
                                                          const uint32_t* __restrict__ base, uint64_t n,
                                                          unsigned long long* __restrict__ out) {
  unsigned long long b = 0, dl = 0, dirty = 0;
  for (uint64_t u = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; u < n;
       u += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t bs = base[u];
    const unsigned long long x = row_delta(row[u].y, deg[u], bs);
    b += bs;
    dl += x;
    dirty += x != 0 ? 1ull : 0ull;
  }
  b = wave_sum<unsigned long long>(b);
  dl = wave_sum<unsigned long long>(dl);
  dirty = wave_sum<unsigned long long>(dirty);
  if (lane_id() == 0) {
    if (b) atomicAdd(&out[0], b);
    if (dl) atomicAdd(&out[1], dl);
    if (dirty) atomicAdd(&out[2], dirty);
  }
}

// Row compaction: every row whose delta is at least `threshold` times
// max(8, base) drops its tombstones in place.  Each lane tests one of 64
// consecutive rows; the wave then squeezes the chosen rows one after
// another, 64 entries a step: a live entry moves to its rank among the
// row's live entries, which never lies past the entry itself, so no write
// reaches a slot that a later step still reads.  The row's length and
// base become its live count.  acc[0] += rows compacted, acc[1] +=
// entries dropped.
__global__ __launch_bounds__(kBlock) void k_squeeze_rows(uint2* __restrict__ row, const uint32_t* __restrict__ deg,
                                                        uint32_t* __restrict__ base, uint32_t* __restrict__ arena,
                                                        uint64_t n, double threshold,
                                                        unsigned long long* __restrict__ acc) {
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  const uint64_t wid = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave;
  const uint32_t lane = lane_id();
  unsigned long long rows = 0, dropped = 0;
  for (uint64_t u0 = wid * kWave; u0 < n; u0 += waves * kWave) {
    const uint64_t u = u0 + lane;
    uint2 rb = make_uint2(0u, 0u);
    bool pick = false;
    if (u < n) {
      rb = row[u];
      const uint32_t bs = base[u];
      const unsigned long long x = row_delta(rb.y, deg[u], bs);
      const double density = static_cast<double>(x) / static_cast<double>(max(8u, bs));
      pick = x != 0 && !(density < threshold);
    }
    unsigned long long m = __ballot(pick);
    while (m) {
      const int l = __ffsll(static_cast<long long>(m)) - 1;
      m &= m - 1;
      const uint32_t beg = __shfl(rb.x, l, kWave), len = __shfl(rb.y, l, kWave);
      uint32_t w = 0;
      for (uint32_t i0 = 0; i0 < len; i0 += kWave) {
        const uint32_t i = i0 + lane;
        const uint32_t x = i < len ? arena[beg + i] : kTomb;
        const bool keep = !(x & kTomb);
        const unsigned long long km = __ballot(keep);
        if (keep) arena[beg + w + __popcll(km & ((1ull << lane) - 1ull))] = x;
        w += __popcll(km);
      }
      if (static_cast<int>(lane) == l) {
        row[u].y = w;
        base[u] = w;
      }
      ++rows;
      dropped += len - w;
    }
  }
  if (lane == 0 && rows) {
    atomicAdd(&acc[0], rows);
    atomicAdd(&acc[1], dropped);
  }
}

struct RowDeltaSums {
  uint64_t base = 0, delta = 0, dirty = 0;
};

static RowDeltaSums row_delta_sums(vgx_graph* g, int d) {
  RowDeltaSums r;
  if (!g->n) return r;
  hipStream_t s = g->stream;
  const Adjacency& a = g->adj[d];
  DevBuf<unsigned long long>& out = g->pk.rowsum;
  out.ensure(4);
  VGX_HIP(hipMemsetAsync(out.p, 0, 4 * sizeof(unsigned long long), s));
  k_row_delta_sums<<<grid_for(g->n, kBlock, 2048), kBlock, 0, s>>>(a.row.p, a.deg.p, a.base.p, g->n, out.p);
  VGX_LAUNCH_CHECK();
  unsigned long long h[3] = {0, 0, 0};
  VGX_HIP(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  r.base = h[0];
  r.delta = h[1];
  r.dirty = h[2];
  return r;
}

// Consolidated snapshot: the live rows of `src` packed, without tombstones
// and with compaction's capacities (deg + deg/8 + 2), into the fresh graph
// `dst` (same vertex count and directedness).  `src` is only read; all work
// runs on src's stream and ends with one synchronisation per direction.
static void snapshot_into(vgx_graph* src, vgx_graph* dst) {
  const uint64_t n = src->n;
  hipStream_t s = src->stream;
  for (int d = 0; d < src->ndir() && n; ++d) {
    Adjacency& sa = src->adj[d];
    Adjacency& da = dst->adj[d];
    k_compact_caps<<<grid_for(n), kBlock, 0, s>>>(sa.deg.p, n, da.cap.p);
    VGX_LAUNCH_CHECK();
    VGX_HIP(hipMemcpyAsync(da.deg.p, sa.deg.p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    VGX_HIP(hipMemcpyAsync(da.base.p, sa.deg.p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    pack_direction(src, d, dst->pk.off, da.arena, da.row.p, da.cap.p);
    k_set_bump<<<1, 64, 0, s>>>(dst->pk.off.p, n, dst->d_ctr, d);
    VGX_LAUNCH_CHECK();
    uint32_t total = 0;
    VGX_HIP(hipMemcpyAsync(&total, dst->pk.off.p + n, sizeof(total), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    if (total > da.arena.n) throw Error(VGX_DEVICE_ERROR, "snapshot exceeded its arena bound");
    da.acap = da.arena.n;
    da.used = total;
    da.stored = sa.live;
    da.live = sa.live;
    da.relocated = 0;
  }
  dst->version = 1;  // one bulk load, as a freshly built graph has
}

// splitmix64's output function (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014): spreads a vertex id over
// 64 bits so that the sum of the mixes of a row identifies its contents.
__device__ __forceinline__ unsigned long long probe_mix(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Neighbour probe: probe i reads row u_i = (i * 0x9E3779B97F4A7C15 mod 2^64)
// mod n, one wave per probe, and adds mix(live length) + sum of mix(v) over
// its live entries into *digest (a sum mod 2^64, so independent of order
// and of how the rows sit in the arena).
__global__ void k_probe_rows(const uint2* __restrict__ row, const uint32_t* __restrict__ arena, uint64_t n,
                             uint64_t count, unsigned long long* __restrict__ digest) {
  const uint64_t W = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
  const unsigned lane = lane_id();
  unsigned long long h = 0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x / kWave) + threadIdx.x / kWave; i < count;
       i += W) {
    const uint64_t u = (i * 0x9E3779B97F4A7C15ull) % n;
    const uint2 rb = row[u];
    unsigned long long part = 0;
    uint32_t live = 0;
    // Eight loads in flight per lane before any is used: a hub row's scan
    // is bound by load latency, not bandwidth.
    constexpr uint32_t kUnroll = 8;
    for (uint32_t j0 = 0; j0 < rb.y; j0 += kUnroll * kWave) {
      uint32_t x[kUnroll];
#pragma unroll
      for (uint32_t q = 0; q < kUnroll; ++q) {
        const uint32_t j = j0 + q * kWave + lane;
        x[q] = j < rb.y ? arena[rb.x + j] : kTomb;  // past the row: skipped like a tombstone
      }
#pragma unroll
      for (uint32_t q = 0; q < kUnroll; ++q) {
        if (!(x[q] & kTomb)) {
          part += probe_mix(x[q]);
          live += 1;
        }
      }
    }
    part = wave_sum(part);
    live = wave_sum(live);
    h += part + probe_mix(live);
  }
  if (lane == 0 && h) atomicAdd(digest, h);
}

const CsrSnapshot& graph_snapshot(vgx_graph* g, int dir) {
  int d = g->directed ? dir : 0;
  CsrSnapshot& sn = g->snap[d];
  if (sn.version == g->version && sn.n == g->n) return sn;
  pack_direction(g, d, sn.off, sn.adj, nullptr, nullptr);
  uint32_t total = 0;
  if (g->n) VGX_HIP(hipMemcpyAsync(&total, sn.off.p + g->n, sizeof(total), hipMemcpyDeviceToHost, g->stream));
  VGX_HIP(hipStreamSynchronize(g->stream));
  if (total > sn.adj.n) throw Error(VGX_DEVICE_ERROR, "snapshot exceeded its bound");
  sn.version = g->version;
  sn.n = g->n;
  sn.arcs = g->adj[d].live;
  return sn;
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

const char* vgx_last_error(void) { return vgx::t_last_error.c_str(); }

const char* vgx_version(void) { return "velographx-amd 0.1.0 (gfx950)"; }

int vgx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

vgx_status vgx_graph_create(uint64_t vertices, int directed, vgx_graph** out) {
  VGX_API_BEGIN
  if (!out) throw Error(VGX_INVALID_ARGUMENT, "out is null");
  if (vertices > (1ull << 31)) throw Error(VGX_INVALID_ARGUMENT, "vertex ids must be below 2^31");
  if (vgx_device_count() == 0) throw Error(VGX_DEVICE_ERROR, "no HIP device visible");
  auto* g = new vgx_graph();
  try {
    VGX_HIP(hipGetDevice(&g->device));
    VGX_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    g->directed = directed != 0;
    g->d_ctr = static_cast<GraphCounters*>(dev_alloc(sizeof(GraphCounters)));
    VGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->h_ctr), sizeof(GraphCounters), hipHostMallocDefault));
    VGX_HIP(hipMemsetAsync(g->d_ctr, 0, sizeof(GraphCounters), g->stream));
    std::memset(g->h_ctr, 0, sizeof(GraphCounters));
    graph_grow_vertices(g, vertices);
    for (int d = 0; d < g->ndir(); ++d) {
      g->adj[d].arena.ensure(1u << 16);
      g->adj[d].acap = g->adj[d].arena.n;
    }
    VGX_HIP(hipStreamSynchronize(g->stream));
  } catch (...) {
    vgx_graph_destroy(g);
    throw;
  }
  *out = g;
  VGX_API_END
}

vgx_status vgx_graph_destroy(vgx_graph* g) {
  VGX_API_BEGIN
  if (!g) return VGX_OK;
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  if (g->d_ctr) dev_free(g->d_ctr);
  if (g->h_ctr) (void)hipHostFree(g->h_ctr);
  hipStream_t s = g->stream;
  delete g;
  if (s) (void)hipStreamDestroy(s);
  VGX_API_END
}

vgx_status vgx_graph_bulk_load(vgx_graph* g, const uint32_t* src, const uint32_t* dst, uint64_t m) {
  VGX_API_BEGIN
  if (!g || (m && (!src || !dst))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if ((g->directed ? m : 2 * m) >= (1ull << 31)) throw Error(VGX_INVALID_ARGUMENT, "edge list too large");
  DeviceGuard dg(g->device);
  hipStream_t s = g->stream;
  DevBuf<uint32_t> buf;
  buf.ensure(2 * m + 2);
  if (m) {
    VGX_HIP(hipMemcpyAsync(buf.p, src, m * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    VGX_HIP(hipMemcpyAsync(buf.p + m, dst, m * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  }
  bulk_load_host(g, src, dst, buf.p, buf.p + m, m, 1);
  VGX_API_END
}

vgx_status vgx_graph_bulk_load_pairs(vgx_graph* g, const uint32_t* pairs, uint64_t m) {
  VGX_API_BEGIN
  if (!g || (m && !pairs)) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if ((g->directed ? m : 2 * m) >= (1ull << 31)) throw Error(VGX_INVALID_ARGUMENT, "edge list too large");
  DeviceGuard dg(g->device);
  hipStream_t s = g->stream;
  DevBuf<uint32_t> buf;
  buf.ensure(2 * m + 2);
  if (m) VGX_HIP(hipMemcpyAsync(buf.p, pairs, m * 2 * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  bulk_load_host(g, pairs, pairs + 1, buf.p, buf.p + 1, m, 2);
  VGX_API_END
}

vgx_status vgx_graph_bulk_load_device(vgx_graph* g, const void* src_dev, const void* dst_dev, uint64_t m,
                                      uint64_t stride, uint32_t id_bytes, void* stream) {
  VGX_API_BEGIN
  if (!g || (m && (!src_dev || !dst_dev))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (id_bytes != 4 && id_bytes != 8) throw Error(VGX_INVALID_ARGUMENT, "ids must be 4 or 8 bytes wide");
  if (stride == 0) throw Error(VGX_INVALID_ARGUMENT, "stride must be positive");
  if ((g && !g->directed ? 2 * m : m) >= (1ull << 31)) throw Error(VGX_INVALID_ARGUMENT, "edge list too large");
  DeviceGuard dg(g->device);
  hipStream_t ps = static_cast<hipStream_t>(stream);
  if (id_bytes == 4)
    bulk_load_device_t(g, static_cast<const uint32_t*>(src_dev), static_cast<const uint32_t*>(dst_dev), m, stride, ps);
  else
    bulk_load_device_t(g, static_cast<const int64_t*>(src_dev), static_cast<const int64_t*>(dst_dev), m, stride, ps);
  VGX_API_END
}

vgx_status vgx_graph_bulk_load_csr_device(vgx_graph* g, const void* indptr_dev, const void* indices_dev,
                                          uint64_t vertices, uint64_t nnz, uint32_t index_bytes, void* stream) {
  VGX_API_BEGIN
  if (!g || !indptr_dev || (nnz && !indices_dev)) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (index_bytes != 4 && index_bytes != 8) throw Error(VGX_INVALID_ARGUMENT, "indices must be 4 or 8 bytes wide");
  if (vertices > (1ull << 31)) throw Error(VGX_INVALID_ARGUMENT, "vertex ids must be below 2^31");
  if ((!g->directed ? 2 * nnz : nnz) >= (1ull << 31)) throw Error(VGX_INVALID_ARGUMENT, "CSR too large");
  DeviceGuard dg(g->device);
  hipStream_t ps = static_cast<hipStream_t>(stream);
  if (index_bytes == 4)
    bulk_load_csr_device_t(g, static_cast<const int32_t*>(indptr_dev), static_cast<const int32_t*>(indices_dev),
                           vertices, nnz, ps);
  else
    bulk_load_csr_device_t(g, static_cast<const int64_t*>(indptr_dev), static_cast<const int64_t*>(indices_dev),
                           vertices, nnz, ps);
  VGX_API_END
}

vgx_status vgx_graph_apply(vgx_graph* g, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k) {
  VGX_API_BEGIN
  if (!g || (k && (!src || !dst || !add))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  DeviceGuard dg(g->device);
  graph_apply(g, src, dst, add, k);
  VGX_API_END
}

vgx_status vgx_graph_apply_device(vgx_graph* g, const uint32_t* src_dev, const uint32_t* dst_dev,
                                  const uint8_t* add_dev, uint64_t k, void* stream) {
  VGX_API_BEGIN
  if (!g || (k && (!src_dev || !dst_dev || !add_dev))) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  DeviceGuard dg(g->device);
  graph_apply(g, src_dev, dst_dev, add_dev, k, true, static_cast<hipStream_t>(stream));
  VGX_API_END
}

vgx_status vgx_graph_device(const vgx_graph* g, int* device) {
  VGX_API_BEGIN
  if (!g || !device) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  *device = g->device;
  VGX_API_END
}

vgx_status vgx_graph_info(const vgx_graph* g, uint64_t* vertex_count, uint64_t* version, int* directed,
                          uint64_t* edge_count_directed) {
  VGX_API_BEGIN
  if (!g) throw Error(VGX_INVALID_ARGUMENT, "null graph");
  if (vertex_count) *vertex_count = g->n;
  if (version) *version = g->version;
  if (directed) *directed = g->directed ? 1 : 0;
  if (edge_count_directed) *edge_count_directed = g->adj[0].live;
  VGX_API_END
}

static std::vector<uint32_t> host_row(vgx_graph* g, int d, uint32_t u) {
  Adjacency& a = g->adj[d];
  uint2 rb;
  VGX_HIP(hipMemcpyAsync(&rb, a.row.p + u, sizeof(rb), hipMemcpyDeviceToHost, g->stream));
  VGX_HIP(hipStreamSynchronize(g->stream));
  std::vector<uint32_t> v(rb.y);
  if (rb.y) {
    VGX_HIP(hipMemcpyAsync(v.data(), a.arena.p + rb.x, rb.y * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
    VGX_HIP(hipStreamSynchronize(g->stream));
  }
  return v;
}

vgx_status vgx_graph_neighbors(vgx_graph* g, uint32_t u, int in_direction, uint32_t* out, uint64_t cap,
                               uint64_t* count) {
  VGX_API_BEGIN
  if (!g || !count) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  DeviceGuard dg(g->device);
  *count = 0;
  if (u >= g->n) return VGX_OK;  // rows beyond the graph are empty (dynamic_graph.hpp:180-181)
  int d = (in_direction && g->directed) ? 1 : 0;
  auto v = host_row(g, d, u);
  uint64_t c = 0;
  for (uint32_t x : v) {
    if (x & kTomb) continue;
    if (out && c < cap) out[c] = x;
    ++c;
  }
  *count = c;
  VGX_API_END
}

vgx_status vgx_graph_has_edge(vgx_graph* g, uint32_t u, uint32_t v, int* present) {
  VGX_API_BEGIN
  if (!g || !present) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  DeviceGuard dg(g->device);
  *present = 0;
  if (u >= g->n || v > kIdMask) return VGX_OK;
  auto r = host_row(g, 0, u);
  auto it = std::lower_bound(r.begin(), r.end(), v, [](uint32_t a, uint32_t b) { return (a & kIdMask) < b; });
  *present = (it != r.end() && (*it & kIdMask) == v && !(*it & kTomb)) ? 1 : 0;
  VGX_API_END
}

vgx_status vgx_graph_degrees(vgx_graph* g, int in_direction, uint32_t* out) {
  VGX_API_BEGIN
  if (!g || (!out && g->n)) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  DeviceGuard dg(g->device);
  int d = (in_direction && g->directed) ? 1 : 0;
  if (g->n) {
    VGX_HIP(hipMemcpyAsync(out, g->adj[d].deg.p, g->n * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
    VGX_HIP(hipStreamSynchronize(g->stream));
  }
  VGX_API_END
}

vgx_status vgx_graph_compact(vgx_graph* g) {
  VGX_API_BEGIN
  if (!g) throw Error(VGX_INVALID_ARGUMENT, "null graph");
  DeviceGuard dg(g->device);
  for (int d = 0; d < g->ndir(); ++d) compact_direction(g, d);
  VGX_API_END
}

vgx_status vgx_graph_storage_info(vgx_graph* g, vgx_graph_storage* out) {
  VGX_API_BEGIN
  if (!g || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  std::memset(out, 0, sizeof(*out));
  for (int d = 0; d < g->ndir(); ++d) {
    const Adjacency& a = g->adj[d];
    out->arena_entries += a.acap;
    out->arena_used += a.used;
    out->tombstones += a.stored - a.live;
    out->relocated_rows += a.relocated;
    out->storage_bytes += a.acap * 4 + g->ncap * (sizeof(uint2) + 8);
  }
  VGX_API_END
}

vgx_status vgx_graph_ensure_vertex(vgx_graph* g, uint32_t v) {
  VGX_API_BEGIN
  if (!g) throw Error(VGX_INVALID_ARGUMENT, "null graph");
  if (v >= (1u << 31) - 1) throw Error(VGX_INVALID_ARGUMENT, "vertex ids must be below 2^31 - 1");
  DeviceGuard dg(g->device);
  if (static_cast<uint64_t>(v) + 1 > g->n) {
    graph_grow_vertices(g, static_cast<uint64_t>(v) + 1);
    g->rows_epoch += 1;
    VGX_HIP(hipStreamSynchronize(g->stream));
  }
  VGX_API_END
}

vgx_status vgx_graph_delta_info(vgx_graph* g, vgx_graph_delta* out) {
  VGX_API_BEGIN
  if (!g || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  DeviceGuard dg(g->device);
  const RowDeltaSums o = row_delta_sums(g, 0);
  const RowDeltaSums i = g->directed ? row_delta_sums(g, 1) : o;
  out->base_arcs = o.base;
  out->delta_arcs = o.delta;
  out->delta_in_arcs = i.delta;
  out->dirty_out_rows = o.dirty;
  out->dirty_in_rows = i.dirty;
  VGX_API_END
}

vgx_status vgx_graph_maybe_compact(vgx_graph* g, double threshold, int* compacted) {
  VGX_API_BEGIN
  if (!g) throw Error(VGX_INVALID_ARGUMENT, "null graph");
  DeviceGuard dg(g->device);
  hipStream_t s = g->stream;
  bool any = false;
  for (int d = 0; d < g->ndir() && g->n; ++d) {
    Adjacency& a = g->adj[d];
    DevBuf<unsigned long long>& acc = g->pk.rowsum;
    acc.ensure(4);
    VGX_HIP(hipMemsetAsync(acc.p, 0, 4 * sizeof(unsigned long long), s));
    const uint64_t groups = (g->n + kWave - 1) / kWave;  // one 64-row group per wave step
    const unsigned blocks = static_cast<unsigned>(std::min<uint64_t>(1024, (groups + 3) / 4));
    k_squeeze_rows<<<blocks, kBlock, 0, s>>>(a.row.p, a.deg.p, a.base.p, a.arena.p, g->n, threshold, acc.p);
    VGX_LAUNCH_CHECK();
    unsigned long long h[2] = {0, 0};
    VGX_HIP(hipMemcpyAsync(h, acc.p, sizeof(h), hipMemcpyDeviceToHost, s));
    VGX_HIP(hipStreamSynchronize(s));
    if (h[1] > a.stored - a.live) throw Error(VGX_DEVICE_ERROR, "row compaction dropped live entries");
    a.stored -= h[1];
    any = any || h[0] != 0;
  }
  if (any) g->rows_epoch += 1;
  if (compacted) *compacted = any ? 1 : 0;
  VGX_API_END
}

vgx_status vgx_graph_export_csr(vgx_graph* g, int in_direction, uint64_t* off, uint32_t* adj) {
  VGX_API_BEGIN
  if (!g || !off) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  DeviceGuard dg(g->device);
  const CsrSnapshot& sn = graph_snapshot(g, in_direction ? 1 : 0);
  std::vector<uint32_t> o(sn.n + 1);
  VGX_HIP(hipMemcpyAsync(o.data(), sn.off.p, (sn.n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
  VGX_HIP(hipStreamSynchronize(g->stream));
  for (uint64_t i = 0; i <= sn.n; ++i) off[i] = o[i];
  if (o[sn.n] && adj) {
    VGX_HIP(hipMemcpyAsync(adj, sn.adj.p, static_cast<size_t>(o[sn.n]) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           g->stream));
    VGX_HIP(hipStreamSynchronize(g->stream));
  }
  VGX_API_END
}

vgx_status vgx_graph_csr_device(vgx_graph* g, int in_direction, const uint32_t** off_dev, const uint32_t** adj_dev,
                                uint64_t* vertices, uint64_t* arcs) {
  VGX_API_BEGIN
  if (!g) throw Error(VGX_INVALID_ARGUMENT, "null graph");
  DeviceGuard dg(g->device);
  const CsrSnapshot& sn = graph_snapshot(g, in_direction ? 1 : 0);
  if (off_dev) *off_dev = sn.off.p;
  if (adj_dev) *adj_dev = sn.adj.p;
  if (vertices) *vertices = sn.n;
  if (arcs) *arcs = sn.arcs;
  VGX_API_END
}

vgx_status vgx_graph_snapshot(vgx_graph* src, vgx_graph** out) {
  VGX_API_BEGIN
  if (!src || !out) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  DeviceGuard dg(src->device);
  vgx_graph* dst = nullptr;
  const vgx_status st = vgx_graph_create(src->n, src->directed ? 1 : 0, &dst);
  if (st != VGX_OK) throw Error(st, std::string(vgx_last_error()));
  try {
    snapshot_into(src, dst);
  } catch (...) {
    vgx_graph_destroy(dst);
    throw;
  }
  *out = dst;
  VGX_API_END
}

vgx_status vgx_graph_probe_neighbors(vgx_graph* g, uint64_t count, uint32_t repeats, double* ns_per_probe,
                                     uint64_t* digest) {
  VGX_API_BEGIN
  if (!g || !ns_per_probe || !digest) throw Error(VGX_INVALID_ARGUMENT, "null argument");
  if (count == 0 || repeats == 0) throw Error(VGX_INVALID_ARGUMENT, "count and repeats must be positive");
  *ns_per_probe = 0.0;
  *digest = 0;
  if (g->n == 0) return VGX_OK;
  DeviceGuard dg(g->device);
  hipStream_t s = g->stream;
  DevBuf<unsigned long long> sums;
  sums.ensure(repeats);
  VGX_HIP(hipMemsetAsync(sums.p, 0, repeats * sizeof(unsigned long long), s));
  struct Events {
    std::vector<hipEvent_t> e;
    ~Events() {
      for (hipEvent_t x : e) (void)hipEventDestroy(x);
    }
  } ev;
  ev.e.resize(2 * static_cast<size_t>(repeats), nullptr);
  for (hipEvent_t& x : ev.e) VGX_HIP(hipEventCreate(&x));
  const unsigned grid = static_cast<unsigned>(std::min<uint64_t>((count + 3) / 4, 1024));
  const Adjacency& a = g->adj[0];
  for (uint32_t r = 0; r < repeats; ++r) {
    VGX_HIP(hipEventRecord(ev.e[2 * r], s));
    k_probe_rows<<<grid, kBlock, 0, s>>>(a.row.p, a.arena.p, g->n, count, sums.p + r);
    VGX_LAUNCH_CHECK();
    VGX_HIP(hipEventRecord(ev.e[2 * r + 1], s));
  }
  std::vector<unsigned long long> h(repeats);
  VGX_HIP(hipMemcpyAsync(h.data(), sums.p, repeats * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  VGX_HIP(hipStreamSynchronize(s));
  std::vector<double> ns(repeats);
  for (uint32_t r = 0; r < repeats; ++r) {
    float ms = 0.f;
    VGX_HIP(hipEventElapsedTime(&ms, ev.e[2 * r], ev.e[2 * r + 1]));
    ns[r] = 1e6 * ms / static_cast<double>(count);
    if (h[r] != h[0]) throw Error(VGX_DEVICE_ERROR, "probe digest changed between repeats");
  }
  std::sort(ns.begin(), ns.end());
  *ns_per_probe = ns[ns.size() / 2];
  *digest = h[0];
  VGX_API_END
}

}  // extern "C"
