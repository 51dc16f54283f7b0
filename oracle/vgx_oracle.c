/*
 * CPU oracle for the traversal hot path -- TEST INFRASTRUCTURE ONLY.
 * See vgx_oracle.h for what each function restates and how it is pinned.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this code, and only as the checker / CPU baseline.
 */
#include "vgx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define INF ORC_UNREACHABLE

/* ---------------------------------------------------------------- helpers */

static int row_contains(const uint64_t* off, const uint32_t* adj, uint32_t n,
                        uint32_t u, uint32_t v) {
  /* SegmentedCsr::contains (dynamic_graph.hpp:190-193) / has_edge (:583-587):
   * binary search in the sorted row; u outside the graph has no edges. */
  if (u >= n) return 0;
  uint64_t lo = off[u], hi = off[u + 1];
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (adj[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo < off[u + 1] && adj[lo] == v;
}

/* Open-addressing set of 64-bit keys (the role of ReusableKeySet,
 * incremental/bfs.hpp:123-176; the slot hash is a splitmix64 finalizer). */
typedef struct {
  uint64_t* keys;
  uint8_t* used;
  uint64_t mask;
} keyset;

static uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

static void keyset_init(keyset* s, uint64_t expected) {
  uint64_t cap = 8;
  while (cap < expected * 2 + 1) cap <<= 1;
  s->keys = (uint64_t*)calloc(cap, sizeof(uint64_t));
  s->used = (uint8_t*)calloc(cap, 1);
  s->mask = cap - 1;
}

static void keyset_free(keyset* s) {
  free(s->keys);
  free(s->used);
}

static int keyset_insert(keyset* s, uint64_t key) {
  uint64_t slot = mix64(key) & s->mask;
  while (s->used[slot]) {
    if (s->keys[slot] == key) return 0;
    slot = (slot + 1) & s->mask;
  }
  s->used[slot] = 1;
  s->keys[slot] = key;
  return 1;
}

static int keyset_contains(const keyset* s, uint64_t key) {
  uint64_t slot = mix64(key) & s->mask;
  while (s->used[slot]) {
    if (s->keys[slot] == key) return 1;
    slot = (slot + 1) & s->mask;
  }
  return 0;
}

/* growable vertex list */
typedef struct {
  uint32_t* v;
  uint64_t size, cap;
} vlist;

static void vl_push(vlist* l, uint32_t x) {
  if (l->size == l->cap) {
    l->cap = l->cap ? l->cap * 2 : 64;
    l->v = (uint32_t*)realloc(l->v, l->cap * sizeof(uint32_t));
  }
  l->v[l->size++] = x;
}

/* ------------------------------------------------------------------ BFS */

uint64_t orc_bfs(uint32_t n, const uint64_t* off, const uint32_t* adj,
                 uint32_t source, uint32_t* dist) {
  /* BasicIncrementalBFS::recompute (bfs.hpp:100-120): FIFO queue BFS. */
  for (uint32_t i = 0; i < n; ++i) dist[i] = INF;
  if (source >= n) return 0;
  uint32_t* q = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  uint64_t head = 0, tail = 0;
  dist[source] = 0;
  q[tail++] = source;
  while (head < tail) {
    uint32_t u = q[head++];
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      uint32_t v = adj[e];
      if (dist[v] == INF) {
        dist[v] = dist[u] + 1;
        q[tail++] = v;
      }
    }
  }
  free(q);
  return tail;
}

/* ------------------------------------------------------ incremental BFS */

typedef struct {
  uint32_t ndist; /* dist_.size() */
  uint32_t* dist;
  uint32_t source;
  uint8_t* affected;
  uint32_t* lost;
  uint32_t* spc;
  uint64_t reachable;
  uint64_t last_affected;
  vlist affected_list, invalidate, touched;
  /* pre-batch graph */
  const uint64_t *pre_out_off, *pre_in_off;
  const uint32_t *pre_out_adj, *pre_in_adj;
} ibfs;

static int is_sp(const ibfs* s, uint32_t u, uint32_t v) {
  /* is_shortest_parent (bfs.hpp:206-209) */
  return u < s->ndist && v < s->ndist && s->dist[u] != INF && s->dist[v] != INF &&
         s->dist[u] + 1 == s->dist[v];
}

static uint32_t shortest_parent_count(ibfs* s, uint32_t v) {
  /* bfs.hpp:211-220, cached per batch, in-neighbours of the pre-batch graph */
  if (v >= s->ndist || v == s->source || s->dist[v] == INF) return 0;
  if (s->spc[v] != 0) return s->spc[v];
  uint32_t count = 0;
  for (uint64_t e = s->pre_in_off[v]; e < s->pre_in_off[v + 1]; ++e) {
    if (is_sp(s, s->pre_in_adj[e], v)) ++count;
  }
  s->spc[v] = count;
  return count;
}

static void record_parent_loss(ibfs* s, uint32_t v) {
  /* bfs.hpp:229-240 */
  if (v >= s->ndist || v == s->source || s->dist[v] == INF || s->affected[v]) return;
  if (s->lost[v] == 0) vl_push(&s->touched, v);
  ++s->lost[v];
  uint32_t support = shortest_parent_count(s, v);
  if (support != 0 && s->lost[v] >= support) {
    s->affected[v] = 1;
    vl_push(&s->affected_list, v);
    vl_push(&s->invalidate, v);
    ++s->last_affected;
  }
}

static uint64_t edge_key(int directed, uint32_t u, uint32_t v) {
  if (!directed && v < u) {
    uint32_t t = u;
    u = v;
    v = t;
  }
  return ((uint64_t)u << 32) | (uint64_t)v;
}

static void propagate_decreases(uint32_t* dist, uint64_t* reachable, const uint64_t* off,
                                const uint32_t* adj, vlist* q) {
  /* bfs.hpp:341-354 */
  uint64_t head = 0;
  while (head < q->size) {
    uint32_t u = q->v[head++];
    if (dist[u] == INF) continue;
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      uint32_t v = adj[e];
      if (dist[u] + 1 < dist[v]) {
        if (dist[v] == INF) ++*reachable;
        dist[v] = dist[u] + 1;
        vl_push(q, v);
      }
    }
  }
}

typedef struct {
  uint32_t d, v;
} hitem;

static int hless(hitem a, hitem b) { return a.d < b.d || (a.d == b.d && a.v < b.v); }

static void heap_push(hitem* h, uint64_t* size, hitem x) {
  uint64_t i = (*size)++;
  h[i] = x;
  while (i > 0) {
    uint64_t p = (i - 1) / 2;
    if (!hless(h[i], h[p])) break;
    hitem t = h[i];
    h[i] = h[p];
    h[p] = t;
    i = p;
  }
}

static hitem heap_pop(hitem* h, uint64_t* size) {
  hitem top = h[0];
  h[0] = h[--*size];
  uint64_t i = 0;
  for (;;) {
    uint64_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < *size && hless(h[l], h[m])) m = l;
    if (r < *size && hless(h[r], h[m])) m = r;
    if (m == i) break;
    hitem t = h[i];
    h[i] = h[m];
    h[m] = t;
    i = m;
  }
  return top;
}

int orc_incbfs_apply(uint32_t n_pre, const uint64_t* pre_out_off, const uint32_t* pre_out_adj,
                     const uint64_t* pre_in_off, const uint32_t* pre_in_adj,
                     uint32_t n_post, const uint64_t* post_out_off, const uint32_t* post_out_adj,
                     const uint64_t* post_in_off, const uint32_t* post_in_adj,
                     int directed, uint32_t source, double fallback_fraction,
                     uint64_t k, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                     const uint32_t* dist_pre, uint64_t reachable_pre,
                     uint32_t* dist_post, orc_incbfs_diag* diag) {
  memset(diag, 0, sizeof(*diag));
  /* bfs.hpp:29-33: an empty batch changes nothing. */
  if (k == 0) {
    for (uint32_t i = 0; i < n_pre; ++i) dist_post[i] = dist_pre[i];
    diag->reachable_count = reachable_pre;
    return 0;
  }
  uint32_t nmax = n_post > n_pre ? n_post : n_pre;
  ibfs s;
  memset(&s, 0, sizeof(s));
  s.ndist = n_pre;
  s.dist = (uint32_t*)malloc((size_t)(nmax + 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < nmax; ++i) s.dist[i] = i < n_pre ? dist_pre[i] : INF;
  s.source = source;
  s.affected = (uint8_t*)calloc(nmax + 1, 1);
  s.lost = (uint32_t*)calloc(nmax + 1, sizeof(uint32_t));
  s.spc = (uint32_t*)calloc(nmax + 1, sizeof(uint32_t));
  s.reachable = reachable_pre;
  s.pre_out_off = pre_out_off;
  s.pre_out_adj = pre_out_adj;
  s.pre_in_off = pre_in_off;
  s.pre_in_adj = pre_in_adj;

  /* bfs.hpp:38-50: final state per edge key, scanning the batch backwards. */
  keyset seen, delkeys;
  keyset_init(&seen, k);
  keyset_init(&delkeys, k);
  uint32_t* add_u = (uint32_t*)malloc(k * sizeof(uint32_t));
  uint32_t* add_v = (uint32_t*)malloc(k * sizeof(uint32_t));
  uint32_t* del_u = (uint32_t*)malloc(k * sizeof(uint32_t));
  uint32_t* del_v = (uint32_t*)malloc(k * sizeof(uint32_t));
  uint64_t nadd = 0, ndel = 0;
  for (uint64_t i = k; i-- > 0;) {
    uint64_t key = edge_key(directed, src[i], dst[i]);
    if (!keyset_insert(&seen, key)) continue;
    if (add[i]) {
      add_u[nadd] = src[i];
      add_v[nadd++] = dst[i];
    } else {
      del_u[ndel] = src[i];
      del_v[ndel++] = dst[i];
      keyset_insert(&delkeys, key);
    }
  }

  /* bfs.hpp:52-63: existing deletions and unique deletion candidates. */
  uint32_t* ex_u = (uint32_t*)malloc((ndel + 1) * sizeof(uint32_t));
  uint32_t* ex_v = (uint32_t*)malloc((ndel + 1) * sizeof(uint32_t));
  uint64_t nex = 0;
  vlist cand = {0};
  for (uint64_t i = 0; i < ndel; ++i) {
    uint32_t u = del_u[i], v = del_v[i];
    if (!row_contains(pre_out_off, pre_out_adj, n_pre, u, v)) continue;
    ex_u[nex] = u;
    ex_v[nex++] = v;
    if (is_sp(&s, u, v)) vl_push(&cand, v);
    if (!directed && is_sp(&s, v, u)) vl_push(&cand, u);
  }
  {
    /* sort + unique */
    uint8_t* mark = (uint8_t*)calloc(nmax + 1, 1);
    for (uint64_t i = 0; i < cand.size; ++i) {
      if (!mark[cand.v[i]]) {
        mark[cand.v[i]] = 1;
        ++diag->deletion_candidates;
      }
    }
    free(mark);
  }

  /* bfs.hpp:222-259: support-count invalidation on the pre-batch graph. */
  int fallback_needed = 0;
  if (cand.size != 0) {
    double lim = (double)n_pre * fallback_fraction;
    uint64_t limit = (uint64_t)lim;
    if (limit < 1) limit = 1;
    for (uint64_t i = 0; i < nex; ++i) {
      uint32_t u = ex_u[i], v = ex_v[i];
      if (is_sp(&s, u, v)) record_parent_loss(&s, v);
      if (!directed && is_sp(&s, v, u)) record_parent_loss(&s, u);
    }
    uint64_t head = 0;
    int aborted = 0;
    while (head < s.invalidate.size) {
      if (s.last_affected > limit) {
        aborted = 1;
        break;
      }
      uint32_t u = s.invalidate.v[head++];
      if (u >= s.ndist || s.dist[u] == INF) continue;
      for (uint64_t e = pre_out_off[u]; e < pre_out_off[u + 1]; ++e) {
        uint32_t v = pre_out_adj[e];
        if (v >= s.ndist || s.dist[v] != s.dist[u] + 1) continue;
        if (keyset_contains(&delkeys, edge_key(directed, u, v))) continue;
        record_parent_loss(&s, v);
      }
    }
    fallback_needed = aborted || s.last_affected > limit;
  }
  diag->affected_vertices = s.last_affected;

  /* bfs.hpp:72-81: apply, resize, fallback to a full recompute. */
  s.ndist = n_post > s.ndist ? n_post : s.ndist;
  if (fallback_needed) {
    diag->reachable_count = orc_bfs(n_post, post_out_off, post_out_adj, source, dist_post);
    diag->used_full_recompute = 1;
  } else {
    if (s.affected_list.size != 0) {
      /* repair_affected, bfs.hpp:271-319 */
      uint64_t na = s.affected_list.size;
      uint32_t* old = (uint32_t*)malloc(na * sizeof(uint32_t));
      for (uint64_t i = 0; i < na; ++i) old[i] = s.dist[s.affected_list.v[i]];
      for (uint64_t i = 0; i < na; ++i) {
        uint32_t v = s.affected_list.v[i];
        if (s.dist[v] != INF) {
          s.dist[v] = INF;
          --s.reachable;
        }
      }
      uint64_t hcap = na + 16, hsize = 0;
      hitem* heap = (hitem*)malloc(hcap * sizeof(hitem));
      for (uint64_t i = 0; i < na; ++i) {
        uint32_t v = s.affected_list.v[i];
        uint32_t best = INF;
        for (uint64_t e = post_in_off[v]; e < post_in_off[v + 1]; ++e) {
          uint32_t p = post_in_adj[e];
          if (p >= s.ndist || s.affected[p] || s.dist[p] == INF) continue;
          if (s.dist[p] + 1 < best) best = s.dist[p] + 1;
        }
        if (best != INF) {
          if (s.dist[v] == INF) ++s.reachable;
          s.dist[v] = best;
          if (hsize == hcap) {
            hcap *= 2;
            heap = (hitem*)realloc(heap, hcap * sizeof(hitem));
          }
          heap_push(heap, &hsize, (hitem){best, v});
        }
      }
      while (hsize) {
        hitem top = heap_pop(heap, &hsize);
        uint32_t u = top.v;
        if (u >= s.ndist || top.d != s.dist[u]) continue;
        for (uint64_t e = post_out_off[u]; e < post_out_off[u + 1]; ++e) {
          uint32_t v = post_out_adj[e];
          if (!s.affected[v]) continue;
          uint32_t c = top.d + 1;
          if (c < s.dist[v]) {
            if (s.dist[v] == INF) ++s.reachable;
            s.dist[v] = c;
            if (hsize == hcap) {
              hcap *= 2;
              heap = (hitem*)realloc(heap, hcap * sizeof(hitem));
            }
            heap_push(heap, &hsize, (hitem){c, v});
          }
        }
      }
      free(heap);
      vlist q = {0};
      for (uint64_t i = 0; i < na; ++i) {
        uint32_t v = s.affected_list.v[i];
        if (s.dist[v] != INF && s.dist[v] < old[i]) vl_push(&q, v);
      }
      propagate_decreases(s.dist, &s.reachable, post_out_off, post_out_adj, &q);
      free(q.v);
      free(old);
    }
    /* bfs.hpp:91-96: relax the final insertions (relax_edge :332-339). */
    vlist q = {0};
    for (uint64_t i = 0; i < nadd; ++i) {
      for (int dir = 0; dir < (directed ? 1 : 2); ++dir) {
        uint32_t u = dir ? add_v[i] : add_u[i];
        uint32_t v = dir ? add_u[i] : add_v[i];
        if (u >= s.ndist || v >= s.ndist || s.dist[u] == INF) continue;
        if (s.dist[u] + 1 < s.dist[v]) {
          if (s.dist[v] == INF) ++s.reachable;
          s.dist[v] = s.dist[u] + 1;
          vl_push(&q, v);
        }
      }
    }
    propagate_decreases(s.dist, &s.reachable, post_out_off, post_out_adj, &q);
    free(q.v);
    for (uint32_t i = 0; i < n_post; ++i) dist_post[i] = s.dist[i];
    diag->reachable_count = s.reachable;
  }

  keyset_free(&seen);
  keyset_free(&delkeys);
  free(add_u);
  free(add_v);
  free(del_u);
  free(del_v);
  free(ex_u);
  free(ex_v);
  free(cand.v);
  free(s.affected_list.v);
  free(s.invalidate.v);
  free(s.touched.v);
  free(s.dist);
  free(s.affected);
  free(s.lost);
  free(s.spc);
  return 0;
}

/* ------------------------------------------------------------- PageRank */

int orc_pagerank_full(uint32_t n, const uint64_t* in_off, const uint32_t* in_adj,
                      const uint32_t* outdeg, double damping, uint64_t max_iterations,
                      double tol, double* values, orc_pr_info* info) {
  /* full_solve, pagerank.hpp:277-321 */
  memset(info, 0, sizeof(*info));
  if (n == 0) {
    info->converged = 1;
    return 0;
  }
  double* next = (double*)malloc((size_t)n * sizeof(double));
  for (uint32_t i = 0; i < n; ++i) values[i] = 1.0 / (double)n;
  const double teleport = (1.0 - damping) / (double)n;
  for (uint64_t it = 0; it < max_iterations; ++it) {
    double dangling_mass = 0.0;
    for (uint32_t u = 0; u < n; ++u)
      if (outdeg[u] == 0) dangling_mass += values[u];
    const double dangling_share = damping * dangling_mass / (double)n;
    /* each next[v] is one sequential sum over v's in-row, so splitting the
     * vertices over threads gives the same bits as one thread */
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint32_t v = 0; v < n; ++v) {
      double incoming = 0.0;
      for (uint64_t e = in_off[v]; e < in_off[v + 1]; ++e) {
        uint32_t u = in_adj[e];
        if (outdeg[u] != 0) incoming += values[u] / (double)outdeg[u];
      }
      next[v] = teleport + dangling_share;
      next[v] += damping * incoming;
    }
    info->residual_l1 = 0.0;
    info->residual_linf = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      double delta = fabs(next[i] - values[i]);
      info->residual_l1 += delta;
      if (delta > info->residual_linf) info->residual_linf = delta;
    }
    memcpy(values, next, (size_t)n * sizeof(double));
    info->iterations = it + 1;
    if (info->residual_linf <= tol) {
      info->converged = 1;
      break;
    }
  }
  free(next);
  return 0;
}

static uint64_t count_dangling(uint32_t n, const uint32_t* outdeg) {
  uint64_t c = 0;
  for (uint32_t u = 0; u < n; ++u) c += outdeg[u] == 0;
  return c;
}

static void pr_recompute(uint32_t n, const uint64_t* in_off, const uint32_t* in_adj,
                         const uint32_t* outdeg, double damping, double* rank,
                         orc_pr_apply_diag* diag) {
  /* recompute(), pagerank.hpp:193-202, default 200 iterations, 1e-12 */
  orc_pr_info info;
  orc_pagerank_full(n, in_off, in_adj, outdeg, damping, 200, 1e-12, rank, &info);
  diag->dangling_vertices = count_dangling(n, outdeg);
  diag->last_repaired_vertices = n;
  diag->last_repair_iterations = info.iterations;
  diag->last_residual_l1 = info.residual_l1;
  diag->last_residual_linf = info.residual_linf;
  diag->last_full_recompute_converged = info.converged;
  diag->used_full_recompute = 1;
}

int orc_pagerank_apply(uint32_t n_prev, uint64_t dangling_before,
                       uint32_t n, const uint64_t* out_off, const uint32_t* out_adj,
                       const uint64_t* in_off, const uint32_t* in_adj, int directed,
                       double damping, uint64_t k, const uint32_t* src, const uint32_t* dst,
                       uint64_t local_iterations, double tol, double full_fallback_fraction,
                       double* rank, orc_pr_apply_diag* diag) {
  /* pagerank.hpp:47-191; the caller has already applied the batch.  diag is
   * in/out: it carries the object's previous diagnostics, and fields the
   * reference leaves untouched on a path stay as they were. */
  diag->used_full_recompute = 0;
  diag->dangling_vertices = dangling_before;
  if (k == 0) {
    diag->last_repaired_vertices = 0;
    diag->last_repair_iterations = 0;
    diag->last_residual_l1 = 0.0;
    diag->last_residual_linf = 0.0;
    return 0;
  }
  uint32_t* outdeg = (uint32_t*)malloc(((size_t)n + 1) * sizeof(uint32_t));
  for (uint32_t u = 0; u < n; ++u) outdeg[u] = (uint32_t)(out_off[u + 1] - out_off[u]);
  if (n == 0) {
    diag->last_repaired_vertices = 0;
    diag->last_repair_iterations = 0;
    diag->last_residual_l1 = 0.0;
    diag->last_residual_linf = 0.0;
    diag->last_full_recompute_converged = 1;
    diag->dangling_vertices = 0;
    free(outdeg);
    return 0;
  }
  /* pagerank.hpp:91-100 */
  int need_full = (n != n_prev) || dangling_before != 0;
  for (uint64_t i = 0; i < k && !need_full; ++i) {
    uint32_t a = src[i];
    if (a < n && outdeg[a] == 0) need_full = 1;
    if (!directed && dst[i] != src[i]) {
      uint32_t b = dst[i];
      if (b < n && outdeg[b] == 0) need_full = 1;
    }
  }
  if (need_full) {
    pr_recompute(n, in_off, in_adj, outdeg, damping, rank, diag);
    free(outdeg);
    return 0;
  }
  diag->dangling_vertices = 0;

  uint8_t* active = (uint8_t*)calloc(n, 1);
  uint64_t active_count = 0;
#define ACTIVATE(arr, cnt, x)            \
  do {                                   \
    uint32_t _x = (x);                   \
    if (_x < n && !(arr)[_x]) {          \
      (arr)[_x] = 1;                     \
      ++(cnt);                           \
    }                                    \
  } while (0)
  for (uint64_t i = 0; i < k; ++i) {
    ACTIVATE(active, active_count, src[i]);
    ACTIVATE(active, active_count, dst[i]);
    if (src[i] < n)
      for (uint64_t e = out_off[src[i]]; e < out_off[src[i] + 1]; ++e)
        ACTIVATE(active, active_count, out_adj[e]);
    if (!directed && dst[i] < n)
      for (uint64_t e = out_off[dst[i]]; e < out_off[dst[i] + 1]; ++e)
        ACTIVATE(active, active_count, out_adj[e]);
  }
  if (active_count == 0) {
    diag->last_repaired_vertices = 0;
    diag->last_repair_iterations = 0;
    diag->last_residual_l1 = 0.0;
    diag->last_residual_linf = 0.0;
    free(active);
    free(outdeg);
    return 0;
  }
  double fl = full_fallback_fraction * (double)n;
  uint64_t fallback_limit = (uint64_t)(fl > 1.0 ? fl : 1.0);
  if (active_count >= fallback_limit) {
    pr_recompute(n, in_off, in_adj, outdeg, damping, rank, diag);
    free(active);
    free(outdeg);
    return 0;
  }
  const double base = (1.0 - damping) / (double)n;
  uint8_t* ever = (uint8_t*)malloc(n);
  memcpy(ever, active, n);
  uint8_t* next_active = (uint8_t*)malloc(n);
  double* next_rank = (double*)malloc((size_t)n * sizeof(double));
  diag->last_repair_iterations = 0;
  diag->last_residual_l1 = 0.0;
  diag->last_residual_linf = 0.0;
  diag->last_full_recompute_converged = 0;
  for (uint64_t it = 0; it < local_iterations && active_count != 0; ++it) {
    ++diag->last_repair_iterations;
    memcpy(next_rank, rank, (size_t)n * sizeof(double));
    memset(next_active, 0, n);
    uint64_t next_count = 0;
    double l1 = 0.0, linf = 0.0;
    for (uint32_t v = 0; v < n; ++v) {
      if (!active[v]) continue;
      double incoming = 0.0;
      for (uint64_t e = in_off[v]; e < in_off[v + 1]; ++e) {
        uint32_t u = in_adj[e];
        if (outdeg[u] != 0) incoming += rank[u] / (double)outdeg[u];
      }
      double updated = base + damping * incoming;
      double delta = fabs(updated - rank[v]);
      next_rank[v] = updated;
      l1 += delta;
      if (delta > linf) linf = delta;
      if (delta > tol) {
        for (uint64_t e = out_off[v]; e < out_off[v + 1]; ++e) {
          uint32_t w = out_adj[e];
          if (!next_active[w]) {
            next_active[w] = 1;
            ++next_count;
            ever[w] = 1;
          }
        }
      }
    }
    memcpy(rank, next_rank, (size_t)n * sizeof(double));
    uint8_t* t = active;
    active = next_active;
    next_active = t;
    active_count = next_count;
    diag->last_residual_l1 = l1;
    diag->last_residual_linf = linf;
    uint64_t repaired = 0;
    for (uint32_t v = 0; v < n; ++v) repaired += ever[v] != 0;
    if (repaired >= fallback_limit) {
      pr_recompute(n, in_off, in_adj, outdeg, damping, rank, diag);
      free(active);
      free(next_active);
      free(ever);
      free(next_rank);
      free(outdeg);
      return 0;
    }
  }
  uint64_t repaired = 0;
  for (uint32_t v = 0; v < n; ++v) repaired += ever[v] != 0;
  diag->last_repaired_vertices = repaired;
#undef ACTIVATE
  free(active);
  free(next_active);
  free(ever);
  free(next_rank);
  free(outdeg);
  return 0;
}

int orc_pagerank_static(uint32_t n, const uint64_t* out_off, const uint32_t* out_adj,
                        double damping, uint64_t max_iterations, double tolerance,
                        double* rank) {
  /* src/algorithms.cpp:56-81: push form, L1 stop. */
  if (n == 0) return 0;
  double* next = (double*)malloc((size_t)n * sizeof(double));
  for (uint32_t i = 0; i < n; ++i) rank[i] = 1.0 / (double)n;
  for (uint64_t iter = 0; iter < max_iterations; ++iter) {
    double dangling = 0.0;
    for (uint32_t u = 0; u < n; ++u)
      if (out_off[u + 1] == out_off[u]) dangling += rank[u];
    const double base = (1.0 - damping) / (double)n + damping * dangling / (double)n;
    for (uint32_t i = 0; i < n; ++i) next[i] = base;
    for (uint32_t u = 0; u < n; ++u) {
      uint64_t d = out_off[u + 1] - out_off[u];
      if (d == 0) continue;
      const double c = damping * rank[u] / (double)d;
      for (uint64_t e = out_off[u]; e < out_off[u + 1]; ++e) next[out_adj[e]] += c;
    }
    double err = 0.0;
    for (uint32_t i = 0; i < n; ++i) err += fabs(next[i] - rank[i]);
    memcpy(rank, next, (size_t)n * sizeof(double));
    if (err < tolerance) break;
  }
  free(next);
  return 0;
}

/* ------------------------------------------------------- canonical CSR */

static int cmp_u32(const void* a, const void* b) {
  const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}

static uint64_t sort_unique_row(uint32_t* r, uint64_t len) {
  if (len < 2) return len;
  if (len <= 32) {
    for (uint64_t i = 1; i < len; ++i) {
      uint32_t x = r[i];
      uint64_t j = i;
      while (j > 0 && r[j - 1] > x) {
        r[j] = r[j - 1];
        --j;
      }
      r[j] = x;
    }
  } else {
    qsort(r, len, sizeof(uint32_t), cmp_u32);
  }
  uint64_t w = 1;
  for (uint64_t i = 1; i < len; ++i)
    if (r[i] != r[w - 1]) r[w++] = r[i];
  return w;
}

uint64_t orc_csr_build(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, int directed,
                       uint64_t* off, uint32_t* adj) {
  /* bulk_load_edges, dynamic_graph.hpp:502-529: loops dropped (:508), both
   * arcs when undirected (:512); SegmentedCsr::build :79-111: sorted rows,
   * duplicates removed (:85-86). */
  memset(off, 0, ((size_t)n + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; ++i) {
    const uint32_t u = src[i], v = dst[i];
    if (u == v) continue;
    off[u + 1]++;
    if (!directed) off[v + 1]++;
  }
  for (uint32_t u = 0; u < n; ++u) off[u + 1] += off[u];
  uint64_t* cur = (uint64_t*)malloc(((size_t)n + 1) * sizeof(uint64_t));
  memcpy(cur, off, ((size_t)n + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; ++i) {
    const uint32_t u = src[i], v = dst[i];
    if (u == v) continue;
    adj[cur[u]++] = v;
    if (!directed) adj[cur[v]++] = u;
  }
  /* cur[u] becomes the de-duplicated row length */
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t u = 0; u < n; ++u) cur[u] = sort_unique_row(adj + off[u], off[u + 1] - off[u]);
  uint64_t w = 0;
  for (uint32_t u = 0; u < n; ++u) {
    const uint64_t b = off[u], len = cur[u];
    off[u] = w;
    if (w != b) memmove(adj + w, adj + b, len * sizeof(uint32_t));
    w += len;
  }
  off[n] = w;
  free(cur);
  return w;
}

typedef struct {
  uint64_t key; /* u << 32 | v */
  uint64_t order;
  uint8_t add;
} arc_op;

static int cmp_arc_op(const void* a, const void* b) {
  const arc_op* x = (const arc_op*)a;
  const arc_op* y = (const arc_op*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->order < y->order ? -1 : (x->order > y->order);
}

uint64_t orc_csr_apply(uint32_t n_old, const uint64_t* off, const uint32_t* adj, int directed, uint64_t k,
                       const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint32_t n_new,
                       uint64_t* off_new, uint32_t* adj_new) {
  /* DynamicGraph::apply, dynamic_graph.hpp:543-549: apply_unversioned per op
   * in order (:717-724: loops ignored, both arcs when undirected), so for
   * every arc only its last op decides its final presence. */
  arc_op* ops = (arc_op*)malloc((2 * k + 1) * sizeof(arc_op));
  uint64_t no = 0;
  for (uint64_t i = 0; i < k; ++i) {
    const uint32_t u = src[i], v = dst[i];
    if (u == v) continue;
    ops[no].key = ((uint64_t)u << 32) | v;
    ops[no].order = i;
    ops[no].add = add[i] ? 1 : 0;
    ++no;
    if (!directed) {
      ops[no].key = ((uint64_t)v << 32) | u;
      ops[no].order = i;
      ops[no].add = add[i] ? 1 : 0;
      ++no;
    }
  }
  qsort(ops, no, sizeof(arc_op), cmp_arc_op);
  /* keep the last op of every arc, and only where it changes the arc */
  uint64_t nf = 0;
  for (uint64_t i = 0; i < no; ++i) {
    if (i + 1 < no && ops[i + 1].key == ops[i].key) continue;
    const uint32_t u = (uint32_t)(ops[i].key >> 32), v = (uint32_t)ops[i].key;
    const int present = row_contains(off, adj, n_old, u, v);
    if ((ops[i].add != 0) != (present != 0)) ops[nf++] = ops[i];
  }
  uint64_t w = 0, j = 0;
  for (uint32_t u = 0; u < n_new; ++u) {
    off_new[u] = w;
    const uint64_t b = u < n_old ? off[u] : 0, e = u < n_old ? off[u + 1] : 0;
    if (j >= nf || (uint32_t)(ops[j].key >> 32) != u) {
      memcpy(adj_new + w, adj + b, (e - b) * sizeof(uint32_t));
      w += e - b;
      continue;
    }
    /* merge the row with its changes (both sorted by id) */
    uint64_t x = b;
    while (j < nf && (uint32_t)(ops[j].key >> 32) == u) {
      const uint32_t v = (uint32_t)ops[j].key;
      while (x < e && adj[x] < v) adj_new[w++] = adj[x++];
      if (ops[j].add) {
        adj_new[w++] = v;
      } else {
        ++x; /* adj[x] == v: the removed arc */
      }
      ++j;
    }
    while (x < e) adj_new[w++] = adj[x++];
  }
  off_new[n_new] = w;
  free(ops);
  return w;
}

/* ----------------------------------------------- personalized PageRank */

int orc_ppr_sweeps(uint32_t n, const uint64_t* off, const uint32_t* idx, const double* p, uint32_t K,
                   const double* B, double d, uint64_t iterations, double* Y) {
  if (K == 0 || K > 64) return -1;
  const size_t cells = (size_t)n * K;
  double* T = (double*)malloc(cells * sizeof(double) + 1);
  double* cur = Y;
  double* nxt = T;
  memcpy(cur, B, cells * sizeof(double));
  for (uint64_t it = 0; it < iterations; ++it) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t v = 0; v < n; ++v) {
      double acc[64];
      for (uint32_t k = 0; k < K; ++k) acc[k] = 0.0;
      for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
        const double* y = cur + (size_t)idx[e] * K;
        const double pe = p[e];
        for (uint32_t k = 0; k < K; ++k) acc[k] += pe * y[k];
      }
      for (uint32_t k = 0; k < K; ++k) nxt[(size_t)v * K + k] = B[(size_t)v * K + k] + d * acc[k];
    }
    double* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (cur != Y) memcpy(Y, cur, cells * sizeof(double));
  free(T);
  return 0;
}
