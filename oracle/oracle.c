/*
 * CPU oracle for the graph-kernel path: a plain-C restatement of the
 * reference's semantics, used ONLY as a checker (by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg). Nothing in the
 * product links, loads or calls it.
 *
 * Each function cites the reference code it restates (paths relative to the
 * reference repository). The reference's own CPU arithmetic lives in the
 * crates.io dependency grafeo-adapters =0.6.3 (Cargo.lock:459-469), which is
 * not vendored; the kernels and host drivers in crates/grafeo-cuda-algos fully
 * specify what the GPU path computes, and that is what is restated here.
 * How each function is pinned (reference known answers, closed forms, brute
 * force, networkx/scipy fixtures) is listed in tests/test_oracle.py.
 *
 * Two kinds of functions:
 *  - *_ref: literal restatements (same loops, same round structure);
 *  - *_fast: the same results by a faster method (degree-ordered DAG), for
 *    graphs where the literal form takes too long; each is checked against its
 *    *_ref twin on random graphs by the CPU tests.
 * OpenMP parallelizes the outer loops (oracle_set_threads).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void oracle_set_threads(int threads) { omp_set_num_threads(threads > 0 ? threads : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------ CSR -- */

/* validate_csr (crates/grafeo-cuda-core/src/csr_export.rs:71-120). Returns 0 if
 * valid, 6 (InvalidCsr) otherwise with the message in `msg`. A node whose end
 * offset runs past the targets (the reference would panic slicing) is skipped;
 * the decrease that must follow is reported instead. */
int oracle_validate_csr(const uint32_t* off, size_t n_off, const uint32_t* tgt, size_t n_tgt,
                        char* msg, size_t cap) {
  if (n_off == 0) {
    snprintf(msg, cap, "offsets must have num_nodes + 1 entries, got an empty array");
    return 6;
  }
  if (n_off > 0xFFFFFFFFull || n_tgt > 0xFFFFFFFFull) {
    snprintf(msg, cap, "graph too large for u32 indices: %zu offsets, %zu targets", n_off, n_tgt);
    return 6;
  }
  if (off[0] != 0) {
    snprintf(msg, cap, "offsets[0] must be 0, got %u", off[0]);
    return 6;
  }
  size_t n = n_off - 1;
  if ((size_t)off[n] != n_tgt) {
    snprintf(msg, cap, "offsets[%zu] is %u but targets has %zu entries", n, off[n], n_tgt);
    return 6;
  }
  for (size_t node = 0; node < n; ++node) {
    uint32_t s = off[node], e = off[node + 1];
    if (e < s) {
      snprintf(msg, cap, "offsets decrease at node %zu: %u then %u", node, s, e);
      return 6;
    }
    if (e > n_tgt) continue;
    for (uint32_t p = s; p < e; ++p) {
      if (tgt[p] >= n) {
        snprintf(msg, cap, "node %zu has neighbor %u, but there are only %zu nodes", node, tgt[p], n);
        return 6;
      }
      if (p > s && tgt[p - 1] >= tgt[p]) {
        snprintf(msg, cap, "neighbor list of node %zu is not strictly increasing (%u then %u)", node,
                 tgt[p - 1], tgt[p]);
        return 6;
      }
    }
  }
  return 0;
}

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y;
}

static size_t sort_unique_u64(uint64_t* k, size_t len) {
  qsort(k, len, sizeof(uint64_t), cmp_u64);
  size_t w = 0;
  for (size_t i = 0; i < len; ++i)
    if (w == 0 || k[i] != k[w - 1]) k[w++] = k[i];
  return w;
}

static void keys_to_csr(const uint64_t* k, size_t len, uint32_t n, uint32_t* off, uint32_t* tgt) {
  memset(off, 0, sizeof(uint32_t) * ((size_t)n + 1));
  for (size_t i = 0; i < len; ++i) off[(k[i] >> 32) + 1]++;
  for (size_t i = 1; i <= n; ++i) off[i] += off[i - 1];
  for (size_t i = 0; i < len; ++i) tgt[i] = (uint32_t)(k[i] & 0xFFFFFFFFu);
}

/* from_edge_pairs (crates/grafeo-cuda-core/src/csr_export.rs:202-239): push
 * (s,d), and (d,s) if symmetric and s != d; sort; dedup; prefix-count
 * offsets. `tgt` needs room for 2*E entries; returns the entry count. */
size_t oracle_from_edge_pairs(const uint32_t* pairs, size_t E, uint32_t n, int symmetric,
                              uint32_t* off, uint32_t* tgt) {
  uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * (2 * E + 1));
  size_t len = 0;
  for (size_t i = 0; i < E; ++i) {
    uint64_t s = pairs[2 * i], d = pairs[2 * i + 1];
    k[len++] = (s << 32) | d;
    if (symmetric && s != d) k[len++] = (d << 32) | s;
  }
  len = sort_unique_u64(k, len);
  keys_to_csr(k, len, n, off, tgt);
  free(k);
  return len;
}

/* from_edge_pairs_oriented (crates/grafeo-cuda-core/src/csr_export.rs:299-350):
 * drop loops; canonical (min,max) pairs, sorted and deduplicated; undirected
 * degrees; orient u->v if deg u < deg v or (equal and u < v); sort. */
size_t oracle_from_edge_pairs_oriented(const uint32_t* pairs, size_t E, uint32_t n, uint32_t* off,
                                       uint32_t* tgt) {
  uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * (E + 1));
  size_t len = 0;
  for (size_t i = 0; i < E; ++i) {
    uint64_t s = pairs[2 * i], d = pairs[2 * i + 1];
    if (s == d) continue;
    uint64_t a = s < d ? s : d, b = s < d ? d : s;
    k[len++] = (a << 32) | b;
  }
  len = sort_unique_u64(k, len);
  uint32_t* deg = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
  for (size_t i = 0; i < len; ++i) {
    deg[k[i] >> 32]++;
    deg[k[i] & 0xFFFFFFFFu]++;
  }
  for (size_t i = 0; i < len; ++i) {
    uint64_t u = k[i] >> 32, v = k[i] & 0xFFFFFFFFu;
    if (deg[u] < deg[v] || (deg[u] == deg[v] && u < v)) k[i] = (u << 32) | v;
    else k[i] = (v << 32) | u;
  }
  qsort(k, len, sizeof(uint64_t), cmp_u64);
  keys_to_csr(k, len, n, off, tgt);
  free(deg);
  free(k);
  return len;
}

/* ------------------------------------------------------ triangle count -- */

/* triangle_count (crates/grafeo-cuda-algos/kernels/triangle_counting.cu:246-315
 * with warp_intersect :88-122): for every u and every v in N(u) with v > u,
 * the number of w > v present in both N(u) and N(v). Merge form of the same
 * count. */
uint64_t oracle_triangle_count_ref(const uint32_t* off, const uint32_t* tgt, uint32_t n) {
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int64_t uu = 0; uu < (int64_t)n; ++uu) {
    uint32_t u = (uint32_t)uu;
    for (uint32_t i = off[u]; i < off[u + 1]; ++i) {
      uint32_t v = tgt[i];
      if (v <= u) continue;
      uint32_t j = off[u], je = off[u + 1], k = off[v], ke = off[v + 1];
      while (j < je && tgt[j] <= v) ++j;
      while (k < ke && tgt[k] <= v) ++k;
      while (j < je && k < ke) {
        if (tgt[j] == tgt[k]) { ++total; ++j; ++k; }
        else if (tgt[j] < tgt[k]) ++j;
        else ++k;
      }
    }
  }
  return total;
}

/* triangle_count_oriented (triangle_counting.cu:331-376): the sum over CSR
 * entries u->v of |N(u) intersect N(v)|. */
uint64_t oracle_triangle_count_oriented_ref(const uint32_t* off, const uint32_t* tgt, uint32_t n) {
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int64_t uu = 0; uu < (int64_t)n; ++uu) {
    uint32_t u = (uint32_t)uu;
    for (uint32_t i = off[u]; i < off[u + 1]; ++i) {
      uint32_t v = tgt[i];
      uint32_t j = off[u], je = off[u + 1], k = off[v], ke = off[v + 1];
      while (j < je && k < ke) {
        if (tgt[j] == tgt[k]) { ++total; ++j; ++k; }
        else if (tgt[j] < tgt[k]) ++j;
        else ++k;
      }
    }
  }
  return total;
}

/* Degree-ordered DAG of a CSR: rank(v) sorts vertices by (degree, id); each
 * vertex keeps the ranks of its neighbours of higher rank, sorted. Self-loops
 * drop out (a vertex is not above itself). */
typedef struct {
  uint32_t n;
  uint32_t* off;  /* n + 1, indexed by rank */
  uint32_t* tgt;  /* ranks */
} dag_t;

static uint32_t* g_sort_deg;
static int cmp_by_degree(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  if (g_sort_deg[x] != g_sort_deg[y]) return g_sort_deg[x] < g_sort_deg[y] ? -1 : 1;
  return x < y ? -1 : x > y;
}
static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}

static dag_t build_dag(const uint32_t* off, const uint32_t* tgt, uint32_t n, uint32_t** rank_out) {
  uint32_t* deg = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
  uint32_t* rank = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
  for (uint32_t v = 0; v < n; ++v) {
    deg[v] = off[v + 1] - off[v];
    order[v] = v;
  }
  g_sort_deg = deg;
  qsort(order, n, sizeof(uint32_t), cmp_by_degree);
  for (uint32_t i = 0; i < n; ++i) rank[order[i]] = i;
  dag_t d;
  d.n = n;
  d.off = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t i = off[u]; i < off[u + 1]; ++i)
      if (rank[tgt[i]] > rank[u]) d.off[rank[u] + 1]++;
  for (uint32_t i = 1; i <= n; ++i) d.off[i] += d.off[i - 1];
  d.tgt = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)d.off[n] + 1));
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t uu = 0; uu < (int64_t)n; ++uu) {
    uint32_t u = (uint32_t)uu, w = d.off[rank[u]];
    for (uint32_t i = off[u]; i < off[u + 1]; ++i)
      if (rank[tgt[i]] > rank[u]) d.tgt[w++] = rank[tgt[i]];
    qsort(d.tgt + d.off[rank[u]], w - d.off[rank[u]], sizeof(uint32_t), cmp_u32);
  }
  free(deg);
  free(order);
  if (rank_out) *rank_out = rank;
  else free(rank);
  return d;
}

static void free_dag(dag_t* d) {
  free(d->off);
  free(d->tgt);
}

/* Same count as oracle_triangle_count_ref for a symmetric CSR (each triangle
 * once, from the DAG's lowest vertex); checked against it by the tests. */
uint64_t oracle_triangle_count_fast(const uint32_t* off, const uint32_t* tgt, uint32_t n) {
  dag_t d = build_dag(off, tgt, n, NULL);
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int64_t uu = 0; uu < (int64_t)n; ++uu) {
    uint32_t u = (uint32_t)uu;
    for (uint32_t i = d.off[u]; i < d.off[u + 1]; ++i) {
      uint32_t v = d.tgt[i];
      uint32_t j = i + 1, je = d.off[u + 1], k = d.off[v], ke = d.off[v + 1];
      while (j < je && k < ke) {
        if (d.tgt[j] == d.tgt[k]) { ++total; ++j; ++k; }
        else if (d.tgt[j] < d.tgt[k]) ++j;
        else ++k;
      }
    }
  }
  free_dag(&d);
  return total;
}

/* The count of oracle_triangle_count_fast in two timed halves, for the CPU
 * baseline's bounded sample: oracle_tc_prepare builds the degree-ordered DAG
 * (an opaque handle), oracle_tc_count_rows runs the same merge over the DAG
 * rows r with r % stride == phase only (the phases of one stride add up to the
 * full count), oracle_tc_release frees the DAG. */
void* oracle_tc_prepare(const uint32_t* off, const uint32_t* tgt, uint32_t n) {
  dag_t* d = (dag_t*)malloc(sizeof(dag_t));
  *d = build_dag(off, tgt, n, NULL);
  return d;
}

uint64_t oracle_tc_count_rows(const void* handle, uint32_t stride, uint32_t phase) {
  const dag_t* d = (const dag_t*)handle;
  const uint32_t n = d->n;
  if (stride == 0) stride = 1;
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
  for (int64_t rr = phase; rr < (int64_t)n; rr += stride) {
    uint32_t u = (uint32_t)rr;
    for (uint32_t i = d->off[u]; i < d->off[u + 1]; ++i) {
      uint32_t v = d->tgt[i];
      uint32_t j = i + 1, je = d->off[u + 1], k = d->off[v], ke = d->off[v + 1];
      while (j < je && k < ke) {
        if (d->tgt[j] == d->tgt[k]) { ++total; ++j; ++k; }
        else if (d->tgt[j] < d->tgt[k]) ++j;
        else ++k;
      }
    }
  }
  return total;
}

void oracle_tc_release(void* handle) {
  dag_t* d = (dag_t*)handle;
  free_dag(d);
  free(d);
}

/* ---------------------------------------------------------------- BFS -- */

/* bfs::bfs (crates/grafeo-cuda-algos/src/bfs.rs:42-131, kernels/bfs.cu:54-83):
 * level-synchronous BFS along the CSR lists; -1 unreachable; reachable =
 * count of distances >= 0; max_depth = last non-empty level. Returns 4
 * (Kernel error) for a source out of range. */
int oracle_bfs(const uint32_t* off, const uint32_t* tgt, uint32_t n, uint32_t source, int32_t* dist,
               uint32_t* reachable, int32_t* max_depth) {
  if (source >= n) return 4;
  for (uint32_t v = 0; v < n; ++v) dist[v] = -1;
  uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * n);
  size_t head = 0, tail = 0;
  dist[source] = 0;
  q[tail++] = source;
  int32_t depth = 0;
  while (head < tail) {
    uint32_t u = q[head++];
    if (dist[u] > depth) depth = dist[u];
    for (uint32_t i = off[u]; i < off[u + 1]; ++i) {
      uint32_t v = tgt[i];
      if (dist[v] < 0) {
        dist[v] = dist[u] + 1;
        q[tail++] = v;
      }
    }
  }
  *reachable = (uint32_t)tail;
  *max_depth = depth;
  free(q);
  return 0;
}

/* ----------------------------------------------------------- PageRank -- */

/* pagerank::pagerank (crates/grafeo-cuda-algos/src/pagerank.rs:46-199,
 * kernels/pagerank.cu:19-134) in FP64: r0 = 1/n; each iteration
 * D = sum of r over nodes of degree 0, r'[v] = (1-d)/n + d*(sum over u in N(v)
 * with deg u > 0 of r[u]/deg u + D/n), diff = sum |r - r'|, iterations += 1,
 * swap, stop if diff < tol. `ranks` receives the latest iterate. */
void oracle_pagerank_f64(const uint32_t* off, const uint32_t* tgt, uint32_t n, double damping,
                         uint32_t max_iterations, double tol, double* ranks, uint32_t* iterations,
                         int* converged) {
  *iterations = 0;
  *converged = 0;
  if (n == 0) {
    *converged = 1;
    return;
  }
  double* a = (double*)malloc(sizeof(double) * n);
  double* b = (double*)malloc(sizeof(double) * n);
  double* c = (double*)malloc(sizeof(double) * n);
  for (uint32_t v = 0; v < n; ++v) a[v] = 1.0 / (double)n;
  for (uint32_t it = 0; it < max_iterations; ++it) {
    double dangling = 0.0;
#pragma omp parallel for reduction(+ : dangling)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
      uint32_t deg = off[v + 1] - off[v];
      c[v] = deg ? a[v] / (double)deg : 0.0;
      if (deg == 0) dangling += a[v];
    }
    double diff = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : diff)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
      double sum = 0.0;
      for (uint32_t i = off[v]; i < off[v + 1]; ++i) sum += c[tgt[i]];
      b[v] = (1.0 - damping) / (double)n + damping * (sum + dangling / (double)n);
      diff += fabs(a[v] - b[v]);
    }
    *iterations += 1;
    double* t = a;
    a = b;
    b = t;
    if (diff < tol) {
      *converged = 1;
      break;
    }
  }
  memcpy(ranks, a, sizeof(double) * n);
  free(a);
  free(b);
  free(c);
}

/* ------------------------------------------------------------- k-truss -- */

/* k_truss::k_truss (crates/grafeo-cuda-algos/src/k_truss.rs:42-161,
 * kernels/k_truss.cu:46-124), literal round structure: each round zeroes the
 * support, then for every active entry u->v with v > u merges the FULL lists
 * of u and v skipping removed entries, stores the count on u->v and on the
 * first active entry of v's list equal to u; then removes every active entry
 * with support < k-2. Rounds stop when one removes nothing; returns the
 * number of rounds including that one (0 if m == 0; k >= 3 is the caller's
 * check). `removed` (m bytes) receives the final mask. */
uint32_t oracle_k_truss_ref(const uint32_t* off, const uint32_t* tgt, uint32_t n, uint32_t k,
                            uint8_t* removed) {
  uint32_t m = off[n];
  if (m == 0) return 0;
  uint32_t threshold = k - 2;
  uint32_t* support = (uint32_t*)malloc(sizeof(uint32_t) * m);
  memset(removed, 0, m);
  uint32_t iterations = 0;
  for (;;) {
    memset(support, 0, sizeof(uint32_t) * m);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t uu = 0; uu < (int64_t)n; ++uu) {
      uint32_t u = (uint32_t)uu;
      for (uint32_t idx = off[u]; idx < off[u + 1]; ++idx) {
        if (removed[idx]) continue;
        uint32_t v = tgt[idx];
        if (v <= u) continue;
        uint32_t count = 0, j = off[u], je = off[u + 1], q = off[v], qe = off[v + 1];
        while (j < je && q < qe) {
          if (removed[j]) { ++j; continue; }
          if (removed[q]) { ++q; continue; }
          if (tgt[j] == tgt[q]) { ++count; ++j; ++q; }
          else if (tgt[j] < tgt[q]) ++j;
          else ++q;
        }
        support[idx] = count;
        for (uint32_t r = off[v]; r < off[v + 1]; ++r) {
          if (tgt[r] == u && !removed[r]) {
            support[r] = count;
            break;
          }
        }
      }
    }
    uint32_t count = 0;
    for (uint32_t e = 0; e < m; ++e) {
      if (!removed[e] && support[e] < threshold) {
        removed[e] = 1;
        ++count;
      }
    }
    ++iterations;
    if (count == 0) break;
  }
  free(support);
  return iterations;
}

static uint32_t find_entry(const uint32_t* off, const uint32_t* tgt, uint32_t row, uint32_t x) {
  uint32_t lo = off[row], hi = off[row + 1];
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (tgt[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return (lo < off[row + 1] && tgt[lo] == x) ? lo : UINT32_MAX;
}

/* Same result as oracle_k_truss_ref for a SYMMETRIC CSR, with each round's
 * support computed by enumerating every active triangle once (merge over the
 * degree-ordered DAG) instead of merging full lists; round 1 adds, as the
 * literal merge does, one for an active self-loop at each endpoint. Checked
 * against oracle_k_truss_ref by the tests. */
uint32_t oracle_k_truss_fast(const uint32_t* off, const uint32_t* tgt, uint32_t n, uint32_t k,
                             uint8_t* removed) {
  uint32_t m = off[n];
  if (m == 0) return 0;
  uint32_t threshold = k - 2;
  uint32_t* rank = NULL;
  dag_t d = build_dag(off, tgt, n, &rank);
  uint32_t* inv = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
  for (uint32_t v = 0; v < n; ++v) inv[rank[v]] = v;
  uint32_t mo = d.off[n];
  /* entry index in the original CSR of each DAG entry */
  uint32_t* orig = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)mo + 1));
  uint32_t* rev = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)mo + 1));
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ru = 0; ru < (int64_t)n; ++ru) {
    uint32_t u = inv[ru];
    for (uint32_t i = d.off[ru]; i < d.off[ru + 1]; ++i) {
      uint32_t v = inv[d.tgt[i]];
      orig[i] = find_entry(off, tgt, u, v);
      rev[i] = find_entry(off, tgt, v, u);
    }
  }
  uint8_t* loop = (uint8_t*)calloc((size_t)n + 1, 1);
  for (uint32_t v = 0; v < n; ++v) loop[v] = find_entry(off, tgt, v, v) != UINT32_MAX;
  uint8_t* dead = (uint8_t*)calloc((size_t)mo + 1, 1);
  uint32_t* support = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)mo + 1));
  memset(removed, 0, m);
  uint32_t iterations = 0;
  for (;;) {
    memset(support, 0, sizeof(uint32_t) * ((size_t)mo + 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ru = 0; ru < (int64_t)n; ++ru) {
      for (uint32_t i = d.off[ru]; i < d.off[ru + 1]; ++i) {
        if (dead[i]) continue;
        uint32_t rv = d.tgt[i];
        uint32_t j = i + 1, je = d.off[ru + 1], q = d.off[rv], qe = d.off[rv + 1];
        while (j < je && q < qe) {
          if (d.tgt[j] == d.tgt[q]) {
            if (!dead[j] && !dead[q]) {
#pragma omp atomic
              support[i]++;
#pragma omp atomic
              support[j]++;
#pragma omp atomic
              support[q]++;
            }
            ++j;
            ++q;
          } else if (d.tgt[j] < d.tgt[q]) ++j;
          else ++q;
        }
      }
    }
    uint32_t count = 0;
    if (iterations == 0) {
      for (uint32_t v = 0; v < n; ++v) {
        if (loop[v]) {
          removed[find_entry(off, tgt, v, v)] = 1;
          ++count;
        }
      }
    }
    for (uint32_t ru = 0; ru < n; ++ru) {
      for (uint32_t i = d.off[ru]; i < d.off[ru + 1]; ++i) {
        if (dead[i]) continue;
        uint32_t s = support[i];
        if (iterations == 0) s += loop[inv[ru]] + loop[inv[d.tgt[i]]];
        if (s < threshold) {
          dead[i] = 1;
          removed[orig[i]] = 1;
          removed[rev[i]] = 1;
          count += 2;
        }
      }
    }
    ++iterations;
    if (count == 0) break;
  }
  free(support);
  free(dead);
  free(loop);
  free(orig);
  free(rev);
  free(inv);
  free(rank);
  free_dag(&d);
  return iterations;
}

/* Trussness of every undirected edge of a symmetric CSR: the largest k whose
 * k-truss holds the edge, 2 for an edge in no triangle. There is no reference
 * function for it; it must agree with k_truss::k_truss
 * (crates/grafeo-cuda-algos/src/k_truss.rs:42-161, restated above by
 * oracle_k_truss_ref): the k-truss is exactly the edges with trussness >= k.
 * Sequential peeling (Wang and Cheng, "Truss decomposition in massive
 * networks", PVLDB 5(9), 2012): in phase k every live edge with at most k-2
 * triangles is removed, one at a time, with trussness k; each removal
 * destroys the triangles the edge closes with two live edges, taking one
 * support from each of those two edges. Self-loops are ignored. `truss` gets
 * one value per entry u -> v with u < v, in CSR order; returns that count, or
 * -1 when the CSR is not symmetric. */
int64_t oracle_truss_numbers(const uint32_t* off, const uint32_t* tgt, uint32_t n, uint32_t* truss) {
  uint32_t m = off[n];
  int64_t count = 0;
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t e = off[u]; e < off[u + 1]; ++e) count += tgt[e] > u;
  if (count == 0) return 0;
  /* edge id of every entry: forward entries in CSR order, reverse entries via
   * the forward entry, self-loops -1 */
  int64_t* id = (int64_t*)malloc(sizeof(int64_t) * m);
  uint32_t* eu = (uint32_t*)malloc(sizeof(uint32_t) * count);
  uint32_t* ev = (uint32_t*)malloc(sizeof(uint32_t) * count);
  int64_t next = 0;
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t e = off[u]; e < off[u + 1]; ++e) {
      id[e] = -1;
      if (tgt[e] > u) {
        eu[next] = u;
        ev[next] = tgt[e];
        id[e] = next++;
      }
    }
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t e = off[u]; e < off[u + 1]; ++e)
      if (tgt[e] != u) {
        uint32_t f = find_entry(off, tgt, tgt[e], u);
        if (f == UINT32_MAX) { /* not symmetric */
          free(ev);
          free(eu);
          free(id);
          return -1;
        }
        if (tgt[e] < u) id[e] = id[f];
      }
  int64_t* sup = (int64_t*)calloc(count, sizeof(int64_t));
  uint8_t* gone = (uint8_t*)calloc(count, 1);
  /* support: common neighbours other than the endpoints */
  for (int64_t i = 0; i < count; ++i) {
    uint32_t u = eu[i], v = ev[i];
    uint32_t a = off[u], b = off[v];
    while (a < off[u + 1] && b < off[v + 1]) {
      if (tgt[a] < tgt[b]) ++a;
      else if (tgt[a] > tgt[b]) ++b;
      else {
        if (tgt[a] != u && tgt[a] != v) ++sup[i];
        ++a;
        ++b;
      }
    }
  }
  int64_t* stack = (int64_t*)malloc(sizeof(int64_t) * count);
  int64_t live = count;
  for (int64_t k = 2; live > 0; ++k) {
    int64_t top = 0;
    for (int64_t i = 0; i < count; ++i)
      if (!gone[i] && sup[i] <= k - 2) stack[top++] = i;
    while (top > 0) {
      int64_t i = stack[--top];
      if (gone[i]) continue;
      gone[i] = 1;
      truss[i] = (uint32_t)k;
      --live;
      uint32_t u = eu[i], v = ev[i];
      uint32_t a = off[u], b = off[v];
      while (a < off[u + 1] && b < off[v + 1]) {
        if (tgt[a] < tgt[b]) ++a;
        else if (tgt[a] > tgt[b]) ++b;
        else {
          uint32_t w = tgt[a];
          if (w != u && w != v) {
            int64_t e1 = id[a], e2 = id[b];
            if (!gone[e1] && !gone[e2]) {
              /* an edge already at or below k-2 is on the stack; one that
               * reaches k-2 now joins it */
              if (--sup[e1] == k - 2) stack[top++] = e1;
              if (--sup[e2] == k - 2) stack[top++] = e2;
            }
          }
          ++a;
          ++b;
        }
      }
    }
  }
  free(stack);
  free(gone);
  free(sup);
  free(ev);
  free(eu);
  free(id);
  return count;
}
