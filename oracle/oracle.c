/*
 * CPU oracle for the SSSP / min-plus hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker; never linked into or called by the product library.
 *
 * Build: gcc -O2 -std=c99 -fno-fast-math -ffp-contract=off -msse2 -mfpmath=sse
 * (see oracle/Makefile).  All float arithmetic is a single rounded fp32 add
 * per relaxation (nd = du + w), exactly as the reference's CPU checkers do.
 *
 * Parity pins (what fixes each function independently of its own code) are
 * listed in tests/test_oracle.py; each function below cites the reference
 * routine it restates.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_U32_INF 0xFFFFFFFFu

/* ------------------------------------------------------------------------ */
/* Binary min-heap of (key, vertex), ordered lexicographically, which is the
 * order std::priority_queue<std::pair<K,int>, ..., std::greater<>> pops. */
typedef struct {
  float d;
  int32_t v;
} heap_f32_item;
typedef struct {
  uint32_t d;
  int32_t v;
} heap_u32_item;

#define DEFINE_HEAP(NAME, ITEM)                                              \
  typedef struct {                                                           \
    ITEM* a;                                                                 \
    int64_t size, cap;                                                       \
  } NAME;                                                                    \
  static int NAME##_less(ITEM x, ITEM y) {                                   \
    return x.d < y.d || (x.d == y.d && x.v < y.v);                           \
  }                                                                          \
  static int NAME##_push(NAME* h, ITEM it) {                                 \
    if (h->size == h->cap) {                                                 \
      int64_t nc = h->cap ? h->cap * 2 : 1024;                               \
      ITEM* na = (ITEM*)realloc(h->a, (size_t)nc * sizeof(ITEM));            \
      if (!na) return -1;                                                    \
      h->a = na;                                                             \
      h->cap = nc;                                                           \
    }                                                                        \
    int64_t i = h->size++;                                                   \
    while (i > 0) {                                                          \
      int64_t p = (i - 1) >> 1;                                              \
      if (!NAME##_less(it, h->a[p])) break;                                  \
      h->a[i] = h->a[p];                                                     \
      i = p;                                                                 \
    }                                                                        \
    h->a[i] = it;                                                            \
    return 0;                                                                \
  }                                                                          \
  static ITEM NAME##_pop(NAME* h) {                                          \
    ITEM top = h->a[0];                                                      \
    ITEM last = h->a[--h->size];                                             \
    int64_t i = 0;                                                           \
    for (;;) {                                                               \
      int64_t c = 2 * i + 1;                                                 \
      if (c >= h->size) break;                                               \
      if (c + 1 < h->size && NAME##_less(h->a[c + 1], h->a[c])) ++c;         \
      if (!NAME##_less(h->a[c], last)) break;                                \
      h->a[i] = h->a[c];                                                     \
      i = c;                                                                 \
    }                                                                        \
    if (h->size > 0) h->a[i] = last;                                         \
    return top;                                                              \
  }

DEFINE_HEAP(heap_f32, heap_f32_item)
DEFINE_HEAP(heap_u32, heap_u32_item)

static uint32_t sat_add_u32(uint32_t a, uint32_t b) {
  /* JAS/hip_adds_sssp.cpp:205-208: wrap-around saturates to INF. */
  uint32_t s = a + b;
  return s < a ? ORC_U32_INF : s;
}

/* Restates DS/testing/verify_bf_delta_vs_dijkstra.cpp:203-242 (and
 * ONE/test_sssp_hip.cpp:68-95): lazy-deletion binary heap, nd = du + w in
 * fp32, relax when nd < dist[v]. */
int orc_dijkstra_f32(int64_t n, const int64_t* rowptr, const int32_t* col,
                     const float* w, int64_t src, float* dist) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  for (int64_t i = 0; i < n; ++i) dist[i] = INFINITY;
  heap_f32 h = {0, 0, 0};
  dist[src] = 0.0f;
  heap_f32_item s = {0.0f, (int32_t)src};
  if (heap_f32_push(&h, s)) return -2;
  while (h.size > 0) {
    heap_f32_item t = heap_f32_pop(&h);
    if (t.d != dist[t.v]) continue;
    for (int64_t e = rowptr[t.v]; e < rowptr[t.v + 1]; ++e) {
      const int32_t v = col[e];
      const float nd = t.d + w[e];
      if (nd < dist[v]) {
        dist[v] = nd;
        heap_f32_item it = {nd, v};
        if (heap_f32_push(&h, it)) {
          free(h.a);
          return -2;
        }
      }
    }
  }
  free(h.a);
  return 0;
}

/* Same algorithm with uint32 labels and the saturating add of the reference's
 * uint32 ADDS path (JAS/hip_adds_sssp.cpp:200-213). */
int orc_dijkstra_u32(int64_t n, const int64_t* rowptr, const int32_t* col,
                     const uint32_t* w, int64_t src, uint32_t* dist) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  for (int64_t i = 0; i < n; ++i) dist[i] = ORC_U32_INF;
  heap_u32 h = {0, 0, 0};
  dist[src] = 0u;
  heap_u32_item s = {0u, (int32_t)src};
  if (heap_u32_push(&h, s)) return -2;
  while (h.size > 0) {
    heap_u32_item t = heap_u32_pop(&h);
    if (t.d != dist[t.v]) continue;
    for (int64_t e = rowptr[t.v]; e < rowptr[t.v + 1]; ++e) {
      const int32_t v = col[e];
      const uint32_t nd = sat_add_u32(t.d, w[e]);
      if (nd < dist[v]) {
        dist[v] = nd;
        heap_u32_item it = {nd, v};
        if (heap_u32_push(&h, it)) {
          free(h.a);
          return -2;
        }
      }
    }
  }
  free(h.a);
  return 0;
}

/* The reference has no predecessor output in C++; its only predecessor logic
 * is PT/.../minplus_bellman_ford.py:105-126.  The canonical choice used for
 * bit-exact comparison is the smallest-index parent that achieves dist[v]. */
void orc_canonical_pred_f32(int64_t n, const int64_t* rowptr, const int32_t* col,
                            const float* w, int64_t src, const float* dist,
                            int32_t* pred) {
  for (int64_t i = 0; i < n; ++i) pred[i] = -1;
  for (int64_t u = 0; u < n; ++u) {
    if (isinf(dist[u])) continue;
    for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
      const int32_t v = col[e];
      if (v == src) continue;
      if (dist[u] + w[e] == dist[v] && (pred[v] < 0 || u < pred[v]))
        pred[v] = (int32_t)u;
    }
  }
}

void orc_canonical_pred_u32(int64_t n, const int64_t* rowptr, const int32_t* col,
                            const uint32_t* w, int64_t src, const uint32_t* dist,
                            int32_t* pred) {
  for (int64_t i = 0; i < n; ++i) pred[i] = -1;
  for (int64_t u = 0; u < n; ++u) {
    if (dist[u] == ORC_U32_INF) continue;
    for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
      const int32_t v = col[e];
      if (v == src || dist[v] == ORC_U32_INF) continue;
      if (sat_add_u32(dist[u], w[e]) == dist[v] && (pred[v] < 0 || u < pred[v]))
        pred[v] = (int32_t)u;
    }
  }
}

/* Predecessor invariant dist[v] == dist[p] (+) w(p,v).  Any parallel edge
 * p->v may witness it. */
int64_t orc_check_pred_f32(int64_t n, const int64_t* rowptr, const int32_t* col,
                           const float* w, int64_t src, const float* dist,
                           const int32_t* pred) {
  int64_t bad = 0;
  if (!(dist[src] == 0.0f) || pred[src] != -1) ++bad;
  for (int64_t v = 0; v < n; ++v) {
    if (v == src) continue;
    if (isinf(dist[v])) {
      if (pred[v] != -1) ++bad;
      continue;
    }
    const int32_t p = pred[v];
    if (p < 0 || p >= n) {
      ++bad;
      continue;
    }
    int ok = 0;
    for (int64_t e = rowptr[p]; e < rowptr[p + 1] && !ok; ++e)
      if (col[e] == v && dist[p] + w[e] == dist[v]) ok = 1;
    if (!ok) ++bad;
  }
  return bad;
}

int64_t orc_check_pred_u32(int64_t n, const int64_t* rowptr, const int32_t* col,
                           const uint32_t* w, int64_t src, const uint32_t* dist,
                           const int32_t* pred) {
  int64_t bad = 0;
  if (dist[src] != 0u || pred[src] != -1) ++bad;
  for (int64_t v = 0; v < n; ++v) {
    if (v == src) continue;
    if (dist[v] == ORC_U32_INF) {
      if (pred[v] != -1) ++bad;
      continue;
    }
    const int32_t p = pred[v];
    if (p < 0 || p >= n) {
      ++bad;
      continue;
    }
    int ok = 0;
    for (int64_t e = rowptr[p]; e < rowptr[p + 1] && !ok; ++e)
      if (col[e] == v && sat_add_u32(dist[p], w[e]) == dist[v]) ok = 1;
    if (!ok) ++bad;
  }
  return bad;
}

/* Restates the CSR Bellman-Ford loops: BF/src/bf_hip_CSR.cpp:60-95 (check
 * every iteration), BF/src/bf_hip_no_checks_CSR.cpp:60-90 (never check) and
 * BF/src/bf_var_checks.cpp:86-160 (check when iteration % k == 0).  One
 * iteration is the sparse product A (x) d with the SpGEMM skip rules of
 * MP/src/minplus_sparse_hip.cpp:93-99 (absent/+INF/NaN ignored) followed by the
 * element-wise min merge; "changed" means some d[v] strictly decreased. */
int orc_bf_jacobi_in_f32(int64_t n, const int64_t* rowptr, const int32_t* col,
                         const float* w, int64_t src, int max_iters,
                         int check_interval, float* dist, int* iterations_used,
                         int* converged) {
  if (n <= 0 || src < 0 || src >= n || check_interval < 0) return -1;
  if (max_iters < 0) max_iters = (int)(n - 1);
  float* prev = (float*)malloc((size_t)n * sizeof(float));
  if (!prev) return -2;
  for (int64_t i = 0; i < n; ++i) dist[i] = INFINITY;
  dist[src] = 0.0f;
  *iterations_used = 0;
  *converged = 0;
  for (int it = 1; it <= max_iters; ++it) {
    memcpy(prev, dist, (size_t)n * sizeof(float));
    int changed = 0;
    for (int64_t v = 0; v < n; ++v) {
      float best = INFINITY;
      int have = 0;
      for (int64_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
        const float du = prev[col[e]];
        const float wv = w[e];
        if (isinf(du) && du > 0) continue;   /* absent in the sparse vector */
        if (isinf(wv) && wv > 0) continue;   /* explicit +INF entry ignored */
        if (isnan(du) || isnan(wv)) continue;
        const float c = du + wv;
        if (isnan(c)) continue;
        if (!have || c < best) best = c;
        have = 1;
      }
      if (have && best < prev[v]) {
        dist[v] = best;
        changed = 1;
      }
    }
    *iterations_used = it;
    const int check = check_interval > 0 && (it % check_interval) == 0;
    if (check && !changed) {
      *converged = 1;
      break;
    }
  }
  free(prev);
  return 0;
}

/* Restates BF/src/bellman_ford_hip.cpp:113-161 with the tile arithmetic of
 * MP/src/minplus_hip.cpp:48-117: relaxed[v] = fold over k ascending of
 * acc = (A[v,k] + d[k] < acc) ? A[v,k] + d[k] : acc, from acc = INF; then
 * d[v] = relaxed[v] < d[v] ? relaxed[v] : d[v] (relax_dist_kernel :43-59). */
int orc_bf_dense_f32(const float* A, int n, int src, int max_iters, float* dist,
                     int* iterations_used, int* converged) {
  if (!A || n <= 0 || src < 0 || src >= n) return -1;
  if (max_iters < 0) max_iters = n - 1;
  float* relaxed = (float*)malloc((size_t)n * sizeof(float));
  if (!relaxed) return -2;
  for (int i = 0; i < n; ++i) dist[i] = INFINITY;
  dist[src] = 0.0f;
  *iterations_used = 0;
  *converged = 0;
  for (int it = 1; it <= max_iters; ++it) {
    for (int v = 0; v < n; ++v) {
      float acc = INFINITY;
      const float* row = A + (size_t)v * (size_t)n;
      for (int k = 0; k < n; ++k) {
        const float c = row[k] + dist[k];
        acc = c < acc ? c : acc;
      }
      relaxed[v] = acc;
    }
    int changed = 0;
    for (int v = 0; v < n; ++v) {
      const float nv = relaxed[v] < dist[v] ? relaxed[v] : dist[v];
      if (nv < dist[v]) {
        dist[v] = nv;
        changed = 1;
      }
    }
    *iterations_used = it;
    if (!changed) {
      *converged = 1;
      break;
    }
  }
  free(relaxed);
  return 0;
}

/* Restates BF/src/bf_gauss_seidel.cpp:37-62 and the driver loop :112-135:
 * exactly max_iters sweeps (no convergence test); within a sweep target v
 * reads updated values for u < v and previous values for u > v; the diagonal
 * u == v is skipped. */
int orc_bf_gauss_seidel_f32(const float* A, int n, int src, int max_iters,
                            float* dist, int* iterations_used) {
  if (!A || n <= 0 || src < 0 || src >= n) return -1;
  if (max_iters < 0) max_iters = n - 1;
  float* prev = (float*)malloc((size_t)n * sizeof(float));
  if (!prev) return -2;
  for (int i = 0; i < n; ++i) dist[i] = INFINITY;
  dist[src] = 0.0f;
  for (int it = 0; it < max_iters; ++it) {
    memcpy(prev, dist, (size_t)n * sizeof(float));
    for (int v = 0; v < n; ++v) {
      float best = prev[v];
      const float* row = A + (size_t)v * (size_t)n;
      for (int u = 0; u < v; ++u) {
        const float c = dist[u] + row[u];
        best = c < best ? c : best;
      }
      for (int u = v + 1; u < n; ++u) {
        const float c = prev[u] + row[u];
        best = c < best ? c : best;
      }
      dist[v] = best;
    }
  }
  *iterations_used = max_iters;
  free(prev);
  return 0;
}

/* Closed form of delta-stepping's iteration count: bucket_index is
 * floorf(d * (1.0f/delta)) (DS/src/delta_stepping_hip_CSR.cpp:233-239, with
 * inv_delta computed once at :664), and the bucket loop (:701-780) processes
 * each bucket holding at least one final distance exactly once. */
static int cmp_i64(const void* a, const void* b) {
  const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

int64_t orc_ds_bucket_count_f32(int64_t n, const float* dist, float delta) {
  const float inv = 1.0f / delta;
  const int64_t kNoBucket = 0x3fffffff;
  int64_t* b = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  if (!b) return -1;
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (!isfinite(dist[i])) continue;
    const float bf = floorf(dist[i] * inv);
    int64_t bi;
    if (bf < 0.0f) bi = 0;
    else if (bf >= (float)kNoBucket) bi = kNoBucket - 1;
    else bi = (int64_t)bf;
    b[k++] = bi;
  }
  qsort(b, (size_t)k, sizeof(int64_t), cmp_i64);
  int64_t distinct = 0;
  for (int64_t i = 0; i < k; ++i)
    if (i == 0 || b[i] != b[i - 1]) ++distinct;
  free(b);
  return distinct;
}

/* Restates MP/src/minplus_hip.cpp:48-117: acc starts at +INF, k ascending,
 * acc = cand < acc ? cand : acc (NaN candidates never win). */
void orc_minplus_gemm_f32(const float* A, const float* B, float* C, int M, int N,
                          int K, int lda, int ldb, int ldc) {
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      float acc = INFINITY;
      for (int k = 0; k < K; ++k) {
        const float c = A[(size_t)i * lda + k] + B[(size_t)k * ldb + j];
        acc = c < acc ? c : acc;
      }
      C[(size_t)i * ldc + j] = acc;
    }
}

/* Restates the semantics of MP/src/minplus_sparse_hip.cpp:93-99 (skip NaN and
 * +INF inputs and candidates), :142-168 (min-reduce duplicates) and
 * :385-397 (one output entry per distinct column with a finite candidate).
 * The GPU output order is unspecified (hpp:59-60); the oracle sorts. */
static int valid_entry(float x) { return !isnan(x) && !(isinf(x) && x > 0); }

int orc_minplus_spgemm_f32(int64_t a_rows, int64_t a_cols, const int64_t* a_rowptr,
                           const int32_t* a_col, const float* a_val,
                           int64_t b_cols, const int64_t* b_rowptr,
                           const int32_t* b_col, const float* b_val,
                           int64_t* out_rowptr, int32_t* out_col, float* out_val,
                           int64_t* out_nnz) {
  (void)a_cols;
  float* acc = (float*)malloc((size_t)(b_cols > 0 ? b_cols : 1) * sizeof(float));
  unsigned char* seen = (unsigned char*)calloc((size_t)(b_cols > 0 ? b_cols : 1), 1);
  int32_t* touched = (int32_t*)malloc((size_t)(b_cols > 0 ? b_cols : 1) * sizeof(int32_t));
  if (!acc || !seen || !touched) {
    free(acc);
    free(seen);
    free(touched);
    return -2;
  }
  int64_t nnz = 0;
  out_rowptr[0] = 0;
  for (int64_t i = 0; i < a_rows; ++i) {
    int64_t nt = 0;
    for (int64_t p = a_rowptr[i]; p < a_rowptr[i + 1]; ++p) {
      const float av = a_val[p];
      if (!valid_entry(av)) continue;
      const int32_t k = a_col[p];
      for (int64_t q = b_rowptr[k]; q < b_rowptr[k + 1]; ++q) {
        const float bv = b_val[q];
        if (!valid_entry(bv)) continue;
        const float c = av + bv;
        if (!valid_entry(c)) continue;
        const int32_t j = b_col[q];
        if (!seen[j]) {
          seen[j] = 1;
          acc[j] = c;
          touched[nt++] = j;
        } else if (c < acc[j]) {
          acc[j] = c;
        }
      }
    }
    /* sort touched columns ascending (insertion sort is fine for the small
     * rows the oracle is used on; fall back to qsort for long rows) */
    if (nt > 32) {
      int64_t* tmp = (int64_t*)malloc((size_t)nt * sizeof(int64_t));
      if (!tmp) {
        free(acc);
        free(seen);
        free(touched);
        return -2;
      }
      for (int64_t t = 0; t < nt; ++t) tmp[t] = touched[t];
      qsort(tmp, (size_t)nt, sizeof(int64_t), cmp_i64);
      for (int64_t t = 0; t < nt; ++t) touched[t] = (int32_t)tmp[t];
      free(tmp);
    } else {
      for (int64_t a = 1; a < nt; ++a) {
        int32_t x = touched[a];
        int64_t b = a - 1;
        while (b >= 0 && touched[b] > x) {
          touched[b + 1] = touched[b];
          --b;
        }
        touched[b + 1] = x;
      }
    }
    for (int64_t t = 0; t < nt; ++t) {
      const int32_t j = touched[t];
      if (out_col) {
        out_col[nnz] = j;
        out_val[nnz] = acc[j];
      }
      ++nnz;
      seen[j] = 0;
    }
    out_rowptr[i + 1] = nnz;
  }
  *out_nnz = nnz;
  free(acc);
  free(seen);
  free(touched);
  return 0;
}

/* Stable counting-sort transpose (the deterministic counterpart of
 * DS/src/delta_stepping_hip_CSR.cpp:558-605, whose atomic cursor leaves the
 * order within a row unspecified). */
void orc_transpose(int64_t rows, int64_t cols, const int64_t* rowptr,
                   const int32_t* col, const uint32_t* w, int64_t* t_rowptr,
                   int32_t* t_col, uint32_t* t_w) {
  for (int64_t c = 0; c <= cols; ++c) t_rowptr[c] = 0;
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) t_rowptr[col[e] + 1]++;
  for (int64_t c = 0; c < cols; ++c) t_rowptr[c + 1] += t_rowptr[c];
  int64_t* cur = (int64_t*)malloc((size_t)(cols > 0 ? cols : 1) * sizeof(int64_t));
  if (!cur) return;
  for (int64_t c = 0; c < cols; ++c) cur[c] = t_rowptr[c];
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      const int64_t pos = cur[col[e]]++;
      t_col[pos] = (int32_t)r;
      t_w[pos] = w[e];
    }
  free(cur);
}

/* ---- box-restricted multi-source queries (routing nets) --------------------
 * The batched multi-source workload has no reference counterpart; it runs the
 * same Dijkstra (DS/testing/verify_bf_delta_vs_dijkstra.cpp:203-242) on the
 * subgraph induced by a grid rectangle (vertex v = row * grid_cols + col,
 * rows [r0, r1), cols [c0, c1)), every source starting at 0.  Outputs are
 * indexed by the box's row-major local index; pred holds global vertex ids
 * (-1 for sources and unreached vertices). */
static int in_box(int64_t v, int32_t grid_cols, int32_t r0, int32_t c0, int32_t r1, int32_t c1,
                  int64_t* local) {
  const int64_t r = v / grid_cols, c = v % grid_cols;
  if (r < r0 || r >= r1 || c < c0 || c >= c1) return 0;
  *local = (r - r0) * (int64_t)(c1 - c0) + (c - c0);
  return 1;
}

static int64_t box_global(int64_t local, int32_t grid_cols, int32_t r0, int32_t c0, int32_t c1) {
  const int64_t bw = c1 - c0;
  return (r0 + local / bw) * (int64_t)grid_cols + c0 + local % bw;
}

int orc_dijkstra_box_f32(int64_t n, const int64_t* rowptr, const int32_t* col, const float* w,
                         int32_t grid_cols, int32_t r0, int32_t c0, int32_t r1, int32_t c1,
                         const int32_t* sources, int32_t nsrc, float* dist) {
  if (n <= 0 || grid_cols <= 0 || r0 < 0 || c0 < 0 || r1 <= r0 || c1 <= c0 || c1 > grid_cols ||
      (int64_t)r1 * grid_cols > n || nsrc <= 0)
    return -1;
  const int64_t area = (int64_t)(r1 - r0) * (c1 - c0);
  for (int64_t i = 0; i < area; ++i) dist[i] = INFINITY;
  heap_f32 h = {0, 0, 0};
  for (int32_t k = 0; k < nsrc; ++k) {
    int64_t l;
    if (!in_box(sources[k], grid_cols, r0, c0, r1, c1, &l)) {
      free(h.a);
      return -1;
    }
    dist[l] = 0.0f;
    heap_f32_item s = {0.0f, (int32_t)l};
    if (heap_f32_push(&h, s)) return -2;
  }
  while (h.size > 0) {
    heap_f32_item t = heap_f32_pop(&h);
    if (t.d != dist[t.v]) continue;
    const int64_t u = box_global(t.v, grid_cols, r0, c0, c1);
    for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
      int64_t l;
      if (!in_box(col[e], grid_cols, r0, c0, r1, c1, &l)) continue;
      const float nd = t.d + w[e];
      if (nd < dist[l]) {
        dist[l] = nd;
        heap_f32_item it = {nd, (int32_t)l};
        if (heap_f32_push(&h, it)) {
          free(h.a);
          return -2;
        }
      }
    }
  }
  free(h.a);
  return 0;
}

int orc_dijkstra_box_u32(int64_t n, const int64_t* rowptr, const int32_t* col, const uint32_t* w,
                         int32_t grid_cols, int32_t r0, int32_t c0, int32_t r1, int32_t c1,
                         const int32_t* sources, int32_t nsrc, uint32_t* dist) {
  if (n <= 0 || grid_cols <= 0 || r0 < 0 || c0 < 0 || r1 <= r0 || c1 <= c0 || c1 > grid_cols ||
      (int64_t)r1 * grid_cols > n || nsrc <= 0)
    return -1;
  const int64_t area = (int64_t)(r1 - r0) * (c1 - c0);
  for (int64_t i = 0; i < area; ++i) dist[i] = ORC_U32_INF;
  heap_u32 h = {0, 0, 0};
  for (int32_t k = 0; k < nsrc; ++k) {
    int64_t l;
    if (!in_box(sources[k], grid_cols, r0, c0, r1, c1, &l)) {
      free(h.a);
      return -1;
    }
    dist[l] = 0u;
    heap_u32_item s = {0u, (int32_t)l};
    if (heap_u32_push(&h, s)) return -2;
  }
  while (h.size > 0) {
    heap_u32_item t = heap_u32_pop(&h);
    if (t.d != dist[t.v]) continue;
    const int64_t u = box_global(t.v, grid_cols, r0, c0, c1);
    for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
      int64_t l;
      if (!in_box(col[e], grid_cols, r0, c0, r1, c1, &l)) continue;
      const uint32_t nd = sat_add_u32(t.d, w[e]);
      if (nd < dist[l]) {
        dist[l] = nd;
        heap_u32_item it = {nd, (int32_t)l};
        if (heap_u32_push(&h, it)) {
          free(h.a);
          return -2;
        }
      }
    }
  }
  free(h.a);
  return 0;
}

/* Canonical predecessor inside the box: the smallest global id u (in the box)
 * with an edge u->v and dist[u] (+) w == dist[v]; -1 for sources and for
 * unreached vertices.  wbits: 32-bit weights, f32 when is_f32 != 0. */
void orc_canonical_pred_box(int64_t n, const int64_t* rowptr, const int32_t* col,
                            const uint32_t* wbits, int32_t is_f32, int32_t grid_cols, int32_t r0,
                            int32_t c0, int32_t r1, int32_t c1, const int32_t* sources,
                            int32_t nsrc, const uint32_t* distbits, int32_t* pred) {
  (void)n;
  const int64_t area = (int64_t)(r1 - r0) * (c1 - c0);
  for (int64_t i = 0; i < area; ++i) pred[i] = -1;
  for (int64_t lu = 0; lu < area; ++lu) {
    const int64_t u = box_global(lu, grid_cols, r0, c0, c1);
    for (int64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
      int64_t lv;
      if (!in_box(col[e], grid_cols, r0, c0, r1, c1, &lv)) continue;
      int hit;
      if (is_f32) {
        float du, dv, we;
        memcpy(&du, &distbits[lu], 4);
        memcpy(&dv, &distbits[lv], 4);
        memcpy(&we, &wbits[e], 4);
        hit = isfinite(du) && isfinite(dv) && (float)(du + we) == dv;
      } else {
        hit = distbits[lu] != ORC_U32_INF && distbits[lv] != ORC_U32_INF &&
              sat_add_u32(distbits[lu], wbits[e]) == distbits[lv];
      }
      if (hit && (pred[lv] < 0 || u < pred[lv])) pred[lv] = (int32_t)u;
    }
  }
  for (int32_t k = 0; k < nsrc; ++k) {
    int64_t l;
    if (in_box(sources[k], grid_cols, r0, c0, r1, c1, &l)) pred[l] = -1;
  }
}
