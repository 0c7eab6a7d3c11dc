/*
 * CPU oracle for the SSSP / min-plus hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * the tests, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py.  The product library never links or calls it.
 *
 * Every function restates one algorithm of the reference (path:line cited in
 * oracle.c) in plain C99 with IEEE fp32 arithmetic (-O2 -fno-fast-math
 * -ffp-contract=off, SSE2 only), so that float sums are bit-identical to a
 * straightforward left fold.
 *
 * Conventions
 *   - CSR arrays: int64 row offsets, int32 column indices, 32-bit weights.
 *   - "out" orientation: row u lists edges u->v.  "in" orientation: row v
 *     lists edges u->v (the reference's Bellman-Ford / delta-stepping form).
 *   - f32 INF is +inf.  u32 INF is 0xFFFFFFFF with saturating addition.
 *   - pred: -1 for the source and for unreachable vertices.
 */
#ifndef MI355_ORACLE_H
#define MI355_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Dijkstra with a lazy-deletion binary heap over an outgoing CSR. */
int orc_dijkstra_f32(int64_t n, const int64_t* rowptr, const int32_t* col,
                     const float* w, int64_t src, float* dist);
int orc_dijkstra_u32(int64_t n, const int64_t* rowptr, const int32_t* col,
                     const uint32_t* w, int64_t src, uint32_t* dist);

/* Canonical predecessor: the smallest-index vertex p with an edge p->v such
 * that dist[v] == dist[p] (+) w.  -1 for src and for unreachable vertices. */
void orc_canonical_pred_f32(int64_t n, const int64_t* rowptr, const int32_t* col,
                            const float* w, int64_t src, const float* dist,
                            int32_t* pred);
void orc_canonical_pred_u32(int64_t n, const int64_t* rowptr, const int32_t* col,
                            const uint32_t* w, int64_t src, const uint32_t* dist,
                            int32_t* pred);

/* Checks dist[src]==0, pred[src]==-1 and, for every reached v != src, that an
 * edge pred[v]->v exists with dist[pred[v]] (+) w == dist[v].  Returns the
 * number of violating vertices. */
int64_t orc_check_pred_f32(int64_t n, const int64_t* rowptr, const int32_t* col,
                           const float* w, int64_t src, const float* dist,
                           const int32_t* pred);
int64_t orc_check_pred_u32(int64_t n, const int64_t* rowptr, const int32_t* col,
                           const uint32_t* w, int64_t src, const uint32_t* dist,
                           const int32_t* pred);

/* Jacobi Bellman-Ford over an INCOMING CSR: d <- min(d, A (x) d), convergence
 * tested every check_interval iterations (0 = never).  max_iters < 0 means
 * n-1.  Writes iterations_used and converged. */
int orc_bf_jacobi_in_f32(int64_t n, const int64_t* rowptr, const int32_t* col,
                         const float* w, int64_t src, int max_iters,
                         int check_interval, float* dist, int* iterations_used,
                         int* converged);

/* Dense Bellman-Ford over an incoming matrix A[v*n+u] = w(u->v), INF = no edge. */
int orc_bf_dense_f32(const float* A, int n, int src, int max_iters, float* dist,
                     int* iterations_used, int* converged);

/* Fixed-iteration ordered Gauss-Seidel sweeps over the same dense matrix. */
int orc_bf_gauss_seidel_f32(const float* A, int n, int src, int max_iters,
                            float* dist, int* iterations_used);

/* Number of buckets a delta-stepping run processes: the count of distinct
 * floorf(dist * (1.0f/delta)) over finite dist. */
int64_t orc_ds_bucket_count_f32(int64_t n, const float* dist, float delta);

/* C[i,j] = min_k A[i,k] + B[k,j], row-major with leading dimensions. */
void orc_minplus_gemm_f32(const float* A, const float* B, float* C, int M, int N,
                          int K, int lda, int ldb, int ldc);

/* Sparse min-plus product of CSR matrices.  Output rows are sorted by column.
 * Two-call protocol: call with out_col == NULL to get *out_nnz and fill
 * out_rowptr; then call again with allocated out_col/out_val. */
int orc_minplus_spgemm_f32(int64_t a_rows, int64_t a_cols, const int64_t* a_rowptr,
                           const int32_t* a_col, const float* a_val,
                           int64_t b_cols, const int64_t* b_rowptr,
                           const int32_t* b_col, const float* b_val,
                           int64_t* out_rowptr, int32_t* out_col, float* out_val,
                           int64_t* out_nnz);

/* Transpose a CSR (stable: rows of the result list columns in ascending
 * order of the original row). */
void orc_transpose(int64_t rows, int64_t cols, const int64_t* rowptr,
                   const int32_t* col, const uint32_t* w, int64_t* t_rowptr,
                   int32_t* t_col, uint32_t* t_w);


/* Multi-source Dijkstra on the subgraph induced by a grid rectangle (vertex
 * v = row * grid_cols + col, rows [r0, r1), cols [c0, c1)); dist indexed by
 * the box's row-major local index. */
int orc_dijkstra_box_f32(int64_t n, const int64_t* rowptr, const int32_t* col, const float* w,
                         int32_t grid_cols, int32_t r0, int32_t c0, int32_t r1, int32_t c1,
                         const int32_t* sources, int32_t nsrc, float* dist);
int orc_dijkstra_box_u32(int64_t n, const int64_t* rowptr, const int32_t* col, const uint32_t* w,
                         int32_t grid_cols, int32_t r0, int32_t c0, int32_t r1, int32_t c1,
                         const int32_t* sources, int32_t nsrc, uint32_t* dist);
/* Smallest-global-id parent inside the box; -1 for sources / unreached. */
void orc_canonical_pred_box(int64_t n, const int64_t* rowptr, const int32_t* col,
                            const uint32_t* wbits, int32_t is_f32, int32_t grid_cols, int32_t r0,
                            int32_t c0, int32_t r1, int32_t c1, const int32_t* sources,
                            int32_t nsrc, const uint32_t* distbits, int32_t* pred);

#ifdef __cplusplus
}
#endif
#endif
