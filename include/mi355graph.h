/*
 * mi355graph -- host-side graph ingest, file formats and seeded synthetic
 * workloads for the MI355X SSSP engine (libmi355graph.so, plain C ABI).
 *
 * A host graph is an opaque handle owning a CSR in one orientation:
 *   int64 rowptr[rows+1], int32 col[nnz], 32-bit weights[nnz]
 * whose weight words are either IEEE fp32 or uint32 (MI355G_W_F32/_U32).
 * Pointers returned by mi355g_view stay valid until mi355g_free.
 *
 * Every generator is a pure function of its arguments: it draws from a
 * counter-based SplitMix64 stream keyed by (seed, stream id, index), so the
 * same call yields the same bytes on any machine (the routing-grid lognormal
 * weights go through libm exp/log; see mi355g_gen_routing_grid).
 *
 * All calls return 0 on success or a negative MI355G_E* code; the message of
 * the last failure on the calling thread is available from mi355g_last_error.
 */
#ifndef MI355_GRAPH_H
#define MI355_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MI355G_W_F32 = 0, MI355G_W_U32 = 1 };
enum { MI355G_OUT = 0, MI355G_IN = 1 };
enum {
  MI355G_OK = 0,
  MI355G_EINVAL = -1,
  MI355G_ENOMEM = -2,
  MI355G_EIO = -3,
  MI355G_EFORMAT = -4
};

typedef struct mi355g_graph mi355g_graph;

typedef struct {
  int64_t rows;
  int64_t cols;
  int64_t nnz;
  int32_t weight_type;  /* MI355G_W_F32 or MI355G_W_U32 */
  int32_t orientation;  /* MI355G_OUT or MI355G_IN */
  const int64_t* rowptr;
  const int32_t* col;
  const uint32_t* w;    /* weight words (reinterpret as float for W_F32) */
  /* RIPSCSR1 bookkeeping (declared / loaded edge counts of the source file) */
  uint64_t declared_edges;
  uint64_t loaded_edges;
} mi355g_view_t;

const char* mi355g_last_error(void);
int mi355g_view(const mi355g_graph* g, mi355g_view_t* out);
void mi355g_free(mi355g_graph* g);

/* Build from caller arrays (copied). */
int mi355g_from_csr(int64_t rows, int64_t cols, int64_t nnz, const int64_t* rowptr,
                    const int32_t* col, const uint32_t* w, int32_t weight_type,
                    int32_t orientation, mi355g_graph** out);
/* Build from an edge list (u->v, w); rows sorted by column, duplicates kept
 * (dedup_min = 0) or reduced by min (dedup_min = 1). */
int mi355g_from_edges(int64_t n, int64_t m, const int32_t* src, const int32_t* dst,
                      const uint32_t* w, int32_t weight_type, int32_t orientation,
                      int32_t dedup_min, mi355g_graph** out);
/* Stable transpose: OUT <-> IN. */
int mi355g_transpose(const mi355g_graph* g, mi355g_graph** out);

/* ---- file formats ---------------------------------------------------- */
/* RIPSCSR1 (BF/src/gr_to_csr.cpp:34-57, RT/interchange_to_csr.cpp:1350-1388):
 * incoming orientation, int64 rowptr, int32 colind, f32 values. */
int mi355g_read_ripscsr1(const char* path, mi355g_graph** out);
int mi355g_write_ripscsr1(const mi355g_graph* g, const char* path);
/* RIPSADJ1 (BF/src/gr_to_adjacency.cpp:35-37, 180-213): dense incoming matrix
 * A[v*n+u] = w(u->v), INF = no edge; duplicates reduced by min.  The reader
 * reports n (matrix == NULL) or fills matrix[0..n*n). */
int mi355g_write_ripsadj1(const mi355g_graph* g, const char* path);
int mi355g_read_ripsadj1(const char* path, float* matrix, int64_t capacity, int64_t* n);
/* Dense row-major matrix -> CSR of its finite entries (orientation MI355G_OUT:
 * adj[u*n+v] = w(u->v), as dense_to_csr_outgoing in ONE/sssp_hip.cpp:320-348;
 * MI355G_IN: A[v*n+u] = w(u->v)). */
int mi355g_from_dense(const float* adj, int64_t n, int32_t orientation, int32_t skip_diagonal,
                      int32_t reject_negative, mi355g_graph** out);
/* DIMACS .gr text (BF/src/gr_to_csr.cpp:124-235): 1-based ids, rows sorted by
 * column, duplicate (u,v) reduced by min; result has incoming orientation. */
int mi355g_read_dimacs(const char* path, mi355g_graph** out);
/* Galois binary .gr v1 (ADDSF/csr_graph.cpp:153-309): outgoing orientation,
 * u64 row ends, u32 destinations, optional f32 edge data. */
int mi355g_read_galois_gr(const char* path, mi355g_graph** out);
int mi355g_write_galois_gr(const mi355g_graph* g, const char* path);

/* ---- seeded synthetic workloads (outgoing orientation) ------------------ */
/* rows x cols 4-neighbour grid, both directions, independent uint32 weights
 * uniform in [wmin, wmax] per directed edge. */
int mi355g_gen_grid4(int32_t rows, int32_t cols, uint32_t wmin, uint32_t wmax,
                     uint64_t seed, mi355g_graph** out);
/* R-MAT(a,b,c) with 2^scale vertices and edgefactor*2^scale draws; self loops
 * and duplicate undirected pairs removed, then symmetrised with one shared
 * fp32 weight (u24+1)*2^-24 in (0,1] per undirected edge; vertex ids
 * scrambled by a seeded permutation when permute != 0. */
int mi355g_gen_rmat(int32_t scale, int32_t edgefactor, double a, double b, double c,
                    uint64_t seed, int32_t permute, mi355g_graph** out);
/* side x side routing grid: 4-neighbour undirected edges each deleted with
 * probability p_delete, each of the two diagonals of every unit cell added
 * with probability p_diag, symmetric, uint32 weights max(1, ceil(exp(mu +
 * sigma*Z))) with Z standard normal (Box-Muller). */
int mi355g_gen_routing_grid(int32_t side, double p_delete, double p_diag, double mu,
                            double sigma, uint64_t seed, mi355g_graph** out);
/* Chain 0->1->...->n-1 (weight wchain) plus avg_deg random out-edges per
 * vertex (no self loops) with weights uniform in [wmin, wmax] (f32 kind:
 * fp32 uniform; u32 kind: integers). */
int mi355g_gen_chain_random(int64_t n, int32_t avg_deg, double wmin, double wmax,
                            int32_t weight_type, uint64_t seed, mi355g_graph** out);
/* count seeded sources among vertices with out-degree > 0. */
int mi355g_pick_sources(const mi355g_graph* g, int32_t count, uint64_t seed,
                        int32_t* sources);
/* SHA-256 of the CSR bytes (rowptr, col, weights, little endian) as 64 hex
 * chars + NUL. */
int mi355g_sha256(const mi355g_graph* g, char* hex65);

/* Counter-based generator primitives, exposed so workloads built elsewhere
 * (query boxes, min-plus matrices) draw from the same streams. */
uint64_t mi355g_rng_u64(uint64_t seed, uint64_t stream, uint64_t index);
/* fp32 in (0,1]: (top 24 bits + 1) * 2^-24, exact in fp32. */
float mi355g_rng_unit_f32(uint64_t seed, uint64_t stream, uint64_t index);
/* Fill buf[0..count) with unit_f32(seed, stream, first+i). */
void mi355g_fill_unit_f32(uint64_t seed, uint64_t stream, uint64_t first,
                          int64_t count, float* buf);

#ifdef __cplusplus
}
#endif
#endif
