/*
 * mi355sssp -- MI355X (gfx950) SSSP and min-plus engine, C ABI.
 *
 * Library: libmi355sssp.so (hand-written wave64 HIP kernels + C++ host
 * drivers).  Every entry point returns an mi355_status (0 = success); on
 * failure mi355_last_error() holds a message for the calling thread.  No C++
 * exception crosses this boundary.  Streams are passed as opaque pointers
 * (hipStream_t; NULL = the default stream).  Unless stated otherwise a call
 * is blocking: it synchronises its stream before returning, like the
 * reference functions it replaces.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository root).  The C++ headers in include/compat/
 * restate those reference signatures on top of this ABI.
 */
#ifndef MI355_SSSP_H
#define MI355_SSSP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ---------------------------------------------------------- */
typedef enum {
  MI355_OK = 0,
  MI355_EINVAL = -1,     /* bad argument (reference: std::invalid_argument / hipErrorInvalidValue) */
  MI355_ERANGE = -2,     /* source out of range (reference: std::out_of_range) */
  MI355_EOVERFLOW = -3,  /* size does not fit the index type (std::overflow_error) */
  MI355_EHIP = -4,       /* HIP runtime error (std::runtime_error / hipError_t) */
  MI355_ENOMEM = -5,     /* device allocation failed */
  MI355_ECAPACITY = -6,  /* a queue / hash / pool capacity was exceeded */
  MI355_EDEVICE = -7     /* a kernel reported an internal error */
} mi355_status;

const char* mi355_last_error(void);
/* The hipError_t of the last MI355_EHIP failure on this thread (0 if none). */
int mi355_last_hip_error(void);

/* ---- device plumbing (for FFI callers without their own allocator) ---- */
int mi355_device_count(int* count);
int mi355_set_device(int device);
int mi355_device_malloc(void** ptr, size_t bytes);
int mi355_device_free(void* ptr);
int mi355_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int mi355_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int mi355_memset_d8(void* dst, int value, size_t bytes, void* stream);
int mi355_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
/* dst[i] = src[idx[i]] for count 8-byte records (device pointers; idx 32-bit).
 * No reference counterpart: the random-access reference point of the HBM
 * traffic counters.  Asynchronous. */
int mi355_device_gather_u64(void* dst, const void* src, const uint32_t* idx, int64_t count, void* stream);
int mi355_stream_create(void** stream);
int mi355_stream_destroy(void* stream);
int mi355_stream_sync(void* stream);

/* ---- CSR descriptors --------------------------------------------------- */
enum { MI355_W_F32 = 0, MI355_W_U32 = 1 };
enum { MI355_OUT = 0, MI355_IN = 1 };        /* row u lists u->v / row v lists u->v */
enum { MI355_HOST = 0, MI355_DEVICE = 1 };   /* where the arrays live */

typedef struct {
  int64_t rows;
  int64_t cols;
  int64_t nnz;
  const void* rowptr;   /* rows+1 offsets, int64 or 32-bit per offset_bits */
  const void* col;      /* nnz 32-bit column indices */
  const void* w;        /* nnz 32-bit weights (fp32 or uint32) */
  int32_t offset_bits;  /* 64 or 32 */
  int32_t weight_type;  /* MI355_W_F32 / MI355_W_U32 */
  int32_t orientation;  /* MI355_OUT / MI355_IN */
  int32_t memory;       /* MI355_HOST / MI355_DEVICE */
} mi355_csr_desc;

/* ---- graph handle -------------------------------------------------------
 * A device-resident graph in the engine's layout: outgoing CSR with int32
 * offsets and interleaved (dst, weight) 8-byte edge records, plus cached
 * degree statistics.  Creating one validates the input (range, monotone
 * offsets; finite non-negative weights unless allow_negative) and, for
 * incoming input, transposes it on the device.  Replaces the per-call
 * validation + transpose of DS/src/delta_stepping_hip_CSR.cpp:546-605, 659-667
 * and the per-call uploads of ONE/sssp_hip.cpp:125-160. */
typedef struct mi355_graph mi355_graph;

typedef struct {
  int64_t n;
  int64_t m;
  int32_t weight_type;
  int32_t max_degree;
  double avg_degree;
  double avg_weight;        /* mean of positive finite weights */
  int64_t large_vertices;   /* out-degree above the wave-split threshold */
  int64_t large_chunks;     /* 2048-edge chunks those vertices split into */
  int32_t has_negative;     /* any negative weight */
  int32_t device;
  int32_t vertex_order;     /* MI355_ORDER_INPUT or MI355_ORDER_DEGREE (the order used) */
  int32_t reserved;
} mi355_graph_info;

/* Vertex order of a graph handle.  DEGREE renumbers the vertices internally
 * by descending out-degree (ties keep input order) so that the labels of the
 * vertices most edges point to share cache lines; sources, outputs and
 * parents stay in input ids and results are identical.  AUTO picks DEGREE
 * for graphs of at least 4096 vertices whose maximum degree is at least 32x
 * the mean (R-MAT), INPUT otherwise (grids keep their geometry). */
enum { MI355_ORDER_AUTO = 0, MI355_ORDER_INPUT = 1, MI355_ORDER_DEGREE = 2 };

int mi355_graph_create(const mi355_csr_desc* desc, int32_t allow_negative, void* stream,
                       mi355_graph** out);
/* Same with an explicit vertex order (mi355_graph_create uses MI355_ORDER_AUTO). */
int mi355_graph_create_ex(const mi355_csr_desc* desc, int32_t allow_negative, int32_t vertex_order,
                          void* stream, mi355_graph** out);
int mi355_graph_destroy(mi355_graph* g);
int mi355_graph_info_get(const mi355_graph* g, mi355_graph_info* info);
/* Copy the engine's outgoing CSR back (int64 rowptr, int32 col, 32-bit w), in
 * the input vertex numbering. */
int mi355_graph_download(const mi355_graph* g, int64_t* rowptr, int32_t* col, uint32_t* w);

/* ---- single-source shortest paths on a graph handle -------------------- */
enum {
  MI355_ALG_BF = 0,         /* frontier Bellman-Ford, in place (chaotic) */
  MI355_ALG_BF_JACOBI = 1,  /* frontier Bellman-Ford with exact Jacobi rounds */
  MI355_ALG_DELTA = 2,      /* bucket-synchronous delta-stepping (light/heavy) */
  MI355_ALG_ADDS = 3        /* asynchronous dynamic delta-stepping, persistent */
};
/* dist output conventions for INF */
enum { MI355_INF_IEEE = 0, MI355_INF_FLT_MAX = 1 };

typedef struct {
  int32_t algorithm;
  int32_t max_iters;        /* <0: algorithm default (BF: n-1; DS/ADDS: unbounded) */
  int32_t check_interval;   /* BF_JACOBI: test convergence when it % k == 0; 0 = never */
  float delta;              /* DS / ADDS bucket width; <= 0: heuristic 16*wbar/dbar */
  int32_t inf_mode;         /* MI355_INF_IEEE or MI355_INF_FLT_MAX (f32 outputs) */
  int32_t dynamic_delta;    /* ADDS: feedback delta per phase (0 = plain ADDS: fixed delta) */
  int32_t use_mlp;          /* ADDS: 1: the loaded delta model picks delta, active buckets and
                               worker share per phase; 2: the phase schedule of
                               mi355_adds_schedule_set does */
  int32_t worker_blocks;    /* ADDS: persistent workgroups (0 = auto) */
  int32_t rounds_per_sync;  /* host sync cadence (0 = adaptive) */
  int32_t active_buckets;   /* ADDS: buckets served at once, 1..32 (0 = 1) */
  int32_t pull;             /* ADDS: how a finished bucket's heavy edges are relaxed.
                               0: pushed from the bucket's vertices (the reference's ADDS);
                               1: pulled by the unsettled vertices instead when the estimated
                               push work exceeds 2^pull_margin x the estimated pull work
                               (use_mlp = 1 with a model that has a pull output: the
                               network's margin for the phase) */
  float pull_margin;        /* ADDS, pull = 1: log2 of the push / pull work ratio above which
                               a bucket is pulled (0: whenever pulling is estimated cheaper) */
  float light_scale;        /* ADDS: light edges (relaxed when a vertex is processed; the rest
                               when its bucket is finished) are those of weight <= light_scale x
                               the first bucket width; 0: 1, or under use_mlp = 1 the network's
                               choice */
  int32_t reserved[3];
} mi355_sssp_options;

typedef struct {
  int32_t iterations_used;  /* BF: rounds; DS: buckets; ADDS: phases */
  int32_t converged;
  int32_t rounds;           /* kernel rounds executed (all algorithms) */
  int32_t syncs;            /* host synchronisations */
  int64_t edges_relaxed;    /* edge visits (work, not unique edges) */
  int64_t vertices_processed;
  int64_t reached;          /* vertices with finite distance */
  int64_t reached_edges;    /* sum of out-degrees of reached vertices (m_r) */
  float initial_delta;
  float final_delta;
  float device_ms;          /* hipEvent time of the GPU work */
  float kernel_ms;          /* hipEvent time of the persistent ADDS kernel alone (0 for the
                               algorithms that run as a sequence of launches) */
  int32_t has_negative_cycle;
  int32_t policy;           /* ADDS phase policy used: 0 fixed delta, 1 feedback, 2 learned
                               network (0 when use_mlp is set but the graph lies outside the
                               model's training range), 3 explicit schedule */
  int32_t pulled_phases;    /* ADDS: finished buckets whose heavy edges were pulled */
  int32_t reserved;
} mi355_sssp_stats;

void mi355_sssp_options_default(mi355_sssp_options* opt);

/* Shortest paths from source.  d_dist: device array of n 32-bit distances
 * (fp32 or uint32 per the graph's weight type); d_pred: device array of n
 * int32 predecessors or NULL.  Predecessor of v is the smallest-index vertex
 * p with dist[v] == dist[p] (+) w(p,v); -1 for the source and unreachable
 * vertices.  per_round_frontier (host, may be NULL, length cap_rounds)
 * receives the frontier size after each BF round. */
int mi355_sssp(mi355_graph* g, int32_t source, const mi355_sssp_options* opt, void* d_dist,
               int32_t* d_pred, mi355_sssp_stats* stats, void* stream);

/* Same, leaving distances as packed 64-bit labels (dist key << 32 | pred) in
 * the graph's internal buffer (no output conversion); used by timing loops.
 * mi355_sssp_labels_to_output converts them afterwards. */
int mi355_sssp_labels(mi355_graph* g, int32_t source, const mi355_sssp_options* opt,
                      mi355_sssp_stats* stats, void* stream);
int mi355_sssp_labels_to_output(mi355_graph* g, int32_t source, int32_t inf_mode, void* d_dist,
                                int32_t* d_pred, void* stream);
/* Record per-round frontier sizes of the next BF / BF_JACOBI run on g: that
 * run writes |F_0| .. |F_rounds| (at most cap entries) into the host array
 * `sizes`, which must stay valid until it returns; the registration is used
 * by that one run.  No counterpart in the reference (a diagnostic of this
 * engine's frontier rounds). */
int mi355_sssp_round_log(mi355_graph* g, int32_t* sizes, int32_t cap);

/* ---- reference entry points ------------------------------------------- */

/* Replaces minplus_gemm_f32 (MP/src/minplus_hip.hpp:24-33, minplus_hip.cpp:122-173):
 * C[i,j] = min_k A[i,k] + B[k,j], row-major, device pointers, asynchronous.
 * MI355_EINVAL for null pointers, non-positive sizes or ld < dims. */
int mi355_minplus_gemm_f32(const float* dA, const float* dB, float* dC, int M, int N, int K,
                           int lda, int ldb, int ldc, void* stream);

/* Batched variant (no reference counterpart; the batched min-plus workload of
 * the routing engine): count independent products.  The pointer and size
 * arrays are HOST arrays of device pointers / sizes.  Each item takes its size
 * class's (ceil(log2(max(M, N)))) autotuned tile, else the tile the static
 * rule picks for the whole batch's size; each tile shape is one launch over a
 * flat list of (item, tile) entries.  The calling thread keeps the last
 * batch's plan on the device: a call with the same pointers and sizes (and
 * tuning) only launches; any other call first waits for the kernels of the
 * previous one.  Asynchronous. */
int mi355_minplus_gemm_f32_batched(int count, const float* const* dA, const float* const* dB,
                                   float* const* dC, const int* M, const int* N, const int* K,
                                   const int* lda, const int* ldb, const int* ldc, void* stream);

/* No counterpart in the reference: the batched product above for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  Same arguments and results; the batch's descriptors and work list
 * go into a fixed 4 MiB array per GPU that the library holds until the
 * process ends, so no hipMalloc / hipFree runs.  At most 4096 products and
 * about 500K output tiles per call (MI355_ECAPACITY beyond).  Every call on a
 * GPU reuses that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race and must be ordered by the caller.  No plan is kept
 * between calls.  Asynchronous. */
int mi355_minplus_gemm_f32_batched_noalloc(int count, const float* const* dA, const float* const* dB,
                                           float* const* dC, const int* M, const int* N, const int* K,
                                           const int* lda, const int* ldb, const int* ldc, void* stream);

/* Project autotuner for the min-plus tile (no reference counterpart; the
 * reference hard-codes 64x64x32 tiles, MP/src/minplus_hip.cpp:144-151).  Times
 * every tile variant on an n x n x n product for each n in sizes and keeps the
 * fastest per size class (ceil(log2 n)).  timings_ms (host, may be NULL)
 * receives count x mi355_minplus_tile_count() times.  Blocking. */
int mi355_minplus_tile_count(void);
const char* mi355_minplus_tile_name(int index);
int mi355_minplus_autotune(const int* sizes, int count, double* timings_ms, void* stream);
/* Read / install the 32-entry tuning table (size class -> tile index, -1 =
 * built-in rule), e.g. to persist a tuning run. */
int mi355_minplus_tuning_get(int* table32);
int mi355_minplus_tuning_set(const int* table32);

/* Replaces minplus_sparse::minplus_spgemm_csr_f32 / _large_rows / _tuned<HC,BS>
 * (MP/src/minplus_sparse_hip.hpp:64-150, minplus_sparse_hip.cpp:410-764).
 * A, B: device CSR (int64 rowptr, int32 col, fp32 values).  Outputs are
 * allocated here with hipMalloc and owned by the caller (hipFree); an empty
 * product gives a zero rowptr and NULL colind/values.  hash_capacity bounds
 * the unique columns of one output row (MI355_ECAPACITY beyond it, outputs
 * freed); block_size is the workgroup size hint (64, 128 or 256). */
int mi355_minplus_spgemm_csr_f32(int64_t a_rows, int64_t a_cols, int64_t a_nnz,
                                 const int64_t* a_rowptr, const int32_t* a_col,
                                 const float* a_val, int64_t b_rows, int64_t b_cols,
                                 int64_t b_nnz, const int64_t* b_rowptr, const int32_t* b_col,
                                 const float* b_val, int hash_capacity, int block_size,
                                 int64_t** c_rowptr, int32_t** c_col, float** c_val,
                                 int64_t* c_nnz, void* stream);

/* Element-wise min of two sparse n x 1 distance vectors (device CSR, int64
 * rowptr, fp32 values, absent = +INF): the merge step d <- min(d, r) of the
 * reference's CSR Bellman-Ford, bf_csr_detail::merge_distances_on_device
 * (header missing from the reference; used at BF/src/bf_hip_CSR.cpp:74-76).
 * Output allocated here (caller frees); changed (may be NULL) = 1 when some
 * entry of b is strictly below a's.  Blocking. */
int mi355_sparse_vector_min_merge(int64_t rows, const int64_t* a_rowptr, const float* a_val,
                                  const int64_t* b_rowptr, const float* b_val,
                                  int64_t** c_rowptr, int32_t** c_col, float** c_val,
                                  int64_t* c_nnz, int* changed, void* stream);

/* Replaces bellman_ford_minplus_hip (BF/src/bellman_ford_hip.hpp:14-59,
 * bellman_ford_hip.cpp:64-264): dense incoming adjacency A[v*n+u] = w(u->v),
 * INF = no edge, Jacobi rounds until no change or max_iters (<0: n-1).
 * adjacency lives in `memory` (MI355_HOST / MI355_DEVICE); dist is host. */
int mi355_bf_dense_f32(const float* adjacency, int32_t memory, int n, int source,
                       int max_iters, float* dist, int* iterations_used, int* converged,
                       void* stream);

/* Replaces bellman_ford_minplus_hip_gauss_seidel (BF/src/bf_gauss_seidel.cpp:76-260):
 * exactly max_iters (<0: n-1) ordered sweeps; has_negative_cycle is never set. */
int mi355_bf_gauss_seidel_f32(const float* adjacency, int32_t memory, int n, int source,
                              int max_iters, float* dist, int* iterations_used, void* stream);

/* Progress record with the layout of BellmanFordCsrProgress
 * (BF/src/bf_hip_CSR.hpp:9-14). */
typedef struct {
  int iteration;
  int max_iters;
  unsigned char convergence_checked;
  unsigned char changed;
} mi355_bf_progress;
typedef void (*mi355_bf_progress_cb)(const mi355_bf_progress* progress, void* user_data);

/* Statistics of BellmanFordCsrVarChecksResult (BF/src/bf_var_checks.hpp:9-27). */
typedef struct {
  int iterations_used;
  int converged;
  int check_interval;
  int checks_performed;
  int changed_checks;
  int unchanged_checks;
  double initialization_ms;
  double bellman_ford_loop_ms;
  double relaxation_ms;
  double checked_merge_ms;
  double unchecked_merge_ms;
  double checked_iteration_ms;
  double unchecked_iteration_ms;
  double final_copy_ms;
  double total_runtime_ms;
  double estimated_check_overhead_ms;
} mi355_bf_csr_stats;

/* Replaces bellman_ford_minplus_hip_csr (BF/src/bf_hip_CSR.hpp:41-66, check every
 * iteration: check_interval = 1), bellman_ford_minplus_hip_csr_no_checks
 * (BF/src/bf_hip_no_checks_CSR.hpp:17-42: check_interval = 0) and
 * bellman_ford_minplus_hip_csr_var_checks[_with_stats]
 * (BF/src/bf_var_checks.hpp:33-71: check_interval = k).  A is the INCOMING CSR
 * (row v, column u = w(u->v)), host or device.  dist: host array of rows floats.
 * The progress callback runs on the calling thread, once per iteration, in
 * order, before the call returns. */
int mi355_bf_csr_f32(const mi355_csr_desc* A, int source, int max_iters, int check_interval,
                     float* dist, mi355_bf_csr_stats* stats, mi355_bf_progress_cb cb,
                     void* user_data, void* stream);

/* Replaces delta_stepping_minplus_hip_csr (DS/src/delta_stepping_hip_CSR.hpp:15-48,
 * delta_stepping_hip_CSR.cpp:648-815): incoming CSR, finite delta > 0,
 * non-negative weights; iterations_used = buckets processed, converged when no
 * bucket remains.  max_iters < 0: unbounded. */
int mi355_delta_stepping_csr_f32(const mi355_csr_desc* A, int source, float delta,
                                 int max_iters, float* dist, int* iterations_used,
                                 int* converged, mi355_bf_progress_cb cb, void* user_data,
                                 void* stream);

/* Options / result of SSSPOptions / SSSPResult (ONE/sssp_hip.hpp:21-47). */
typedef struct {
  int max_iters;
  int num_buckets;
  int bucket_capacity_per_bucket;
  float initial_delta;
  int dynamic_delta;
  int target_low_work;
  int target_high_work;
} mi355_one_options;
typedef struct {
  int iterations_used;
  int converged;
  int has_negative_cycle;
  unsigned long long work_count;
  float milliseconds;
  float final_delta;
} mi355_one_result;

/* Replaces adds_sssp_hip (ONE/sssp_hip.hpp:61-64, sssp_hip.cpp:428-595) and
 * bellman_ford_sssp_hip (ONE/sssp_hip.hpp:56-59, sssp_hip.cpp:350-426):
 * outgoing host CSR with int offsets; dist is a host array of n floats.
 * algorithm: MI355_ALG_ADDS or MI355_ALG_BF (the latter also detects a
 * negative cycle reachable from the source). */
int mi355_one_sssp_f32(int32_t algorithm, int n, int m, const int* row_offsets,
                       const int* col_indices, const float* weights, int source,
                       const mi355_one_options* options, float* dist, mi355_one_result* result,
                       void* stream);

/* Options / RunStats of hip_adds (JAS/hip_adds_sssp.hpp:28-87). */
typedef struct {
  uint32_t worker_blocks;
  uint32_t assignment_chunk;
  uint32_t initial_active_buckets;
  uint32_t max_active_buckets;
  double initial_delta;
  double initial_delta_scale;
  double min_delta;
  double delta_growth;
  double delta_shrink;
  double clip_threshold;
  double utilization_low;
  double utilization_high;
  uint32_t adjust_period_head_switches;
  int32_t dynamic_delta;
  int32_t dynamic_active_buckets;
  int32_t initialize_distances;
  uint64_t pool_words;
  double memory_fraction;
} mi355_adds_options;
typedef struct {
  double initial_delta;
  double final_delta;
  double final_bucket_base;
  uint32_t worker_blocks;
  uint32_t final_active_buckets;
  uint32_t pool_pages;
  uint32_t device_error;
  uint64_t assigned_items;
  uint64_t processed_items;
  uint64_t enqueued_items;
  uint64_t edge_relaxations;
  uint64_t successful_relaxations;
  uint64_t head_switches;
  uint64_t delta_updates;
  uint64_t max_inflight_items;
} mi355_adds_stats;

/* Replaces hip_adds::sssp for float and uint32_t weights
 * (JAS/hip_adds_sssp.hpp:99-113, hip_adds_sssp.cpp:1138-1301): device CSR with
 * uint32 offsets/indices, caller-owned device distances (INF = +inf for float,
 * 0xFFFFFFFF for uint32).  Returns the hip_adds::Status value (0 = Success,
 * 1 InvalidArgument, 2 HipError, 3 OutOfMemory, 4 PoolExhausted,
 * 5 BucketOverflow, 6 DeviceError). */
int mi355_hip_adds_sssp(int32_t weight_type, uint32_t num_vertices, uint32_t num_edges,
                        const uint32_t* d_row_offsets, const uint32_t* d_col_indices,
                        const void* d_weights, uint32_t source, void* d_distances,
                        mi355_adds_stats* stats, const mi355_adds_options* options,
                        void* stream);

/* ---- batched multi-source queries --------------------------------------
 * No reference counterpart (the routing engine's nets; nearest reference code:
 * the multi-column min-plus products of MP/src/minplus_sparse_hip.cpp:410-764
 * and MP/src/minplus_hip.cpp:122-173).  Query q: sources
 * sources[src_offsets[q] .. src_offsets[q+1]) (global vertex ids, all at
 * distance 0) and the rectangle boxes[4q .. 4q+4) = row0, col0, row1, col1
 * (half-open) of a grid graph whose vertex v is row * grid_cols + col; only
 * vertices inside the rectangle are reached.  Outputs per query are the
 * rectangle's vertices in row-major order, queries concatenated: query q at
 * [out_offsets[q], out_offsets[q+1]) (mi355_batch_layout). */
typedef struct {
  int32_t grid_cols;
  int32_t count;
  const int32_t* src_offsets;  /* host, count + 1 */
  const int32_t* sources;      /* host */
  const int32_t* boxes;        /* host, 4 per query */
} mi355_query_batch;

typedef struct {
  int32_t rounds_per_graph;    /* rounds per captured graph (even; 0 = 24) */
  int32_t blocks_per_cu;       /* relaxation workgroups per CU (0 = 4; the total is rounded
                                  up to a multiple of 8, one share per XCD part) */
  int32_t use_graph;           /* 1: replay a captured HIP graph, 0: plain launches */
  int32_t inf_mode;            /* MI355_INF_IEEE or MI355_INF_FLT_MAX (f32 outputs) */
  int32_t count_edges;         /* count edge visits (stats) */
  int32_t order;               /* label layout: 0 input order, 1 by rectangle position */
  float delta_scale;           /* near-far bucket width = delta_scale x the graph's heuristic
                                  delta0 (0 = 1); negative: Bellman-Ford rounds (no far pile) */
  int32_t reserved;
  int64_t queue_capacity;      /* frontier entries per queue, all parts together, split
                                  in proportion to the parts' areas (0 = auto) */
} mi355_batch_options;

typedef struct {
  int64_t label_slots;         /* sum of rectangle areas */
  int64_t sources;
  int32_t rounds;              /* rounds with a non-empty frontier */
  int32_t graph_replays;
  int32_t rounds_per_graph;
  int32_t far_passes;          /* rounds that moved the far pile (threshold advances) */
  int64_t edges_relaxed;       /* when count_edges */
  int64_t queue_capacity;      /* the capacity asked for (or chosen), before the split */
  float device_ms;             /* initialisation + all rounds (hipEvent) */
  int32_t stopped_queries;     /* mi355_sssp_batched_targets: queries that settled all their
                                  targets and stopped before the batch ended */
} mi355_batch_stats;

void mi355_batch_options_default(mi355_batch_options* o);
/* out_offsets: host array of count + 1 (int64): output offsets per query. */
int mi355_batch_layout(const mi355_graph* g, const mi355_query_batch* qb, int64_t* out_offsets);
/* d_dist: device array of label_slots 32-bit distances (fp32 or uint32 per the
 * graph); d_pred: device int32 parents (global vertex ids, -1 for sources and
 * unreached) or NULL.  Blocking. */
int mi355_sssp_batched(mi355_graph* g, const mi355_query_batch* qb, const mi355_batch_options* o,
                       void* d_dist, int32_t* d_pred, mi355_batch_stats* stats, void* stream);
/* mi355_sssp_batched with targets: query q stops once every vertex in
 * targets[target_offsets[q] .. target_offsets[q+1]) (global ids, inside q's
 * rectangle) is settled, i.e. once the batch's near-far threshold has passed
 * all of their distances.  target_offsets: host array of count + 1 (int32,
 * starting at 0); targets: host int32.  A query with no targets runs to the
 * end.  Outputs of a query that stopped hold the exact distance and parent for
 * every vertex whose distance is below the threshold it stopped at (its
 * targets among them) and the unreached value (INF, parent -1) for the rest.
 * Same batching and options as mi355_sssp_batched; no reference counterpart
 * (the reference's routing_api runs every query to completion and reads the
 * sinks afterwards, GRAPH/routing_api.cpp). */
int mi355_sssp_batched_targets(mi355_graph* g, const mi355_query_batch* qb, const int32_t* target_offsets,
                               const int32_t* targets, const mi355_batch_options* o, void* d_dist,
                               int32_t* d_pred, mi355_batch_stats* stats, void* stream);
/* Paths out of a batched run's outputs, read on the device: target i is
 * vertex target[i] (global id) of query query[i] (host arrays); d_dist /
 * d_pred and out_offsets are that run's outputs and mi355_batch_layout's
 * offsets.  path_offsets (host, count + 1) receives the offsets of each path:
 * its length is the number of vertices from the source to the target, 0 when
 * the target is unreached.  target_dist (host, count, may be NULL) receives
 * each target's 32-bit distance.  path (host, may be NULL): the vertices of
 * every path, source first; path_capacity must hold path_offsets[count]
 * (MI355_ECAPACITY otherwise; call with path = NULL first to size it).
 * Blocking.  No reference counterpart: the routing consumer walks each sink's
 * parents without downloading whole output blocks. */
int mi355_batch_paths(mi355_graph* g, const mi355_query_batch* qb, const int64_t* out_offsets,
                      const void* d_dist, const int32_t* d_pred, const int32_t* query, const int32_t* target,
                      int64_t count, int64_t* path_offsets, uint32_t* target_dist, int32_t* path,
                      int64_t path_capacity, void* stream);
/* Autotuner: times every (order, delta_scale, blocks_per_cu) candidate, order
 * in {0, 1} x delta_scale in {0.5, 1, 2} x blocks_per_cu in {2, 4, 8}, on every
 * sample_stride-th query of the batch and writes the fastest into *best (other
 * fields of *best are kept).  timings_ms (host, may be NULL) receives
 * MI355_BATCH_TUNE_CANDIDATES device times in candidate order. */
#define MI355_BATCH_TUNE_CANDIDATES 18
int mi355_batch_autotune(mi355_graph* g, const mi355_query_batch* qb, int32_t sample_stride,
                         mi355_batch_options* best, double* timings_ms, void* stream);

/* ---- learned delta policy ---------------------------------------------
 * No reference counterpart: the network replaces the fixed per-phase delta
 * heuristics (JAS/hip_adds_sssp.cpp:683-727, ONE/sssp_hip.cpp:564-571,
 * ADDSF/wl.h:349-410).  32 phase features -> 64 hidden (ReLU) -> 4 outputs:
 * log2 of the bucket width relative to the heuristic delta0, the number of
 * buckets served at once, log2 of the share of the persistent workgroups
 * that claim work, and (read once, before the first claim) log2 of the light
 * edges' width relative to the first bucket width.  The engine evaluates
 * bf16-rounded weights with
 * fp32 accumulation; training keeps fp32 master weights. */
typedef struct {
  float w1[64 * 32];       /* hidden x input, row-major */
  float b1[64];
  float w2[4 * 64];        /* output x hidden, row-major */
  float b2[4];
  float feat_mean[32];     /* inputs are normalised (x - mean) * inv_std; not trained */
  float feat_inv_std[32];
  /* Range of the 14 graph features (inputs 0..13) over the training graphs.
   * A graph with a feature outside [lo, hi] (1 % of the range as tolerance)
   * runs with the fixed heuristic delta instead of the network
   * (mi355_sssp_stats.policy = 0); a feature with hi <= lo is not bounded
   * (zero-initialised params bound nothing).  Not trained. */
  float graph_feat_lo[14];
  float graph_feat_hi[14];
} mi355_mlp_params;

typedef struct {
  int32_t epochs;
  int32_t batch;           /* samples per Adam step */
  float lr, beta1, beta2, eps, weight_decay;
  uint64_t seed;           /* sample order of each epoch */
} mi355_mlp_train_options;

void mi355_mlp_train_options_default(mi355_mlp_train_options* o);
/* x: n x 32 features (host), out: n x 4 outputs (host). */
int mi355_mlp_forward(const mi355_mlp_params* p, const float* x, int64_t n, float* out);
/* Mean loss (1/n) sum w_s |o_s - y_s|^2 and its gradient (normalisation
 * fields of grad are zero).  y: n x 4 targets, w: n sample weights or NULL. */
int mi355_mlp_loss_grad(const mi355_mlp_params* p, const float* x, const float* y, const float* w,
                        int64_t n, float* loss, mi355_mlp_params* grad);
/* Adam on the device; p is updated in place; epoch_loss (host, may be NULL)
 * receives the mean training loss of each epoch. */
int mi355_mlp_train(mi355_mlp_params* p, const float* x, const float* y, const float* w, int64_t n,
                    const mi355_mlp_train_options* o, float* epoch_loss);
/* Persistent workgroups an ADDS run launches when worker_blocks is 0. */
int mi355_adds_default_workers(int32_t* workers);
/* Load / drop the model used by mi355_sssp_options.use_mlp (process-wide). */
int mi355_adds_mlp_set(const mi355_mlp_params* p);
int mi355_adds_mlp_clear(void);
/* Explicit ADDS phase schedule for use_mlp = 2 (process-wide): phase i (0 =
 * the source's bucket) runs with entry min(i, count - 1) of table (count x 3
 * floats: bucket width as a multiple of the heuristic delta0, buckets served
 * at once in [1, 8], share of the persistent workgroups that claim in (0, 1]).
 * count 1..64; 0 clears it.  No reference counterpart: the settings the
 * learned policy chooses among, fixed per phase (training-label search). */
int mi355_adds_schedule_set(const float* table, int32_t count);

/* One ADDS run that records, at every phase boundary, MI355_ADDS_TRACE_STRIDE
 * floats: [0,32) the phase features (before the model's normalisation), 32
 * items queued in the window, 33 the delta
 * chosen for the next phase, 34 active buckets, 35 edges relaxed so far,
 * 36 pushes so far, 37 workgroups claiming work, 38 microseconds from the
 * run's start to this boundary.  trace: host array of
 * cap_phases records; phases receives
 * the number of phase boundaries (records beyond cap_phases are dropped). */
#define MI355_ADDS_TRACE_STRIDE 40
int mi355_adds_trace(mi355_graph* g, int32_t source, const mi355_sssp_options* opt, float* trace,
                     int32_t cap_phases, int32_t* phases, mi355_sssp_stats* stats, void* stream);

#ifdef __cplusplus
}
#endif
#endif
