// SSSP drivers on a graph handle.
#pragma once

#include <vector>

#include "adds.hpp"
#include "graph.hpp"

namespace mi355 {

struct RunRequest {
  int algorithm = MI355_ALG_BF;
  int64_t max_iters = -1;  // <0: default
  int check_interval = 1;
  float delta = 0.0f;
  int rounds_per_sync = 0;
  bool want_work = false;
  int dynamic_delta = 1;
  int use_mlp = 0;
  int worker_blocks = 0;
  // Optional warm start: device array of n distances (output convention);
  // labels start from them (source forced to 0), only the source is active.
  const uint32_t* init_dist = nullptr;
  int active_buckets = 0;  // ADDS priority window (0: 1)
  int64_t bucket_capacity = 0;  // ADDS: most live entries per bucket (0: the ring size)
  float* trace = nullptr;  // ADDS phase trace (device, trace_cap x kTraceStride)
  uint32_t trace_cap = 0;
  // ADDS: 0 push every finished bucket's heavy edges; 1 pull them instead when
  // the push work estimate exceeds 2^pull_margin x the pull work estimate
  int pull = 0;
  float pull_margin = 0.0f;
  // light edges: weight <= light_scale x the first bucket width (0: 1, or the
  // network's choice under the MLP policy)
  float light_scale = 0.0f;
  // Output conversion enqueued by the driver right behind its kernels, before
  // it waits for them (ADDS does; the others leave it to the caller): labels
  // to distances / parents in input order and the reached counts (any of the
  // three may be null; all null: no conversion).  RunResult::converted says
  // whether it was enqueued.
  void* out_dist = nullptr;
  int32_t* out_pred = nullptr;
  unsigned long long* out_reach = nullptr;
  int out_inf_mode = 0;
};

struct RunResult {
  int64_t iterations_used = 0;
  bool converged = false;
  int64_t rounds = 0;
  int syncs = 0;
  int64_t edges = 0;
  int64_t vertices = 0;
  float initial_delta = 0.f;
  float final_delta = 0.f;
  float device_ms = 0.f;
  float kernel_ms = 0.f;   // the persistent kernel alone (ADDS)
  int policy = 0;          // ADDS: PolicyMode used
  int64_t pulls = 0;       // ADDS: finished buckets whose heavy edges were pulled
  bool negative_cycle = false;
  std::vector<uint32_t> frontier_log;  // BF: |F_r| for r = 0..rounds
  bool converted = false;  // the driver enqueued the request's output conversion
};

// Runs SSSP leaving packed labels in g.ws.labels.
void run_sssp(mi355_graph& g, uint32_t source, const RunRequest& req, RunResult& res,
              hipStream_t stream);

// Continues the last Bellman-Ford run on g for req.max_iters more rounds.
void run_bf_continue(mi355_graph& g, uint32_t source, const RunRequest& req, RunResult& res,
                     hipStream_t stream);

// Converts labels to distances (+ predecessors); counts reached vertices and
// the sum of their out-degrees when reach != nullptr (device: 2 x u64).
void labels_to_output(mi355_graph& g, uint32_t source, int inf_mode, void* d_dist, int32_t* d_pred,
                      unsigned long long* reach, hipStream_t stream);
// The same with the source given as its internal id.
void labels_to_output_internal(mi355_graph& g, uint32_t src, int inf_mode, void* d_dist, int32_t* d_pred,
                               unsigned long long* reach, hipStream_t stream);

// Delta-stepping driver (delta.hip).
// Streaming device-to-device copy (copy.hip).
void device_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);
// dst[i] = src[idx[i]] for 8-byte records, i < count (copy.hip).
void device_gather8(void* dst, const void* src, const uint32_t* idx, int64_t count, hipStream_t stream);

// Light edges (w <= delta) of every vertex, a prefix of its weight-sorted row;
// cached on the graph for the last delta asked for.
const int* light_split(mi355_graph& g, float delta, hipStream_t stream);
// The same for the in-edge rows of pulls (pull_rows must have run).
const int* pull_light_split(mi355_graph& g, float delta, hipStream_t stream);
void run_delta_stepping(mi355_graph& g, uint32_t source, const RunRequest& req, RunResult& res,
                        hipStream_t stream);

// Device policy object of an ADDS run (adds.hip): fixed, feedback heuristic
// or learned MLP, plus the optional phase trace (nullptr: fixed, no trace).
// *mode_out receives the policy used (PolicyMode): the MLP only when asked for
// and the graph lies in the model's training range.
const AddsPolicy* prepare_adds_policy(mi355_graph& g, const RunRequest& req, hipStream_t stream,
                                      int* mode_out);

// Batched multi-source queries (batch.hip).
void batch_layout(const mi355_query_batch& qb, int64_t n, std::vector<int64_t>& offsets);
// Parent chains of targets out of a batched run's outputs (see mi355_batch_paths)
void batch_paths(const mi355_query_batch& qb, const int64_t* out_offsets, const int32_t* d_pred,
                 const uint32_t* d_dist, const int32_t* query, const int32_t* target, int64_t count,
                 int64_t* path_offsets, uint32_t* target_dist, int32_t* path, int64_t path_capacity,
                 hipStream_t s);
// t_off / targets (host, may be null): per-query target vertices; a query
// stops once all of its targets are settled
void run_batch(mi355_graph& g, const mi355_query_batch& qb, const mi355_batch_options& o, const int32_t* t_off,
               const int32_t* targets, void* d_dist, int32_t* d_pred, mi355_batch_stats& st, hipStream_t stream);
void batch_autotune(mi355_graph& g, const mi355_query_batch& qb, int32_t stride, mi355_batch_options& best,
                    double* timings, hipStream_t stream);

// Process-wide delta model (policy.hip) and its training kernels (mlp.hip).
void set_mlp(const AddsPolicy& p, const float* feat_lo, const float* feat_hi);
void set_schedule(const float* table, int32_t count);
bool get_mlp(AddsPolicy* p, float* feat_lo = nullptr, float* feat_hi = nullptr);
void clear_mlp();
void set_mlp_params(const mi355_mlp_params* p);
void mlp_forward(const mi355_mlp_params* p, const float* h_x, int64_t n, float* h_out, hipStream_t s);
void mlp_loss_grad(const mi355_mlp_params* p, const float* h_x, const float* h_y, const float* h_w,
                   int64_t n, float* loss, mi355_mlp_params* grad, hipStream_t s);
void mlp_train(mi355_mlp_params* p, const float* h_x, const float* h_y, const float* h_w, int64_t n,
               const mi355_mlp_train_options& o, float* h_epoch_loss, hipStream_t s);

// Heuristic bucket width: 16 * mean positive weight / mean degree
// (ONE/sssp_hip.cpp:59-81, JAS/hip_adds_sssp.cpp:1121).
float heuristic_delta(const mi355_graph& g);

}  // namespace mi355
