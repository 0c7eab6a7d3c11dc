// Drop-in for HIP_kernel/faster_GPU_algo/Charlotte_oneshot/sssp_hip.hpp
// (:1-78) over the mi355sssp C ABI.  adds_sssp_hip runs the engine's
// persistent ADDS kernel with the same initial-delta heuristic
// (16 * mean positive weight / mean degree, sssp_hip.cpp:59-81); its
// work_count counts processed vertices (sssp_hip.cpp:286), and a bucket whose
// live entries exceed bucket_capacity_per_bucket fails the call with
// std::runtime_error (sssp_hip.cpp:548-556).  bellman_ford_sssp_hip runs
// in-place frontier Bellman-Ford (work_count: edges relaxed, :196) and reports
// a negative cycle reachable from the source the way its extra detection pass
// does (sssp_hip.cpp:204-220).  Distances are exact (fp32 left fold);
// iterations_used describes the engine's own schedule.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <vector>

#include "mi355_compat_detail.hpp"

struct CsrGraph {
  int num_nodes = 0;
  int num_edges = 0;
  std::vector<int> row_offsets;
  std::vector<int> col_indices;
  std::vector<float> weights;
};

struct SSSPOptions {
  int max_iters = -1;
  int num_buckets = 32;
  int bucket_capacity_per_bucket = 0;
  float initial_delta = 0.0f;
  bool dynamic_delta = true;
  int target_low_work = 8192;
  int target_high_work = 262144;
};

struct SSSPResult {
  std::vector<float> dist;
  int iterations_used = 0;
  bool converged = false;
  bool has_negative_cycle = false;
  unsigned long long work_count = 0;
  float milliseconds = 0.0f;
  float final_delta = 0.0f;
};

namespace mi355_compat {
inline SSSPResult run_one(int algorithm, const CsrGraph& g, int source, const SSSPOptions& o,
                          hipStream_t stream) {
  if (g.num_nodes <= 0) throw std::invalid_argument("graph must have at least one node");
  if ((int)g.row_offsets.size() != g.num_nodes + 1 || (int)g.col_indices.size() != g.num_edges ||
      (int)g.weights.size() != g.num_edges)
    throw std::invalid_argument("CSR array sizes do not match num_nodes / num_edges");
  mi355_one_options opt{o.max_iters, o.num_buckets, o.bucket_capacity_per_bucket, o.initial_delta,
                        o.dynamic_delta ? 1 : 0, o.target_low_work, o.target_high_work};
  SSSPResult r;
  r.dist.resize((size_t)g.num_nodes);
  mi355_one_result res{};
  throw_on_error(mi355_one_sssp_f32(algorithm, g.num_nodes, g.num_edges, g.row_offsets.data(),
                                    g.col_indices.data(), g.weights.data(), source, &opt,
                                    r.dist.data(), &res, stream));
  r.iterations_used = res.iterations_used;
  r.converged = res.converged != 0;
  r.has_negative_cycle = res.has_negative_cycle != 0;
  r.work_count = res.work_count;
  r.milliseconds = res.milliseconds;
  r.final_delta = res.final_delta;
  return r;
}
}  // namespace mi355_compat

// dense adj[u*n + v] = w(u -> v) (outgoing) -> CSR; the diagonal and
// non-finite entries are not edges, a negative off-diagonal weight is an
// error (the contract of ONE/sssp_hip.cpp:320-348).  Two passes: the first
// validates and sizes every row, the second fills the arrays in place.
inline CsrGraph dense_to_csr_outgoing(const std::vector<float>& adjacency, int n) {
  if (n <= 0) throw std::invalid_argument("n must be positive");
  const size_t nn = (size_t)n * (size_t)n;
  if (adjacency.size() != nn) throw std::invalid_argument("adjacency must contain n*n row-major values");
  auto is_edge = [&](size_t i) { return i % ((size_t)n + 1) != 0 && std::isfinite(adjacency[i]); };
  CsrGraph g;
  g.num_nodes = n;
  g.row_offsets.resize((size_t)n + 1);
  size_t total = 0;
  for (size_t i = 0; i < nn; ++i) {
    if (i % (size_t)n == 0) g.row_offsets[i / (size_t)n] = (int)total;
    if (!is_edge(i)) continue;
    if (adjacency[i] < 0.0f) throw std::invalid_argument("negative edge weight in dense adjacency");
    ++total;
  }
  g.row_offsets[(size_t)n] = (int)total;
  g.col_indices.resize(total);
  g.weights.resize(total);
  size_t k = 0;
  for (size_t i = 0; i < nn; ++i) {
    if (!is_edge(i)) continue;
    g.col_indices[k] = (int)(i % (size_t)n);
    g.weights[k] = adjacency[i];
    ++k;
  }
  g.num_edges = (int)total;
  return g;
}

inline SSSPResult bellman_ford_sssp_hip(const CsrGraph& graph, int source,
                                        const SSSPOptions& options = {},
                                        hipStream_t stream = nullptr) {
  return mi355_compat::run_one(MI355_ALG_BF, graph, source, options, stream);
}

inline SSSPResult adds_sssp_hip(const CsrGraph& graph, int source,
                                const SSSPOptions& options = {}, hipStream_t stream = nullptr) {
  return mi355_compat::run_one(MI355_ALG_ADDS, graph, source, options, stream);
}

inline SSSPResult bellman_ford_sssp_hip(const std::vector<float>& adjacency, int n, int source,
                                        const SSSPOptions& options = {},
                                        hipStream_t stream = nullptr) {
  return bellman_ford_sssp_hip(dense_to_csr_outgoing(adjacency, n), source, options, stream);
}

inline SSSPResult adds_sssp_hip(const std::vector<float>& adjacency, int n, int source,
                                const SSSPOptions& options = {}, hipStream_t stream = nullptr) {
  return adds_sssp_hip(dense_to_csr_outgoing(adjacency, n), source, options, stream);
}
