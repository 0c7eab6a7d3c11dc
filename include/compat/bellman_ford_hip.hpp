// Drop-in for BF/src/bellman_ford_hip.hpp (:8-59): dense incoming-matrix
// Bellman-Ford, 8 overloads, over the mi355sssp C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mi355_compat_detail.hpp"

struct BellmanFordResult {
  std::vector<float> dist;
  int iterations_used = 0;
  bool converged = false;
};

namespace mi355_compat {
inline BellmanFordResult run_dense_bf(const float* adj, int memory, int n, int source,
                                      int max_iters, hipStream_t stream) {
  BellmanFordResult r;
  r.dist.resize(n > 0 ? (size_t)n : 0);
  int it = 0, conv = 0;
  throw_on_error(mi355_bf_dense_f32(adj, memory, n, source, max_iters, r.dist.data(), &it, &conv,
                                    stream));
  r.iterations_used = it;
  r.converged = conv != 0;
  return r;
}
}  // namespace mi355_compat

inline BellmanFordResult bellman_ford_minplus_hip(const float* d_adjacency, int n, int source,
                                                  int max_iters, hipStream_t stream) {
  return mi355_compat::run_dense_bf(d_adjacency, MI355_DEVICE, n, source, max_iters, stream);
}
inline BellmanFordResult bellman_ford_minplus_hip(const float* d_adjacency, int n, int source,
                                                  hipStream_t stream) {
  return bellman_ford_minplus_hip(d_adjacency, n, source, n - 1, stream);
}
inline BellmanFordResult bellman_ford_minplus_hip(const float* d_adjacency, int n, int source,
                                                  int max_iters) {
  return bellman_ford_minplus_hip(d_adjacency, n, source, max_iters, nullptr);
}
inline BellmanFordResult bellman_ford_minplus_hip(const float* d_adjacency, int n, int source) {
  return bellman_ford_minplus_hip(d_adjacency, n, source, n - 1, nullptr);
}
inline BellmanFordResult bellman_ford_minplus_hip(const std::vector<float>& adjacency, int n,
                                                  int source, int max_iters, hipStream_t stream) {
  if (n > 0 && adjacency.size() != (size_t)n * (size_t)n)
    throw std::invalid_argument("adjacency must hold n*n values");
  return mi355_compat::run_dense_bf(adjacency.data(), MI355_HOST, n, source, max_iters, stream);
}
inline BellmanFordResult bellman_ford_minplus_hip(const std::vector<float>& adjacency, int n,
                                                  int source, hipStream_t stream) {
  return bellman_ford_minplus_hip(adjacency, n, source, n - 1, stream);
}
inline BellmanFordResult bellman_ford_minplus_hip(const std::vector<float>& adjacency, int n,
                                                  int source, int max_iters) {
  return bellman_ford_minplus_hip(adjacency, n, source, max_iters, nullptr);
}
inline BellmanFordResult bellman_ford_minplus_hip(const std::vector<float>& adjacency, int n,
                                                  int source) {
  return bellman_ford_minplus_hip(adjacency, n, source, n - 1, nullptr);
}
