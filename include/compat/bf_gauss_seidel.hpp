// Header for the reference's Gauss-Seidel Bellman-Ford, which ships without
// one (BF/src/bf_gauss_seidel.cpp:66-260): result struct + 8 overloads over the
// mi355sssp C ABI.  Exactly max_iters (< 0: n-1) ordered sweeps; no
// convergence test; has_negative_cycle is never set (as in the reference).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mi355_compat_detail.hpp"

struct BellmanFordGaussSeidelResult {
  std::vector<float> dist;
  int iterations_used = 0;
  bool converged = false;
  bool has_negative_cycle = false;
};

namespace mi355_compat {
inline BellmanFordGaussSeidelResult run_gs(const float* adj, int memory, int n, int source,
                                           int max_iters, hipStream_t stream) {
  BellmanFordGaussSeidelResult r;
  r.dist.resize(n > 0 ? (size_t)n : 0);
  int it = 0;
  throw_on_error(
      mi355_bf_gauss_seidel_f32(adj, memory, n, source, max_iters, r.dist.data(), &it, stream));
  r.iterations_used = it;
  return r;
}
}  // namespace mi355_compat

inline BellmanFordGaussSeidelResult bellman_ford_minplus_hip_gauss_seidel(
    const float* d_adjacency, int n, int source, int max_iters, hipStream_t stream) {
  return mi355_compat::run_gs(d_adjacency, MI355_DEVICE, n, source, max_iters, stream);
}
inline BellmanFordGaussSeidelResult bellman_ford_minplus_hip_gauss_seidel(
    const float* d_adjacency, int n, int source, hipStream_t stream) {
  return bellman_ford_minplus_hip_gauss_seidel(d_adjacency, n, source, n - 1, stream);
}
inline BellmanFordGaussSeidelResult bellman_ford_minplus_hip_gauss_seidel(
    const float* d_adjacency, int n, int source, int max_iters) {
  return bellman_ford_minplus_hip_gauss_seidel(d_adjacency, n, source, max_iters, nullptr);
}
inline BellmanFordGaussSeidelResult bellman_ford_minplus_hip_gauss_seidel(
    const float* d_adjacency, int n, int source) {
  return bellman_ford_minplus_hip_gauss_seidel(d_adjacency, n, source, n - 1, nullptr);
}
inline BellmanFordGaussSeidelResult bellman_ford_minplus_hip_gauss_seidel(
    const std::vector<float>& adjacency, int n, int source, int max_iters, hipStream_t stream) {
  if (n > 0 && adjacency.size() != (size_t)n * (size_t)n)
    throw std::invalid_argument("adjacency must hold n*n values");
  return mi355_compat::run_gs(adjacency.data(), MI355_HOST, n, source, max_iters, stream);
}
inline BellmanFordGaussSeidelResult bellman_ford_minplus_hip_gauss_seidel(
    const std::vector<float>& adjacency, int n, int source, hipStream_t stream) {
  return bellman_ford_minplus_hip_gauss_seidel(adjacency, n, source, n - 1, stream);
}
inline BellmanFordGaussSeidelResult bellman_ford_minplus_hip_gauss_seidel(
    const std::vector<float>& adjacency, int n, int source, int max_iters) {
  return bellman_ford_minplus_hip_gauss_seidel(adjacency, n, source, max_iters, nullptr);
}
inline BellmanFordGaussSeidelResult bellman_ford_minplus_hip_gauss_seidel(
    const std::vector<float>& adjacency, int n, int source) {
  return bellman_ford_minplus_hip_gauss_seidel(adjacency, n, source, n - 1, nullptr);
}
