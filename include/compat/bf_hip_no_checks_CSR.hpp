// Drop-in for BF/src/bf_hip_no_checks_CSR.hpp (:1-42): exactly max_iters
// iterations (< 0: n-1), no convergence test.  The engine stops relaxing once
// a round changes nothing (the remaining rounds would be no-ops), so the
// distances and iterations_used are those of the reference.
#pragma once

#include "bf_hip_CSR.hpp"

struct BellmanFordCsrNoChecksResult {
  std::vector<float> dist;
  int iterations_used = 0;
};

namespace mi355_compat {
inline BellmanFordCsrNoChecksResult no_checks(BellmanFordCsrResult&& r) {
  BellmanFordCsrNoChecksResult o;
  o.dist = std::move(r.dist);
  o.iterations_used = r.iterations_used;
  return o;
}
}  // namespace mi355_compat

inline BellmanFordCsrNoChecksResult bellman_ford_minplus_hip_csr_no_checks(
    const minplus_sparse::DeviceCsrF32& d_adjacency, int source, int max_iters,
    hipStream_t stream = nullptr, BellmanFordCsrProgressCallback progress_callback = nullptr,
    void* progress_user_data = nullptr) {
  return mi355_compat::no_checks(mi355_compat::run_bf_csr(mi355_compat::device_desc(d_adjacency),
                                                          source, max_iters, 0, stream,
                                                          progress_callback, progress_user_data));
}

inline BellmanFordCsrNoChecksResult bellman_ford_minplus_hip_csr_no_checks(
    const minplus_sparse::DeviceCsrF32& d_adjacency, int source, hipStream_t stream = nullptr) {
  return bellman_ford_minplus_hip_csr_no_checks(d_adjacency, source, (int)d_adjacency.rows - 1,
                                                stream, nullptr, nullptr);
}

inline BellmanFordCsrNoChecksResult bellman_ford_minplus_hip_csr_no_checks(
    const HostCsrF32& adjacency, int source, int max_iters, hipStream_t stream = nullptr,
    BellmanFordCsrProgressCallback progress_callback = nullptr,
    void* progress_user_data = nullptr) {
  return mi355_compat::no_checks(mi355_compat::run_bf_csr(mi355_compat::host_desc(adjacency),
                                                          source, max_iters, 0, stream,
                                                          progress_callback, progress_user_data));
}

inline BellmanFordCsrNoChecksResult bellman_ford_minplus_hip_csr_no_checks(
    const HostCsrF32& adjacency, int source, hipStream_t stream = nullptr) {
  return bellman_ford_minplus_hip_csr_no_checks(adjacency, source, (int)adjacency.rows - 1, stream,
                                                nullptr, nullptr);
}
