// Drop-in for DS/src/delta_stepping_hip_CSR.hpp (:1-67) over the mi355sssp
// C ABI: incoming CSR, finite delta > 0, non-negative finite weights;
// iterations_used = buckets processed; std::out_of_range for a bad source,
// std::overflow_error when rows exceed int, std::invalid_argument otherwise.
#pragma once

#include "bf_hip_CSR.hpp"

using DeltaSteppingCsrProgress = BellmanFordCsrProgress;
using DeltaSteppingCsrProgressCallback = BellmanFordCsrProgressCallback;
using DeltaSteppingCsrResult = BellmanFordCsrResult;

namespace mi355_compat {
inline DeltaSteppingCsrResult run_ds(const mi355_csr_desc& d, int source, float delta,
                                     int max_iters, hipStream_t stream,
                                     DeltaSteppingCsrProgressCallback cb, void* user) {
  DeltaSteppingCsrResult r;
  r.dist.resize(d.rows > 0 ? (size_t)d.rows : 0);
  ProgressBridge bridge{cb, user};
  int it = 0, conv = 0;
  throw_on_error(mi355_delta_stepping_csr_f32(&d, source, delta, max_iters, r.dist.data(), &it,
                                              &conv, cb ? &ProgressBridge::call : nullptr,
                                              cb ? &bridge : nullptr, stream));
  r.iterations_used = it;
  r.converged = conv != 0;
  return r;
}
}  // namespace mi355_compat

inline DeltaSteppingCsrResult delta_stepping_minplus_hip_csr(
    const minplus_sparse::DeviceCsrF32& d_adjacency, int source, float delta, int max_iters,
    hipStream_t stream = nullptr, DeltaSteppingCsrProgressCallback progress_callback = nullptr,
    void* progress_user_data = nullptr) {
  return mi355_compat::run_ds(mi355_compat::device_desc(d_adjacency), source, delta, max_iters,
                              stream, progress_callback, progress_user_data);
}

inline DeltaSteppingCsrResult delta_stepping_minplus_hip_csr(
    const minplus_sparse::DeviceCsrF32& d_adjacency, int source, float delta,
    hipStream_t stream = nullptr) {
  return delta_stepping_minplus_hip_csr(d_adjacency, source, delta, -1, stream, nullptr, nullptr);
}

inline DeltaSteppingCsrResult delta_stepping_minplus_hip_csr(
    const HostCsrF32& adjacency, int source, float delta, int max_iters,
    hipStream_t stream = nullptr, DeltaSteppingCsrProgressCallback progress_callback = nullptr,
    void* progress_user_data = nullptr) {
  return mi355_compat::run_ds(mi355_compat::host_desc(adjacency), source, delta, max_iters, stream,
                              progress_callback, progress_user_data);
}

inline DeltaSteppingCsrResult delta_stepping_minplus_hip_csr(const HostCsrF32& adjacency,
                                                             int source, float delta,
                                                             hipStream_t stream = nullptr) {
  return delta_stepping_minplus_hip_csr(adjacency, source, delta, -1, stream, nullptr, nullptr);
}

inline DeltaSteppingCsrResult delta_stepping_hip_csr(
    const minplus_sparse::DeviceCsrF32& d_adjacency, int source, float delta, int max_iters,
    hipStream_t stream = nullptr, DeltaSteppingCsrProgressCallback progress_callback = nullptr,
    void* progress_user_data = nullptr) {
  return delta_stepping_minplus_hip_csr(d_adjacency, source, delta, max_iters, stream,
                                        progress_callback, progress_user_data);
}

inline DeltaSteppingCsrResult delta_stepping_hip_csr(
    const HostCsrF32& adjacency, int source, float delta, int max_iters,
    hipStream_t stream = nullptr, DeltaSteppingCsrProgressCallback progress_callback = nullptr,
    void* progress_user_data = nullptr) {
  return delta_stepping_minplus_hip_csr(adjacency, source, delta, max_iters, stream,
                                        progress_callback, progress_user_data);
}
