// Drop-in for BF/src/bf_hip_CSR.hpp (:1-66) over the mi355sssp C ABI.
// Matrix convention (unchanged): adjacency row v, column u = w(u -> v).
// Iteration semantics are the reference's Jacobi rounds
// d_k = min(d_{k-1}, A (x) d_{k-1}), checked every iteration; the engine runs
// them as frontier rounds with exact round-by-round equivalence.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "minplus_sparse_hip.hpp"
#include "mi355_compat_detail.hpp"

struct BellmanFordCsrProgress {
  int iteration = 0;
  int max_iters = 0;
  bool convergence_checked = false;
  bool changed = false;
};

using BellmanFordCsrProgressCallback =
    void (*)(const BellmanFordCsrProgress& progress, void* user_data);

struct BellmanFordCsrResult {
  std::vector<float> dist;
  int iterations_used = 0;
  bool converged = false;
};

struct HostCsrF32 {
  minplus_sparse::Offset rows = 0;
  minplus_sparse::Offset cols = 0;
  minplus_sparse::Offset nnz = 0;
  std::vector<minplus_sparse::Offset> rowptr;
  std::vector<minplus_sparse::Index> colind;
  std::vector<float> values;
};

namespace mi355_compat {

struct ProgressBridge {
  BellmanFordCsrProgressCallback cb;
  void* user;
  static void call(const mi355_bf_progress* p, void* self) {
    auto* b = static_cast<ProgressBridge*>(self);
    BellmanFordCsrProgress q;
    q.iteration = p->iteration;
    q.max_iters = p->max_iters;
    q.convergence_checked = p->convergence_checked != 0;
    q.changed = p->changed != 0;
    b->cb(q, b->user);
  }
};

inline mi355_csr_desc device_desc(const minplus_sparse::DeviceCsrF32& a) {
  mi355_csr_desc d{};
  d.rows = a.rows;
  d.cols = a.cols;
  d.nnz = a.nnz;
  d.rowptr = a.rowptr;
  d.col = a.colind;
  d.w = a.values;
  d.offset_bits = 64;
  d.weight_type = MI355_W_F32;
  d.orientation = MI355_IN;
  d.memory = MI355_DEVICE;
  return d;
}

inline mi355_csr_desc host_desc(const HostCsrF32& a) {
  if (a.rowptr.size() != (size_t)a.rows + 1 || a.colind.size() != (size_t)a.nnz ||
      a.values.size() != (size_t)a.nnz)
    throw std::invalid_argument("HostCsrF32 array sizes do not match rows / nnz");
  mi355_csr_desc d{};
  d.rows = a.rows;
  d.cols = a.cols;
  d.nnz = a.nnz;
  d.rowptr = a.rowptr.data();
  d.col = a.colind.data();
  d.w = a.values.data();
  d.offset_bits = 64;
  d.weight_type = MI355_W_F32;
  d.orientation = MI355_IN;
  d.memory = MI355_HOST;
  return d;
}

// check_interval: 1 = every iteration, 0 = never, k = every k-th iteration.
inline BellmanFordCsrResult run_bf_csr(const mi355_csr_desc& d, int source, int max_iters,
                                       int check_interval, hipStream_t stream,
                                       BellmanFordCsrProgressCallback cb, void* user,
                                       mi355_bf_csr_stats* stats = nullptr) {
  BellmanFordCsrResult r;
  r.dist.resize(d.rows > 0 ? (size_t)d.rows : 0);
  ProgressBridge bridge{cb, user};
  mi355_bf_csr_stats st{};
  const int status = mi355_bf_csr_f32(&d, source, max_iters, check_interval, r.dist.data(), &st,
                                      cb ? &ProgressBridge::call : nullptr, cb ? &bridge : nullptr,
                                      stream);
  if (status == MI355_ERANGE) throw std::invalid_argument(mi355_last_error());
  throw_on_error(status);
  r.iterations_used = st.iterations_used;
  r.converged = st.converged != 0;
  if (stats) *stats = st;
  return r;
}

}  // namespace mi355_compat

inline BellmanFordCsrResult bellman_ford_minplus_hip_csr(
    const minplus_sparse::DeviceCsrF32& d_adjacency, int source, int max_iters,
    hipStream_t stream = nullptr, BellmanFordCsrProgressCallback progress_callback = nullptr,
    void* progress_user_data = nullptr) {
  return mi355_compat::run_bf_csr(mi355_compat::device_desc(d_adjacency), source, max_iters, 1,
                                  stream, progress_callback, progress_user_data);
}

inline BellmanFordCsrResult bellman_ford_minplus_hip_csr(
    const minplus_sparse::DeviceCsrF32& d_adjacency, int source, hipStream_t stream = nullptr) {
  return bellman_ford_minplus_hip_csr(d_adjacency, source, (int)d_adjacency.rows - 1, stream,
                                      nullptr, nullptr);
}

inline BellmanFordCsrResult bellman_ford_minplus_hip_csr(
    const HostCsrF32& adjacency, int source, int max_iters, hipStream_t stream = nullptr,
    BellmanFordCsrProgressCallback progress_callback = nullptr,
    void* progress_user_data = nullptr) {
  return mi355_compat::run_bf_csr(mi355_compat::host_desc(adjacency), source, max_iters, 1, stream,
                                  progress_callback, progress_user_data);
}

inline BellmanFordCsrResult bellman_ford_minplus_hip_csr(const HostCsrF32& adjacency, int source,
                                                         hipStream_t stream = nullptr) {
  return bellman_ford_minplus_hip_csr(adjacency, source, (int)adjacency.rows - 1, stream, nullptr,
                                      nullptr);
}
