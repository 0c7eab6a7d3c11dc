// Re-specification of the reference's missing bf_hip_CSR_device_utils.hpp
// (included by BF/src/bf_hip_CSR.cpp:23, bf_hip_no_checks_CSR.cpp:23,
// bf_var_checks.cpp:17, tests/test_check_intervals.cpp:19) from its call
// sites (bf_hip_CSR.cpp:36-99, bf_var_checks.cpp:52-167,
// test_check_intervals.cpp:329-344), implemented over the mi355sssp C ABI.
//
// Distance vectors are n x 1 CSR matrices whose absent entries are +INF;
// sparse_minplus_relax is the SpGEMM A (x) d and merge_distances_on_device the
// element-wise min.  The engine's own CSR Bellman-Ford does not go through
// these (it relaxes frontiers in place); they exist for callers of this API.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bf_hip_CSR.hpp"

namespace bf_csr_detail {

using minplus_sparse::DeviceCsrF32;
using minplus_sparse::Index;
using minplus_sparse::Offset;

inline void check_hip(hipError_t status, const char* message) {
  if (status != hipSuccess)
    throw std::runtime_error(std::string(message) + ": " + hipGetErrorString(status));
}

// Owner of three device arrays (hipMalloc'ed; freed with hipFree) and the
// non-owning view over them.
struct DeviceCsrOwner {
  Offset* rowptr = nullptr;
  Index* colind = nullptr;
  float* values = nullptr;
  DeviceCsrF32 view{};

  DeviceCsrOwner() = default;
  DeviceCsrOwner(const DeviceCsrOwner&) = delete;
  DeviceCsrOwner& operator=(const DeviceCsrOwner&) = delete;
  DeviceCsrOwner(DeviceCsrOwner&& o) noexcept { take(o); }
  DeviceCsrOwner& operator=(DeviceCsrOwner&& o) noexcept {
    if (this != &o) {
      reset();
      take(o);
    }
    return *this;
  }
  ~DeviceCsrOwner() { reset(); }

  void reset() noexcept {
    if (rowptr) (void)hipFree(rowptr);
    if (colind) (void)hipFree(colind);
    if (values) (void)hipFree(values);
    rowptr = nullptr;
    colind = nullptr;
    values = nullptr;
    view = DeviceCsrF32{};
  }

  void adopt(Offset rows, Offset cols, Offset nnz, Offset* rp, Index* ci, float* v) {
    reset();
    rowptr = rp;
    colind = ci;
    values = v;
    view = DeviceCsrF32{rows, cols, nnz, rp, ci, v};
  }

 private:
  void take(DeviceCsrOwner& o) {
    rowptr = o.rowptr;
    colind = o.colind;
    values = o.values;
    view = o.view;
    o.rowptr = nullptr;
    o.colind = nullptr;
    o.values = nullptr;
    o.view = DeviceCsrF32{};
  }
};

inline void validate_device_csr_adjacency(const DeviceCsrF32& a, int source) {
  if (a.rows <= 0) throw std::invalid_argument("CSR graph must contain at least one vertex");
  if (a.rows != a.cols) throw std::invalid_argument("SSSP expects a square CSR adjacency matrix");
  if (a.nnz < 0) throw std::invalid_argument("CSR nnz must be nonnegative");
  if (source < 0 || source >= a.rows) throw std::invalid_argument("source is outside graph");
  if (a.rowptr == nullptr || (a.nnz > 0 && (a.colind == nullptr || a.values == nullptr)))
    throw std::invalid_argument("device CSR arrays must not be null");
}

inline void validate_host_csr_adjacency(const HostCsrF32& a, int source) {
  if (a.rows <= 0) throw std::invalid_argument("CSR graph must contain at least one vertex");
  if (a.rows != a.cols) throw std::invalid_argument("SSSP expects a square CSR adjacency matrix");
  if (source < 0 || source >= a.rows) throw std::invalid_argument("source is outside graph");
  if (a.rowptr.size() != (size_t)a.rows + 1 || a.colind.size() != (size_t)a.nnz ||
      a.values.size() != (size_t)a.nnz)
    throw std::invalid_argument("CSR array sizes do not match rows / nnz");
  if (a.rowptr.front() != 0 || a.rowptr.back() != a.nnz)
    throw std::invalid_argument("CSR rowptr must start at 0 and end at nnz");
  for (size_t r = 0; r + 1 < a.rowptr.size(); ++r)
    if (a.rowptr[r + 1] < a.rowptr[r]) throw std::invalid_argument("CSR rowptr must be monotone");
  for (Index c : a.colind)
    if (c < 0 || c >= a.cols) throw std::invalid_argument("CSR column index out of range");
}

inline DeviceCsrOwner copy_host_csr_to_device(const HostCsrF32& h, hipStream_t stream) {
  DeviceCsrOwner d;
  Offset* rp = nullptr;
  Index* ci = nullptr;
  float* v = nullptr;
  check_hip(hipMalloc(reinterpret_cast<void**>(&rp), (h.rowptr.size()) * sizeof(Offset)), "hipMalloc rowptr");
  d.adopt(h.rows, h.cols, h.nnz, rp, nullptr, nullptr);
  if (h.nnz > 0) {
    check_hip(hipMalloc(reinterpret_cast<void**>(&ci), (size_t)h.nnz * sizeof(Index)), "hipMalloc colind");
    d.colind = ci;
    check_hip(hipMalloc(reinterpret_cast<void**>(&v), (size_t)h.nnz * sizeof(float)), "hipMalloc values");
    d.values = v;
    d.view.colind = ci;
    d.view.values = v;
    check_hip(hipMemcpyAsync(ci, h.colind.data(), (size_t)h.nnz * sizeof(Index),
                             hipMemcpyHostToDevice, stream), "copy colind");
    check_hip(hipMemcpyAsync(v, h.values.data(), (size_t)h.nnz * sizeof(float),
                             hipMemcpyHostToDevice, stream), "copy values");
  }
  check_hip(hipMemcpyAsync(rp, h.rowptr.data(), h.rowptr.size() * sizeof(Offset),
                           hipMemcpyHostToDevice, stream), "copy rowptr");
  check_hip(hipStreamSynchronize(stream), "sync graph copy");
  return d;
}

// rows x 1 vector with the single entry (source, 0) = 0.0f.
inline DeviceCsrOwner make_device_single_source_vector(Offset rows, int source, hipStream_t stream) {
  if (rows <= 0 || source < 0 || source >= rows)
    throw std::invalid_argument("source is outside graph");
  std::vector<Offset> rp((size_t)rows + 1, 0);
  for (Offset r = source + 1; r <= rows; ++r) rp[(size_t)r] = 1;
  HostCsrF32 h;
  h.rows = rows;
  h.cols = 1;
  h.nnz = 1;
  h.rowptr = std::move(rp);
  h.colind = {0};
  h.values = {0.0f};
  return copy_host_csr_to_device(h, stream);
}

// relaxed = A (x) d (sparse min-plus product; absent/+INF/NaN ignored).
inline DeviceCsrOwner sparse_minplus_relax(const DeviceCsrF32& A, const DeviceCsrF32& d,
                                           hipStream_t stream) {
  Offset* rp = nullptr;
  Index* ci = nullptr;
  float* v = nullptr;
  Offset nnz = 0;
  const hipError_t e = minplus_sparse::minplus_spgemm_csr_f32_large_rows(A, d, &rp, &ci, &v, &nnz,
                                                                         stream);
  check_hip(e, "sparse min-plus relaxation");
  DeviceCsrOwner out;
  out.adopt(A.rows, d.cols, nnz, rp, ci, v);
  return out;
}

// next = min(d, r); *changed (when non-null) = some entry of r beats d.
inline DeviceCsrOwner merge_distances_on_device(const DeviceCsrF32& d, const DeviceCsrF32& r,
                                                int* changed, hipStream_t stream) {
  Offset* rp = nullptr;
  Index* ci = nullptr;
  float* v = nullptr;
  Offset nnz = 0;
  int ch = 0;
  mi355_compat::throw_on_error(mi355_sparse_vector_min_merge(
      d.rows, d.rowptr, d.values, r.rowptr, r.values, &rp, &ci, &v, &nnz, &ch, stream));
  if (changed) *changed = ch;
  DeviceCsrOwner out;
  out.adopt(d.rows, 1, nnz, rp, ci, v);
  return out;
}

// Dense host copy of a distance vector (+INF where absent).
inline std::vector<float> copy_dist_vector_to_host(const DeviceCsrF32& d, hipStream_t stream) {
  std::vector<Offset> rp((size_t)d.rows + 1);
  std::vector<float> vals((size_t)d.nnz);
  check_hip(hipMemcpyAsync(rp.data(), d.rowptr, rp.size() * sizeof(Offset), hipMemcpyDeviceToHost,
                           stream), "copy vector rowptr");
  if (d.nnz > 0)
    check_hip(hipMemcpyAsync(vals.data(), d.values, vals.size() * sizeof(float),
                             hipMemcpyDeviceToHost, stream), "copy vector values");
  check_hip(hipStreamSynchronize(stream), "sync vector copy");
  std::vector<float> out((size_t)d.rows, std::numeric_limits<float>::infinity());
  for (Offset r = 0; r < d.rows; ++r)
    for (Offset e = rp[(size_t)r]; e < rp[(size_t)r + 1]; ++e)
      if (vals[(size_t)e] < out[(size_t)r]) out[(size_t)r] = vals[(size_t)e];
  return out;
}

}  // namespace bf_csr_detail
