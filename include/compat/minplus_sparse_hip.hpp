// Drop-in for MP/src/minplus_sparse_hip.hpp (:1-152) over the mi355sssp C ABI.
// The tuned template accepts any (HASH_CAP, BLOCK_SIZE) with BLOCK_SIZE in
// {64, 128, 256}; the reference's five explicit instantiations are a subset.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mi355_compat_detail.hpp"

namespace minplus_sparse {

using Offset = std::int64_t;
using Index = int;

inline constexpr int kDefaultHashCapacity = 512;
inline constexpr int kDefaultBlockSize = 64;
inline constexpr int kLargeHashCapacity = 4096;
inline constexpr int kLargeBlockSize = 256;

struct DeviceCsrF32 {
  Offset rows;
  Offset cols;
  Offset nnz;
  const Offset* rowptr;
  const Index* colind;
  const float* values;
};

template <int HASH_CAP, int BLOCK_SIZE>
inline hipError_t minplus_spgemm_csr_f32_tuned(const DeviceCsrF32& A, const DeviceCsrF32& B,
                                               Offset** d_C_rowptr, Index** d_C_colind,
                                               float** d_C_values, Offset* h_C_nnz,
                                               hipStream_t stream = 0) {
  return mi355_compat::to_hip_error(mi355_minplus_spgemm_csr_f32(
      A.rows, A.cols, A.nnz, A.rowptr, A.colind, A.values, B.rows, B.cols, B.nnz, B.rowptr,
      B.colind, B.values, HASH_CAP, BLOCK_SIZE, d_C_rowptr, d_C_colind, d_C_values, h_C_nnz,
      stream));
}

inline hipError_t minplus_spgemm_csr_f32(const DeviceCsrF32& A, const DeviceCsrF32& B,
                                         Offset** d_C_rowptr, Index** d_C_colind,
                                         float** d_C_values, Offset* h_C_nnz,
                                         hipStream_t stream = 0) {
  return minplus_spgemm_csr_f32_tuned<kDefaultHashCapacity, kDefaultBlockSize>(
      A, B, d_C_rowptr, d_C_colind, d_C_values, h_C_nnz, stream);
}

inline hipError_t minplus_spgemm_csr_f32_large_rows(const DeviceCsrF32& A, const DeviceCsrF32& B,
                                                    Offset** d_C_rowptr, Index** d_C_colind,
                                                    float** d_C_values, Offset* h_C_nnz,
                                                    hipStream_t stream = 0) {
  return minplus_spgemm_csr_f32_tuned<kLargeHashCapacity, kLargeBlockSize>(
      A, B, d_C_rowptr, d_C_colind, d_C_values, h_C_nnz, stream);
}

}  // namespace minplus_sparse
