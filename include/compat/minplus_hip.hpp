// Drop-in for MP/src/minplus_hip.hpp (minplus_gemm_f32, :24-33) over the
// mi355sssp C ABI.  Same contract: row-major, device pointers, asynchronous,
// hipErrorInvalidValue for null pointers / non-positive sizes / ld < dims.
#pragma once

#include <hip/hip_runtime.h>

#include "mi355_compat_detail.hpp"

inline hipError_t minplus_gemm_f32(const float* d_A, const float* d_B, float* d_C, int M, int N,
                                   int K, int lda, int ldb, int ldc, hipStream_t stream = nullptr) {
  return mi355_compat::to_hip_error(
      mi355_minplus_gemm_f32(d_A, d_B, d_C, M, N, K, lda, ldb, ldc, stream));
}
