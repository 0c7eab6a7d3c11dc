// C ABI of the sparse distance-vector merge (declared in include/mi355sssp.h).
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"

namespace mi355 {
void sparse_vector_min_merge(int64_t rows, const int64_t* arp, const float* aval,
                             const int64_t* brp, const float* bval, int64_t** c_rowptr,
                             int32_t** c_col, float** c_val, int64_t* c_nnz, int* changed,
                             hipStream_t s);
}

// Errors are reported through the library-wide last-error slot.
extern "C" int mi355_internal_set_error(int code, const char* msg, int hip);

extern "C" int mi355_sparse_vector_min_merge(int64_t rows, const int64_t* a_rowptr,
                                             const float* a_val, const int64_t* b_rowptr,
                                             const float* b_val, int64_t** c_rowptr,
                                             int32_t** c_col, float** c_val, int64_t* c_nnz,
                                             int* changed, void* stream) {
  try {
    mi355::sparse_vector_min_merge(rows, a_rowptr, a_val, b_rowptr, b_val, c_rowptr, c_col, c_val,
                                   c_nnz, changed, reinterpret_cast<hipStream_t>(stream));
    return MI355_OK;
  } catch (const mi355::Error& e) {
    return mi355_internal_set_error(e.code, e.what(), e.hip);
  } catch (const std::exception& e) {
    return mi355_internal_set_error(MI355_EINVAL, e.what(), 0);
  }
}
