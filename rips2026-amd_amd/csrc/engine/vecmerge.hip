// Element-wise min of two sparse n x 1 distance vectors (CSR, absent = +INF).
//
// The reference's CSR Bellman-Ford merges d and A (x) d after every SpGEMM
// through bf_csr_detail::merge_distances_on_device, whose header is missing
// from the reference (called at BF/src/bf_hip_CSR.cpp:74-76,
// bf_hip_no_checks_CSR.cpp:72-74, bf_var_checks.cpp:117-121).  Contract
// inferred from those call sites: the result keeps, per row, the smaller of
// the two entries, and `changed` reports whether any entry of the second
// vector is strictly smaller than the first's (absent counts as +INF).
// Used only by the compatibility header restating that utility API.
#include <cstring>

#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "common.hpp"

namespace mi355 {

namespace {

__device__ inline uint32_t key_of(float x) {
  const uint32_t b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b ^ 0x80000000u);
}
__device__ inline float val_of(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k ^ 0x80000000u) : ~k);
}

__global__ void k_fill(uint32_t* __restrict__ a, uint32_t* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    a[i] = kF32InfKey;
    b[i] = kF32InfKey;
  }
}

// Scatter a vector's entries into dense keys (duplicates reduce by min; NaN
// entries are ignored like absent ones).
__global__ void k_scatter(int64_t rows, const int64_t* __restrict__ rp,
                          const float* __restrict__ val, uint32_t* __restrict__ dense) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
       r += (int64_t)gridDim.x * blockDim.x)
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
      const float x = val[e];
      if (x == x) atomicMin(dense + r, key_of(x));
    }
}

__global__ void k_merge(int64_t rows, const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                        uint32_t* __restrict__ out, int64_t* __restrict__ present,
                        int* __restrict__ changed) {
  int ch = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
       r += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t m = b[r] < a[r] ? b[r] : a[r];
    if (b[r] < a[r]) ch = 1;
    out[r] = m;
    present[r] = m != kF32InfKey ? 1 : 0;
  }
  if (ch) atomicExch(changed, 1);
}

__global__ void k_compact(int64_t rows, const uint32_t* __restrict__ keys,
                          const int64_t* __restrict__ rp, int32_t* __restrict__ col,
                          float* __restrict__ val) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
       r += (int64_t)gridDim.x * blockDim.x)
    if (rp[r + 1] > rp[r]) {
      col[rp[r]] = 0;
      val[rp[r]] = val_of(keys[r]);
    }
}

}  // namespace

void sparse_vector_min_merge(int64_t rows, const int64_t* arp, const float* aval,
                             const int64_t* brp, const float* bval, int64_t** c_rowptr,
                             int32_t** c_col, float** c_val, int64_t* c_nnz, int* changed,
                             hipStream_t s) {
  if (!c_rowptr || !c_col || !c_val || !c_nnz) fail(MI355_EINVAL, "null output pointer");
  *c_rowptr = nullptr;
  *c_col = nullptr;
  *c_val = nullptr;
  *c_nnz = 0;
  if (rows < 0 || !arp || !brp) fail(MI355_EINVAL, "invalid vectors");
  const unsigned grid = std::min<unsigned>(div_up((uint64_t)std::max<int64_t>(rows, 1), 256), 8192u);
  DevBuf<uint32_t> a((size_t)std::max<int64_t>(rows, 1)), b((size_t)std::max<int64_t>(rows, 1)),
      out((size_t)std::max<int64_t>(rows, 1));
  DevBuf<int64_t> present((size_t)rows + 1);
  DevBuf<int> flag(1);
  MI_HIP(hipMemsetAsync(flag.p, 0, 4, s));
  MI_HIP(hipMemsetAsync(present.p + rows, 0, 8, s));
  k_fill<<<grid, 256, 0, s>>>(a.p, b.p, rows);
  k_scatter<<<grid, 256, 0, s>>>(rows, arp, aval, a.p);
  k_scatter<<<grid, 256, 0, s>>>(rows, brp, bval, b.p);
  k_merge<<<grid, 256, 0, s>>>(rows, a.p, b.p, out.p, present.p, flag.p);
  MI_LAUNCH_CHECK();
  int64_t* rp = nullptr;
  MI_HIP(hipMalloc(reinterpret_cast<void**>(&rp), (size_t)(rows + 1) * 8));
  try {
    size_t temp_bytes = 0;
    MI_HIP(rocprim::exclusive_scan(nullptr, temp_bytes, present.p, rp, (int64_t)0, (size_t)rows + 1,
                                   rocprim::plus<int64_t>(), s));
    DevBuf<unsigned char> temp(temp_bytes);
    MI_HIP(rocprim::exclusive_scan(temp.p, temp_bytes, present.p, rp, (int64_t)0, (size_t)rows + 1,
                                   rocprim::plus<int64_t>(), s));
    int64_t nnz = 0;
    int ch = 0;
    MI_HIP(hipMemcpyAsync(&nnz, rp + rows, 8, hipMemcpyDeviceToHost, s));
    MI_HIP(hipMemcpyAsync(&ch, flag.p, 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (changed) *changed = ch;
    if (nnz > 0) {
      int32_t* col = nullptr;
      float* val = nullptr;
      MI_HIP(hipMalloc(reinterpret_cast<void**>(&col), (size_t)nnz * 4));
      if (hipMalloc(reinterpret_cast<void**>(&val), (size_t)nnz * 4) != hipSuccess) {
        (void)hipFree(col);
        fail(MI355_ENOMEM, "cannot allocate merged vector");
      }
      k_compact<<<grid, 256, 0, s>>>(rows, out.p, rp, col, val);
      MI_HIP(hipGetLastError());
      MI_HIP(hipStreamSynchronize(s));
      *c_col = col;
      *c_val = val;
    }
    *c_rowptr = rp;
    *c_nnz = nnz;
  } catch (...) {
    (void)hipFree(rp);
    throw;
  }
}

}  // namespace mi355
