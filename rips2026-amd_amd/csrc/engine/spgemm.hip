// Sparse min-plus product of CSR matrices: C = A (min,+) B.
//
// Replaces minplus_sparse::minplus_spgemm_csr_f32[_large_rows|_tuned<HC,BS>]
// (MP/src/minplus_sparse_hip.cpp:170-764) with the same contract: missing
// entries are +INF, NaN / +INF candidates are not stored (:93-99), duplicate
// columns reduce by min, output columns are unsorted, the callee allocates the
// output (caller frees), an empty product has a zero rowptr and null
// colind/values, and a row with more than hash_capacity distinct columns is an
// error with every output freed.
//
// Layout on the device: one wave per output row (block_size / 64 rows per
// workgroup), each wave with a private open-addressing table in LDS of twice
// the capacity rounded to a power of two, so the load factor stays <= 1/2.
// Keys are claimed with an LDS atomicCAS; values are kept as order-preserving
// 32-bit keys so the min-reduction is one LDS atomicMin (no CAS retry loop).
// A symbolic pass counts columns per row, an exclusive scan (rocPRIM) gives
// the row pointer, and a numeric pass recomputes the rows and compacts each
// table into the output.
#include <cstring>

#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "common.hpp"

namespace mi355 {

namespace {

constexpr int kEmpty = -1;

__device__ inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85ebca6bu;
  x ^= x >> 13;
  x *= 0xc2b2ae35u;
  x ^= x >> 16;
  return x;
}

__device__ inline bool storable(float x) { return x == x && x != __builtin_huge_valf(); }

template <bool NUMERIC>
__global__ void k_spgemm_rows(int64_t rows, const int64_t* __restrict__ arp,
                              const int32_t* __restrict__ acol, const float* __restrict__ aval,
                              const int64_t* __restrict__ brp, const int32_t* __restrict__ bcol,
                              const float* __restrict__ bval, uint32_t table, uint32_t cap,
                              int64_t* __restrict__ counts, const int64_t* __restrict__ crp,
                              int32_t* __restrict__ ccol, float* __restrict__ cval,
                              int* __restrict__ status) {
  extern __shared__ unsigned char smem[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  int* keys = reinterpret_cast<int*>(smem) + (size_t)wave * table;
  uint32_t* vals = reinterpret_cast<uint32_t*>(smem) + (size_t)wpb * table + (size_t)wave * table;
  __shared__ int s_count[4];
  __shared__ int s_abort[4];
  const uint32_t mask = table - 1;
  for (int64_t row = (int64_t)blockIdx.x * wpb + wave; row < rows; row += (int64_t)gridDim.x * wpb) {
    for (uint32_t t = lane; t < table; t += 64) {
      keys[t] = kEmpty;
      if (NUMERIC) vals[t] = 0xFFFFFFFFu;
    }
    if (lane == 0) {
      s_count[wave] = 0;
      s_abort[wave] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int64_t p = arp[row]; p < arp[row + 1]; ++p) {
      const float a = aval[p];
      const int k = acol[p];
      if (s_abort[wave]) break;
      for (int64_t q0 = brp[k]; q0 < brp[k + 1]; q0 += 64) {
        const int64_t q = q0 + lane;
        if (q < brp[k + 1]) {
          const float c = a + bval[q];
          if (storable(c)) {
            const int j = bcol[q];
            uint32_t slot = hash32((uint32_t)j) & mask;
            for (uint32_t probe = 0; probe < table; ++probe) {
              const int old = atomicCAS(keys + slot, kEmpty, j);
              if (old == kEmpty || old == j) {
                if (old == kEmpty && atomicAdd(&s_count[wave], 1) >= (int)cap) s_abort[wave] = 1;
                if (NUMERIC) {
                  uint32_t bits = __float_as_uint(c);
                  const uint32_t key = (bits & 0x80000000u) ? ~bits : (bits ^ 0x80000000u);
                  atomicMin(vals + slot, key);
                }
                break;
              }
              slot = (slot + 1) & mask;
            }
          }
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (s_abort[wave]) {
      if (lane == 0) atomicExch(status, 1);
      if (!NUMERIC && lane == 0) counts[row] = 0;
      continue;
    }
    if (!NUMERIC) {
      if (lane == 0) counts[row] = s_count[wave];
      continue;
    }
    // compact the table in slot order (columns unsorted, as in the reference)
    int64_t out = crp[row];
    for (uint32_t t0 = 0; t0 < table; t0 += 64) {
      const uint32_t t = t0 + lane;
      const bool occ = t < table && keys[t] != kEmpty;
      const unsigned long long m = __ballot(occ);
      if (occ) {
        const int64_t pos = out + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        const uint32_t key = vals[t];
        const uint32_t bits = (key & 0x80000000u) ? (key ^ 0x80000000u) : ~key;
        ccol[pos] = keys[t];
        cval[pos] = __uint_as_float(bits);
      }
      out += __popcll(m);
    }
  }
}

__global__ void k_check_csr(int64_t rows, int64_t cols, int64_t nnz, const int64_t* __restrict__ rp,
                            const int32_t* __restrict__ col, int* __restrict__ bad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int b = 0;
  if (tid == 0 && (rp[0] != 0 || rp[rows] != nnz)) b = 1;
  for (int64_t r = tid; r < rows; r += stride)
    if (rp[r + 1] < rp[r]) b = 1;
  for (int64_t e = tid; e < nnz; e += stride)
    if (col[e] < 0 || (int64_t)col[e] >= cols) b = 1;
  if (b) atomicExch(bad, 1);
}

}  // namespace

void spgemm_minplus(int64_t a_rows, int64_t a_cols, int64_t a_nnz, const int64_t* arp,
                    const int32_t* acol, const float* aval, int64_t b_rows, int64_t b_cols,
                    int64_t b_nnz, const int64_t* brp, const int32_t* bcol, const float* bval,
                    int hash_capacity, int block_size, int64_t** c_rowptr, int32_t** c_col,
                    float** c_val, int64_t* c_nnz, hipStream_t s) {
  if (!c_rowptr || !c_col || !c_val || !c_nnz) fail(MI355_EINVAL, "null output pointer");
  *c_rowptr = nullptr;
  *c_col = nullptr;
  *c_val = nullptr;
  *c_nnz = 0;
  if (a_rows < 0 || a_cols < 0 || b_rows < 0 || b_cols < 0 || a_nnz < 0 || b_nnz < 0)
    fail(MI355_EINVAL, "negative dimension");
  if (a_cols != b_rows) fail(MI355_EINVAL, "A.cols must equal B.rows");
  if (b_cols > INT32_MAX || a_cols > INT32_MAX) fail(MI355_EINVAL, "columns must fit in int");
  if (!arp || !brp || (a_nnz && (!acol || !aval)) || (b_nnz && (!bcol || !bval)))
    fail(MI355_EINVAL, "null input pointer");
  if (hash_capacity <= 0) fail(MI355_EINVAL, "hash capacity must be positive");
  if (block_size != 64 && block_size != 128 && block_size != 256)
    fail(MI355_EINVAL, "block size must be 64, 128 or 256");

  DevBuf<int> flag(2);
  MI_HIP(hipMemsetAsync(flag.p, 0, 8, s));
  if (a_rows > 0) {
    k_check_csr<<<std::min<unsigned>(div_up((uint64_t)std::max(a_rows, a_nnz), 256), 4096u), 256, 0, s>>>(
        a_rows, a_cols, a_nnz, arp, acol, flag.p);
  }
  if (b_rows > 0) {
    k_check_csr<<<std::min<unsigned>(div_up((uint64_t)std::max(b_rows, b_nnz), 256), 4096u), 256, 0, s>>>(
        b_rows, b_cols, b_nnz, brp, bcol, flag.p);
  }
  MI_LAUNCH_CHECK();
  int hbad = 0;
  MI_HIP(hipMemcpyAsync(&hbad, flag.p, 4, hipMemcpyDeviceToHost, s));
  MI_HIP(hipStreamSynchronize(s));
  if (hbad) fail(MI355_EINVAL, "invalid CSR structure or column index");

  int64_t* rp = nullptr;
  MI_HIP(hipMalloc(reinterpret_cast<void**>(&rp), (size_t)(a_rows + 1) * 8));
  try {
    if (a_rows == 0) {
      MI_HIP(hipMemsetAsync(rp, 0, 8, s));
      MI_HIP(hipStreamSynchronize(s));
      *c_rowptr = rp;
      return;
    }
    uint32_t pow2 = 1;
    while (pow2 < (uint32_t)hash_capacity) pow2 <<= 1;
    uint32_t table = pow2 * 2;
    int dev = 0;
    MI_HIP(hipGetDevice(&dev));
    int max_lds = 65536;
    MI_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    int wpb = block_size / 64;
    while (wpb > 1 && (size_t)wpb * table * 8 > (size_t)max_lds) --wpb;
    if ((size_t)table * 8 > (size_t)max_lds) table = pow2;  // load factor up to 1
    if ((size_t)table * 8 > (size_t)max_lds) fail(MI355_EINVAL, "hash capacity exceeds LDS");
    const size_t lds = (size_t)wpb * table * 8;
    const unsigned grid = std::min<unsigned>(div_up((uint64_t)a_rows, wpb), 1u << 20);

    DevBuf<int64_t> counts((size_t)a_rows + 1);
    MI_HIP(hipMemsetAsync(counts.p + a_rows, 0, 8, s));
    k_spgemm_rows<false><<<grid, wpb * 64, lds, s>>>(a_rows, arp, acol, aval, brp, bcol, bval, table,
                                                     (uint32_t)hash_capacity, counts.p, nullptr,
                                                     nullptr, nullptr, flag.p + 1);
    MI_LAUNCH_CHECK();
    size_t temp_bytes = 0;
    MI_HIP(rocprim::exclusive_scan(nullptr, temp_bytes, counts.p, rp, (int64_t)0, (size_t)a_rows + 1,
                                   rocprim::plus<int64_t>(), s));
    DevBuf<unsigned char> temp(temp_bytes);
    MI_HIP(rocprim::exclusive_scan(temp.p, temp_bytes, counts.p, rp, (int64_t)0, (size_t)a_rows + 1,
                                   rocprim::plus<int64_t>(), s));
    int64_t nnz = 0;
    int overflow = 0;
    MI_HIP(hipMemcpyAsync(&nnz, rp + a_rows, 8, hipMemcpyDeviceToHost, s));
    MI_HIP(hipMemcpyAsync(&overflow, flag.p + 1, 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (overflow) fail(MI355_ECAPACITY, "an output row exceeds the hash capacity");
    if (nnz == 0) {
      *c_rowptr = rp;
      return;
    }
    int32_t* col = nullptr;
    float* val = nullptr;
    MI_HIP(hipMalloc(reinterpret_cast<void**>(&col), (size_t)nnz * 4));
    if (hipMalloc(reinterpret_cast<void**>(&val), (size_t)nnz * 4) != hipSuccess) {
      (void)hipFree(col);
      fail(MI355_ENOMEM, "cannot allocate SpGEMM values");
    }
    k_spgemm_rows<true><<<grid, wpb * 64, lds, s>>>(a_rows, arp, acol, aval, brp, bcol, bval, table,
                                                    (uint32_t)hash_capacity, nullptr, rp, col, val,
                                                    flag.p + 1);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      (void)hipFree(col);
      (void)hipFree(val);
      MI_HIP(e1 != hipSuccess ? e1 : e2);
    }
    *c_rowptr = rp;
    *c_col = col;
    *c_val = val;
    *c_nnz = nnz;
  } catch (...) {
    (void)hipFree(rp);
    throw;
  }
}

}  // namespace mi355
