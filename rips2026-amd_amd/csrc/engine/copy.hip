// Device-to-device copy as a plain streaming kernel: 16-byte loads and stores,
// grid-stride, bytes not a multiple of 16 handled one byte per thread.  It is
// the reference point for HBM traffic counters (a copy of B bytes reads and
// writes B bytes), and the engine's device copy entry point.  The random
// 8-byte gather beside it is the reference point for the traffic counters
// of random accesses (what one 8-byte load from an address not in L2 moves).
#include "common.hpp"
#include "sssp.hpp"

namespace mi355 {

namespace {

__global__ __launch_bounds__(kBlock) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                   int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ void k_copy1(unsigned char* __restrict__ dst, const unsigned char* __restrict__ src,
                        int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// dst[i] = src[idx[i]]: 8-byte records at the given indices.
__global__ __launch_bounds__(kBlock) void k_gather8(unsigned long long* __restrict__ dst,
                                                    const unsigned long long* __restrict__ src,
                                                    const uint32_t* __restrict__ idx, int64_t count) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = src[idx[i]];
}

}  // namespace

void device_gather8(void* dst, const void* src, const uint32_t* idx, int64_t count, hipStream_t s) {
  if (count <= 0) return;
  if (!dst || !src || !idx) fail(MI355_EINVAL, "null gather pointer");
  const unsigned grid = (unsigned)std::min<int64_t>(div_up((uint64_t)count, kBlock), (int64_t)device_props().cus * 16);
  k_gather8<<<grid, kBlock, 0, s>>>(static_cast<unsigned long long*>(dst),
                                    static_cast<const unsigned long long*>(src), idx, count);
  MI_LAUNCH_CHECK();
}

void device_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  if (!dst || !src) fail(MI355_EINVAL, "null copy pointer");
  const bool aligned = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u) == 0;
  const int64_t n16 = aligned ? (int64_t)(bytes / 16) : 0;
  if (n16 > 0) {
    const unsigned grid = (unsigned)std::min<int64_t>(div_up((uint64_t)n16, kBlock), (int64_t)device_props().cus * 16);
    k_copy16<<<grid, kBlock, 0, s>>>(static_cast<uint4*>(dst), static_cast<const uint4*>(src), n16);
    MI_LAUNCH_CHECK();
  }
  const int64_t done = n16 * 16, rest = (int64_t)bytes - done;
  if (rest > 0) {
    k_copy1<<<(unsigned)div_up((uint64_t)rest, kBlock), kBlock, 0, s>>>(
        static_cast<unsigned char*>(dst) + done, static_cast<const unsigned char*>(src) + done, rest);
    MI_LAUNCH_CHECK();
  }
}

}  // namespace mi355
