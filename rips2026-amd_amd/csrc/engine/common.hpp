// Shared host/device definitions of the MI355X SSSP engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "mi355sssp.h"

namespace mi355 {

// ---- errors ---------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  int hip;
  Error(int c, const std::string& m, int h = 0) : std::runtime_error(m), code(c), hip(h) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

[[noreturn]] inline void fail_hip(hipError_t e, const char* expr, const char* file, int line) {
  const int code = (e == hipErrorOutOfMemory) ? MI355_ENOMEM : MI355_EHIP;
  throw Error(code,
              std::string("HIP error ") + hipGetErrorString(e) + " at " + file + ":" +
                  std::to_string(line) + " (" + expr + ")",
              (int)e);
}

#define MI_HIP(expr)                                                        \
  do {                                                                      \
    hipError_t _mi_e = (expr);                                              \
    if (_mi_e != hipSuccess) ::mi355::fail_hip(_mi_e, #expr, __FILE__, __LINE__); \
  } while (0)

#define MI_LAUNCH_CHECK() MI_HIP(hipGetLastError())

// ---- device buffers -------------------------------------------------------
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { reset(count); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  void reset(size_t count) {
    release();
    if (count) MI_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  // grow-only reuse
  void ensure(size_t count) {
    if (count > n) reset(count);
  }
  void release() noexcept {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  T* get() const { return p; }
};

// ---- label encoding -------------------------------------------------------
// A vertex label is one 64-bit word: (dist key << 32) | pred.  The dist key is
// an order-preserving 32-bit image of the distance, so an unsigned 64-bit
// atomicMin keeps the smallest distance and, among equal distances, the
// smallest predecessor index.  pred 0xFFFFFFFF means "none".
//   fp32:   key = bits ^ 0x80000000 for x >= +0, ~bits for x < 0 (sign set)
//   uint32: key = value (INF = 0xFFFFFFFF, saturating add)
constexpr uint32_t kNoPred = 0xFFFFFFFFu;
constexpr uint32_t kF32InfKey = 0xFF800000u;   // key of +inf
constexpr uint32_t kF32ZeroKey = 0x80000000u;  // key of +0.0f
constexpr uint32_t kU32Inf = 0xFFFFFFFFu;

__host__ __device__ inline uint32_t f32_to_key(float x) {
  const uint32_t b = __builtin_bit_cast(uint32_t, x);
  return (b & 0x80000000u) ? ~b : (b ^ 0x80000000u);
}
__host__ __device__ inline float key_to_f32(uint32_t k) {
  const uint32_t b = (k & 0x80000000u) ? (k ^ 0x80000000u) : ~k;
  return __builtin_bit_cast(float, b);
}

struct WF32 {
  static constexpr int kType = MI355_W_F32;
  static constexpr uint32_t kInfKey = kF32InfKey;
  static constexpr uint32_t kZeroKey = kF32ZeroKey;
  // One IEEE round-to-nearest fp32 add, exactly the left fold d + w.
  __device__ static inline uint32_t relax(uint32_t dkey, uint32_t wbits) {
    return f32_to_key(key_to_f32(dkey) + __uint_as_float(wbits));
  }
  __device__ static inline float bucket_value(uint32_t dkey) { return key_to_f32(dkey); }
  // distance word in the output convention (fp32 bits; +inf or FLT_MAX = INF)
  __device__ static inline uint32_t output_to_key(uint32_t bits) {
    const float x = __uint_as_float(bits);
    return (x >= 3.4028234663852886e38f) ? kF32InfKey : f32_to_key(x);
  }
};

struct WU32 {
  static constexpr int kType = MI355_W_U32;
  static constexpr uint32_t kInfKey = kU32Inf;
  static constexpr uint32_t kZeroKey = 0u;
  __device__ static inline uint32_t relax(uint32_t d, uint32_t w) {
    const uint32_t s = d + w;
    return s < d ? kU32Inf : s;
  }
  __device__ static inline float bucket_value(uint32_t d) { return (float)d; }
  __device__ static inline uint32_t output_to_key(uint32_t d) { return d; }
};

inline uint64_t pack_label(uint32_t key, uint32_t pred) { return ((uint64_t)key << 32) | pred; }

// ---- launch geometry --------------------------------------------------------
constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
// Out-degree above which a vertex is split into kChunk-edge pieces processed
// by whole waves (R-MAT hubs); below it, 64 frontier vertices share a wave and
// their edges are packed across the lanes.
constexpr int kLargeDegree = 1024;
constexpr int kChunk = 2048;

struct DeviceProps {
  int device = 0;
  int cus = 256;
  int max_blocks_per_cu = 8;
};
DeviceProps device_props();

__host__ __device__ inline unsigned div_up(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace mi355
