// Exact division of 32-bit unsigned integers by a divisor fixed at launch,
// as a multiply-high, a subtract, an add and two shifts (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI
// 1994, Figure 4.1): with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// q = (t + ((x - t) >> s1)) >> s2 where t = mulhi(m, x), s1 = min(l, 1),
// s2 = max(l - 1, 0), for every 32-bit x and 1 <= d < 2^32.  The GPU has no
// integer divide instruction, so x / d is otherwise a long instruction
// sequence.  Host and device code alike.
#pragma once

#include <cstdint>

namespace mi355 {

struct FastDiv {
  uint32_t d, m, s1, s2;
};

inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while (l < 32 && (uint64_t{1} << l) < d) ++l;
  const uint64_t m = ((uint64_t{1} << 32) * ((uint64_t{1} << l) - d)) / d + 1;
  return FastDiv{d, (uint32_t)m, l < 1 ? l : 1u, l > 0 ? l - 1 : 0u};
}

// (the 64-bit product's high word compiles to one v_mul_hi_u32)
#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint32_t fastdiv(uint32_t x, const FastDiv& f) {
  const uint32_t t = (uint32_t)(((uint64_t)f.m * x) >> 32);
  return (t + ((x - t) >> f.s1)) >> f.s2;
}

}  // namespace mi355
