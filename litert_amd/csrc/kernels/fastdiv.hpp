// fastdiv.hpp - division of 32-bit unsigned integers by a divisor fixed before
// a launch, as a multiply-high and two shifts (T. Granlund, P. Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, fig. 4.1:
// exact for every n < 2^32 and 1 <= d < 2^32). The render kernels map packed
// pixel indices to (view, super-block row, column) with it: the divisors are
// wave-uniform, so the quotient stays on the scalar unit, where the
// compiler's expansion of a division keeps reciprocal constants in vector
// registers for the whole launch.
#pragma once
#include <stdint.h>

namespace lrtd {

struct FastDiv {
  uint32_t d = 1, m = 1, s1 = 0, s2 = 0;
  __host__ __device__ uint32_t div(uint32_t n) const {
    const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
    return (t + ((n - t) >> s1)) >> s2;
  }
};

// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, shifts min(l, 1) and l - 1.
inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  uint32_t l = 0;
  while (l < 32 && (1ull << l) < d) l++;
  f.d = d;
  f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  f.s1 = l ? 1u : 0u;
  f.s2 = l ? l - 1u : 0u;
  return f;
}

}  // namespace lrtd
