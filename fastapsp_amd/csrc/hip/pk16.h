// Packed 16-bit min-plus helpers (two distances per 32-bit register).
//
// Distances are capped at kCap16 = 0x3E00, so every stored value and every sum of
// two lies in [0, 0x7C00]: non-negative IEEE half-precision bit patterns up to +inf,
// which order as the integers do (no NaN, no negative zero; the kernels keep FP16
// denormals).  Sums use v_pk_add_u16; minima use v_pk_minimum3_f16 on the bit
// patterns, which is then exactly the integer minimum.
#pragma once

#include <hip/hip_runtime.h>

namespace fap_hip {

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

// Pairwise sums with the low (resp. high) half of a broadcast to both lanes.
__device__ __forceinline__ unsigned pk_add_lo(unsigned a, unsigned b) {
  const us2 av = __builtin_bit_cast(us2, a);
  return __builtin_bit_cast(unsigned, (us2)(av.xx + __builtin_bit_cast(us2, b)));
}
__device__ __forceinline__ unsigned pk_add_hi(unsigned a, unsigned b) {
  const us2 av = __builtin_bit_cast(us2, a);
  return __builtin_bit_cast(unsigned, (us2)(av.yy + __builtin_bit_cast(us2, b)));
}

// Per 16-bit half: min(acc, x, y) on values in [0, 0x7C00] as half-precision
// minimum (IEEE 754-2019 minimum: exact, associative, no rounding); the compiler
// folds the two into one v_pk_minimum3_f16.
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned pk_min3(unsigned acc, unsigned x, unsigned y) {
  const h2 m = __builtin_elementwise_minimum(
      __builtin_elementwise_minimum(__builtin_bit_cast(h2, acc), __builtin_bit_cast(h2, x)),
      __builtin_bit_cast(h2, y));
  return __builtin_bit_cast(unsigned, m);
}

// Both halves at the largest value a sum can take (2 * kCap16 = half +inf): the
// identity of pk_min3 here.
constexpr unsigned kPk16Top = 0x7C007C00u;

}  // namespace fap_hip
