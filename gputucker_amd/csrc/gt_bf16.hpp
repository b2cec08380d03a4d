// gt_bf16.hpp -- BF16 helpers of the split-BF16 MFMA delta: two floats to a
// BF16 word pair, and the split words a factor row contributes to the K=32
// MFMA operand of each lane group (the plans of bf16r_plan).
#pragma once

#include "gt_internal.hpp"

namespace gt {

typedef __bf16 gt_bf2 __attribute__((ext_vector_type(2)));
typedef float gt_fv2 __attribute__((ext_vector_type(2)));
typedef uint32_t gt_u4 __attribute__((ext_vector_type(4)));

// two floats to a BF16 word pair (a in the low half), round to nearest even
// (v_cvt_pk_bf16_f32)
__device__ inline uint32_t pk_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((gt_fv2{a, b}), gt_bf2));
}
__device__ inline float bf16_low_f(uint32_t p) { return __uint_as_float(p << 16); }
__device__ inline float bf16_high_f(uint32_t p) { return __uint_as_float(p & 0xFFFF0000u); }

// The 16 bytes of lane group g for a factor row whose column j is col(j)
// (zero at and beyond R): kPlanHL [hi(a[4g..4g+3]), lo(a[4g..4g+3])],
// kPlanHH [hi(a[0..3]), hi(a[4..7])], kPlanR10 [hi(a[8..11]), lo(a[8]),
// lo(a[9]), hi(a[8]), hi(a[9])], kPlanZero zeros.
template <int R, typename Col> __device__ inline gt_u4 split_words(Col col, int g) {
  auto hi2 = [&](int j) { return pk_bf16(col(j), col(j + 1)); };
  auto lo2 = [&](int j) {
    const uint32_t h = hi2(j);
    return pk_bf16(col(j) - bf16_low_f(h), col(j + 1) - bf16_high_f(h));
  };
  gt_u4 v = {0u, 0u, 0u, 0u};
  switch (bf16r_plan(R, g)) {
  case kPlanHL:
    if (4 * g < R) v = gt_u4{hi2(4 * g), hi2(4 * g + 2), lo2(4 * g), lo2(4 * g + 2)};
    break;
  case kPlanHH:
    v = gt_u4{hi2(0), hi2(2), hi2(4), hi2(6)};
    break;
  case kPlanR10:
    v = gt_u4{hi2(8), hi2(10), lo2(8), hi2(8)};
    break;
  default:
    break;
  }
  return v;
}

} // namespace gt
