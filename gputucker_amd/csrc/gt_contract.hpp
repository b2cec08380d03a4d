// gt_contract.hpp -- device helpers shared by the Gram kernels: the VALU
// contraction that forms one nonzero's delta, and a wavefront-scope barrier.
#pragma once

#include "gt_internal.hpp"

namespace gt {

__device__ inline __attribute__((always_inline)) void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// delta of mode n for this lane's nonzero.  gi[k] is the global index of mode
// k.  The contraction runs over the other modes ascending; the last of them
// (L) is held in registers, the others are staged in this wave's LDS rows
// (one column per lane, so each lane only reads what it wrote).  For each
// outer multi-index o (mixed radix over the outer modes, last fastest):
//   s = prod_t A_t[i_t][o_t];  delta[r] += Gt[o][jL][r] * (s * A_L[i_L][jL])
// Gt is the mode-n unfolding of the core, laid out [o][jL][r]; its address is
// uniform across the wave, so it is read through scalar loads.
template <int R>
__device__ __forceinline__ void contract(int N, int n, const float *const (&A)[kMaxOrder],
                                         const float *__restrict__ Gt,
                                         const uint32_t (&gi)[kMaxOrder], float *lrows, int lane,
                                         float (&d)[R]) {
  const int L = (n == N - 1) ? N - 2 : N - 1;
  constexpr bool kRegL = R <= 16; // above that A_L's row is staged in LDS too
  float aL[kRegL ? R : 1];
#pragma unroll
  for (int k = 0; k < kMaxOrder; ++k) {
    if (k >= N || k == n) continue;
    const float *row = A[k] + (size_t)gi[k] * R;
    if (kRegL && k == L) {
#pragma unroll
      for (int j = 0; j < (kRegL ? R : 1); ++j) aL[j] = row[j];
    } else {
      const int t = k - (k > n ? 1 : 0); // L lands in slot N-2
#pragma unroll
      for (int j = 0; j < R; ++j) lrows[(t * R + j) * kWave + lane] = row[j];
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) d[r] = 0.f;
  const int M1 = N - 2; // outer modes
  int O = 1;
  for (int t = 0; t < M1; ++t) O *= R;
  int dig[kMaxOrder - 2];
#pragma unroll
  for (int t = 0; t < kMaxOrder - 2; ++t) dig[t] = 0;
  for (int o = 0; o < O; ++o) {
    float s = 1.f;
#pragma unroll
    for (int t = 0; t < kMaxOrder - 2; ++t)
      if (t < M1) s *= lrows[(t * R + dig[t]) * kWave + lane];
    const float *__restrict__ g = Gt + (size_t)o * R * R;
    if constexpr (kRegL) {
#pragma unroll
      for (int jL = 0; jL < R; ++jL) {
        const float K = s * aL[jL];
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = __builtin_fmaf(g[jL * R + r], K, d[r]);
      }
    } else {
      const float *lL = lrows + (size_t)(N - 2) * R * kWave + lane;
      for (int jL = 0; jL < R; ++jL) {
        const float K = s * lL[jL * kWave];
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = __builtin_fmaf(g[jL * R + r], K, d[r]);
      }
    }
#pragma unroll
    for (int t = kMaxOrder - 3; t >= 0; --t) {
      if (t < M1) {
        if (++dig[t] < R) break;
        dig[t] = 0;
      }
    }
  }
}

// factor rows staged in LDS per lane: the outer modes, plus A_L above R = 16
template <int R> __host__ __device__ constexpr int lds_slots(int N) {
  return (R <= 16) ? (N > 2 ? N - 2 : 0) : N - 1;
}

} // namespace gt
