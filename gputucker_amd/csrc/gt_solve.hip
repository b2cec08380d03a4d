// gt_solve.hip -- batched R x R regularised row solves of one mode update.
//
// Reference: update.cuh:167-194 inverts B_i = lambda I + sum dd^T with Eigen's
// partial-pivot LU (MatrixXd::inverse) on the host and forms a_i = C_i^T B_i^-1.
// B_i is symmetric positive definite (lambda > 0), so here each row is solved
// by an FP64 Cholesky factorisation and two triangular solves on the GPU:
// one lane per row for R <= 16 (the factor lives in registers), one wave per
// row above that (the factor lives in LDS).  Rows without nonzeros have
// B = lambda I and C = 0, so their solution is 0 (the reference's result too).
#include "gt_internal.hpp"

namespace gt {
namespace {

// index of B(p, q), p <= q, in the packed upper triangle (row-major)
__host__ __device__ constexpr int up_index(int R, int p, int q) {
  return p * R - p * (p - 1) / 2 + (q - p);
}

template <int R>
__global__ void __launch_bounds__(64) k_solve_lane(SolveArgs a) {
  constexpr int W = gram_width(R);
  const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.rows) return;
  float *out = a.A + row * row_stride(R);
  if (a.row_ptr[row + 1] == a.row_ptr[row]) {
#pragma unroll
    for (int r = 0; r < R; ++r) out[r] = 0.f;
    return;
  }
  const double *g = a.gram + row * W;
  double L[R * (R + 1) / 2]; // packed lower: L(i, j) at i(i+1)/2 + j
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) L[i * (i + 1) / 2 + j] = g[up_index(R, j, i)];
#pragma unroll
  for (int i = 0; i < R; ++i) L[i * (i + 1) / 2 + i] += a.lambda;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    double dj = L[j * (j + 1) / 2 + j];
#pragma unroll
    for (int k = 0; k < j; ++k) dj -= L[j * (j + 1) / 2 + k] * L[j * (j + 1) / 2 + k];
    dj = sqrt(dj);
    L[j * (j + 1) / 2 + j] = dj;
    const double inv = 1.0 / dj;
#pragma unroll
    for (int i = j + 1; i < R; ++i) {
      double v = L[i * (i + 1) / 2 + j];
#pragma unroll
      for (int k = 0; k < j; ++k) v -= L[i * (i + 1) / 2 + k] * L[j * (j + 1) / 2 + k];
      L[i * (i + 1) / 2 + j] = v * inv;
    }
  }
  double y[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    double v = g[R * (R + 1) / 2 + i];
#pragma unroll
    for (int k = 0; k < i; ++k) v -= L[i * (i + 1) / 2 + k] * y[k];
    y[i] = v / L[i * (i + 1) / 2 + i];
  }
#pragma unroll
  for (int i = R - 1; i >= 0; --i) {
    double v = y[i];
#pragma unroll
    for (int k = i + 1; k < R; ++k) v -= L[k * (k + 1) / 2 + i] * y[k];
    y[i] = v / L[i * (i + 1) / 2 + i];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) out[r] = (float)y[r];
}

// One wave per row; lane i owns row i of the factor (R <= 64).
template <int R>
__global__ void __launch_bounds__(64) k_solve_wave(SolveArgs a) {
  constexpr int W = gram_width(R);
  __shared__ double M[R][R + 1];
  __shared__ double y[R];
  const uint64_t row = blockIdx.x;
  const int lane = threadIdx.x;
  float *out = a.A + row * row_stride(R);
  if (a.row_ptr[row + 1] == a.row_ptr[row]) {
    if (lane < R) out[lane] = 0.f;
    return;
  }
  const double *g = a.gram + row * W;
  if (lane < R) {
    for (int j = 0; j <= lane; ++j) M[lane][j] = g[up_index(R, j, lane)];
    M[lane][lane] += a.lambda;
    y[lane] = g[R * (R + 1) / 2 + lane];
  }
  __syncthreads();
  for (int j = 0; j < R; ++j) { // right-looking Cholesky
    if (lane == j) M[j][j] = sqrt(M[j][j]);
    __syncthreads();
    if (lane > j && lane < R) M[lane][j] /= M[j][j];
    __syncthreads();
    if (lane > j && lane < R)
      for (int k = j + 1; k <= lane; ++k) M[lane][k] -= M[lane][j] * M[k][j];
    __syncthreads();
  }
  for (int i = 0; i < R; ++i) { // forward: L y = c
    if (lane == i) y[i] /= M[i][i];
    __syncthreads();
    if (lane > i && lane < R) y[lane] -= M[lane][i] * y[i];
    __syncthreads();
  }
  for (int i = R - 1; i >= 0; --i) { // backward: L^T a = y
    if (lane == i) y[i] /= M[i][i];
    __syncthreads();
    if (lane < i) y[lane] -= M[i][lane] * y[i];
    __syncthreads();
  }
  if (lane < R) out[lane] = (float)y[lane];
}

template <int R> void solve_impl(const SolveArgs &a, hipStream_t s) {
  if (a.rows == 0) return;
  if constexpr (R <= 16) {
    hipLaunchKernelGGL(k_solve_lane<R>, dim3((unsigned)((a.rows + 63) / 64)), dim3(64), 0, s, a);
  } else {
    hipLaunchKernelGGL(k_solve_wave<R>, dim3((unsigned)a.rows), dim3(64), 0, s, a);
  }
}

} // namespace

void launch_solve(int R, const SolveArgs &a, hipStream_t s) {
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    solve_impl<r>(a, s);                                                                         \
    break;
    GT_FOR_EACH_RANK(GT_CASE)
#undef GT_CASE
  default:
    throw std::invalid_argument("rank not built");
  }
  GT_HIP(hipGetLastError());
}

} // namespace gt
