// gt_solve.hip -- batched R x R regularised row solves of one mode update.
//
// Reference: update.cuh:167-194 inverts B_i = lambda I + sum dd^T with Eigen's
// partial-pivot LU (MatrixXd::inverse) on the host and forms a_i = C_i^T B_i^-1.
// B_i is symmetric positive definite (lambda > 0), so here each row is solved
// by an FP64 Cholesky factorisation and two triangular solves on the GPU:
// one lane per row for R <= 16 (the factor lives in registers), one wave per
// row above that (lane i holds row i of the factor in registers up to
// R = 36, in LDS above).  Rows without nonzeros have
// B = lambda I and C = 0, so their solution is 0 (the reference's result too).
//
// With fit_partials set, each solve also yields the row's share of the
// residual, sum_e (x_e - a_i . delta_e)^2 - sum_e x_e^2 = a^T G a - 2 a^T C
// with G = sum delta delta^T = L L^T - lambda I, evaluated in FP64 for a
// rounded to FP32 as stored.  After the last mode of an ALS iteration delta is
// the contraction with every other (already updated) factor, so a_i . delta_e
// is the reconstruction of reconstruction.cuh:58-81 and the fit needs no
// pass over the nonzeros.
#include "gt_bf16.hpp"

namespace gt {
namespace {

// index of B(p, q), p <= q, in the packed upper triangle (row-major)
__host__ __device__ constexpr int up_index(int R, int p, int q) {
  return p * R - p * (p - 1) / 2 + (q - p);
}

template <int R>
__global__ void __launch_bounds__(64) k_solve_lane(SolveArgs a) {
  constexpr int W = gram_width(R);
  constexpr int kPiece = 16;                                   // rows per staged piece
  constexpr int kPieceD2 = kPiece * W / 2;                     // 16-byte units per piece
  constexpr int kCopy = (kPieceD2 + kWave - 1) / kWave;        // load instructions per piece
  __shared__ __align__(16) double buf[2][kCopy * kWave * 2];   // two pieces, whole instructions
  const int lane = threadIdx.x;
  const uint64_t row0 = (uint64_t)blockIdx.x * kWave;
  const uint64_t row = row0 + lane;
  // The wave's 64 Gram rows are contiguous.  They are read 16 rows at a time
  // as one block of 16-byte loads (whole cache lines, every byte used once)
  // straight into LDS (global_load_lds: no registers held while in flight),
  // two pieces in flight, and the 16 lanes that own a piece's rows copy
  // theirs out.  A lane-per-row read would touch 64 lines per load
  // instruction and refetch them as they are evicted.
  // the row's degree is loaded first (unconditionally, from an in-bounds row)
  // and compared after the staging, so its latency overlaps the staging
  const uint64_t rc = row < a.rows ? row : 0;
  const uint64_t rp_lo = a.row_ptr[rc], rp_hi = a.row_ptr[rc + 1];
  auto rows_in = [&](int k) -> int {
    const uint64_t pr = row0 + (uint64_t)k * kPiece;
    return pr < a.rows ? (int)(a.rows - pr < (uint64_t)kPiece ? a.rows - pr : kPiece) : 0;
  };
  // A piece with rows issues exactly kCopy loads (lanes past its end reload
  // its last 16 bytes into LDS slots nobody reads), so the wait for piece k
  // while piece k + 1 is in flight is a fixed count.
  auto issue = [&](int k) {
    const int n2 = rows_in(k) * W / 2;
    if (n2 == 0) return;
    const double2 *src = reinterpret_cast<const double2 *>(a.gram + (row0 + (uint64_t)k * kPiece) * W);
#pragma unroll
    for (int c = 0; c < kCopy; ++c) {
      // lane l of copy c lands at byte 1024 c + 16 l of the buffer
      const int i = lane + c * kWave < n2 ? lane + c * kWave : n2 - 1;
      __builtin_amdgcn_global_load_lds(
          src + i, (__attribute__((address_space(3))) void *)(buf[k & 1] + c * kWave * 2), 16, 0,
          0);
    }
  };
  // s_waitcnt immediates: vmcnt(kCopy) or vmcnt(0), other counters untouched
  constexpr int kWaitNext = (kCopy & 0xF) | ((kCopy >> 4) << 14) | (7 << 4) | (15 << 8);
  constexpr int kWaitAll = (7 << 4) | (15 << 8);
  static_assert(kCopy < 64, "vmcnt holds at most 63");
  double L[R * (R + 1) / 2]; // packed lower: L(i, j) at i(i+1)/2 + j
  double cv[R];
  issue(0);
  issue(1);
#pragma unroll
  for (int k = 0; k < kWave / kPiece; ++k) {
    const int n = rows_in(k);
    // piece k has landed: at most the next piece's loads are still in flight
    if (k + 1 < kWave / kPiece && rows_in(k + 1) > 0)
      __builtin_amdgcn_s_waitcnt(kWaitNext);
    else
      __builtin_amdgcn_s_waitcnt(kWaitAll);
    __syncthreads();
    if (lane / kPiece == k && lane % kPiece < n) {
      const double *g = buf[k & 1] + (lane % kPiece) * W;
#pragma unroll
      for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) L[i * (i + 1) / 2 + j] = g[up_index(R, j, i)];
#pragma unroll
      for (int i = 0; i < R; ++i) cv[i] = g[R * (R + 1) / 2 + i];
    }
    __syncthreads();
    if (k + 2 < kWave / kPiece) issue(k + 2);
  }
  const bool empty = rp_hi == rp_lo;
  double contrib = 0.0;
  if (row < a.rows) {
    float *out = a.A + row * row_stride(R);
    if (empty) {
#pragma unroll
      for (int r = 0; r < R; ++r) out[r] = 0.f;
      if (a.As) {
#pragma unroll
        for (int w = 0; w < kSplitWords; ++w) a.As[row * kSplitWords + w] = 0u;
      }
    } else {
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
        double v = cv[i];
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
      float af[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        af[r] = (float)y[r];
        out[r] = af[r];
      }
      if constexpr (R <= 15) {
        if (a.As) { // the row's split BF16 copy for the MFMA delta
          auto col = [&](int j) { return j < R ? af[j] : 0.f; };
#pragma unroll
          for (int g = 0; g < 4; ++g)
            *reinterpret_cast<gt_u4 *>(a.As + row * kSplitWords + 4 * g) = split_words<R>(col, g);
        }
      }
      if (a.fit_partials) { // a^T (L L^T) a - lambda |a|^2 - 2 a^T c
#pragma unroll
        for (int i = 0; i < R; ++i) {
          double v = 0.0;
#pragma unroll
          for (int k = i; k < R; ++k) v += L[k * (k + 1) / 2 + i] * (double)af[k];
          const double ai = (double)af[i];
          contrib += v * v - a.lambda * ai * ai - 2.0 * ai * cv[i];
        }
      }
    }
  }
  if (a.fit_partials) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off);
    if (threadIdx.x == 0) a.fit_partials[blockIdx.x] = contrib;
  }
}

// Largest rank whose wave solve keeps the factor's rows in registers (at
// most 168 VGPRs, three waves per SIMD); above it the factor lives in LDS.
constexpr int kSolveRegMax = 36;

// A double read from lane `l` of the wave (l wave-uniform), as an SGPR pair.
__device__ inline double readlane_f64(double v, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Orders the wave's LDS writes before its later reads and back.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// floor(64 / R) rows per wave (16 < R <= kSolveRegMax: three up to R = 21,
// two up to R = 32): lane group gi (lanes gi R .. gi R + R - 1) solves row
// NR * block + gi.  Lane i of a group holds row i of B in registers and
// turns it into row i of the Cholesky factor L, right-looking: at step j the
// diagonal comes from the group's lane j, each lane below scales its L(i, j),
// column j goes through LDS to the group's lanes, and each lane updates its
// own row (M(i, k) -= L(i, j) L(k, j), k > j).  Every step is unrolled, so
// the row stays in registers.  The forward solve L y = c then runs on the
// rows in registers (y_k from the group's lane k), and L goes to LDS (packed
// lower triangle) for the backward solve L^T a = y (row k of L read by the
// lanes of its columns) and the fit partial.  Lanes past the last group
// shadow its row R - 1, and groups past the last row (or on an empty row)
// solve the identity; their results are never stored.  Above the diagonal
// the registers hold values no step reads.
template <int R> constexpr int solve_rows_per_wave() { return kWave / R; }

// The value of v in lane `base + k` of the wave, for every lane (base may
// differ between lanes: one read per lane group); with one group, a readlane.
template <int NR> __device__ inline double group_read(double v, int base, int k) {
  if constexpr (NR == 1) {
    return readlane_f64(v, k);
  } else {
    return __shfl(v, base + k);
  }
}

template <int R>
__global__ void __launch_bounds__(64) k_solve_wave(SolveArgs a) {
  constexpr int W = gram_width(R);
  constexpr int NR = solve_rows_per_wave<R>();
  constexpr int LT = R * (R + 1) / 2; // L(k, i) at k (k + 1) / 2 + i
  __shared__ double colj[NR][R];
  __shared__ double Lm[NR][LT];
  __shared__ double part[kWave];
  const int lane = threadIdx.x;
  const bool act = lane / R < NR; // a group's lane, not a shadow
  const int gi = act ? lane / R : NR - 1;
  const int li = act ? lane % R : R - 1;
  const int base = gi * R;
  const uint64_t row = (uint64_t)blockIdx.x * NR + gi;
  const bool live = row < a.rows;
  const uint64_t rr = live ? row : a.rows - 1; // an in-bounds row for the reads
  const bool solve = live && a.row_ptr[rr + 1] != a.row_ptr[rr];
  const double *g = a.gram + rr * W;
  double m[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const double v = solve ? g[k <= li ? up_index(R, k, li) : 0] : 0.0;
    m[k] = k < li ? v : (k == li ? (solve ? v + a.lambda : 1.0) : 0.0);
  }
  const double ci = solve ? g[R * (R + 1) / 2 + li] : 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const double d = sqrt(group_read<NR>(m[j], base, j));
    m[j] = li == j ? d : (li > j ? m[j] / d : m[j]);
    colj[gi][li] = m[j]; // a shadow lane writes what its group's lane R - 1 writes
    wave_lds_sync();
#pragma unroll
    for (int k = j + 1; k < R; ++k) m[k] -= m[j] * colj[gi][k];
    wave_lds_sync(); // every read of column j before column j + 1 is written
  }
  // forward: y_k = (c_k - sum_{i<k} L(k, i) y_i) / L(k, k), lane k's c
  // holding the bracket when step k comes
  double c = ci, y = 0.0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const double yk = group_read<NR>(c, base, k) / group_read<NR>(m[k], base, k);
    if (li == k) y = yk;
    c -= m[k] * yk;
  }
  // L to LDS, row i at i (i + 1) / 2
#pragma unroll
  for (int k = 0; k < R; ++k)
    if (k <= li) Lm[gi][li * (li + 1) / 2 + k] = m[k];
  wave_lds_sync();
  // backward: a_k = (y_k - sum_{i>k} L(i, k) a_i) / L(k, k), descending k;
  // lane i < k subtracts L(k, i) a_k from its y.  With the fit partial, lane i
  // also sums v_i = (L^T a)_i = sum_{k >= i} L(k, i) a_k, a as stored (FP32)
  double v = 0.0;
  float af = 0.f;
#pragma unroll
  for (int k = R - 1; k >= 0; --k) {
    const double ak = group_read<NR>(y, base, k) / Lm[gi][k * (k + 1) / 2 + k];
    const float akf = (float)ak;
    if (li == k) af = akf;
    const double lki = Lm[gi][k * (k + 1) / 2 + (li <= k ? li : k)];
    if (li < k) y -= lki * ak;
    if (li <= k) v += lki * (double)akf;
  }
  if (act && live) a.A[row * row_stride(R) + li] = solve ? af : 0.f;
  if (a.fit_partials) { // a^T (L L^T) a - lambda |a|^2 - 2 a^T c, summed over the group
    double contrib = 0.0;
    if (act && solve) {
      const double ai = (double)af;
      contrib = v * v - a.lambda * ai * ai - 2.0 * ai * ci;
    }
    part[lane] = contrib;
    wave_lds_sync();
    if (act && live && li == 0) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < R; ++i) s += part[base + i];
      a.fit_partials[row] = s;
    }
  }
}

// One wave per row with the factor in LDS (R > kSolveRegMax): lane i owns
// row i; the register form below would need more than 168 VGPRs.
template <int R>
__global__ void __launch_bounds__(64) k_solve_wave_lds(SolveArgs a) {
  constexpr int W = gram_width(R);
  __shared__ double M[R][R + 1];
  __shared__ double y[R];
  const uint64_t row = blockIdx.x;
  const int lane = threadIdx.x;
  float *out = a.A + row * row_stride(R);
  if (a.row_ptr[row + 1] == a.row_ptr[row]) {
    if (lane < R) out[lane] = 0.f;
    if (a.fit_partials && lane == 0) a.fit_partials[row] = 0.0;
    return;
  }
  const double *g = a.gram + row * W;
  double ci = 0.0;
  if (lane < R) {
    for (int j = 0; j <= lane; ++j) M[lane][j] = g[up_index(R, j, lane)];
    M[lane][lane] += a.lambda;
    ci = g[R * (R + 1) / 2 + lane];
    y[lane] = ci;
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
  if (a.fit_partials) { // a^T (L L^T) a - lambda |a|^2 - 2 a^T c, lane i: (L^T a)_i
    double contrib = 0.0;
    if (lane < R) {
      double v = 0.0;
      for (int k = lane; k < R; ++k) v += M[k][lane] * (double)(float)y[k];
      const double ai = (double)(float)y[lane];
      contrib = v * v - a.lambda * ai * ai - 2.0 * ai * ci;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off);
    if (lane == 0) a.fit_partials[row] = contrib;
  }
}

template <int R> uint64_t solve_impl(const SolveArgs &a, hipStream_t s) {
  if (a.rows == 0) return 0;
  if constexpr (R <= 16) {
    const uint64_t grid = (a.rows + 63) / 64;
    hipLaunchKernelGGL(k_solve_lane<R>, dim3((unsigned)grid), dim3(64), 0, s, a);
    return grid;
  } else if constexpr (R <= kSolveRegMax) {
    constexpr int NR = solve_rows_per_wave<R>();
    const uint64_t grid = (a.rows + NR - 1) / NR;
    hipLaunchKernelGGL(k_solve_wave<R>, dim3((unsigned)grid), dim3(64), 0, s, a);
    return a.rows;
  } else {
    hipLaunchKernelGGL(k_solve_wave_lds<R>, dim3((unsigned)a.rows), dim3(64), 0, s, a);
    return a.rows;
  }
}

} // namespace

uint64_t solve_partials(int R, uint64_t rows) {
  return R <= 16 ? (rows + 63) / 64 : rows;
}

uint64_t launch_solve(int R, const SolveArgs &a, hipStream_t s) {
  uint64_t parts = 0;
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    parts = solve_impl<r>(a, s);                                                                 \
    break;
    GT_FOR_EACH_RANK(GT_CASE)
#undef GT_CASE
  default:
    throw std::invalid_argument("rank not built");
  }
  GT_HIP(hipGetLastError());
  return parts;
}

} // namespace gt
