// gt_gram_mfma.hip -- one mode update's delta + Gram pass with the Gram on
// matrix cores, for R <= 15, in three delta variants:
//
//   kValu   delta by the FP32 VALU contraction (the reference's arithmetic,
//           delta.cuh:62-90), one lane per nonzero, then transposed via LDS
//   kBf16   delta by BF16 MFMA, FP32 accumulate
//   kBf16R  kBf16 plus the split-BF16 correction terms (refinement): with
//           a = a_hi + a_lo and G = G_hi + G_lo (each part BF16),
//           a.G ~ a_hi.G_hi + a_lo.G_hi + a_hi.G_lo, all on MFMA
//
// BF16 delta, two stages per 64-nonzero tile.  Let L be the last mode other
// than n and o the multi-index of the remaining "outer" modes:
//   stage 1 (MFMA 16x16x16 bf16):  U[m][(o,r)] = sum_jL A_L[i_L(m)][jL] G[o][jL][r]
//   stage 2 (VALU FP32):           delta[m][r] = sum_o U[m][(o,r)] * s_m[o],
//                                  s_m[o] = prod_t A_t[i_t(m)][o_t]
// The A operand is the FP32 factor row of each nonzero split to BF16 on the
// fly; the B operand (the core, zero-padded to 16 x 16 per o) comes from a
// small table laid out in lane order.
//
// Gram on MFMA (16x16x4 f32).  After the delta stage lane l holds
// v[mb][i] = delta[m][q] with m = 16 mb + 4 (l/16) + i and q = l % 16, with
// x[m] in column q = R and zeros beyond.  For a fixed (mb, i) the four lane
// groups hold four different nonzeros, so one MFMA with v as both operands
// adds sum_m v[m][p] v[m][q] for those four nonzeros: the upper triangle is
// B = sum delta delta^T, column R is C = sum x delta.  Each run of equal rows
// inside the tile gets its own masked MFMAs; partial sums stay in FP32 for
// at most 64 products and are then added into FP64 accumulators, which are
// flushed on row change exactly like the VALU Gram (complete rows to the
// Gram table, chunk-straddling runs to the partial tables).
#include "gt_contract.hpp"

namespace gt {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

enum DeltaKind { kValu = 0, kBf16 = 1, kBf16R = 2 };

__device__ inline short bf16_bits(float x) { // round to nearest even
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ inline float bf16_value(short h) {
  return __uint_as_float(((uint32_t)(uint16_t)h) << 16);
}

constexpr int kTP = 17; // LDS pitch of a delta tile row (odd: fewer bank conflicts)

template <int R>
__device__ inline int packed_index(int p, int q) {
  if (p <= q && q < R) return p * R - p * (p - 1) / 2 + (q - p);
  if (p < R && q == R) return R * (R + 1) / 2 + p;
  return -1;
}

template <int R, int KIND>
__global__ void __launch_bounds__(256) k_gram_mfma(GramArgs a, int wpb, int O) {
  static_assert(R <= 15, "MFMA Gram covers R <= 15");
  constexpr int W = gram_width(R);
  __shared__ float s_tile[4][kWave * kTP];
  __shared__ uint32_t s_row[4][kWave];
  __shared__ uint32_t s_iL[4][kWave];
  __shared__ float s_x[4][kWave];
  extern __shared__ float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint64_t c = (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.nchunks) return;
  const int N = a.order, n = a.mode;
  const int L = (n == N - 1) ? N - 2 : N - 1;
  const int slots = (N > 2 ? N - 2 : 0);
  const size_t per_wave = (size_t)slots * R * kWave + (KIND == kValu ? 0 : (size_t)O * kWave);
  float *lrows = dyn + (size_t)w * per_wave;
  float *souter = lrows + (size_t)slots * R * kWave;
  float *tile = s_tile[w];
  const uint64_t cb = c * a.chunk;
  const uint64_t ce = (cb + a.chunk < a.st.nnz) ? cb + a.chunk : a.st.nnz;
  const int q = lane & 15, g = lane >> 4;

  int pk[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) pk[r] = packed_index<R>(4 * g + r, q);
  double acc64[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t cur = __builtin_amdgcn_readfirstlane(a.st.idx[n][cb]);
  uint64_t run_begin = cb;

  auto flush = [&](uint32_t row, uint64_t b, uint64_t e_end) {
    const uint64_t rb = a.row_ptr[row], re = a.row_ptr[row + 1];
    double *dst;
    if (b == rb && e_end == re)
      dst = a.gram + (size_t)row * W;
    else if (b == cb)
      dst = a.bpart0 + (size_t)c * W;
    else
      dst = a.bpart1 + (size_t)c * W;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (pk[r] >= 0) dst[pk[r]] = acc64[r];
  };

  for (uint64_t tb = cb; tb < ce; tb += kWave) {
    const uint64_t e = tb + lane;
    const bool valid = e < ce;
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    uint32_t gi[kMaxOrder];
#pragma unroll
    for (int k = 0; k < kMaxOrder; ++k) gi[k] = (k < N && valid) ? a.st.idx[k][e] : 0u;
    const float x = valid ? a.st.val[e] : 0.f;
    s_row[w][lane] = gi[n];
    float v[4][4];
    if constexpr (KIND == kValu) {
      float d[R];
      contract<R>(N, n, a.A, a.Gt, gi, lrows, lane, d);
#pragma unroll
      for (int r = 0; r < 16; ++r) tile[lane * kTP + r] = (r < R) ? (valid ? d[r] : 0.f) : (r == R ? x : 0.f);
      wave_sync();
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) v[mb][i] = tile[(mb * 16 + 4 * g + i) * kTP + q];
    } else {
      // outer rows of this lane's nonzero, then s_m[o] for every o
#pragma unroll
      for (int k = 0; k < kMaxOrder; ++k) {
        if (k >= N || k == n || k == L) continue;
        const float *row = a.A[k] + (size_t)gi[k] * R;
        const int t = k - (k > n ? 1 : 0);
#pragma unroll
        for (int j = 0; j < R; ++j) lrows[(t * R + j) * kWave + lane] = row[j];
      }
      s_iL[w][lane] = gi[L];
      s_x[w][lane] = x;
      {
        int dig[kMaxOrder - 2];
#pragma unroll
        for (int t = 0; t < kMaxOrder - 2; ++t) dig[t] = 0;
        for (int o = 0; o < O; ++o) {
          float s = valid ? 1.f : 0.f;
#pragma unroll
          for (int t = 0; t < kMaxOrder - 2; ++t)
            if (t < slots) s *= lrows[(t * R + dig[t]) * kWave + lane];
          souter[o * kWave + lane] = s;
#pragma unroll
          for (int t = kMaxOrder - 3; t >= 0; --t) {
            if (t < slots) {
              if (++dig[t] < R) break;
              dig[t] = 0;
            }
          }
        }
      }
      wave_sync();
      // A operands: row of A_L for nonzero 16 mb + q, columns 4g..4g+3
      s4 ahi[4], alo[4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const float *rowL = a.A[L] + (size_t)s_iL[w][mb * 16 + q] * R;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int k = 4 * g + t;
          const float val = (k < R) ? rowL[k] : 0.f;
          const short h = bf16_bits(val);
          ahi[mb][t] = h;
          alo[mb][t] = bf16_bits(val - bf16_value(h));
        }
      }
      float dacc[4][4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) dacc[mb][i] = 0.f;
      const s4 *__restrict__ gb_hi = reinterpret_cast<const s4 *>(a.Gb);
      const s4 *__restrict__ gb_lo = gb_hi + (size_t)O * kWave;
      for (int o = 0; o < O; ++o) {
        const s4 bh = gb_hi[(size_t)o * kWave + lane];
        s4 bl;
        if constexpr (KIND == kBf16R) bl = gb_lo[(size_t)o * kWave + lane];
        f4 u[4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
          const f4 z = {0.f, 0.f, 0.f, 0.f};
          u[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[mb], bh, z, 0, 0, 0);
        }
        if constexpr (KIND == kBf16R) {
#pragma unroll
          for (int mb = 0; mb < 4; ++mb)
            u[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo[mb], bh, u[mb], 0, 0, 0);
#pragma unroll
          for (int mb = 0; mb < 4; ++mb)
            u[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[mb], bl, u[mb], 0, 0, 0);
        }
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
          const f4 sv = *reinterpret_cast<const f4 *>(&souter[o * kWave + mb * 16 + 4 * g]);
#pragma unroll
          for (int i = 0; i < 4; ++i) dacc[mb][i] = __builtin_fmaf(u[mb][i], sv[i], dacc[mb][i]);
        }
      }
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          v[mb][i] = (q == R) ? s_x[w][mb * 16 + 4 * g + i] : dacc[mb][i];
    }

    // run boundaries inside the tile (uniform)
    const uint32_t myrow = s_row[w][lane];
    const uint32_t prev = lane ? s_row[w][lane - 1] : cur;
    const uint64_t starts = __ballot(lane < cnt && myrow != prev);
    int s = 0;
    while (s < cnt) {
      if ((starts >> s) & 1ull) {
        flush(cur, run_begin, tb + s);
        cur = __builtin_amdgcn_readfirstlane(s_row[w][s]);
        run_begin = tb + s;
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[r] = 0.0;
      }
      const uint64_t above = (s >= 63) ? 0ull : (starts & (~0ull << (s + 1)));
      const int e_run = above ? (int)__builtin_ctzll(above) : cnt;
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const int lo = s > mb * 16 ? s : mb * 16;
        const int hi = e_run < mb * 16 + 16 ? e_run : mb * 16 + 16;
        if (lo >= hi) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (lo + ((i - lo) & 3) >= hi) continue; // no nonzero of this (mb, i) in the run
          const int m = mb * 16 + 4 * g + i;
          const float val = (m >= s && m < e_run) ? v[mb][i] : 0.f;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(val, val, acc, 0, 0, 0);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[r] += (double)acc[r];
      s = e_run;
    }
    wave_sync();
  }
  flush(cur, run_begin, ce);
}

template <int R, int KIND> void launch_impl(const GramArgs &a, int O, hipStream_t st) {
  if (a.nchunks == 0) return;
  const int slots = a.order > 2 ? a.order - 2 : 0;
  const size_t per_wave =
      ((size_t)slots * R * kWave + (KIND == kValu ? 0 : (size_t)O * kWave)) * sizeof(float);
  int wpb = 4;
  while (wpb > 1 && per_wave * wpb > 64 * 1024) --wpb;
  const unsigned grid = (unsigned)((a.nchunks + wpb - 1) / wpb);
  hipLaunchKernelGGL((k_gram_mfma<R, KIND>), dim3(grid), dim3(64 * wpb), per_wave * wpb, st, a,
                     wpb, O);
}

template <int KIND> void dispatch(int R, const GramArgs &a, int O, hipStream_t st) {
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    launch_impl<r, KIND>(a, O, st);                                                              \
    break;
    GT_CASE(1) GT_CASE(2) GT_CASE(3) GT_CASE(4) GT_CASE(5) GT_CASE(6) GT_CASE(7) GT_CASE(8)
    GT_CASE(9) GT_CASE(10) GT_CASE(11) GT_CASE(12) GT_CASE(13) GT_CASE(14) GT_CASE(15)
#undef GT_CASE
  default:
    throw std::invalid_argument("MFMA Gram needs R <= 15");
  }
}

} // namespace

bool gram_mfma_supported(int R) { return R >= 1 && R <= 15; }

void launch_gram_mfma(int R, int precision, const GramArgs &a, int O, hipStream_t s) {
  if (precision == GT_PREC_FP32)
    dispatch<kValu>(R, a, O, s);
  else if (precision == GT_PREC_BF16)
    dispatch<kBf16>(R, a, O, s);
  else
    dispatch<kBf16R>(R, a, O, s);
  GT_HIP(hipGetLastError());
}

} // namespace gt
