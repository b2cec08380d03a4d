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
// Operands come in "slot order" (built on the host per mode): L is the last
// mode other than n, the outer slots t are the remaining other modes
// ascending, o is the mixed-radix index over the outer slots (last fastest).
// The kernel never indexes a register array by a runtime mode number, so all
// gathers of a tile are issued back to back.  Factor rows have a stride of
// R rounded up to 4 floats and are read with 16-byte loads.
//
// VALU delta: delta[r] = sum_o s[o] sum_jL A_L[jL] G[o][jL][r], with
// s[o] = prod_t A_t[o_t]; for N = 3 (one outer slot) the whole contraction
// is unrolled with the core read through scalar loads.
//
// BF16 delta, two stages per 64-nonzero tile:
//   stage 1 (MFMA 16x16x16 bf16):  U[m][(o,r)] = sum_jL A_L[i_L(m)][jL] G[o][jL][r]
//   stage 2 (VALU FP32):           delta[m][r] = sum_o U[m][(o,r)] * s_m[o]
// The A operand is the FP32 factor row of each nonzero split to BF16 on the
// fly; the B operand (the core, zero-padded to 16 x 16 per o) comes from a
// small table laid out in lane order.
//
// Gram on MFMA (16x16x4 f64).  After the delta stage lane l holds
// v[mb][i] = delta[m][q] with m = 16 mb + 4 (l/16) + i and q = l % 16, with
// x[m] in column q = R and zeros beyond.  For a fixed (mb, i) the four lane
// groups hold four different nonzeros, so one MFMA with v as both operands
// adds sum_m v[m][p] v[m][q] for those four nonzeros: the upper triangle is
// B = sum delta delta^T, column R is C = sum x delta.  Each run of equal rows
// inside the tile gets its own masked MFMAs.  The FP32 deltas enter as FP64,
// so every product is exact and every sum is FP64: B is positive definite to
// FP64 rounding even for rows with fewer than R nonzeros, where FP32-rounded
// products would shift the zero eigenvalues of sum dd^T by about lambda.
// The accumulators are flushed on row change (complete rows to the Gram
// table, chunk-straddling runs to the partial tables for the fix-up).
#include "gt_contract.hpp"

namespace gt {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

enum DeltaKind { kValu = 0, kBf16 = 1, kBf16R = 2 };
constexpr int kTP = 17; // LDS pitch of a delta tile row (odd: fewer bank conflicts)

__device__ inline short bf16_bits(float x) { // round to nearest even
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ inline float bf16_value(short h) {
  return __uint_as_float(((uint32_t)(uint16_t)h) << 16);
}

template <int R> __device__ inline int packed_index(int p, int q) {
  if (p <= q && q < R) return p * R - p * (p - 1) / 2 + (q - p);
  if (p < R && q == R) return R * (R + 1) / 2 + p;
  return -1;
}

template <int KIND> __host__ __device__ constexpr size_t lds_floats_per_wave(int slots, int R,
                                                                             int O) {
  return (slots >= 2 ? (size_t)slots * R * kWave : 0) + (KIND == kValu ? 0 : (size_t)O * kWave);
}

template <int R, int KIND>
__global__ void __launch_bounds__(256) k_gram_mfma(GramArgs a, int wpb, int O) {
  static_assert(R <= 15, "MFMA Gram covers R <= 15");
  constexpr int W = gram_width(R);
  constexpr int RS = row_stride(R);
  __shared__ float s_tile[4][kWave * kTP];
  __shared__ uint32_t s_row[4][kWave];
  __shared__ uint32_t s_iL[4][kWave];
  __shared__ float s_x[4][kWave];
  extern __shared__ float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int slots = a.slots;
  // the core operand table (BF16 kinds) goes to LDS once per workgroup
  s4 *s_gb = reinterpret_cast<s4 *>(dyn + (size_t)wpb * lds_floats_per_wave<KIND>(slots, R, O));
  if constexpr (KIND != kValu) {
    const int tab = (KIND == kBf16R ? 2 : 1) * O * kWave;
    const s4 *src = reinterpret_cast<const s4 *>(a.Gb);
    for (int i = threadIdx.x; i < tab; i += blockDim.x) s_gb[i] = src[i];
    __syncthreads();
  }
  const uint64_t c = (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.nchunks) return;
  float *lrows = dyn + (size_t)w * lds_floats_per_wave<KIND>(slots, R, O);
  float *souter = lrows + (slots >= 2 ? (size_t)slots * R * kWave : 0);
  float *tile = s_tile[w];
  const uint64_t cb = c * a.chunk;
  const uint64_t ce = (cb + a.chunk < a.st.nnz) ? cb + a.chunk : a.st.nnz;
  const int q = lane & 15, g = lane >> 4;

  // v_mfma_f64_16x16x4f64 leaves D[p][q] in lane 16 (p % 4) + q, register
  // p / 4 (rows interleaved across the lane groups)
  int pk[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) pk[r] = packed_index<R>(g + 4 * r, q);
  d4 acc64 = {0.0, 0.0, 0.0, 0.0};
  uint32_t cur = __builtin_amdgcn_readfirstlane(a.idx_row[cb]);
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
    const uint64_t ee = valid ? e : cb; // in-bounds position for padding lanes
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    // ---- coalesced stream reads, then every gather of the tile
    const uint32_t row = a.idx_row[ee];
    const uint32_t iL = a.idx_L[ee];
    uint32_t io[kSlots];
#pragma unroll
    for (int t = 0; t < kSlots; ++t) io[t] = (t < slots) ? a.idx_O[t][ee] : 0u;
    const float x = valid ? a.st.val[ee] : 0.f;
    s_row[w][lane] = valid ? row : 0xFFFFFFFFu;
    float v[4][4];
    if constexpr (KIND == kValu) {
      float d[R];
      contract_slots<R>(slots, a.A_L, iL, a.A_O, io, a.Gt, lrows, lane, d);
#pragma unroll
      for (int r = 0; r < 16; ++r)
        tile[lane * kTP + r] = (r < R) ? (valid ? d[r] : 0.f) : (r == R ? x : 0.f);
      wave_sync();
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) v[mb][i] = tile[(mb * 16 + 4 * g + i) * kTP + q];
    } else {
      // outer rows of this lane's nonzero
      float aO[RS];
      if (slots == 1) load_row<R>(a.A_O[0], io[0], aO);
      if (slots >= 2) {
#pragma unroll
        for (int t = 0; t < kSlots; ++t) {
          if (t >= slots) continue;
          float rt[RS];
          load_row<R>(a.A_O[t], io[t], rt);
#pragma unroll
          for (int j = 0; j < R; ++j) lrows[(t * R + j) * kWave + lane] = rt[j];
        }
      }
      // s_m[o] for this lane's nonzero, into LDS for the stage-2 layout
      s_iL[w][lane] = iL;
      s_x[w][lane] = x;
      if (slots == 0) {
        souter[lane] = valid ? 1.f : 0.f;
      } else if (slots == 1) {
#pragma unroll
        for (int o = 0; o < R; ++o) souter[o * kWave + lane] = valid ? aO[o] : 0.f;
      } else {
        int dig[kSlots];
#pragma unroll
        for (int t = 0; t < kSlots; ++t) dig[t] = 0;
        for (int o = 0; o < O; ++o) {
          float s = valid ? 1.f : 0.f;
#pragma unroll
          for (int t = 0; t < kSlots; ++t)
            if (t < slots) s *= lrows[(t * R + dig[t]) * kWave + lane];
          souter[o * kWave + lane] = s;
#pragma unroll
          for (int t = kSlots - 1; t >= 0; --t) {
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
        f4 val = {0.f, 0.f, 0.f, 0.f};
        if (4 * g < RS)
          val = *reinterpret_cast<const f4 *>(a.A_L + (size_t)s_iL[w][mb * 16 + q] * RS + 4 * g);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const short h = bf16_bits(val[t]);
          ahi[mb][t] = h;
          alo[mb][t] = bf16_bits(val[t] - bf16_value(h));
        }
      }
      float dacc[4][4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) dacc[mb][i] = 0.f;
      auto stage = [&](int o) {
        const s4 bh = s_gb[o * kWave + lane];
        s4 bl;
        if constexpr (KIND == kBf16R) bl = s_gb[(O + o) * kWave + lane];
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
      };
      if (slots == 0) {
        stage(0);
      } else if (slots == 1) { // N = 3: O = R, unrolled so the operand reads pipeline
#pragma unroll
        for (int o = 0; o < R; ++o) stage(o);
      } else {
        for (int o = 0; o < O; ++o) stage(o);
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
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const int lo = s > mb * 16 ? s : mb * 16;
        const int hi = e_run < mb * 16 + 16 ? e_run : mb * 16 + 16;
        if (lo >= hi) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (lo + ((i - lo) & 3) >= hi) continue; // no nonzero of this (mb, i) in the run
          const int m = mb * 16 + 4 * g + i;
          const double val = (m >= s && m < e_run) ? (double)v[mb][i] : 0.0;
          acc64 = __builtin_amdgcn_mfma_f64_16x16x4f64(val, val, acc64, 0, 0, 0);
        }
      }
      s = e_run;
    }
    wave_sync();
  }
  flush(cur, run_begin, ce);
}

template <int R, int KIND> void launch_impl(const GramArgs &a, int O, hipStream_t st) {
  if (a.nchunks == 0) return;
  const size_t per_wave = lds_floats_per_wave<KIND>(a.slots, R, O) * sizeof(float);
  const size_t table = KIND == kValu ? 0 : (size_t)(KIND == kBf16R ? 2 : 1) * O * kWave * 8;
  int wpb = 4;
  while (wpb > 1 && per_wave * wpb + table > 64 * 1024) --wpb;
  const unsigned grid = (unsigned)((a.nchunks + wpb - 1) / wpb);
  hipLaunchKernelGGL((k_gram_mfma<R, KIND>), dim3(grid), dim3(64 * wpb), per_wave * wpb + table,
                     st, a, wpb, O);
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
