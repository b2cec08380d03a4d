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
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
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
  extern __shared__ __align__(16) float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int slots = a.slots;
  // the core operand tables (BF16 kinds) go to LDS once per workgroup:
  // kBf16: K16 hi table; kBf16R: K32 table, then (R > 10) the K16 lo table
  uint4 *s_tab = reinterpret_cast<uint4 *>(dyn + (size_t)wpb * lds_floats_per_wave<KIND>(slots, R, O));
  const s4 *t16 = reinterpret_cast<const s4 *>(s_tab);
  const s8 *t32 = reinterpret_cast<const s8 *>(s_tab);
  const s4 *t16lo = reinterpret_cast<const s4 *>(s_tab + (size_t)O * kWave);
  if constexpr (KIND != kValu) {
    const uint4 *src = reinterpret_cast<const uint4 *>(a.Gb);
    if constexpr (KIND == kBf16) {
      const int n = (int)(gb_k16_words(O) / 8);
      for (int i = threadIdx.x; i < n; i += blockDim.x) s_tab[i] = src[i];
    } else {
      const uint4 *k32 = src + gb_k16_words(O) / 8;
      const int n = (int)(gb_k32_words(O) / 8);
      for (int i = threadIdx.x; i < n; i += blockDim.x) s_tab[i] = k32[i];
      if constexpr (!bf16r_single_mfma(R)) {
        const uint4 *lo = k32 + n;
        const int n2 = (int)(gb_k16_words(O) / 8);
        for (int i = threadIdx.x; i < n2; i += blockDim.x) s_tab[n + i] = lo[i];
      }
    }
    __syncthreads();
  }
  const uint64_t c = a.chunk_begin + (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.chunk_end) return;
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
  // Does the chunk's first row start before it, or its last row end after it?
  // Two row-pointer reads per chunk; inside the chunk every other run is
  // complete, so flushes need no memory reads.
  const bool head_open = a.row_ptr[cur] < cb;
  const uint32_t last_row = __builtin_amdgcn_readfirstlane(a.idx_row[ce - 1]);
  const bool tail_open = a.row_ptr[last_row + 1] > ce;

  auto flush = [&](uint32_t row, uint64_t b, uint64_t e_end) {
    const bool open = (b == cb && head_open) || (e_end == ce && tail_open);
    double *dst;
    if (!open)
      dst = a.gram + (size_t)row * W;
    else if (b == cb)
      dst = a.bpart0 + (size_t)c * W;
    else
      dst = a.bpart1 + (size_t)c * W;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (pk[r] >= 0) dst[pk[r]] = acc64[r];
  };

  // stream indices of the next tile are prefetched behind the current
  // tile's gathers, so each tile waits for one memory round trip, not three
  uint32_t n_row, n_iL, n_io[kSlots];
  float n_x;
  auto load_idx = [&](uint64_t tb) {
    const uint64_t e = tb + lane;
    const uint64_t ee = e < ce ? e : cb; // in-bounds position for padding lanes
    n_row = a.idx_row[ee];
    n_iL = a.idx_L[ee];
#pragma unroll
    for (int t = 0; t < kSlots; ++t) n_io[t] = (t < slots) ? a.idx_O[t][ee] : 0u;
    n_x = e < ce ? a.st.val[ee] : 0.f;
  };
  load_idx(cb);

  for (uint64_t tb = cb; tb < ce; tb += kWave) {
    const bool valid = tb + lane < ce;
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    const uint32_t row = n_row, iL = n_iL;
    uint32_t io[kSlots];
#pragma unroll
    for (int t = 0; t < kSlots; ++t) io[t] = n_io[t];
    const float x = n_x;
    s_row[w][lane] = valid ? row : 0xFFFFFFFFu;
    float v[4][4];
    if constexpr (KIND == kValu) {
      float d[R];
      if (tb + kWave < ce) load_idx(tb + kWave);
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
      s_iL[w][lane] = iL;
      s_x[w][lane] = x;
      wave_sync();
      // every gather of the tile: the A_L row words of the 16 nonzeros each
      // MFMA block needs (per the lane group's plan), then this lane's outer rows
      const int plan = (KIND == kBf16R) ? bf16r_plan(R, g) : (int)kPlanHL;
      const int off0 = plan == kPlanHH ? 0 : (plan == kPlanR10 ? 8 : 4 * g);
      const bool has0 = plan != kPlanZero && off0 < RS;
      const bool has1 = plan == kPlanHH && RS >= 8;
      f4 f0[4], f1[4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const float *rowp = a.A_L + (size_t)s_iL[w][mb * 16 + q] * RS;
        f0[mb] = f4{0.f, 0.f, 0.f, 0.f};
        f1[mb] = f4{0.f, 0.f, 0.f, 0.f};
        if (has0) f0[mb] = *reinterpret_cast<const f4 *>(rowp + off0);
        if (has1) f1[mb] = *reinterpret_cast<const f4 *>(rowp + 4);
      }
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
      if (tb + kWave < ce) load_idx(tb + kWave);
      // s_m[o] for this lane's nonzero, into LDS for the stage-2 layout
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
      // A operands: kBf16 hi words (K16); kBf16R the plan's K32 words, plus
      // hi words for the a_hi.G_lo MFMA when R > 10
      s4 ahi[4];
      s8 a32[4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        s4 lo0, hi1;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const short h = bf16_bits(f0[mb][t]);
          ahi[mb][t] = h;
          lo0[t] = bf16_bits(f0[mb][t] - bf16_value(h));
          hi1[t] = bf16_bits(f1[mb][t]);
        }
        if constexpr (KIND == kBf16R) {
          s4 up = lo0;
          if (plan == kPlanHH) up = hi1;
          if (plan == kPlanR10) up = s4{lo0[0], lo0[1], ahi[mb][0], ahi[mb][1]};
          a32[mb] = s8{ahi[mb][0], ahi[mb][1], ahi[mb][2], ahi[mb][3], up[0], up[1], up[2], up[3]};
        }
      }
      wave_sync();
      float dacc[4][4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) dacc[mb][i] = 0.f;
      auto stage = [&](int o) {
        f4 u[4];
        const f4 z = {0.f, 0.f, 0.f, 0.f};
        if constexpr (KIND == kBf16) {
          const s4 bh = t16[o * kWave + lane];
#pragma unroll
          for (int mb = 0; mb < 4; ++mb)
            u[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[mb], bh, z, 0, 0, 0);
        } else {
          const b8 b = __builtin_bit_cast(b8, t32[o * kWave + lane]);
#pragma unroll
          for (int mb = 0; mb < 4; ++mb)
            u[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a32[mb]), b, z,
                                                            0, 0, 0);
          if constexpr (!bf16r_single_mfma(R)) {
            const s4 bl = t16lo[o * kWave + lane];
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
              u[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[mb], bl, u[mb], 0, 0, 0);
          }
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
        cur = __builtin_amdgcn_readlane(myrow, s);
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
  if (lane == 0) a.chunk_tail[c] = tail_open ? last_row : 0xFFFFFFFFu;
}

template <int R, int KIND> void launch_impl(const GramArgs &a, int O, hipStream_t st) {
  if (a.chunk_end <= a.chunk_begin) return;
  const size_t per_wave = lds_floats_per_wave<KIND>(a.slots, R, O) * sizeof(float);
  const size_t table = KIND == kValu    ? 0
                       : KIND == kBf16  ? gb_k16_words(O) * 2
                                        : (gb_k32_words(O) + (bf16r_single_mfma(R) ? 0 : gb_k16_words(O))) * 2;
  if constexpr (KIND != kValu) {
    if (per_wave + table > 64 * 1024) { // core tables too large for LDS: FP32 VALU delta
      launch_impl<R, kValu>(a, O, st);
      return;
    }
  }
  int wpb = 4;
  while (wpb > 1 && per_wave * wpb + table > 64 * 1024) --wpb;
  const unsigned grid = (unsigned)((a.chunk_end - a.chunk_begin + wpb - 1) / wpb);
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
