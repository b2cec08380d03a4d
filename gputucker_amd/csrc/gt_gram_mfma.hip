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
// Gram on MFMA (v_mfma_f64_4x4x4f64).  The deltas of a 64-nonzero tile and
// the values (v = [delta, x], zero-padded to whole 4-column tiles) go to LDS
// column-major; per group of 16 nonzeros each 4x4 block pair (ta <= tc) of
// the upper triangle of sum v v^T takes one MFMA, the four 4x4x4 blocks of
// the instruction holding four nonzeros each.  Column R of the result is
// C = sum x delta.  Each run of equal rows inside the tile gets its own
// masked MFMAs.  The FP32 values enter as FP64, so every product is exact
// and every sum is FP64: B is positive definite to FP64 rounding even for
// rows with fewer than R nonzeros, where FP32-rounded products would shift
// the zero eigenvalues of sum dd^T by about lambda.  The accumulators are
// flushed on row change (complete rows to the Gram table, chunk-straddling
// runs to the partial tables for the fix-up).
#include "gt_contract.hpp"

namespace gt {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float fv2 __attribute__((ext_vector_type(2)));
typedef uint32_t u2 __attribute__((ext_vector_type(2)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

enum DeltaKind { kValu = 0, kBf16 = 1, kBf16R = 2 };
// The delta tile of 64 nonzeros sits in LDS column-major, tileT[col][m]
// (col < 16: delta columns < R, x in column R, zeros above), with a column
// pitch of 72 floats so the Gram operand reads hit distinct banks.
constexpr int kDP = 72;

// DPP row rotation (within each row of 16 lanes) of an FP64 value
template <int CTRL> __device__ inline double dpp_f64(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_update_dpp(0u, (uint32_t)u, CTRL, 0xF, 0xF, false);
  const uint32_t hi = __builtin_amdgcn_update_dpp(0u, (uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
constexpr int kDppRowRor4 = 0x124, kDppRowRor8 = 0x128;

// two floats to a BF16 word pair (a in the low half), round to nearest even
__device__ inline uint32_t pk_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((fv2{a, b}), bf2));
}
__device__ inline float bf16_low_f(uint32_t p) { return __uint_as_float(p << 16); }
__device__ inline float bf16_high_f(uint32_t p) { return __uint_as_float(p & 0xFFFF0000u); }

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
  __shared__ float s_tile[4][16 * kDP];
  __shared__ uint32_t s_row[4][kWave];
  __shared__ uint32_t s_iL[4][kWave];
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

  // Gram on v_mfma_f64_4x4x4f64: four 4x4x4 blocks; block beta takes A[i][k]
  // and B[k][j] from lane 16k + 4beta + i (+ j) and leaves D[i][j] in lane
  // 16i + 4beta + j.  Lane 16k + 4beta + xx supplies, for column tile t, the
  // value X_t = tile[4t + xx][m] of nonzero m = 16mb + 4k + beta, so one MFMA
  // with A = X_ta, B = X_tc adds the (ta, tc) block of sum over those 16
  // nonzeros of v v^T (v = [delta, x]), one partial per beta; the four
  // partials are added at flush.  Only the T(T+1)/2 block pairs ta <= tc of
  // the upper triangle are formed.  Products of FP32 values in FP64 are
  // exact, every sum is FP64.
  constexpr int T = (R + 4) / 4; // column tiles covering columns 0..R
  constexpr int NP = T * (T + 1) / 2;
  const int gk = lane >> 4, gbeta = (lane >> 2) & 3, gx = lane & 3;
  double acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) acc[p] = 0.0;
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
    // add the four block partials (lanes 4 apart inside each row of 16);
    // quad beta then stores the block pairs p with p % 4 == beta
    const int fi = lane >> 4, fj = lane & 3;
    int p = 0;
#pragma unroll
    for (int ta = 0; ta < T; ++ta)
#pragma unroll
      for (int tc = ta; tc < T; ++tc, ++p) {
        double v = acc[p];
        v += dpp_f64<kDppRowRor8>(v);
        v += dpp_f64<kDppRowRor4>(v);
        const int P = 4 * ta + fi, Q = 4 * tc + fj;
        const int at = (P <= Q) ? packed_index<R>(P, Q) : -1;
        if ((p & 3) == gbeta && at >= 0) dst[at] = v;
      }
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
    if constexpr (KIND == kValu) {
      float d[R];
      if (tb + kWave < ce) load_idx(tb + kWave);
      contract_slots<R>(slots, a.A_L, iL, a.A_O, io, a.Gt, lrows, lane, d);
#pragma unroll
      for (int r = 0; r < 4 * T; ++r)
        tile[r * kDP + lane] = (r < R) ? (valid ? d[r] : 0.f) : (r == R ? x : 0.f);
    } else {
      s_iL[w][lane] = iL;
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
      // hi words for the a_hi.G_lo MFMA when R > 10.  v_cvt_pk_bf16_f32
      // rounds two floats to nearest even into one word pair.
      u2 ahi[4];
      u4 a32[4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const uint32_t h01 = pk_bf16(f0[mb][0], f0[mb][1]);
        const uint32_t h23 = pk_bf16(f0[mb][2], f0[mb][3]);
        ahi[mb] = u2{h01, h23};
        if constexpr (KIND == kBf16R) {
          const uint32_t l01 = pk_bf16(f0[mb][0] - bf16_low_f(h01), f0[mb][1] - bf16_high_f(h01));
          const uint32_t l23 = pk_bf16(f0[mb][2] - bf16_low_f(h23), f0[mb][3] - bf16_high_f(h23));
          uint32_t up0 = l01, up1 = l23;
          if constexpr (bf16r_single_mfma(R)) {
            if (plan == kPlanHH) {
              up0 = pk_bf16(f1[mb][0], f1[mb][1]);
              up1 = pk_bf16(f1[mb][2], f1[mb][3]);
            }
            if (plan == kPlanR10) up1 = h01;
          }
          a32[mb] = u4{h01, h23, up0, up1};
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
            u[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, ahi[mb]), bh, z,
                                                              0, 0, 0);
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
              u[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, ahi[mb]), bl,
                                                                u[mb], 0, 0, 0);
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
      // lane (q, g) holds delta[16mb + 4g + i][q]: columns to LDS, then x
      // (zero for padding lanes) over column R
      if (q < 4 * T) {
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
          *reinterpret_cast<f4 *>(&tile[q * kDP + mb * 16 + 4 * g]) =
              f4{dacc[mb][0], dacc[mb][1], dacc[mb][2], dacc[mb][3]};
      }
      tile[R * kDP + lane] = x;
    }
    wave_sync();
    float X[4][T]; // Gram operands of this lane (see above)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int t = 0; t < T; ++t) X[mb][t] = tile[(4 * t + gx) * kDP + mb * 16 + 4 * gk + gbeta];

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
        for (int p = 0; p < NP; ++p) acc[p] = 0.0;
      }
      const uint64_t above = (s >= 63) ? 0ull : (starts & (~0ull << (s + 1)));
      const int e_run = above ? (int)__builtin_ctzll(above) : cnt;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const int lo = s > mb * 16 ? s : mb * 16;
        const int hi = e_run < mb * 16 + 16 ? e_run : mb * 16 + 16;
        if (lo >= hi) continue;
        const int m = mb * 16 + 4 * gk + gbeta;
        const bool in = m >= s && m < e_run;
        double xa[T], xb[T];
#pragma unroll
        for (int t = 0; t < T; ++t) {
          xb[t] = (double)X[mb][t];
          xa[t] = (double)(in ? X[mb][t] : 0.f);
        }
        int p = 0;
#pragma unroll
        for (int ta = 0; ta < T; ++ta)
#pragma unroll
          for (int tc = ta; tc < T; ++tc, ++p)
            acc[p] = __builtin_amdgcn_mfma_f64_4x4x4f64(xa[ta], xb[tc], acc[p], 0, 0, 0);
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
