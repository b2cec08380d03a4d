// gt_gram_mfma.hip -- one mode update's delta + Gram pass with the Gram on
// matrix cores, for R <= 15, in three delta variants:
//
//   kValu   delta by the FP32 VALU contraction (the reference's arithmetic,
//           delta.cuh:62-90), one lane per nonzero
//   kBf16   delta by BF16 MFMA, FP32 accumulate
//   kBf16R  kBf16 plus the split-BF16 correction terms (refinement): with
//           a = a_hi + a_lo and G = G_hi + G_lo (each part BF16),
//           a.G ~ a_hi.G_hi + a_lo.G_hi + a_hi.G_lo, all on MFMA
//
// Operands come in "slot order" (built on the host per mode): L is the last
// mode other than n, the outer slots t are the remaining other modes
// ascending, o is the mixed-radix index over the outer slots (last fastest).
// The kernels never index a register array by a runtime mode number.
//
// BF16 delta, two stages per 64-nonzero tile:
//   stage 1 (MFMA):      U[m][(o,r)] = sum_jL A_L[i_L(m)][jL] G[o][jL][r]
//   stage 2 (VALU FP32): delta[m][r] = sum_o U[m][(o,r)] * s_m[o]
// with s_m[o] = prod_t A_t[i_t(m)][o_t].  The nonzero-side operand of stage
// 1 is read ready-made from a split copy of A_L (k_split_factor: per row and
// lane group the 16 bytes of BF16 words the group's K=32 slots take, see
// bf16r_plan), so a tile spends no arithmetic on it; the core operand comes
// from a small table in LDS laid out in lane order.  The gathered
// operands of the next tile and the stream indices of the tile after it are
// loaded while the current tile is computed.
//
// Gram on MFMA (v_mfma_f64_4x4x4f64).  The deltas of a tile and the values
// (v = [delta, x], zero-padded to whole 4-column tiles) go to LDS
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
#include <utility>

#include "gt_bf16.hpp"
#include "gt_contract.hpp"

namespace gt {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef uint32_t u2 __attribute__((ext_vector_type(2)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

enum DeltaKind { kValu = 0, kBf16 = 1, kBf16R = 2 };

// The delta tile of 64 nonzeros sits in LDS column-major, tile[col][m]
// (col < 16: delta columns < R, x in column R, zeros above), with a column
// pitch of 72 floats so the Gram operand reads hit distinct banks.
constexpr int kDP = 72;

// DPP row rotation (within each row of 16 lanes) of an FP64 value
template <int CTRL> __device__ inline double dpp_f64(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  // all rows and banks enabled and every source lane valid: no old value needed
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xF, 0xF, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
constexpr int kDppRowRor4 = 0x124, kDppRowRor8 = 0x128;

template <int R> __device__ inline int packed_index(int p, int q) {
  if (p <= q && q < R) return p * R - p * (p - 1) / 2 + (q - p);
  if (p < R && q == R) return R * (R + 1) / 2 + p;
  return -1;
}

// The static LDS of the Gram kernels (delta tiles and row ids of four waves;
// the FP32 kernel keeps two of each per wave), which the dynamic part must
// leave room for within the device's LDS per workgroup (lds_limit(): 160 KiB
// on gfx950).  A kernel variant that keeps its core table in LDS is chosen
// only when the LDS still holds kMinLdsWavesPerCu waves per CU: below that
// the waves cannot hide the gathers' latency, and the variant that reads the
// core operand from L2 at three waves per SIMD is faster.
constexpr int kMinLdsWavesPerCu = 8;
constexpr size_t kStaticLdsValu = 4 * 2 * (16 * kDP * sizeof(float) + kWave * sizeof(uint32_t));
constexpr size_t kStaticLdsBf16 = 4 * (16 * kDP * sizeof(float) + kWave * sizeof(uint32_t));

template <int KIND> __host__ __device__ constexpr size_t lds_floats_per_wave(int slots, int R,
                                                                             int O) {
  return (slots >= 2 ? (size_t)slots * R * kWave : 0) + (KIND == kValu ? 0 : (size_t)O * kWave);
}

// ------------------------------------------------------------ Gram of a chunk
// One wave's Gram accumulation over its chunk [cb, ce) of the mode-n stream.
// v_mfma_f64_4x4x4f64 holds four 4x4x4 blocks; block beta takes A[i][k] and
// B[k][j] from lane 16k + 4beta + i (+ j) and leaves D[i][j] in lane
// 16i + 4beta + j.  Lane 16k + 4beta + xx supplies, for column tile t, the
// value X_t = tile[4t + xx][m] of nonzero m = 16mb + 4k + beta, so one MFMA
// with A = X_ta, B = X_tc adds the (ta, tc) block of sum v v^T over those 16
// nonzeros, one partial per beta; the four partials are added at flush.
template <int R> struct ChunkGram {
  static constexpr int W = gram_width(R);
  static constexpr int T = (R + 4) / 4; // column tiles covering columns 0..R
  static constexpr int NP = T * (T + 1) / 2;
  static constexpr int NS = (NP + 3) / 4; // block pairs stored per lane at flush
  const GramArgs &a;
  uint64_t c, cb, ce;
  int lane, gk, gbeta, gx;
  bool head_open, tail_open;
  uint32_t cur, last_row;
  uint64_t run_begin;
  double acc[NP];
  // Gram-row offset this lane stores for pair 4k + gbeta, or -1: lane
  // 16i + 4beta + j holds entry (4ta + i, 4tc + j) of pair (ta, tc)
  int st_at[NS];

  __device__ ChunkGram(const GramArgs &args, uint64_t chunk, uint64_t b, uint64_t e, int l)
      : a(args), c(chunk), cb(b), ce(e), lane(l), gk(l >> 4), gbeta((l >> 2) & 3), gx(l & 3) {
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] = 0.0;
    const int fi = lane >> 4, fj = lane & 3;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      int at = -1, p = 0;
#pragma unroll
      for (int ta = 0; ta < T; ++ta)
#pragma unroll
        for (int tc = ta; tc < T; ++tc, ++p) {
          const int P = 4 * ta + fi, Q = 4 * tc + fj;
          if (p == 4 * k + gbeta && P <= Q) at = packed_index<R>(P, Q);
        }
      st_at[k] = at;
    }
    cur = __builtin_amdgcn_readfirstlane(a.idx_row[cb]);
    run_begin = cb;
    // Does the chunk's first row start before it, or its last row end after
    // it?  Two row-pointer reads per chunk; every other run of the chunk is
    // a whole row, so flushes need no memory reads.
    head_open = a.row_ptr[cur] < cb;
    last_row = __builtin_amdgcn_readfirstlane(a.idx_row[ce - 1]);
    tail_open = a.row_ptr[last_row + 1] > ce;
  }

  __device__ void flush(uint64_t e_end) {
    const bool open = (run_begin == cb && head_open) || (e_end == ce && tail_open);
    double *dst;
    if (!open)
      dst = a.gram + (size_t)cur * W;
    else if (run_begin == cb)
      dst = a.bpart0 + (size_t)c * W;
    else
      dst = a.bpart1 + (size_t)c * W;
    // add the four block partials (lanes 4 apart inside each row of 16),
    // every pair's step before the next step so the pairs' DPP chains
    // interleave; quad beta then stores the block pairs p with p % 4 == beta
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] += dpp_f64<kDppRowRor8>(acc[p]);
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] += dpp_f64<kDppRowRor4>(acc[p]);
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double v = acc[4 * k];
#pragma unroll
      for (int b = 1; b < 4; ++b)
        if (4 * k + b < NP) v = (gbeta == b) ? acc[4 * k + b] : v;
      if (st_at[k] >= 0) dst[st_at[k]] = v;
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] = 0.0;
  }

  // Adds tile [tb, tb + cnt) (values in `tile`, row ids in `srow`); the
  // tile must be visible to the whole wave.
  __device__ void add_tile(const float *tile, const uint32_t *srow, uint64_t tb, int cnt) {
    float X[4][T];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int t = 0; t < T; ++t) X[mb][t] = tile[(4 * t + gx) * kDP + mb * 16 + 4 * gk + gbeta];
    const uint32_t myrow = srow[lane];
    const uint32_t prev = lane ? srow[lane - 1] : cur;
    const uint64_t starts = __ballot(lane < cnt && myrow != prev);
    if (starts == 0 && cnt == kWave) {
      // the whole tile continues the current run (the common case inside
      // long rows): no masks.  The condition is wave-uniform, a scalar branch.
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        double xb[T];
#pragma unroll
        for (int t = 0; t < T; ++t) xb[t] = (double)X[mb][t];
        int p = 0;
#pragma unroll
        for (int ta = 0; ta < T; ++ta)
#pragma unroll
          for (int tc = ta; tc < T; ++tc, ++p)
            acc[p] = __builtin_amdgcn_mfma_f64_4x4x4f64(xb[ta], xb[tc], acc[p], 0, 0, 0);
      }
      return;
    }
    int s = 0;
    while (s < cnt) {
      if ((starts >> s) & 1ull) {
        flush(tb + s);
        cur = __builtin_amdgcn_readlane(myrow, s);
        run_begin = tb + s;
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
        // (a select, not a branch: MFMA ignores EXEC, so no MFMA may sit under
        // a condition the compiler could turn into an EXEC mask)
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
  }

  __device__ void finish() {
    flush(ce);
    if (lane == 0) a.chunk_tail[c] = tail_open ? last_row : 0xFFFFFFFFu;
  }
};

// Test hook: the deltas of tile [tb, tb + cnt) (columns < R of `tile`) to
// a.delta_out in stream order.  a.delta_out is a kernel argument, so the
// branch is uniform; it is null on the ALS path.
template <int R>
__device__ inline void dump_tile(const GramArgs &a, const float *tile, uint64_t tb, int cnt,
                                 int lane) {
  if (a.delta_out != nullptr && lane < cnt) {
#pragma unroll
    for (int r = 0; r < R; ++r) a.delta_out[(tb + lane) * R + r] = tile[r * kDP + lane];
  }
}

// ------------------------------------------------------------ FP32 VALU delta
// A chunk with NS outer slots and R^N <= 4096: two tiles per delta pass
// (contract_pair), the Gram of each tile added in turn.
template <int R, int NS>
__device__ inline void valu_pairs(const GramArgs &a, ChunkGram<R> &gram, uint64_t cb, uint64_t ce,
                                  int lane, float (*tiles)[16 * kDP],
                                  uint32_t (*rows)[kWave]) {
  constexpr int T = ChunkGram<R>::T;
  uint32_t p_row[2], p_iL[2], p_io[2][kSlots];
  float p_x[2];
  auto load_pair = [&](uint64_t tb) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint64_t e = tb + h * kWave + lane;
      const uint64_t ee = e < ce ? e : cb;
      p_row[h] = a.idx_row[ee];
      p_iL[h] = a.idx_L[ee];
#pragma unroll
      for (int t = 0; t < NS; ++t) p_io[h][t] = a.idx_O[t][ee];
      p_x[h] = e < ce ? a.st.val[ee] : 0.f;
    }
  };
  load_pair(cb);
  for (uint64_t tb = cb; tb < ce; tb += 2 * kWave) {
    const uint32_t row[2] = {p_row[0], p_row[1]};
    const uint32_t iL[2] = {p_iL[0], p_iL[1]};
    uint32_t io[2][kSlots];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < kSlots; ++t) io[h][t] = t < NS ? p_io[h][t] : 0u;
    const float x[2] = {p_x[0], p_x[1]};
    if (tb + 2 * kWave < ce) load_pair(tb + 2 * kWave);
    float d[2][R];
    contract_pair<R, NS>(a.A_L, a.A_O, iL, io, core_operand(a.Gt), d[0], d[1]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bool valid = tb + h * kWave + lane < ce;
      rows[h][lane] = valid ? row[h] : 0xFFFFFFFFu;
#pragma unroll
      for (int r = 0; r < 4 * T; ++r)
        tiles[h][r * kDP + lane] = (r < R) ? (valid ? d[h][r] : 0.f) : (r == R ? x[h] : 0.f);
    }
    wave_sync();
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const uint64_t th = tb + h * kWave;
      if (th >= ce) break;
      const int cnt = (ce - th < (uint64_t)kWave) ? (int)(ce - th) : kWave;
      dump_tile<R>(a, tiles[h], th, cnt, lane);
      gram.add_tile(tiles[h], rows[h], th, cnt);
    }
    wave_sync();
  }
  gram.finish();
}

template <int R>
__global__ void __launch_bounds__(256) k_gram_valu(GramArgs a, int wpb) {
  static_assert(R <= 15, "MFMA Gram covers R <= 15");
  constexpr int T = ChunkGram<R>::T;
  __shared__ float s_tile[4][2][16 * kDP];
  __shared__ uint32_t s_row[4][2][kWave];
  extern __shared__ __align__(16) float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  const int slots = a.slots;
  const uint64_t c = a.chunk_begin + (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.chunk_end) return;
  float *lrows = dyn + (size_t)w * lds_floats_per_wave<kValu>(slots, R, 1);
  float *tile = s_tile[w][0];
  const uint64_t cb = c * a.chunk;
  const uint64_t ce = (cb + a.chunk < a.st.nnz) ? cb + a.chunk : a.st.nnz;
  ChunkGram<R> gram(a, c, cb, ce, lane);

  // orders 3 to 5 with small enough R^N: the unrolled two-tile contraction
  if constexpr (pair_unrolled<R, 1>()) {
    if (slots == 1) {
      valu_pairs<R, 1>(a, gram, cb, ce, lane, s_tile[w], s_row[w]);
      return;
    }
  }
  if constexpr (pair_unrolled<R, 2>()) {
    if (slots == 2) {
      valu_pairs<R, 2>(a, gram, cb, ce, lane, s_tile[w], s_row[w]);
      return;
    }
  }
  if constexpr (pair_unrolled<R, 3>()) {
    if (slots == 3) {
      valu_pairs<R, 3>(a, gram, cb, ce, lane, s_tile[w], s_row[w]);
      return;
    }
  }

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
    s_row[w][0][lane] = valid ? row : 0xFFFFFFFFu;
    float d[R];
    if (tb + kWave < ce) load_idx(tb + kWave);
    contract_slots<R>(slots, a.A_L, iL, a.A_O, io, a.Gt, lrows, lane, d);
#pragma unroll
    for (int r = 0; r < 4 * T; ++r)
      tile[r * kDP + lane] = (r < R) ? (valid ? d[r] : 0.f) : (r == R ? x : 0.f);
    wave_sync();
    dump_tile<R>(a, tile, tb, cnt, lane);
    gram.add_tile(tile, s_row[w][0], tb, cnt);
    wave_sync();
  }
  gram.finish();
}

// ------------------------------------------------------------ BF16 MFMA delta
// Three waves per SIMD (at most 168 VGPRs) measured fastest for R <= 10 on
// the order-3 workload: the tile pipeline needs the extra wave to cover
// gather latency; four waves spill.  Above R = 10 the second MFMA's
// operands need the registers of a third wave.
template <int R, int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(R <= 10 ? 3 : 2)))
k_gram_bf16(GramArgs a, int wpb, int O) {
  static_assert(R <= 15, "MFMA Gram covers R <= 15");
  constexpr int RS = row_stride(R);
  constexpr int T = ChunkGram<R>::T;
  __shared__ float s_tile[4][16 * kDP];
  __shared__ uint32_t s_row[4][kWave];
  extern __shared__ __align__(16) float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  const int slots = a.slots;
  // the core operand tables go to LDS once per workgroup:
  // kBf16: K16 hi table; kBf16R: K32 table, then (R > 10) the K32 lo table
  uint4 *s_tab = reinterpret_cast<uint4 *>(dyn + (size_t)wpb * lds_floats_per_wave<KIND>(slots, R, O));
  const s4 *t16 = reinterpret_cast<const s4 *>(s_tab);
  const s8 *t32 = reinterpret_cast<const s8 *>(s_tab);
  const s8 *t32lo = reinterpret_cast<const s8 *>(s_tab + (size_t)O * kWave);
  {
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
        const int n2 = (int)(gb_k32_words(O) / 8);
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
  ChunkGram<R> gram(a, c, cb, ce, lane);

  // Pipeline: while tile t is computed, the gathers of tile t + 1 and the
  // stream loads of tile t + 2 are in flight.  Positions past the chunk are
  // clamped to its first position, so every load is in bounds and none
  // needs a branch.
  uint32_t n_row, n_iL, n_io[kSlots];
  float n_x;
  auto load_idx = [&](uint64_t tb) {
    const uint64_t e = tb + lane;
    const uint64_t ee = e < ce ? e : cb;
    n_row = a.idx_row[ee];
    n_iL = a.idx_L[ee];
#pragma unroll
    for (int t = 0; t < kSlots; ++t) n_io[t] = (t < slots) ? a.idx_O[t][ee] : 0u;
    n_x = e < ce ? a.st.val[ee] : 0.f;
  };
  // lane (q, g) takes, for block mb, the split words of nonzero 16 mb + q
  u4 g32[4];
  float gO[RS];
  auto gather = [&](uint32_t iL, uint32_t io0) {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const uint32_t r = (uint32_t)__shfl((int)iL, mb * 16 + q);
      g32[mb] = *reinterpret_cast<const u4 *>(a.As_L + ((size_t)r * 4 + g) * 4);
    }
    if (slots == 1) load_row<R>(a.A_O[0], io0, gO);
  };
  load_idx(cb);
  uint32_t c_row = n_row;
  float c_x = n_x;
  uint32_t c_io[kSlots];
#pragma unroll
  for (int t = 0; t < kSlots; ++t) c_io[t] = n_io[t];
  gather(n_iL, n_io[0]);
  load_idx(cb + kWave);

  for (uint64_t tb = cb; tb < ce; tb += kWave) {
    const bool valid = tb + lane < ce;
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    // this tile's operands; then start the next tile's gathers and the
    // stream loads of the tile after it
    u4 a32[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) a32[mb] = g32[mb];
    const uint32_t row = c_row;
    const float x = c_x;
    uint32_t io[kSlots];
#pragma unroll
    for (int t = 0; t < kSlots; ++t) io[t] = c_io[t];
    s_row[w][lane] = valid ? row : 0xFFFFFFFFu;
    // N = 3: s_m[o] is this lane's outer row, written to LDS before the
    // next gather reuses its registers (padding lanes gathered in-bounds
    // rows; their deltas never enter the Gram)
    if (slots == 1) {
#pragma unroll
      for (int o = 0; o < R; ++o) souter[o * kWave + lane] = gO[o];
    }
    c_row = n_row;
    c_x = n_x;
#pragma unroll
    for (int t = 0; t < kSlots; ++t) c_io[t] = n_io[t];
    gather(n_iL, n_io[0]);
    load_idx(tb + 2 * kWave);

    // s_m[o] for this lane's nonzero, into LDS for the stage-2 layout
    if (slots == 0) {
      souter[lane] = valid ? 1.f : 0.f;
    } else if (slots >= 2) {
#pragma unroll
      for (int t = 0; t < kSlots; ++t) {
        if (t >= slots) continue;
        float rt[RS];
        load_row<R>(a.A_O[t], io[t], rt);
#pragma unroll
        for (int j = 0; j < R; ++j) lrows[(t * R + j) * kWave + lane] = rt[j];
      }
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
    float dacc[4][4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) dacc[mb][i] = 0.f;
    // The core table is the A operand and the split A_L words the B
    // operand, so each MFMA yields U transposed: lane (q, g) holds
    // U[16mb + q][o][4g + i], i < 4, and stage 2 needs one s value per
    // block and o (s_m[o] of the lane's own column nonzero) for four FMAs.
    auto stage = [&](int o) {
      f4 u[4];
      const f4 z = {0.f, 0.f, 0.f, 0.f};
      if constexpr (KIND == kBf16) {
        const s4 bh = t16[o * kWave + lane];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
          u[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
              bh, __builtin_bit_cast(s4, u2{a32[mb][0], a32[mb][1]}), z, 0, 0, 0);
      } else {
        const b8 b = __builtin_bit_cast(b8, t32[o * kWave + lane]);
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
          u[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, __builtin_bit_cast(b8, a32[mb]), z, 0,
                                                          0, 0);
        if constexpr (!bf16r_single_mfma(R)) {
          const b8 bl = __builtin_bit_cast(b8, t32lo[o * kWave + lane]);
#pragma unroll
          for (int mb = 0; mb < 4; ++mb)
            u[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, __builtin_bit_cast(b8, a32[mb]),
                                                            u[mb], 0, 0, 0);
        }
      }
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const float sv = souter[o * kWave + mb * 16 + q];
#pragma unroll
        for (int i = 0; i < 4; ++i) dacc[mb][i] = __builtin_fmaf(u[mb][i], sv, dacc[mb][i]);
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
    // lane (q, g) holds delta[16mb + q][4g + i]: columns to LDS, then x
    // (zero for padding lanes) over column R
    if (g < T) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) tile[(4 * g + i) * kDP + mb * 16 + q] = dacc[mb][i];
    }
    tile[R * kDP + lane] = x;
    wave_sync();
    dump_tile<R>(a, tile, tb, cnt, lane);
    gram.add_tile(tile, s_row[w], tb, cnt);
    wave_sync();
  }
  gram.finish();
}

// ------------------------------------------------------------ BF16, order 5, pairs
// The refined BF16 delta of an order-5 tensor with more outer indices than
// the o-block kernel below takes (R = 6..8: 216 to 512), one MFMA per 16
// nonzeros and two outer indices (the pair table).  The core operand of each
// MFMA comes from the table in global memory (it stays in L2), kPairRing
// steps ahead; s_m[o] = prod_t A_t[i_t][o_t] is formed at its use from the
// outer rows staged in LDS.  The next tile's gathers (split A_L words and the
// outer rows) are issued after this tile's stage loop, so the core loads
// never wait behind them.  (The pair table, 83 KiB at R = 6, does not fit in
// LDS beside enough waves.)
// core operands in flight in the pair loop (a ring of registers, one per step)
constexpr int kPairRing = 2;
// the workgroup of a kernel whose waves share one table copy in LDS: 12
// waves, three per SIMD (the VGPR budget of amdgpu_waves_per_eu(3))
constexpr int kTableWaves = 12;

// LDS floats per wave of k_gram_bf16_o5: the outer rows, the delta tile, the
// tile's row ids
template <int R> constexpr size_t o5_lds_floats_per_wave() {
  return (size_t)3 * R * kWave + 16 * kDP + kWave;
}

template <int R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_gram_bf16_o5(GramArgs a, int wpb) {
  static_assert(bf16r_pair_table(R) && cpow(R, 3) > kObMaxO,
                "one refined MFMA per two outer indices, beyond the o-block kernel");
  constexpr int NS = 3;
  constexpr int RS = row_stride(R);
  constexpr int T = ChunkGram<R>::T;
  constexpr int O = cpow(R, NS);
  constexpr int NB = gb_pairs(O); // MFMAs per 16 nonzeros
  extern __shared__ __align__(16) float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  const s8 *tg = reinterpret_cast<const s8 *>(static_cast<const uint16_t *>(a.Gb) +
                                              gb_k16_words(O) + 2 * gb_k32_words(O));
  const uint64_t c = a.chunk_begin + (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.chunk_end) return;
  float *lrows = dyn + (size_t)w * o5_lds_floats_per_wave<R>(); // [slot][j][nonzero]
  float *tile = lrows + NS * R * kWave;
  uint32_t *s_rowv = reinterpret_cast<uint32_t *>(tile + 16 * kDP);
  const uint64_t cb = c * a.chunk;
  const uint64_t ce = (cb + a.chunk < a.st.nnz) ? cb + a.chunk : a.st.nnz;
  const int q = lane & 15, g = lane >> 4;
  ChunkGram<R> gram(a, c, cb, ce, lane);

  uint32_t n_row, n_iL, n_io[NS];
  float n_x;
  auto load_idx = [&](uint64_t tb) {
    const uint64_t e = tb + lane;
    const uint64_t ee = e < ce ? e : cb;
    n_row = a.idx_row[ee];
    n_iL = a.idx_L[ee];
#pragma unroll
    for (int t = 0; t < NS; ++t) n_io[t] = a.idx_O[t][ee];
    n_x = e < ce ? a.st.val[ee] : 0.f;
  };
  // lane (q, g) takes, for block mb, the split words of nonzero 16 mb + q;
  // each lane gathers its own nonzero's outer rows
  u4 g32[4];
  float ro[NS][RS];
  auto gather = [&]() {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const uint32_t r = (uint32_t)__shfl((int)n_iL, mb * 16 + q);
      g32[mb] = *reinterpret_cast<const u4 *>(a.As_L + ((size_t)r * 4 + g) * 4);
    }
#pragma unroll
    for (int t = 0; t < NS; ++t) load_row<R>(a.A_O[t], n_io[t], ro[t]);
  };
  load_idx(cb);
  uint32_t c_row = n_row;
  float c_x = n_x;
  gather();
  load_idx(cb + kWave);

  for (uint64_t tb = cb; tb < ce; tb += kWave) {
    const bool valid = tb + lane < ce;
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    const uint32_t row = c_row;
    const float x = c_x;
    c_row = n_row;
    c_x = n_x;
    s_rowv[lane] = valid ? row : 0xFFFFFFFFu;
    // this tile's outer rows to LDS, before the next gather reuses them
#pragma unroll
    for (int t = 0; t < NS; ++t)
#pragma unroll
      for (int j = 0; j < R; ++j) lrows[(t * R + j) * kWave + lane] = ro[t][j];
    wave_sync();
    float dacc[4][4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) dacc[mb][i] = 0.f;
    // MFMA p covers outer indices 2p (output rows 0..7) and 2p + 1 (rows
    // 8..15): lane (q, g) gets U[16mb + q][o][r] with o = 2p + g / 2 and
    // r = 4 (g % 2) + i, so it scales by s_m[o] of its own o.  The digits
    // of both outer indices of a step are wave-uniform (scalar work); a
    // lane only selects its half's.
    const bool odd = g >= 2;
    auto pair_step = [&](int p, const s8 &bw) {
      const b8 b = __builtin_bit_cast(b8, bw);
      f4 u[4];
      const f4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
        u[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, __builtin_bit_cast(b8, g32[mb]), z, 0, 0,
                                                        0);
      // past O the core words are zero: any in-range outer index will do
      const int oe = min(2 * p, O - 1), oo = min(2 * p + 1, O - 1);
      const float *l[NS];
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        // both offsets are computed in scalar registers and hidden from the
        // optimiser: otherwise it folds the select into the digits of a
        // per-lane 2p + half, with vector divisions by R at every step
        int offe = (t * R + oe / cpow(R, NS - 1 - t) % R) * kWave;
        int offo = (t * R + oo / cpow(R, NS - 1 - t) % R) * kWave;
        asm volatile("" : "+s"(offe), "+s"(offo));
        l[t] = lrows + q + (odd ? offo : offe);
      }
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        float sv = l[0][mb * 16];
#pragma unroll
        for (int t = 1; t < NS; ++t) sv *= l[t][mb * 16];
#pragma unroll
        for (int i = 0; i < 4; ++i) dacc[mb][i] = __builtin_fmaf(u[mb][i], sv, dacc[mb][i]);
      }
    };
    {
      constexpr int D = kPairRing;
      static_assert(NB % D == 0, "the pair table has a multiple of kPairRing pairs");
      s8 ring[D];
#pragma unroll
      for (int k = 0; k < D; ++k) ring[k] = tg[k * kWave + lane];
#pragma unroll 1
      for (int p = 0; p < NB; p += D) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const s8 bw = ring[k];
          if (p + k + D < NB) ring[k] = tg[(p + k + D) * kWave + lane];
          pair_step(p + k, bw);
        }
      }
    }
    // even and odd outer indices sit in lanes 32 apart: add the halves
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) dacc[mb][i] = add_xor32(dacc[mb][i]);
    // the next tile's gathers and the stream loads of the tile after it
    gather();
    load_idx(tb + 2 * kWave);
    // lane (q, g) holds delta[16mb + q][4g + i] for g < 2 (columns 8 and up
    // are zero): columns to LDS, then x
    if (g < T) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) tile[(4 * g + i) * kDP + mb * 16 + q] = g < 2 ? dacc[mb][i] : 0.f;
    }
    tile[R * kDP + lane] = x;
    wave_sync();
    dump_tile<R>(a, tile, tb, cnt, lane);
    gram.add_tile(tile, s_rowv, tb, cnt);
    wave_sync();
  }
  gram.finish();
}

// ------------------------------------------------------------ BF16, o-blocks
// The refined BF16 delta of an order-4 or order-5 tensor with at most
// kObMaxO outer indices (R <= 10), one MFMA per 16 nonzeros, 16 outer
// indices and output column r: A = block (ob, r) of the o-block table (row q
// holds the refined words of G[16 ob + q][.][r]), B = the split A_L words of
// 16 nonzeros, so lane (q, g) gets U[16 mb + q][16 ob + 4g + i][r], i < 4.
// Stage 2 scales those by s_m[o]: each lane forms s for its own nonzero and
// the block's 16 outer indices (o is wave-uniform and the loop over o is
// unrolled in full, so the outer rows are register reads), stores them to a
// per-wave LDS block [nonzero][o] and reads back the four of its MFMA
// layout with one 16-byte load per block of 16 nonzeros.  Each lane group
// accumulates over its own outer indices; a butterfly over the four groups
// adds them at the end of the tile.  Beside k_gram_bf16_o5 (one MFMA per two
// outer indices, output rows (o, r) padded to 8 columns) this runs R/8 of the
// MFMAs and scaling FMAs (the same at R = 8) and about half the s work.  The
// core operand comes from a copy of the table in LDS (TL, 40 KiB at R = 5,
// order 5) or from L2.
// pitch of the per-wave s block in floats: 16 outer indices and 4 of padding,
// so the 16-byte reads of a block of 16 nonzeros spread over the banks
constexpr int kObSP = 20;
template <int R, int NS> constexpr size_t ob_lds_floats_per_wave() {
  // the s block [64][kObSP] and the delta tile [16][kDP] share a buffer;
  // then the tile's row ids
  return (size_t)(kWave * kObSP > 16 * kDP ? kWave * kObSP : 16 * kDP) + kWave;
}
template <int R, int NS> constexpr size_t ob_table_bytes() {
  return (size_t)gb_oblocks(cpow(R, NS)) * R * kWave * 16;
}

// s of a lane's nonzero for outer indices 16 OB .. 16 OB + 15 (zero past O,
// where the table's rows are zero too), from its outer rows in registers
template <int R, int NS, int OB>
__device__ inline __attribute__((always_inline)) void ob_svalues(
    const float (&ro)[NS][row_stride(R)], f4 (&sv)[4]) {
  constexpr int O = cpow(R, NS);
#pragma unroll
  for (int k4 = 0; k4 < 4; ++k4)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 16 * OB + 4 * k4 + i;
      float sp = 0.f;
      if (o < O) {
        sp = ro[0][o / cpow(R, NS - 1) % R];
#pragma unroll
        for (int t = 1; t < NS; ++t) sp *= ro[t][o / cpow(R, NS - 1 - t) % R];
      }
      sv[k4][i] = sp;
    }
}
// the same for a wave-uniform block index: one branch per block, each with
// its digits fixed at compile time (an unrolled loop over the blocks would
// let the scheduler hoist every block's work and spill)
template <int R, int NS, int... OB>
__device__ inline __attribute__((always_inline)) void ob_svalues_at(
    int ob, const float (&ro)[NS][row_stride(R)], f4 (&sv)[4],
    std::integer_sequence<int, OB...>) {
  ((ob == OB ? ob_svalues<R, NS, OB>(ro, sv) : void()), ...);
}

template <int R, int NS, bool TL>
__global__ void __launch_bounds__(64 * kTableWaves) __attribute__((amdgpu_waves_per_eu(3)))
k_gram_bf16_ob(GramArgs a, int wpb) {
  constexpr int O = cpow(R, NS);
  constexpr int NOB = gb_oblocks(O);
  static_assert(bf16r_single_mfma(R) && (NS == 2 || NS == 3) && O <= kObMaxO,
                "orders 4-5, one refined MFMA per block, at most kObMaxO outer indices");
  constexpr int RS = row_stride(R);
  constexpr int T = ChunkGram<R>::T;
  constexpr size_t kBuf = ob_lds_floats_per_wave<R, NS>() - kWave;
  extern __shared__ __align__(16) float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  const s8 *tg = reinterpret_cast<const s8 *>(static_cast<const uint16_t *>(a.Gb) +
                                              gb_ob_offset(R, O));
  constexpr size_t kTabFloats = TL ? ob_table_bytes<R, NS>() / sizeof(float) : 0;
  if constexpr (TL) { // the o-block table to LDS, once per workgroup
    uint4 *dst = reinterpret_cast<uint4 *>(dyn);
    const uint4 *src = reinterpret_cast<const uint4 *>(tg);
    for (int i = threadIdx.x; i < NOB * R * kWave; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
  }
  const s8 *tab = TL ? reinterpret_cast<const s8 *>(dyn) : tg;
  const uint64_t c = a.chunk_begin + (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.chunk_end) return;
  float *buf = dyn + kTabFloats + (size_t)w * ob_lds_floats_per_wave<R, NS>();
  float *tile = buf; // after the stage loop; the s block before it
  uint32_t *s_rowv = reinterpret_cast<uint32_t *>(buf + kBuf);
  const uint64_t cb = c * a.chunk;
  const uint64_t ce = (cb + a.chunk < a.st.nnz) ? cb + a.chunk : a.st.nnz;
  const int q = lane & 15, g = lane >> 4;
  ChunkGram<R> gram(a, c, cb, ce, lane);

  uint32_t n_row, n_iL, n_io[NS];
  float n_x;
  auto load_idx = [&](uint64_t tb) {
    const uint64_t e = tb + lane;
    const uint64_t ee = e < ce ? e : cb;
    n_row = a.idx_row[ee];
    n_iL = a.idx_L[ee];
#pragma unroll
    for (int t = 0; t < NS; ++t) n_io[t] = a.idx_O[t][ee];
    n_x = e < ce ? a.st.val[ee] : 0.f;
  };
  // lane (q, g) takes, for block mb, the split words of nonzero 16 mb + q;
  // each lane gathers its own nonzero's outer rows
  u4 g32[4];
  float ro[NS][RS];
  auto gather = [&]() {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const uint32_t r = (uint32_t)__shfl((int)n_iL, mb * 16 + q);
      g32[mb] = *reinterpret_cast<const u4 *>(a.As_L + ((size_t)r * 4 + g) * 4);
    }
#pragma unroll
    for (int t = 0; t < NS; ++t) load_row<R>(a.A_O[t], n_io[t], ro[t]);
  };
  load_idx(cb);
  uint32_t c_row = n_row;
  float c_x = n_x;
  gather();
  load_idx(cb + kWave);

  for (uint64_t tb = cb; tb < ce; tb += kWave) {
    const bool valid = tb + lane < ce;
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    const uint32_t row = c_row;
    const float x = c_x;
    c_row = n_row;
    c_x = n_x;
    s_rowv[lane] = valid ? row : 0xFFFFFFFFu;
    float dacc[4][R];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int r = 0; r < R; ++r) dacc[mb][r] = 0.f;
#pragma unroll 1
    for (int ob = 0; ob < NOB; ++ob) {
      // the outer rows as values of this iteration: otherwise every block's
      // products are hoisted out of the loop, O registers that spill
#pragma unroll
      for (int t = 0; t < NS; ++t)
#pragma unroll
        for (int j = 0; j < R; ++j) asm volatile("" : "+v"(ro[t][j]));
      f4 sv[4];
      ob_svalues_at<R, NS>(ob, ro, sv, std::make_integer_sequence<int, NOB>{});
      wave_sync(); // the previous block's reads of the s block are done
#pragma unroll
      for (int k4 = 0; k4 < 4; ++k4)
        *reinterpret_cast<f4 *>(buf + lane * kObSP + 4 * k4) = sv[k4];
      wave_sync();
      // lane (q, g) scales outer indices 16 ob + 4g + i of nonzero 16 mb + q
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
        sv[mb] = *reinterpret_cast<const f4 *>(buf + (mb * 16 + q) * kObSP + 4 * g);
      // column r's four MFMAs issue before column r - 1's scaling, so the
      // scaling never waits on a result just issued
      const s8 *tb_ob = tab + (size_t)ob * R * kWave + lane;
      f4 up[4];
#pragma unroll
      for (int r = 0; r <= R; ++r) {
        f4 un[4];
        if (r < R) {
          const b8 b = __builtin_bit_cast(b8, tb_ob[r * kWave]);
          const f4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int mb = 0; mb < 4; ++mb)
            un[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, __builtin_bit_cast(b8, g32[mb]), z,
                                                             0, 0, 0);
        }
        if (r > 0) {
#pragma unroll
          for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int i = 0; i < 4; ++i)
              dacc[mb][r - 1] = __builtin_fmaf(up[mb][i], sv[mb][i], dacc[mb][r - 1]);
        }
        if (r < R) {
#pragma unroll
          for (int mb = 0; mb < 4; ++mb) up[mb] = un[mb];
        }
      }
    }
    // the next tile's gathers and the stream loads of the tile after it
    gather();
    load_idx(tb + 2 * kWave);
    // add the four lane groups' partial sums (disjoint outer indices)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        dacc[mb][r] = add_xor16(dacc[mb][r]);
        dacc[mb][r] = add_xor32(dacc[mb][r]);
      }
    wave_sync(); // the s block's reads are done before the tile overwrites it
    // lane (q, g) writes the deltas of nonzero 16 g + q = lane, then x over
    // column R and zeros up to the Gram's last column tile
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const float v = g == 0 ? dacc[0][r] : g == 1 ? dacc[1][r] : g == 2 ? dacc[2][r] : dacc[3][r];
      tile[r * kDP + lane] = v;
    }
    tile[R * kDP + lane] = x;
#pragma unroll
    for (int r = R + 1; r < 4 * T; ++r) tile[r * kDP + lane] = 0.f;
    wave_sync();
    dump_tile<R>(a, tile, tb, cnt, lane);
    gram.add_tile(tile, s_rowv, tb, cnt);
    wave_sync();
  }
  gram.finish();
}

// ------------------------------------------------------------ BF16, order 4, Kronecker
// The refined BF16 delta of an order-4 tensor at R = 4 or 8 with two modes
// contracted on the matrix cores (table layout in gt_internal.hpp).  With
// outer slots 0 and 1 (o0, o1) and L the last other mode,
//   stage 1 (MFMA): V[m][(o0, r)] = sum_{o1, jL} G[o0][o1][jL][r] k_m[o1 R + jL],
//                   k_m[o1 R + jL] = A_O1[i_O1(m)][o1] * A_L[i_L(m)][jL] (FP32),
//   stage 2 (VALU): delta[m][r] = sum_{o0} A_O0[i_O0(m)][o0] V[m][(o0, r)].
// k is split as k_hi + k_lo (BF16) per lane on the fly; one MFMA per 16
// nonzeros covers 16 output rows (o0, r) and 16 (HL) or 32 (HH) entries of k.
// Stage 2 is R FMAs per output, against O R = R^3 for the o-block kernel.
// Each lane gathers its own nonzero's three rows; they go to this wave's LDS
// (outer rows permuted so that the values a lane group needs are
// contiguous) and each lane reads the four nonzeros of its MFMA columns.
// The table sits in LDS for a workgroup of 12 waves.
template <int R> constexpr int kr_rc() { return 16 / R; } // o = RC * block + lane-group part
template <int R> constexpr int kr_sigma(int o) {          // LDS position of outer value o
  return (o % kr_rc<R>()) * kr_nhl(R) + o / kr_rc<R>();
}
template <int R> constexpr size_t kr_lds_floats_per_wave() {
  // the three rows of 64 nonzeros, or the delta tile; then the row ids
  return (size_t)((3 * R * kWave > 16 * kDP) ? 3 * R * kWave : 16 * kDP) + kWave;
}

template <int R>
__global__ void __launch_bounds__(64 * kTableWaves) __attribute__((amdgpu_waves_per_eu(3)))
k_gram_bf16_kr(GramArgs a, int wpb) {
  static_assert(R == 4 || R == 8, "the Kronecker kernel covers order 4 at R = 4 and 8");
  constexpr int NHL = kr_nhl(R), NHH = kr_nhh(R), NMF = NHL + NHH, NKB = kr_nkb(R, 4);
  constexpr int NP = (NHL + 1) / 2; // 8-word operand blocks per block of 16 nonzeros
  constexpr int T = ChunkGram<R>::T;
  constexpr int RW = 3 * R; // floats per nonzero in the row area: A_L, A_O1, A_O0
  constexpr size_t kBuf = kr_lds_floats_per_wave<R>() - kWave;
  constexpr size_t kTabFloats = gb_kr_words(R, 4) / 2;
  extern __shared__ __align__(16) float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(a.Gb) +
                                                       gb_kr_offset(R, 4, R * R));
    uint4 *dst = reinterpret_cast<uint4 *>(dyn);
    for (int i = threadIdx.x; i < NKB * NMF * kWave; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
  }
  const s8 *tab = reinterpret_cast<const s8 *>(dyn) + lane;
  const uint64_t c = a.chunk_begin + (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.chunk_end) return;
  float *buf = dyn + kTabFloats + (size_t)w * kr_lds_floats_per_wave<R>();
  float *tile = buf; // after stage 2; the row area before it
  uint32_t *s_rowv = reinterpret_cast<uint32_t *>(buf + kBuf);
  const uint64_t cb = c * a.chunk;
  const uint64_t ce = (cb + a.chunk < a.st.nnz) ? cb + a.chunk : a.st.nnz;
  const int q = lane & 15, g = lane >> 4;
  const int jl0 = (4 * g) % R;         // this lane group's first jL
  const int oc = (4 * g) / R;          // its part of o0 and o1
  ChunkGram<R> gram(a, c, cb, ce, lane);

  uint32_t n_row, n_iL, n_io0, n_io1;
  float n_x;
  auto load_idx = [&](uint64_t tb) {
    const uint64_t e = tb + lane;
    const uint64_t ee = e < ce ? e : cb;
    n_row = a.idx_row[ee];
    n_iL = a.idx_L[ee];
    n_io0 = a.idx_O[0][ee];
    n_io1 = a.idx_O[1][ee];
    n_x = e < ce ? a.st.val[ee] : 0.f;
  };
  float rL[R], rO0[R], rO1[R]; // this lane's nonzero's rows (the next tile's, in flight)
  auto gather = [&]() {
    load_row<R>(a.A_L, n_iL, rL);
    load_row<R>(a.A_O[0], n_io0, rO0);
    load_row<R>(a.A_O[1], n_io1, rO1);
  };
  load_idx(cb);
  uint32_t c_row = n_row;
  float c_x = n_x;
  gather();
  load_idx(cb + kWave);

  for (uint64_t tb = cb; tb < ce; tb += kWave) {
    const bool valid = tb + lane < ce;
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    const uint32_t row = c_row;
    const float x = c_x;
    c_row = n_row;
    c_x = n_x;
    s_rowv[lane] = valid ? row : 0xFFFFFFFFu;
    {
      float *my = buf + lane * RW;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        my[j] = rL[j];
        my[R + kr_sigma<R>(j)] = rO1[j];
        my[2 * R + kr_sigma<R>(j)] = rO0[j];
      }
    }
    wave_sync();
    // Two halves of 32 nonzeros (blocks mb = 2h, 2h + 1), so the operands of
    // one half fill the registers: per pair p of HL MFMAs the factor-side
    // words are [lo(k[2p]), hi(k[2p]), hi(k[2p + 1]), lo(k[2p + 1])]
    float dacc[4][4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      // no instruction of one half moves into the other: the scheduler would
      // hoist the second half's operands over the first half's MFMAs, and
      // both halves' registers would be live at once
      __builtin_amdgcn_sched_barrier(0);
      gt_u4 V[2][NP][2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const float *rm = buf + ((2 * h + b) * 16 + q) * RW;
        const gt_f4 aL = *reinterpret_cast<const gt_f4 *>(rm + jl0);
        float ao[NHL];
#pragma unroll
        for (int mu = 0; mu < NHL; ++mu) ao[mu] = rm[R + oc * NHL + mu];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          uint32_t hw[2][2], lw[2][2];
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            const int mu = 2 * p + e;
            if (mu < NHL) {
              float k[4];
#pragma unroll
              for (int i = 0; i < 4; ++i) k[i] = ao[mu] * aL[i];
#pragma unroll
              for (int d = 0; d < 2; ++d) {
                const uint32_t hh = pk_bf16(k[2 * d], k[2 * d + 1]);
                hw[e][d] = hh;
                lw[e][d] = pk_bf16(k[2 * d] - bf16_low_f(hh), k[2 * d + 1] - bf16_high_f(hh));
              }
            } else {
              hw[e][0] = hw[e][1] = lw[e][0] = lw[e][1] = 0u;
            }
          }
          V[b][p][0] = gt_u4{lw[0][0], lw[0][1], hw[0][0], hw[0][1]};
          V[b][p][1] = gt_u4{hw[1][0], hw[1][1], lw[1][0], lw[1][1]};
        }
      }
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) dacc[2 * h + b][i] = 0.f;
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb) {
        f4 D[2];
#pragma unroll
        for (int b = 0; b < 2; ++b) D[b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int mu = 0; mu < NMF; ++mu) {
          const b8 t = __builtin_bit_cast(b8, tab[(kb * NMF + mu) * kWave]);
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            b8 bop;
            if (mu < NHL) { // HL: block p = mu / 2, first or second half
              bop = __builtin_bit_cast(b8, V[b][mu / 2][mu & 1]);
            } else { // HH: the middle of block p = mu - NHL
              const int p = mu - NHL;
              bop = __builtin_bit_cast(
                  b8, (gt_u4{V[b][p][0][2], V[b][p][0][3], V[b][p][1][0], V[b][p][1][1]}));
            }
            D[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(t, bop, D[b], 0, 0, 0);
          }
        }
        // rows 16 kb + 4g + i are (o0, r) = (RC kb + oc, jl0 + i): scale by A_O0[o0]
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const float wv = buf[((2 * h + b) * 16 + q) * RW + 2 * R + oc * NKB + kb];
#pragma unroll
          for (int i = 0; i < 4; ++i)
            dacc[2 * h + b][i] = __builtin_fmaf(D[b][i], wv, dacc[2 * h + b][i]);
        }
      }
    }
    // the next tile's rows and the stream loads of the tile after it (issued
    // after the stage loop, whose operands fill the registers)
    gather();
    load_idx(tb + 2 * kWave);
    // lane groups with the same jl0 hold the same columns for other o0: add
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if constexpr (R == 4) dacc[mb][i] = add_xor16(dacc[mb][i]);
        dacc[mb][i] = add_xor32(dacc[mb][i]);
      }
    wave_sync(); // every read of the row area before the tile overwrites it
    // lane (q, g < R / 4) holds delta[16mb + q][4g + i]; then x over column R
    // and zeros up to the Gram's last column tile
    if (g < R / 4) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) tile[(4 * g + i) * kDP + mb * 16 + q] = dacc[mb][i];
    }
    tile[R * kDP + lane] = x;
#pragma unroll
    for (int r = R + 1; r < 4 * T; ++r) tile[r * kDP + lane] = 0.f;
    wave_sync();
    dump_tile<R>(a, tile, tb, cnt, lane);
    gram.add_tile(tile, s_rowv, tb, cnt);
    wave_sync();
  }
  gram.finish();
}

// ------------------------------------------------------------ BF16, order 5, Kronecker
// The order-5 form of the two-mode refined BF16 delta, at R = 5 (layout in
// gt_internal.hpp).  With outer slots o0, o1, o2 and L the last other mode:
//   stage 1 (MFMA): V[m][(o01, r)] = sum_{o2, jL} G[o01][o2][jL][r] k_m[o2 R + jL],
//                   k_m = A_O2[i_O2(m)] (x) A_L[i_L(m)], split per nonzero,
//   stage 2 (VALU): delta[m][r] = sum_{o01} s_m[o01] V[m][(o01, r)],
//                   s_m[o01] = A_O0[i_O0(m)][o0] * A_O1[i_O1(m)][o1],
// R FMAs per output where the o-block kernel takes R^3.  The 25 entries of k
// are dense in the two HL MFMAs (a lane's (o2, jL) pairs come from a small
// per-lane offset table); output rows pad r to 8 so a lane's rows are
// (o01, r) = (2 kb + g / 2, 4 (g % 2) + i).  Each lane stages its own
// nonzero's A_L and A_O2 rows and its 25 s values (permuted: even o01, then
// odd) in this wave's LDS.
template <int R> constexpr int kr5_rw() { // floats per nonzero in the row area
  return 2 * R + 2 * kr_nkb(R, 5);
}
template <int R> constexpr size_t kr5_lds_floats_per_wave() {
  return (size_t)((kr5_rw<R>() * kWave > 16 * kDP) ? kr5_rw<R>() * kWave : 16 * kDP) + kWave;
}

template <int R>
__global__ void __launch_bounds__(64 * kTableWaves) __attribute__((amdgpu_waves_per_eu(3)))
k_gram_bf16_kr5(GramArgs a, int wpb) {
  static_assert(R == 5, "the order-5 Kronecker kernel covers R = 5");
  constexpr int NHL = kr_nhl(R), NHH = kr_nhh(R), NMF = NHL + NHH, NKB = kr_nkb(R, 5);
  static_assert(NHL == 2 && NHH == 1, "one operand block of two HL and one HH MFMA");
  constexpr int T = ChunkGram<R>::T;
  constexpr int RW = kr5_rw<R>();   // A_L, A_O2, s (even o01, then odd)
  constexpr int SO = 2 * R;         // offset of the s values
  constexpr size_t kBuf = kr5_lds_floats_per_wave<R>() - kWave;
  constexpr size_t kTabFloats = gb_kr_words(R, 5) / 2;
  extern __shared__ __align__(16) float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(a.Gb) +
                                                       gb_kr_offset(R, 5, R * R * R));
    uint4 *dst = reinterpret_cast<uint4 *>(dyn);
    for (int i = threadIdx.x; i < NKB * NMF * kWave; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
  }
  const s8 *tab = reinterpret_cast<const s8 *>(dyn) + lane;
  const uint64_t c = a.chunk_begin + (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.chunk_end) return;
  float *buf = dyn + kTabFloats + (size_t)w * kr5_lds_floats_per_wave<R>();
  float *tile = buf; // after stage 2; the row area before it
  uint32_t *s_rowv = reinterpret_cast<uint32_t *>(buf + kBuf);
  const uint64_t cb = c * a.chunk;
  const uint64_t ce = (cb + a.chunk < a.st.nnz) ? cb + a.chunk : a.st.nnz;
  const int q = lane & 15, g = lane >> 4;
  // this lane group's entries e = 16 mu + 4g + i: LDS offsets of A_L[jL] and
  // A_O2[o2] in a nonzero's row area (an entry past R^2 reads entry 0; its
  // core words are zero)
  uint32_t eoff[NHL][4];
#pragma unroll
  for (int mu = 0; mu < NHL; ++mu)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e0 = 16 * mu + 4 * g + i, e = e0 < R * R ? e0 : 0;
      eoff[mu][i] = (uint32_t)(e % R) | ((uint32_t)(R + e / R) << 16);
    }
  const int sp = (g >> 1) * NKB; // this lane's s values: o01 = 2 kb + g / 2
  ChunkGram<R> gram(a, c, cb, ce, lane);

  uint32_t n_row, n_iL, n_io[3];
  float n_x;
  auto load_idx = [&](uint64_t tb) {
    const uint64_t e = tb + lane;
    const uint64_t ee = e < ce ? e : cb;
    n_row = a.idx_row[ee];
    n_iL = a.idx_L[ee];
#pragma unroll
    for (int t = 0; t < 3; ++t) n_io[t] = a.idx_O[t][ee];
    n_x = e < ce ? a.st.val[ee] : 0.f;
  };
  float rL[row_stride(R)], ro[3][row_stride(R)]; // this lane's nonzero's rows
  auto gather = [&]() {
    load_row<R>(a.A_L, n_iL, rL);
#pragma unroll
    for (int t = 0; t < 3; ++t) load_row<R>(a.A_O[t], n_io[t], ro[t]);
  };
  load_idx(cb);
  uint32_t c_row = n_row;
  float c_x = n_x;
  gather();
  load_idx(cb + kWave);

  for (uint64_t tb = cb; tb < ce; tb += kWave) {
    const bool valid = tb + lane < ce;
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    const uint32_t row = c_row;
    const float x = c_x;
    c_row = n_row;
    c_x = n_x;
    s_rowv[lane] = valid ? row : 0xFFFFFFFFu;
    {
      float *my = buf + lane * RW;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        my[j] = rL[j];
        my[R + j] = ro[2][j];
      }
#pragma unroll
      for (int o = 0; o < R * R; ++o) my[SO + (o % 2) * NKB + o / 2] = ro[0][o / R] * ro[1][o % R];
      if constexpr (R * R % 2 == 1) my[SO + NKB + NKB - 1] = 0.f; // the odd half's pad
    }
    wave_sync();
    // [lo(k[entries of mu 0]), hi(...), hi(k[mu 1]), lo(...)] per block of 16
    gt_u4 V[4][2];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const float *rm = buf + (mb * 16 + q) * RW;
      uint32_t hw[2][2], lw[2][2];
#pragma unroll
      for (int mu = 0; mu < NHL; ++mu) {
        float k[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) k[i] = rm[eoff[mu][i] >> 16] * rm[eoff[mu][i] & 0xFFFFu];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          const uint32_t hh = pk_bf16(k[2 * d], k[2 * d + 1]);
          hw[mu][d] = hh;
          lw[mu][d] = pk_bf16(k[2 * d] - bf16_low_f(hh), k[2 * d + 1] - bf16_high_f(hh));
        }
      }
      V[mb][0] = gt_u4{lw[0][0], lw[0][1], hw[0][0], hw[0][1]};
      V[mb][1] = gt_u4{hw[1][0], hw[1][1], lw[1][0], lw[1][1]};
    }
    float dacc[4][4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) dacc[mb][i] = 0.f;
#pragma unroll 1
    for (int kb = 0; kb < NKB; ++kb) {
      f4 D[4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) D[mb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int mu = 0; mu < NMF; ++mu) {
        const b8 t = __builtin_bit_cast(b8, tab[(kb * NMF + mu) * kWave]);
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
          const gt_u4 v = mu < NHL ? V[mb][mu]
                                   : gt_u4{V[mb][0][2], V[mb][0][3], V[mb][1][0], V[mb][1][1]};
          D[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(t, __builtin_bit_cast(b8, v), D[mb], 0,
                                                          0, 0);
        }
      }
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const float wv = buf[(mb * 16 + q) * RW + SO + sp + kb];
#pragma unroll
        for (int i = 0; i < 4; ++i) dacc[mb][i] = __builtin_fmaf(D[mb][i], wv, dacc[mb][i]);
      }
    }
    // the next tile's rows and the stream loads of the tile after it
    gather();
    load_idx(tb + 2 * kWave);
    // lane groups g and g + 2 hold the same columns for the other parity of o01
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) dacc[mb][i] = add_xor32(dacc[mb][i]);
    wave_sync(); // every read of the row area before the tile overwrites it
    // lane (q, 0) holds delta[16mb + q][0..3], lane (q, 1) delta[16mb + q][4];
    // then x over column R and zeros up to the Gram's last column tile
    if (g < 2) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (4 * g + i < R) tile[(4 * g + i) * kDP + mb * 16 + q] = dacc[mb][i];
    }
    tile[R * kDP + lane] = x;
#pragma unroll
    for (int r = R + 1; r < 4 * T; ++r) tile[r * kDP + lane] = 0.f;
    wave_sync();
    dump_tile<R>(a, tile, tb, cnt, lane);
    gram.add_tile(tile, s_rowv, tb, cnt);
    wave_sync();
  }
  gram.finish();
}

// ------------------------------------------------------------ split factor
// Row i of the split copy holds, for each lane group g, the 16 bytes of BF16
// words that group's K=32 slots take (bf16r_plan), so the gather of one
// nonzero's A_L row is one 16-byte load per lane.  The first 8 bytes of a
// kPlanHL group are the hi words of its four columns, which the K=16 forms
// use; groups whose columns are padding get zeros or words the core table
// multiplies by zero.
template <int R>
__global__ void __launch_bounds__(256) k_split_factor(const float *__restrict__ A,
                                                      uint32_t *__restrict__ As, uint64_t rows) {
  constexpr int RS = row_stride(R);
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t i = t >> 2;
  const int g = (int)(t & 3);
  if (i >= rows) return;
  const float *row = A + i * RS;
  const u4 v = split_words<R>([&](int j) { return j < R ? row[j] : 0.f; }, g);
  *reinterpret_cast<u4 *>(As + t * 4) = v;
}

template <int R> void launch_valu(const GramArgs &a, hipStream_t st) {
  const size_t per_wave = lds_floats_per_wave<kValu>(a.slots, R, 1) * sizeof(float);
  int wpb = 4;
  while (wpb > 1 && per_wave * wpb + kStaticLdsValu > lds_limit()) --wpb;
  if (per_wave * wpb + kStaticLdsValu > lds_limit())
    throw std::runtime_error("FP32 Gram: outer factor rows do not fit in LDS");
  if (a.chunk_end <= a.chunk_begin) return;
  const unsigned grid = (unsigned)((a.chunk_end - a.chunk_begin + wpb - 1) / wpb);
  hipLaunchKernelGGL((k_gram_valu<R>), dim3(grid), dim3(64 * wpb), per_wave * wpb, st, a, wpb);
}

// Launches k_gram_bf16_o5 over the chunks of `a` (if any); false where it is
// not built for R (below R = 6 the o-block kernel takes order 5).
template <int R> bool launch_o5(const GramArgs &a, hipStream_t st, bool any) {
  if constexpr (!(bf16r_pair_table(R) && cpow(R, 3) > kObMaxO)) {
    return false;
  } else {
    const int wpb = 4;
    const size_t lds = (size_t)wpb * o5_lds_floats_per_wave<R>() * sizeof(float);
    if (lds > lds_limit()) return false;
    if (any) {
      const unsigned grid = (unsigned)((a.chunk_end - a.chunk_begin + wpb - 1) / wpb);
      hipLaunchKernelGGL((k_gram_bf16_o5<R>), dim3(grid), dim3(64 * wpb), lds, st, a, wpb);
    }
    return true;
  }
}

// Launches k_gram_bf16_ob over the chunks of `a` (if any) and returns the
// gt_delta_path it takes: the table in LDS beside twelve waves a workgroup
// unless a.alt_source or it does not fit, else from L2; -1 where the kernel
// is not built for R and NS.
template <int R, int NS> int launch_ob(const GramArgs &a, hipStream_t st, bool any) {
  constexpr int O = cpow(R, NS);
  if constexpr (!(bf16r_single_mfma(R) && R >= 2 && O <= kObMaxO)) {
    return -1;
  } else {
    const size_t per_wave = ob_lds_floats_per_wave<R, NS>() * sizeof(float);
    const bool tl =
        !a.alt_source && ob_table_bytes<R, NS>() + kTableWaves * per_wave <= lds_limit();
    const int wpb = tl ? kTableWaves : 4;
    const size_t lds = (tl ? ob_table_bytes<R, NS>() : 0) + (size_t)wpb * per_wave;
    if (lds > lds_limit()) return -1;
    if (any) {
      const unsigned grid = (unsigned)((a.chunk_end - a.chunk_begin + wpb - 1) / wpb);
      if (tl)
        hipLaunchKernelGGL((k_gram_bf16_ob<R, NS, true>), dim3(grid), dim3(64 * wpb), lds, st, a,
                           wpb);
      else
        hipLaunchKernelGGL((k_gram_bf16_ob<R, NS, false>), dim3(grid), dim3(64 * wpb), lds, st, a,
                           wpb);
    }
    return tl ? GT_DELTA_BF16R_OB : GT_DELTA_BF16R_OB_L2;
  }
}

// Launches k_gram_bf16_kr over the chunks of `a` (if any); -1 where it is
// not built for R or the table and twelve waves do not fit in LDS.
template <int R> int launch_kr(const GramArgs &a, hipStream_t st, bool any) {
  if constexpr (!bf16r_kr_table(R, 4)) {
    return -1;
  } else {
    const size_t lds = gb_kr_words(R, 4) * sizeof(uint16_t) +
                       (size_t)kTableWaves * kr_lds_floats_per_wave<R>() * sizeof(float);
    if (lds > lds_limit()) return -1;
    if (any) {
      const unsigned grid =
          (unsigned)((a.chunk_end - a.chunk_begin + kTableWaves - 1) / kTableWaves);
      hipLaunchKernelGGL((k_gram_bf16_kr<R>), dim3(grid), dim3(64 * kTableWaves), lds, st, a,
                         kTableWaves);
    }
    return GT_DELTA_BF16R_KR;
  }
}

// Launches k_gram_bf16_kr5 over the chunks of `a` (if any); -1 where it is
// not built for R or the table and twelve waves do not fit in LDS.
template <int R> int launch_kr5(const GramArgs &a, hipStream_t st, bool any) {
  if constexpr (!bf16r_kr_table(R, 5)) {
    return -1;
  } else {
    const size_t lds = gb_kr_words(R, 5) * sizeof(uint16_t) +
                       (size_t)kTableWaves * kr5_lds_floats_per_wave<R>() * sizeof(float);
    if (lds > lds_limit()) return -1;
    if (any) {
      const unsigned grid =
          (unsigned)((a.chunk_end - a.chunk_begin + kTableWaves - 1) / kTableWaves);
      hipLaunchKernelGGL((k_gram_bf16_kr5<R>), dim3(grid), dim3(64 * kTableWaves), lds, st, a,
                         kTableWaves);
    }
    return GT_DELTA_BF16R_KR;
  }
}

// Picks the delta + Gram kernel for rank R and the requested delta kind, and
// launches it over the chunks of `a` (if any); returns the gt_delta_path
// that ran.  A BF16 request falls back to the FP32 delta only where no MFMA
// variant is built (plain BF16 whose table leaves too few waves per CU,
// refined BF16 at orders of 6 and up or order 5 above R = 8); the returned
// path says so.
template <int R, int KIND> int launch_impl(const GramArgs &a, int O, hipStream_t st) {
  const bool any = a.chunk_end > a.chunk_begin;
  if constexpr (KIND == kValu) {
    launch_valu<R>(a, st);
    return GT_DELTA_FP32;
  } else {
    if constexpr (KIND == kBf16R) {
      // orders 4-5: o-blocks up to kObMaxO outer indices, then (order 5,
      // R <= 8) pairs of outer indices
      if (a.As_L != nullptr && (a.slots == 2 || a.slots == 3) && O == cpow(R, a.slots)) {
        // order 4 at R = 4, 8 and order 5 at R = 5: two modes on the matrix
        // cores; the other source (a.alt_source) is the o-block kernel with
        // its table in LDS
        if (bf16r_kr_table(R, a.slots + 2)) {
          if (!a.alt_source) {
            const int p = a.slots == 2 ? launch_kr<R>(a, st, any) : launch_kr5<R>(a, st, any);
            if (p >= 0) return p;
          } else {
            GramArgs b = a;
            b.alt_source = 0;
            const int p = a.slots == 2 ? launch_ob<R, 2>(b, st, any) : launch_ob<R, 3>(b, st, any);
            if (p >= 0) return p;
          }
        }
        const int p = a.slots == 2 ? launch_ob<R, 2>(a, st, any) : launch_ob<R, 3>(a, st, any);
        if (p >= 0) return p;
        if (a.slots == 3 && launch_o5<R>(a, st, any)) return GT_DELTA_BF16R_L2;
      }
    }
    const size_t per_wave = lds_floats_per_wave<KIND>(a.slots, R, O) * sizeof(float);
    const size_t table =
        KIND == kBf16 ? gb_k16_words(O) * 2
                      : (gb_k32_words(O) * (bf16r_single_mfma(R) ? 1 : 2)) * 2;
    // waves per CU that the LDS holds with the core table in it, 4 per workgroup
    const size_t wg4 = per_wave * 4 + table + kStaticLdsBf16;
    const bool table_fits =
        wg4 <= lds_limit() && (int)(lds_limit() / wg4) * 4 >= kMinLdsWavesPerCu;
    if (!table_fits || a.As_L == nullptr) { // no MFMA variant for this shape
      launch_valu<R>(a, st);
      return GT_DELTA_FP32;
    }
    const int wpb = 4;
    const unsigned grid = (unsigned)((a.chunk_end - a.chunk_begin + wpb - 1) / wpb);
    if (any)
      hipLaunchKernelGGL((k_gram_bf16<R, KIND>), dim3(grid), dim3(64 * wpb),
                         per_wave * wpb + table, st, a, wpb, O);
    return KIND == kBf16 ? GT_DELTA_BF16 : GT_DELTA_BF16R;
  }
}

template <int KIND> int dispatch(int R, const GramArgs &a, int O, hipStream_t st) {
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    return launch_impl<r, KIND>(a, O, st);
    GT_CASE(1) GT_CASE(2) GT_CASE(3) GT_CASE(4) GT_CASE(5) GT_CASE(6) GT_CASE(7) GT_CASE(8)
    GT_CASE(9) GT_CASE(10) GT_CASE(11) GT_CASE(12) GT_CASE(13) GT_CASE(14) GT_CASE(15)
#undef GT_CASE
  default:
    throw std::invalid_argument("MFMA Gram needs R <= 15");
  }
}

} // namespace

bool gram_mfma_supported(int R) { return R >= 1 && R <= 15; }

int launch_gram_mfma(int R, int precision, const GramArgs &a, int O, hipStream_t s) {
  int path;
  if (precision == GT_PREC_FP32)
    path = dispatch<kValu>(R, a, O, s);
  else if (precision == GT_PREC_BF16)
    path = dispatch<kBf16>(R, a, O, s);
  else
    path = dispatch<kBf16R>(R, a, O, s);
  GT_HIP(hipGetLastError());
  return path;
}

void launch_split_factor(int R, const float *A, uint64_t rows, uint32_t *As, hipStream_t s) {
  if (rows == 0) return;
  const unsigned grid = (unsigned)((rows * 4 + 255) / 256);
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    hipLaunchKernelGGL(k_split_factor<r>, dim3(grid), dim3(256), 0, s, A, As, rows);            \
    break;
    GT_CASE(1) GT_CASE(2) GT_CASE(3) GT_CASE(4) GT_CASE(5) GT_CASE(6) GT_CASE(7) GT_CASE(8)
    GT_CASE(9) GT_CASE(10) GT_CASE(11) GT_CASE(12) GT_CASE(13) GT_CASE(14) GT_CASE(15)
#undef GT_CASE
  default:
    throw std::invalid_argument("split factor needs R <= 15");
  }
  GT_HIP(hipGetLastError());
}

} // namespace gt
