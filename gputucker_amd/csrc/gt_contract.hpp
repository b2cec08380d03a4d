// gt_contract.hpp -- device helpers shared by the kernels: the VALU
// contraction that forms one nonzero's delta, 16-byte row loads, and a
// wavefront-scope barrier.
#pragma once

#include "gt_internal.hpp"

namespace gt {

constexpr int kSlots = kMaxOrder - 2;

// x[l] + x[l ^ 32] and x[l] + x[l ^ 16] through the gfx950 lane swaps
// (v_permlane32_swap / v_permlane16_swap: VALU, no LDS round trip).  Each
// returns the same sum in both lanes of a pair (IEEE addition commutes).
__device__ inline __attribute__((always_inline)) float add_xor32(float x) {
  const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false,
                                                  false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
__device__ inline __attribute__((always_inline)) float add_xor16(float x) {
  const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false,
                                                  false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}

__device__ inline __attribute__((always_inline)) void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The wavefront's index in its workgroup, as a wave-uniform value: without
// readfirstlane the compiler treats threadIdx.x / 64 as per-lane, and every
// chunk bound, tile position and run boundary derived from it would live in
// vector registers, with 64-bit VALU arithmetic and EXEC-masked branches.
__device__ inline __attribute__((always_inline)) int wave_index() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
}

typedef float gt_f4 __attribute__((ext_vector_type(4)));

// row i of a factor (stride row_stride(R), zero padded) into registers:
// 16-byte loads for whole groups of four columns, then 8- and 4-byte loads
// for the rest, so no load writes a register the caller never reads (such a
// register may be reused while the load is in flight, forcing a wait for
// every outstanding load); the padding entries are set to zero.
typedef float gt_f2 __attribute__((ext_vector_type(2)));
template <int R>
__device__ inline __attribute__((always_inline)) void load_row(const float *__restrict__ base,
                                                               uint32_t i,
                                                               float (&out)[row_stride(R)]) {
  const float *row = base + (size_t)i * row_stride(R);
  const gt_f4 *p = reinterpret_cast<const gt_f4 *>(row);
#pragma unroll
  for (int v = 0; v < R / 4; ++v) {
    const gt_f4 t = p[v];
    out[4 * v + 0] = t[0];
    out[4 * v + 1] = t[1];
    out[4 * v + 2] = t[2];
    out[4 * v + 3] = t[3];
  }
  constexpr int b = R / 4 * 4;
  if constexpr (R % 4 >= 2) {
    const gt_f2 t = *reinterpret_cast<const gt_f2 *>(row + b);
    out[b] = t[0];
    out[b + 1] = t[1];
  }
  if constexpr (R % 4 == 1) out[b] = row[b];
  if constexpr (R % 4 == 3) out[b + 2] = row[b + 2];
#pragma unroll
  for (int j = R; j < row_stride(R); ++j) out[j] = 0.f;
}

// The core as contract_slots and contract_pair read it: through the constant
// address space (scalar loads), from a pointer offset by a zero the compiler
// cannot see through.  Taken inside a caller's loop, this keeps the core's
// loads in the loop; without it the compiler may hoist them out as
// loop-invariant, into up to 256 vector registers (one wave per SIMD).
typedef __attribute__((address_space(4))) const float gt_cfloat;
__device__ inline __attribute__((always_inline)) gt_cfloat *core_operand(const float *Gt) {
  uint32_t off = 0;
  asm volatile("" : "+s"(off));
  return (gt_cfloat *)(Gt + off);
}

// LDS floats per wave that contract_slots needs: the outer rows when there
// are two or more outer slots (or any, above R = 16), plus A_L above R = 16.
template <int R> __host__ __device__ constexpr int contract_lds_rows(int slots) {
  return (R <= 16) ? (slots >= 2 ? slots : 0) : slots + 1;
}

// delta of one mode for this lane's nonzero, with the operands in slot order:
// A_L / iL is the last mode other than n, A_O[t] / io[t] the remaining other
// modes ascending (slots = N - 2 of them).  For each outer multi-index o
// (mixed radix over the outer slots, last fastest):
//   s = prod_t A_t[i_t][o_t];   delta[r] += Gt[o][jL][r] * (s * A_L[i_L][jL])
// Gt is the mode-n unfolding of the core, [o][jL][r]; its address is uniform
// across the wave, so it is read through scalar loads.  All gathers of the
// lane are issued before any arithmetic.  N = 3 (one outer slot) with
// R <= 16 is fully unrolled out of registers; otherwise rows that need a
// runtime index are staged in this wave's LDS (one column per lane, so each
// lane reads only what it wrote).
template <int R>
__device__ inline __attribute__((always_inline)) void
contract_slots(int slots, const float *__restrict__ A_L, uint32_t iL,
               const float *const (&A_O)[kSlots], const uint32_t (&io)[kSlots],
               const float *__restrict__ Gt_, float *lrows, int lane, float (&d)[R]) {
  constexpr int RS = row_stride(R);
  constexpr bool kReg = R <= 16;
  gt_cfloat *__restrict__ Gt = core_operand(Gt_);
  float aL[kReg ? RS : 4];
  float aO[kReg ? RS : 4];
  if constexpr (kReg) {
    load_row<R>(A_L, iL, aL);
    if (slots == 1) load_row<R>(A_O[0], io[0], aO);
  }
  if (!kReg || slots >= 2) {
#pragma unroll
    for (int t = 0; t < kSlots; ++t) {
      if (t >= slots) continue;
      const float *row = A_O[t] + (size_t)io[t] * RS;
#pragma unroll
      for (int j = 0; j < R; ++j) lrows[(t * R + j) * kWave + lane] = row[j];
    }
    if (!kReg) {
      const float *row = A_L + (size_t)iL * RS;
#pragma unroll
      for (int j = 0; j < R; ++j) lrows[(slots * R + j) * kWave + lane] = row[j];
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) d[r] = 0.f;
  if constexpr (kReg) {
    if (slots == 0) {
#pragma unroll
      for (int jL = 0; jL < R; ++jL)
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = __builtin_fmaf(Gt[jL * R + r], aL[jL], d[r]);
      return;
    }
    if (slots == 1) {
#pragma unroll
      for (int o = 0; o < R; ++o) {
        const float s = aO[o];
#pragma unroll
        for (int jL = 0; jL < R; ++jL) {
          const float K = s * aL[jL];
#pragma unroll
          for (int r = 0; r < R; ++r) d[r] = __builtin_fmaf(Gt[(o * R + jL) * R + r], K, d[r]);
        }
      }
      return;
    }
  }
  int O = 1;
  for (int t = 0; t < slots; ++t) O *= R;
  int dig[kSlots];
#pragma unroll
  for (int t = 0; t < kSlots; ++t) dig[t] = 0;
  for (int o = 0; o < O; ++o) {
    float s = 1.f;
#pragma unroll
    for (int t = 0; t < kSlots; ++t)
      if (t < slots) s *= lrows[(t * R + dig[t]) * kWave + lane];
    gt_cfloat *__restrict__ gp = Gt + (size_t)o * R * R;
    if constexpr (kReg) {
#pragma unroll
      for (int jL = 0; jL < R; ++jL) {
        const float K = s * aL[jL];
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = __builtin_fmaf(gp[jL * R + r], K, d[r]);
      }
    } else {
      const float *lL = lrows + (size_t)slots * R * kWave + lane;
      // four jL per iteration: their 4R core values go out as one batch of
      // scalar loads (measured at R = 32: 16 % faster than one or two, and
      // than eight)
#pragma unroll 4
      for (int jL = 0; jL < R; ++jL) {
        const float K = s * lL[jL * kWave];
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = __builtin_fmaf(gp[jL * R + r], K, d[r]);
      }
    }
#pragma unroll
    for (int t = kSlots - 1; t >= 0; --t) {
      if (t < slots) {
        if (++dig[t] < R) break;
        dig[t] = 0;
      }
    }
  }
}

// The per-row Gram accumulation of a chunk for the wide ranks (R > 15), one
// block of Gram entries per lane.  The entries are those of the (R+1)-column
// matrix [delta | x] against itself: (p, c) with p <= c < R is B(p, c), (p, R)
// is C(p) = sum x delta_p.  Lane l owns rows p0 = 2i and p0 + 1 against the K
// columns c0 .. c0 + K - 1 of one block; the blocks of row pair i start at
// column 2i and tile the row pair's columns up to R, and K is the least width
// whose blocks fit in one wave (K = 5 at R = 32: 64 blocks).  Per nonzero a
// lane reads K + 2 values and adds 2K FP64 products of the FP32 deltas (K + 2
// reads and conversions for 2K entries, where one entry per lane takes two of
// each); entries below the diagonal or past column R are formed and dropped.
// Each entry is summed over the nonzeros in stream order, so the sums are
// those of one accumulator per entry.  The wave walks each tile's nonzeros in
// order and flushes on every row change: complete rows to the Gram table,
// the chunk's first and last runs, when they cross its edges, to the partial
// tables of the fix-up.  The tile is kWave rows of pitch RP >= R + 1 in LDS,
// deltas in columns < R and x in column R; `srow` holds the tiles' row ids.
template <int R> struct GramBlocks {
  static constexpr int blocks(int K) {
    int n = 0;
    for (int i = 0; 2 * i < R; ++i) n += (R + 1 - 2 * i + K - 1) / K;
    return n;
  }
  static constexpr int width() {
    int K = 1;
    while (blocks(K) > kWave) ++K;
    return K;
  }
  static constexpr int K = width();
  static constexpr int NB = blocks(K);
};
// nonzeros per step of BlockGram's whole-tile loop
constexpr int kBlockGramUnroll = 4;

template <int R, int RP> struct BlockGram {
  static_assert(RP >= R + 1, "the tile holds columns 0..R");
  static constexpr int W = gram_width(R);
  static constexpr int K = GramBlocks<R>::K;
  const GramArgs &a;
  uint64_t c, cb, ce;
  int lane;
  int p0, c0;      // the lane's block: rows p0, p0 + 1, columns c0 .. c0 + K - 1 (p0 = -1: none)
  int rd[2], cd[K]; // the tile columns read for them, clamped to column R
  double acc[2][K];
  uint32_t cur, last_row;
  uint64_t run_begin;
  bool head_open, tail_open;

  __device__ BlockGram(const GramArgs &args, uint64_t chunk, uint64_t b, uint64_t e, int l)
      : a(args), c(chunk), cb(b), ce(e), lane(l) {
    p0 = -1;
    c0 = 0;
    int n = 0;
    for (int i = 0; 2 * i < R; ++i) {
      const int nb = (R + 1 - 2 * i + K - 1) / K;
      if (lane >= n && lane < n + nb) {
        p0 = 2 * i;
        c0 = 2 * i + (lane - n) * K;
      }
      n += nb;
    }
    const int pr = p0 < 0 ? 0 : p0;
    rd[0] = pr;
    rd[1] = pr + 1 <= R ? pr + 1 : R;
#pragma unroll
    for (int k = 0; k < K; ++k) cd[k] = c0 + k <= R ? c0 + k : R;
#pragma unroll
    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
      for (int k = 0; k < K; ++k) acc[rr][k] = 0.0;
    cur = __builtin_amdgcn_readfirstlane(a.idx_row[cb]);
    run_begin = cb;
    // Runs other than the chunk's first and last are whole rows.
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
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
      const int p = p0 + rr;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int cc = c0 + k;
        if (p0 >= 0 && p < R && cc >= p && cc <= R) {
          const int at = cc < R ? p * R - p * (p - 1) / 2 + (cc - p) : R * (R + 1) / 2 + p;
          dst[at] = acc[rr][k];
        }
        acc[rr][k] = 0.0;
      }
    }
  }

  __device__ void add_row(const float *tr) {
    const double r0 = (double)tr[rd[0]], r1 = (double)tr[rd[1]];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double ck = (double)tr[cd[k]];
      acc[0][k] += r0 * ck;
      acc[1][k] += r1 * ck;
    }
  }

  // Adds tile [tb, tb + cnt); tile and srow must be visible to the whole wave.
  __device__ void add_tile(const float *tile, const uint32_t *srow, uint64_t tb, int cnt) {
    // A whole tile that continues the current run (the common case inside
    // long rows) needs no row checks: the lane's columns are read through
    // per-lane pointers with the nonzero's offset as an immediate, so the
    // loop spends no arithmetic on addresses.  Same products, same order of
    // the sums as the general loop below.
    if (cnt == kWave && __ballot(srow[lane] != cur) == 0) {
      // LDS pointers by type: through generic ones the compiler may read
      // some columns with flat loads, which wait on both memory counters
      typedef __attribute__((address_space(3))) const float lds_cfloat;
      lds_cfloat *bp[2], *bq[K];
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) bp[rr] = (lds_cfloat *)(tile + rd[rr]);
#pragma unroll
      for (int k = 0; k < K; ++k) bq[k] = (lds_cfloat *)(tile + cd[k]);
      for (int j0 = 0; j0 < kWave; j0 += kBlockGramUnroll) {
#pragma unroll
        for (int j = 0; j < kBlockGramUnroll; ++j) {
          const double r0 = (double)bp[0][j * RP], r1 = (double)bp[1][j * RP];
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const double ck = (double)bq[k][j * RP];
            acc[0][k] += r0 * ck;
            acc[1][k] += r1 * ck;
          }
        }
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) bp[rr] += kBlockGramUnroll * RP;
#pragma unroll
        for (int k = 0; k < K; ++k) bq[k] += kBlockGramUnroll * RP;
      }
      return;
    }
    for (int j = 0; j < cnt; ++j) {
      const uint32_t row = __builtin_amdgcn_readfirstlane(srow[j]);
      if (row != cur) {
        flush(tb + j);
        cur = row;
        run_begin = tb + j;
      }
      add_row(tile + j * RP);
    }
  }

  __device__ void finish() {
    flush(ce);
    if (lane == 0) a.chunk_tail[c] = tail_open ? last_row : 0xFFFFFFFFu;
  }
};

constexpr int cpow(int b, int e) { return e == 0 ? 1 : b * cpow(b, e - 1); }

// Whether contract_pair's fully unrolled form is used for NS outer slots:
// R^N multiply-adds per nonzero (N = NS + 2), twice over for the pair, must
// stay a loop body the instruction cache holds (order 3 up to R = 16, order
// 4 up to R = 8, order 5 up to R = 5).
template <int R, int NS> constexpr bool pair_unrolled() {
  return cpow(R, NS + 2) <= 4096;
}

// contract_slots for NS outer slots (R^N <= 4096, so fully unrolled out of
// registers) on two nonzeros per lane: every core value read through a scalar
// load feeds two FMAs, so the wave waits on half as many scalar loads per
// nonzero.  Each delta is formed in the same order as contract_slots forms
// it: s = prod_t A_t[i_t][o_t] in slot order, then Gt[o][jL][r] * (s *
// A_L[i_L][jL]) accumulated over o (slot 0 slowest), jL and r.
template <int R, int NS>
__device__ inline __attribute__((always_inline)) void
contract_pair(const float *__restrict__ A_L, const float *const (&A_O)[kSlots],
              const uint32_t (&iL)[2], const uint32_t (&io)[2][kSlots],
              gt_cfloat *__restrict__ Gt, float (&d0)[R], float (&d1)[R]) {
  static_assert(pair_unrolled<R, NS>(), "unrolled contraction covers R^N <= 4096");
  constexpr int RS = row_stride(R);
  float aL0[RS], aL1[RS], aO0[NS][RS], aO1[NS][RS];
  load_row<R>(A_L, iL[0], aL0);
#pragma unroll
  for (int t = 0; t < NS; ++t) load_row<R>(A_O[t], io[0][t], aO0[t]);
  load_row<R>(A_L, iL[1], aL1);
#pragma unroll
  for (int t = 0; t < NS; ++t) load_row<R>(A_O[t], io[1][t], aO1[t]);
#pragma unroll
  for (int r = 0; r < R; ++r) d0[r] = d1[r] = 0.f;
#pragma unroll
  for (int o = 0; o < cpow(R, NS); ++o) {
    float s0 = aO0[0][o / cpow(R, NS - 1)], s1 = aO1[0][o / cpow(R, NS - 1)];
#pragma unroll
    for (int t = 1; t < NS; ++t) {
      const int dg = o / cpow(R, NS - 1 - t) % R;
      s0 *= aO0[t][dg];
      s1 *= aO1[t][dg];
    }
#pragma unroll
    for (int jL = 0; jL < R; ++jL) {
      const float K0 = s0 * aL0[jL], K1 = s1 * aL1[jL];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float g = Gt[(o * R + jL) * R + r];
        d0[r] = __builtin_fmaf(g, K0, d0[r]);
        d1[r] = __builtin_fmaf(g, K1, d1[r]);
      }
    }
  }
}

} // namespace gt
