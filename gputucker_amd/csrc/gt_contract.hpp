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

// The per-row Gram accumulation of a chunk with one lane per Gram entry (the
// wide ranks, R > 15): lane q owns entries q, q + 64, ... of the packed row
// (upper triangle of B, then C = sum x delta) and walks each tile's
// nonzeros in order with FP64 products of the FP32 deltas, flushing on every
// row change: complete rows to the Gram table, the chunk's first and last
// runs, when they cross its edges, to the partial tables of the fix-up.  The
// tile is kWave rows of pitch RP in LDS, deltas in columns < R and x in
// column R; `srow` holds the tiles' row ids.
// nonzeros per step of LaneGram's whole-tile loop
constexpr int kLaneGramUnroll = 16;

template <int R, int RP> struct LaneGram {
  static constexpr int W = gram_width(R);
  static constexpr int NE = gram_entries(R);
  static constexpr int E = (NE + kWave - 1) / kWave; // entries per lane
  const GramArgs &a;
  uint64_t c, cb, ce;
  int lane;
  uint32_t off[E]; // the entry's two tile columns, p | q << 16 (p = q = 0 past NE)
  double acc[E];
  uint32_t cur, last_row;
  uint64_t run_begin;
  bool head_open, tail_open;

  __device__ LaneGram(const GramArgs &args, uint64_t chunk, uint64_t b, uint64_t e, int l)
      : a(args), c(chunk), cb(b), ce(e), lane(l) {
    // entry ownership: q < R(R+1)/2 -> B(p, p2), p <= p2; else C(p) = delta_p * x
#pragma unroll
    for (int qi = 0; qi < E; ++qi) {
      const int q = lane + qi * kWave;
      int p = 0, p2 = 0;
      if (q < R * (R + 1) / 2) {
        int rem = q;
        while (rem >= R - p) {
          rem -= R - p;
          ++p;
        }
        p2 = p + rem;
      } else if (q < NE) {
        p = q - R * (R + 1) / 2;
        p2 = R;
      }
      off[qi] = (uint32_t)p | ((uint32_t)p2 << 16);
      acc[qi] = 0.0;
    }
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
    for (int qi = 0; qi < E; ++qi)
      if (lane + qi * kWave < NE) dst[lane + qi * kWave] = acc[qi];
  }

  // Adds tile [tb, tb + cnt); tile and srow must be visible to the whole wave.
  __device__ void add_tile(const float *tile, const uint32_t *srow, uint64_t tb, int cnt) {
    // A whole tile that continues the current run (the common case inside
    // long rows) needs no row checks: each entry's two columns are read
    // through per-lane pointers with the nonzero's offset as an immediate,
    // so the loop spends no arithmetic on addresses.  Same products, same
    // order of the sums as the general loop below.
    if (cnt == kWave && __ballot(srow[lane] != cur) == 0) {
      const float *bp[E], *bq[E];
#pragma unroll
      for (int qi = 0; qi < E; ++qi) {
        bp[qi] = tile + (off[qi] & 0xFFFFu);
        bq[qi] = tile + (off[qi] >> 16);
      }
      for (int j0 = 0; j0 < kWave; j0 += kLaneGramUnroll) {
#pragma unroll
        for (int j = 0; j < kLaneGramUnroll; ++j)
#pragma unroll
          for (int qi = 0; qi < E; ++qi)
            acc[qi] += (double)bp[qi][j * RP] * (double)bq[qi][j * RP];
#pragma unroll
        for (int qi = 0; qi < E; ++qi) {
          bp[qi] += kLaneGramUnroll * RP;
          bq[qi] += kLaneGramUnroll * RP;
        }
      }
      return;
    }
    for (int j = 0; j < cnt; ++j) {
      const uint32_t row = __builtin_amdgcn_readfirstlane(srow[j]);
      if (row != cur) {
        flush(tb + j);
        cur = row;
        run_begin = tb + j;
#pragma unroll
        for (int qi = 0; qi < E; ++qi) acc[qi] = 0.0;
      }
      const float *tr = tile + j * RP;
#pragma unroll
      for (int qi = 0; qi < E; ++qi)
        acc[qi] += (double)tr[off[qi] & 0xFFFFu] * (double)tr[off[qi] >> 16];
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
