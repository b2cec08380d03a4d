// gt_gram.hip -- the FP32 (VALU) path of one mode update: delta and the
// per-row Gram accumulation fused in one pass over the mode's row-sorted
// nonzero stream, the fix-up of rows that straddle chunks, and the COO-order
// contraction used for the fit (and as a delta/xhat test hook).
//
// Reference: computing_delta_kernel (delta.cuh:33-92) computes
//   delta[e][r] = sum_{j: j_n = r} G[j] prod_{k != n} A_k[i_k][j_k]
// one thread per nonzero, and ComputingBC (update.cuh:35-101) then sums
// B_i += delta delta^T, C_i += x delta on the host in double.  Here both
// happen in one kernel: a wave owns a chunk of the stream (64 * tiles
// nonzeros, rows ascending), each lane computes one nonzero's delta per tile,
// the tile goes to LDS, and the lanes then switch roles -- each lane owns a
// block of Gram entries (two rows of the upper triangle of B, or of C,
// against a few columns) and walks the tile's nonzeros in order, flushing its
// FP64 accumulators whenever the row changes.  delta never reaches HBM.  Products of the FP32 deltas are formed exactly in FP64 and
// summed in FP64 (the reference rounds each product to value_t first).
#include <atomic>

#include "gt_contract.hpp"

namespace gt {
namespace {

template <int R> struct GramLayout {
  static constexpr int W = gram_width(R);                // row stride
  static constexpr int NE = gram_entries(R);
  static constexpr int E = (NE + kWave - 1) / kWave;     // entries per lane
  static constexpr int RP = (R % 2 == 0) ? R + 1 : R + 2; // odd tile pitch; column R holds x
};

// LDS floats per wave of k_gram_fp32: the delta tile (kWave x RP, column R
// holding x), the tile's row ids, and the rows contract_slots stages.
template <int R> constexpr size_t fp32_lds_floats_per_wave(int slots) {
  return (size_t)kWave * GramLayout<R>::RP + kWave + (size_t)contract_lds_rows<R>(slots) * R * kWave;
}

// Waves per workgroup for a kernel whose waves each take `per_wave` bytes of
// LDS: the most waves per CU that the device's LDS holds, in workgroups of
// 4, 2 or 1 waves (ties to the larger workgroup); 0 if not even one fits.
int waves_per_block(size_t per_wave) {
  const size_t limit = lds_limit();
  int best = 0, best_waves = 0;
  for (int w : {4, 2, 1}) {
    const size_t wg = per_wave * w;
    if (wg > limit) continue;
    const int waves = (int)(limit / wg) * w;
    if (waves > best_waves) {
      best = w;
      best_waves = waves;
    }
  }
  return best;
}

template <int R>
__global__ void __launch_bounds__(256) k_gram_fp32(GramArgs a, int wpb) {
  constexpr int RP = GramLayout<R>::RP;
  extern __shared__ float s_dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  const uint64_t c = a.chunk_begin + (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.chunk_end) return;
  float *tile = s_dyn + (size_t)w * fp32_lds_floats_per_wave<R>(a.slots);
  uint32_t *s_row = reinterpret_cast<uint32_t *>(tile + kWave * RP);
  float *lrows = tile + kWave * RP + kWave;
  const uint64_t cb = c * a.chunk;
  const uint64_t ce = (cb + a.chunk < a.st.nnz) ? cb + a.chunk : a.st.nnz;
  const uint32_t *__restrict__ rows = a.idx_row;
  BlockGram<R, RP> gram(a, c, cb, ce, lane);

  for (uint64_t tb = cb; tb < ce; tb += kWave) {
    const uint64_t e = tb + lane;
    const bool valid = e < ce;
    const uint64_t ee = valid ? e : cb; // in-bounds position for padding lanes
    const uint32_t iL = a.idx_L[ee];
    uint32_t io[kSlots];
#pragma unroll
    for (int t = 0; t < kSlots; ++t) io[t] = (t < a.slots) ? a.idx_O[t][ee] : 0u;
    const float x = valid ? a.st.val[ee] : 0.f;
    float d[R];
    contract_slots<R>(a.slots, a.A_L, iL, a.A_O, io, a.Gt, lrows, lane, d);
#pragma unroll
    for (int r = 0; r < R; ++r) tile[lane * RP + r] = d[r];
    tile[lane * RP + R] = x;
    s_row[lane] = rows[ee];
    wave_sync();
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    if (a.delta_out != nullptr && lane < cnt) { // test hook, uniform branch
#pragma unroll
      for (int r = 0; r < R; ++r) a.delta_out[(tb + lane) * R + r] = tile[lane * RP + r];
    }
    gram.add_tile(tile, s_row, tb, cnt);
    wave_sync();
  }
  gram.finish();
}

// Fix-up of rows that straddle chunks, in two levels so that a row spread
// over thousands of chunks is summed by many waves.  chunk_tail[c] names the
// row whose run continues from chunk c into chunk c + 1.  The partial of a
// row's run in chunk c sits in bpart0[c] when the run is the chunk's first
// (the row starts at or before the chunk's first position), else in
// bpart1[c].
//
// Level 1, one workgroup per piece of 64 chunks: for each row whose run goes on
// from chunk c into later chunks of the same piece, the partials of those
// later chunks (always bpart0) are added into c's slot.
// Level 2, one workgroup per piece: for each row that starts in chunk c and
// continues past it, c's slot (now the row's sum over its first piece) plus
// bpart0 of the first chunk of each later piece the row covers (the row's
// sum over that piece) is the row's Gram entry.  Every sum runs in a fixed
// order, so results are deterministic.
constexpr int kPiece = 64; // chunks per first-level piece (one per lane)

// out[j] = sum of bp[k][lane + 64 j] over k = first, first + step, ..., <= last,
// for every entry lane + 64 j < NE: all of the lane's entries in one pass,
// four independent chains each, so a wave waits for one memory round trip per
// four chunks however many entries a row has
template <int NE>
__device__ inline void sum_bpart0(const double *bp, uint64_t first, uint64_t last, uint64_t step,
                                  int W, int lane, double (&out)[(NE + kWave - 1) / kWave]) {
  constexpr int NQ = (NE + kWave - 1) / kWave;
  double s[NQ][4];
#pragma unroll
  for (int j = 0; j < NQ; ++j) s[j][0] = s[j][1] = s[j][2] = s[j][3] = 0.0;
  uint64_t k = first;
  for (; k + 3 * step <= last; k += 4 * step) {
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const int q = lane + kWave * j;
      if (q < NE) {
#pragma unroll
        for (int u = 0; u < 4; ++u) s[j][u] += bp[(k + u * step) * W + q];
      }
    }
  }
  for (; k <= last; k += step) {
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      const int q = lane + kWave * j;
      if (q < NE) s[j][0] += bp[k * W + q];
    }
  }
#pragma unroll
  for (int j = 0; j < NQ; ++j) out[j] = (s[j][0] + s[j][1]) + (s[j][2] + s[j][3]);
}

// Each fix-up workgroup has kFixWaves waves.  All waves find the same row
// segments from the chunk tails; independent segments are dealt to the
// waves in turn, and a row spread over many pieces is summed by all waves
// together (wave w takes pieces w, w + kFixWaves, ...; wave 0 adds the
// per-wave sums in wave order).
constexpr int kFixWaves = 4;
constexpr uint64_t kLongRow = 8; // pieces above which all waves share a row

template <int R>
__global__ void __launch_bounds__(kWave * kFixWaves) k_gram_fix_pieces(GramArgs a) {
  constexpr int W = gram_width(R), NE = gram_entries(R), NQ = (NE + kWave - 1) / kWave;
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  const uint64_t c = (uint64_t)blockIdx.x * kPiece + lane;
  const uint32_t none = 0xFFFFFFFFu;
  const uint32_t t = c < a.nchunks ? a.chunk_tail[c] : none;               // c -> c + 1
  const uint32_t pv = (c > 0 && c < a.nchunks) ? a.chunk_tail[c - 1] : none; // c - 1 -> c
  uint64_t starts = __ballot(t != none && (lane == 0 || pv != t));
  int k = 0;
  while (starts) {
    const int j = __builtin_ctzll(starts);
    starts &= starts - 1;
    if ((k++ % kFixWaves) != w) continue; // this segment belongs to another wave
    const uint32_t r = __builtin_amdgcn_readlane(t, j);
    const uint64_t bal = __ballot(pv == r); // chunks the row runs into from the one before
    const uint64_t into = j >= kPiece - 1 ? 0ull : bal >> (j + 1);
    const int e = j + (int)__builtin_ctzll(~into);
    if (e == j) continue;
    const uint64_t cj = (uint64_t)blockIdx.x * kPiece + j;
    double *first = (a.row_ptr[r] <= cj * a.chunk ? a.bpart0 : a.bpart1) + cj * W;
    double sum[NQ];
    sum_bpart0<NE>(a.bpart0, cj + 1, cj + (e - j), 1, W, lane, sum);
#pragma unroll
    for (int jq = 0; jq < NQ; ++jq)
      if (lane + kWave * jq < NE) first[lane + kWave * jq] += sum[jq];
  }
}

template <int R>
__global__ void __launch_bounds__(kWave * kFixWaves) k_gram_fix_rows(GramArgs a) {
  constexpr int W = gram_width(R), NE = gram_entries(R), NQ = (NE + kWave - 1) / kWave;
  __shared__ double red[kFixWaves][NE];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  const uint64_t c = (uint64_t)blockIdx.x * kPiece + lane;
  const uint32_t none = 0xFFFFFFFFu;
  const uint32_t t = c < a.nchunks ? a.chunk_tail[c] : none;
  const bool begins = t != none && a.row_ptr[t] >= c * a.chunk; // the row starts in chunk c
  uint64_t mask = __ballot(begins);
  int k = 0;
  while (mask) {
    const int j = __builtin_ctzll(mask);
    mask &= mask - 1;
    const uint32_t r = __builtin_amdgcn_readlane(t, j);
    const uint64_t cj = (uint64_t)blockIdx.x * kPiece + j;
    const uint64_t rb = a.row_ptr[r];
    const uint64_t last = (a.row_ptr[r + 1] - 1) / a.chunk;
    const uint64_t p0 = cj / kPiece, pl = last / kPiece;
    const double *first = (rb == cj * a.chunk ? a.bpart0 : a.bpart1) + cj * W;
    if (pl - p0 > kLongRow) { // the same decision in every wave
      double sum[NQ];
      sum_bpart0<NE>(a.bpart0, (p0 + 1 + w) * kPiece, pl * kPiece, kPiece * kFixWaves, W, lane,
                     sum);
#pragma unroll
      for (int jq = 0; jq < NQ; ++jq)
        if (lane + kWave * jq < NE) red[w][lane + kWave * jq] = sum[jq];
      __syncthreads();
      if (w == 0) {
        for (int q = lane; q < NE; q += kWave)
          a.gram[(size_t)r * W + q] =
              first[q] + ((red[0][q] + red[1][q]) + (red[2][q] + red[3][q]));
      }
      __syncthreads();
    } else {
      if ((k++ % kFixWaves) != w) continue;
      double sum[NQ];
#pragma unroll
      for (int jq = 0; jq < NQ; ++jq) sum[jq] = 0.0;
      if (pl > p0) sum_bpart0<NE>(a.bpart0, (p0 + 1) * kPiece, pl * kPiece, kPiece, W, lane, sum);
#pragma unroll
      for (int jq = 0; jq < NQ; ++jq) {
        const int q = lane + kWave * jq;
        if (q < NE) a.gram[(size_t)r * W + q] = pl > p0 ? first[q] + sum[jq] : first[q];
      }
    }
  }
}

template <int R>
__global__ void __launch_bounds__(256) k_coo_contract(CooArgs a) {
  __shared__ double red[256]; // blockDim.x entries: 64, 128 or 256
  extern __shared__ float s_rows_dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  float *lrows = s_rows_dyn + (size_t)w * contract_lds_rows<R>(a.slots) * R * kWave;
  double err = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < a.nnz; base += stride) {
    const uint64_t e = base + threadIdx.x;
    const bool valid = e < a.nnz;
    const uint64_t ee = valid ? e : 0;
    const uint32_t iL = a.idx_L[ee];
    uint32_t io[kSlots];
#pragma unroll
    for (int t = 0; t < kSlots; ++t) io[t] = (t < a.slots) ? a.idx_O[t][ee] : 0u;
    float d[R];
    contract_slots<R>(a.slots, a.A_L, iL, a.A_O, io, a.Gt, lrows, lane, d);
    if (valid) {
      if (a.delta_out) {
#pragma unroll
        for (int r = 0; r < R; ++r) a.delta_out[e * R + r] = d[r];
      }
      if (a.xhat_out || a.err_partials) {
        const float *an = a.A_n + (size_t)a.idx_n[e] * row_stride(R);
        float xh = 0.f;
#pragma unroll
        for (int r = 0; r < R; ++r) xh = __builtin_fmaf(an[r], d[r], xh);
        if (a.xhat_out) a.xhat_out[e] = xh;
        const double df = (double)a.val[e] - (double)xh;
        err += df * df;
      }
    }
  }
  if (a.err_partials) {
    red[threadIdx.x] = err;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.err_partials[blockIdx.x] = red[0];
  }
}

__global__ void k_reduce_f64(const double *__restrict__ in, int n, double *__restrict__ out) {
  __shared__ double red[256];
  double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; // independent chains
  int i = threadIdx.x;
  for (; i + 7 * 256 < n; i += 8 * 256) {
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] += in[i + k * 256];
  }
  for (; i < n; i += 256) s[0] += in[i];
  red[threadIdx.x] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

template <int R> void gram_fp32_impl(const GramArgs &a, hipStream_t s) {
  if (a.chunk_end <= a.chunk_begin) return;
  const size_t per_wave = fp32_lds_floats_per_wave<R>(a.slots) * sizeof(float);
  const int wpb = waves_per_block(per_wave);
  if (wpb == 0)
    throw std::invalid_argument("FP32 Gram: one wave's LDS (" + std::to_string(per_wave) +
                                " bytes) exceeds the device's LDS per workgroup");
  const unsigned grid = (unsigned)((a.chunk_end - a.chunk_begin + wpb - 1) / wpb);
  hipLaunchKernelGGL(k_gram_fp32<R>, dim3(grid), dim3(64 * wpb), per_wave * wpb, s, a, wpb);
}

template <int R> void gram_fixup_impl(const GramArgs &a, hipStream_t s) {
  if (a.nchunks <= 1) return;
  const unsigned grid = (unsigned)((a.nchunks + kPiece - 1) / kPiece);
  hipLaunchKernelGGL(k_gram_fix_pieces<R>, dim3(grid), dim3(kWave * kFixWaves), 0, s, a);
  hipLaunchKernelGGL(k_gram_fix_rows<R>, dim3(grid), dim3(kWave * kFixWaves), 0, s, a);
}

template <int R> int coo_impl(const CooArgs &a, hipStream_t s) {
  const size_t per_wave = (size_t)contract_lds_rows<R>(a.slots) * R * kWave * sizeof(float);
  const int wpb = per_wave == 0 ? 4 : waves_per_block(per_wave + 2048 / 4);
  if (wpb == 0)
    throw std::invalid_argument("reconstruction: one wave's LDS exceeds the device's LDS");
  const uint64_t threads = 64 * (uint64_t)wpb;
  uint64_t g = (a.nnz + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_coo_contract<R>, dim3((unsigned)g), dim3((unsigned)threads),
                     per_wave * wpb, s, a);
  return (int)g;
}

std::atomic<size_t> g_lds_limit{64 * 1024};

} // namespace

size_t lds_limit() { return g_lds_limit.load(std::memory_order_relaxed); }
void set_lds_limit(size_t bytes) {
  if (bytes) g_lds_limit.store(bytes, std::memory_order_relaxed);
}

void launch_gram_fp32(int R, const GramArgs &a, hipStream_t s) {
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    gram_fp32_impl<r>(a, s);                                                                     \
    break;
    GT_FOR_EACH_RANK(GT_CASE)
#undef GT_CASE
  default:
    throw std::invalid_argument("rank not built");
  }
  GT_HIP(hipGetLastError());
}

void launch_gram_fixup(int R, const GramArgs &a, hipStream_t s) {
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    gram_fixup_impl<r>(a, s);                                                                    \
    break;
    GT_FOR_EACH_RANK(GT_CASE)
#undef GT_CASE
  default:
    throw std::invalid_argument("rank not built");
  }
  GT_HIP(hipGetLastError());
}

int launch_coo_contract(int R, const CooArgs &a, hipStream_t s) {
  int g = 0;
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    g = coo_impl<r>(a, s);                                                                       \
    break;
    GT_FOR_EACH_RANK(GT_CASE)
#undef GT_CASE
  default:
    throw std::invalid_argument("rank not built");
  }
  GT_HIP(hipGetLastError());
  return g;
}

void reduce_f64(const double *in, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_f64, dim3(1), dim3(256), 0, s, in, n, out);
  GT_HIP(hipGetLastError());
}

} // namespace gt
