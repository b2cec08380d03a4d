// gt_gram_wide.hip -- the refined BF16 MFMA delta for the wide ranks of an
// order-3 tensor (16 <= R <= 32), fused with the per-row Gram, and the split
// BF16 copy of a factor it gathers.
//
// With L the last mode other than n and O the remaining one (s_m[o] =
// A_O[i_O(m)][o]), the delta of nonzero m is, as in delta.cuh:62-90,
//   delta[m][r] = sum_o s_m[o] sum_jL A_L[i_L(m)][jL] G[o][jL][r].
// Stage 1 (MFMA): for each o, U_o[m][r] = sum_jL A_L[m][jL] G[o][jL][r] on
// v_mfma_f32_16x16x32_bf16, the core as the A operand (16 output columns r
// per tile, MT = ceil(R/16) tiles), the factor rows as the B operand (16
// nonzeros), K = jL (32 slots, zero above R), refined: with a = a_hi + a_lo
// and G = G_hi + G_lo (each part BF16), three accumulating MFMAs add
// a_hi.G_hi, a_lo.G_hi and a_hi.G_lo.  Lane (q, g) then holds
// U_o[16mb + q][16mt + 4g + i], i < 4.
// Stage 2 (VALU FP32): delta[m][r] += U_o[m][r] s_m[o], one s value per block
// for the lane's own nonzero column, read from this wave's LDS.
// The core operand (128 KiB for R = 32, more than LDS holds beside the
// waves) is read from global memory, where it stays in L2, one outer index
// ahead.  A tile's stage-1 and stage-2 work (768 MFMAs and 1024 FMAs per lane
// at R = 32) is long beside its gathers, so the gathers are not prefetched
// across tiles: the other waves of the SIMD run while one waits.  The tile
// of deltas then goes through LDS to the FP64 Gram with a block of entries
// per lane (BlockGram), which reuses the LDS of the s values.
// Shared factor rows: where a few rows of the outer factor hold most
// nonzeros, the partition groups each row's nonzeros by them, and a
// 16-nonzero block sharing one such row h takes U = A_L M_h with
// M_h = sum_o A_O[h][o] G[o] (formed once per mode update by k_hot_tables):
// one refined MFMA set per output tile and no stage 2, instead of one per
// outer index.  Where the mode's rows are so long that they hold long runs
// of every row of A_L, the rows are grouped by their A_L row l instead, and
// a block sharing l takes A_O N_l with N_l[o][r] = sum_jL A_L[l][jL]
// G[o][jL][r], gathering the outer factor's split rows.
#include "gt_bf16.hpp"
#include "gt_contract.hpp"

namespace gt {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

template <int R> struct WideLayout {
  static constexpr int MT = wide_tiles(R);           // 16-column output tiles
  static constexpr int PO = wide_pack(R);            // outer indices per packed last tile
  static constexpr int MTO = wide_tiles_per_o(R);    // tiles formed per outer index
  static constexpr int RP = ((R + 1) + 3) / 4 * 4;   // delta tile pitch: 16-byte rows, x at R
  static constexpr int S_FLOATS = R * kWave;         // s values, [o][nonzero]
  static constexpr int T_FLOATS = kWave * RP;        // delta tile, [nonzero][column]
  static constexpr int BUF = S_FLOATS > T_FLOATS ? S_FLOATS : T_FLOATS;
  static constexpr int PER_WAVE = BUF + kWave;       // + the tile's row ids
};

// HOT: the instance for a mode with shared factor rows (a.hot_kind, 0 for
// none); the plain instance carries none of their code.
template <int R, int HOT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_gram_bf16w(GramArgs a, int wpb) {
  static_assert(R >= 16 && R <= 32, "the wide kernel covers 16 <= R <= 32");
  static_assert(HOT == 0 || wide_pack(R) == 1, "shared rows need an unpacked last tile");
  using Lay = WideLayout<R>;
  constexpr int MT = Lay::MT, MTO = Lay::MTO, PO = Lay::PO, RP = Lay::RP, RS = row_stride(R);
  constexpr int RPr = 16 / PO; // rows per outer index in the packed last tile
  extern __shared__ __align__(16) float dyn[];
  const int lane = threadIdx.x & (kWave - 1), w = wave_index();
  const uint64_t c = a.chunk_begin + (uint64_t)blockIdx.x * wpb + w;
  if (c >= a.chunk_end) return;
  float *buf = dyn + (size_t)w * Lay::PER_WAVE; // s values, then the delta tile
  uint32_t *s_row = reinterpret_cast<uint32_t *>(buf + Lay::BUF);
  const uint64_t cb = c * a.chunk;
  const uint64_t ce = (cb + a.chunk < a.st.nnz) ? cb + a.chunk : a.st.nnz;
  const int q = lane & 15, g = lane >> 4;
  // core operand words, [o][mt][lane][hi, lo] (16 bytes each), then the
  // packed last tile's [ob][lane][hi, lo]
  const u4 *tab = reinterpret_cast<const u4 *>(a.Gb);
  const u4 *tab_p = tab + (size_t)R * MTO * kWave * 2;
  // lane (q, g) of a packed tile holds rows 4g + i: outer index PO ob + o_sub,
  // column 16 (MT - 1) + c_sub + i
  const int o_sub = (4 * g) / RPr, c_sub = (4 * g) % RPr;
  BlockGram<R, RP> gram(a, c, cb, ce, lane);

  for (uint64_t tb = cb; tb < ce; tb += kWave) {
    const uint64_t e = tb + lane;
    const bool valid = e < ce;
    const int cnt = (ce - tb < (uint64_t)kWave) ? (int)(ce - tb) : kWave;
    const uint64_t ee = valid ? e : cb; // in-bounds position for padding lanes
    const uint32_t row = a.idx_row[ee];
    const uint32_t iL = a.idx_L[ee];
    const uint32_t io = a.idx_O[0][ee];
    const float x = valid ? a.st.val[ee] : 0.f;
    // A block of 16 nonzeros that share one hot outer row h has the delta
    // sum_jL A_L[jL] M_h[jL][r], M_h the core contracted with that row (or,
    // sharing one row l of A_L, sum_o A_O[o] N_l[o][r]): one refined MFMA
    // set per output tile instead of one per outer index.  Bit mb of hot4
    // (wave-uniform): block mb is such a block, with table hslot[mb].
    uint32_t hot4 = 0;
    uint32_t hslot[4] = {0u, 0u, 0u, 0u};
    if constexpr (HOT != 0) {
      const uint32_t key = HOT == kHotOuter ? io : iL;
      const uint32_t sl = !valid ? kNoHotSlot : HOT == kHotOuter ? a.hot_slot[io] : iL;
      const uint32_t key_first = (uint32_t)__shfl((int)key, lane & 48);
      const uint64_t same = __ballot(sl != kNoHotSlot && key == key_first);
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        if (((same >> (16 * mb)) & 0xFFFFull) == 0xFFFFull) hot4 |= 1u << mb;
        hslot[mb] = __builtin_amdgcn_readlane(sl, 16 * mb);
      }
    }
    // lane (q, g) takes, for block mb, the split words of nonzero 16 mb + q:
    // hi and lo of columns 8g..8g+7 of its A_L row (its A_O row in a block
    // that shares its A_L row)
    u4 ah[4], al[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      uint32_t src = iL;
      const uint32_t *base = a.As_L;
      if constexpr (HOT == kHotLast) {
        if ((hot4 >> mb) & 1u) { // (uniform)
          src = io;
          base = a.As_O;
        }
      }
      const uint32_t r = (uint32_t)__shfl((int)src, mb * 16 + q);
      const u4 *p = reinterpret_cast<const u4 *>(base + (size_t)r * kSplitWordsWide + 8 * g);
      ah[mb] = p[0];
      al[mb] = p[1];
    }
    if (hot4 != 0xFu) { // the outer rows' s values, unless every block is hot
      float ro[RS];
      load_row<R>(a.A_O[0], io, ro);
#pragma unroll
      for (int o = 0; o < R; ++o) buf[o * kWave + lane] = ro[o];
    }
    s_row[lane] = valid ? row : 0xFFFFFFFFu;
    f4 dacc[4][MT]; // [mb][MT - 1]: the packed last tile, when PO > 1
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) dacc[mb][mt] = f4{0.f, 0.f, 0.f, 0.f};
    if constexpr (HOT != 0) {
      if (hot4 != 0) { // the hot blocks: U = A_L M_h (A_O N_l) is the delta itself
        const u4 *htab = reinterpret_cast<const u4 *>(a.hot_tab);
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
          if (!((hot4 >> mb) & 1u)) continue;
          const b8 bh = __builtin_bit_cast(b8, ah[mb]), bl = __builtin_bit_cast(b8, al[mb]);
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) {
            const size_t at = (((size_t)hslot[mb] * MT + mt) * kWave + lane) * 2;
            const b8 th = __builtin_bit_cast(b8, htab[at]), tl = __builtin_bit_cast(b8, htab[at + 1]);
            f4 u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(th, bh, (f4{0.f, 0.f, 0.f, 0.f}), 0, 0, 0);
            u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(th, bl, u, 0, 0, 0);
            u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(tl, bh, u, 0, 0, 0);
            dacc[mb][mt] = u;
          }
        }
      }
    }
    u4 nh[MTO], nl[MTO];
#pragma unroll
    for (int mt = 0; mt < MTO; ++mt) {
      nh[mt] = tab[((size_t)mt * kWave + lane) * 2];
      nl[mt] = tab[((size_t)mt * kWave + lane) * 2 + 1];
    }
    wave_sync();
    u4 ph = u4{0u, 0u, 0u, 0u}, pl = ph; // the packed tile's core operand of this block
    const int o_end = hot4 == 0xFu ? 0 : R;
#pragma unroll 1
    for (int o = 0; o < o_end; ++o) {
      u4 gh[MTO], gl[MTO];
#pragma unroll
      for (int mt = 0; mt < MTO; ++mt) {
        gh[mt] = nh[mt];
        gl[mt] = nl[mt];
      }
      if (o + 1 < R) { // the next outer index's core operand, from L2
#pragma unroll
        for (int mt = 0; mt < MTO; ++mt) {
          nh[mt] = tab[(((size_t)(o + 1) * MTO + mt) * kWave + lane) * 2];
          nl[mt] = tab[(((size_t)(o + 1) * MTO + mt) * kWave + lane) * 2 + 1];
        }
      }
      // the packed tile of block o / PO: its operand is loaded at the block's
      // first outer index and used at its last
      const bool p_first = PO > 1 && o % PO == 0;
      const bool p_last = PO > 1 && (o % PO == PO - 1 || o == R - 1);
      if (p_first) {
        ph = tab_p[((size_t)(o / PO) * kWave + lane) * 2];
        pl = tab_p[((size_t)(o / PO) * kWave + lane) * 2 + 1];
      }
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        if ((hot4 >> mb) & 1u) continue; // (uniform)
        const float sv = buf[o * kWave + mb * 16 + q];
        const b8 bh = __builtin_bit_cast(b8, ah[mb]), bl = __builtin_bit_cast(b8, al[mb]);
#pragma unroll
        for (int mt = 0; mt < MTO; ++mt) {
          const b8 th = __builtin_bit_cast(b8, gh[mt]), tl = __builtin_bit_cast(b8, gl[mt]);
          f4 u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(th, bh, (f4{0.f, 0.f, 0.f, 0.f}), 0, 0, 0);
          u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(th, bl, u, 0, 0, 0);
          u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(tl, bh, u, 0, 0, 0);
#pragma unroll
          for (int i = 0; i < 4; ++i) dacc[mb][mt][i] = __builtin_fmaf(u[i], sv, dacc[mb][mt][i]);
        }
        if constexpr (PO > 1) {
          if (p_last) { // (a uniform branch: MFMAs never sit under a lane mask)
            const b8 th = __builtin_bit_cast(b8, ph), tl = __builtin_bit_cast(b8, pl);
            f4 u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(th, bh, (f4{0.f, 0.f, 0.f, 0.f}), 0, 0,
                                                           0);
            u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(th, bl, u, 0, 0, 0);
            u = __builtin_amdgcn_mfma_f32_16x16x32_bf16(tl, bh, u, 0, 0, 0);
            // this lane group's rows are outer index PO (o / PO) + o_sub (past R:
            // zero rows of the table, any s)
            const int op = PO * (o / PO) + o_sub;
            const float spv = buf[(op < R ? op : R - 1) * kWave + mb * 16 + q];
#pragma unroll
            for (int i = 0; i < 4; ++i)
              dacc[mb][MT - 1][i] = __builtin_fmaf(u[i], spv, dacc[mb][MT - 1][i]);
          }
        }
      }
    }
    if constexpr (PO > 1) { // the packed tile: add the lane groups of the same columns
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if constexpr (PO == 4) dacc[mb][MT - 1][i] = add_xor16(dacc[mb][MT - 1][i]);
          dacc[mb][MT - 1][i] = add_xor32(dacc[mb][MT - 1][i]);
        }
    }
    wave_sync(); // every s value read before the delta tile overwrites them
    // lane (q, g) holds delta[16mb + q][16mt + 4g + i]: four columns per
    // 16-byte store; the column group holding R gets x over column R after
    float *tile = buf;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        // the packed tile's lane groups past the first outer index hold copies
        const int col = 16 * mt + (PO > 1 && mt == MT - 1 ? c_sub : 4 * g);
        const bool own = PO > 1 && mt == MT - 1 ? o_sub == 0 : true;
        if (own && col < R)
          *reinterpret_cast<f4 *>(tile + (mb * 16 + q) * RP + col) = dacc[mb][mt];
      }
    tile[lane * RP + R] = x;
    wave_sync();
    if (a.delta_out != nullptr && lane < cnt) { // test hook, uniform branch
#pragma unroll
      for (int r = 0; r < R; ++r) a.delta_out[(tb + lane) * R + r] = tile[lane * RP + r];
    }
    gram.add_tile(tile, s_row, tb, cnt);
    wave_sync();
  }
  gram.finish();
}

// Row i of the wide split copy: for each lane group g, the BF16 words of
// hi(a[8g..8g+7]) (four words) then lo(a[8g..8g+7]), zero at and beyond R,
// lo = bf16(a - hi).  One thread per (row, group).
template <int R>
__global__ void __launch_bounds__(256) k_split_wide(const float *__restrict__ A,
                                                    uint32_t *__restrict__ As, uint64_t rows) {
  constexpr int RS = row_stride(R);
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t i = t >> 2;
  const int g = (int)(t & 3);
  if (i >= rows) return;
  const float *row = A + i * RS;
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = (8 * g + k < R) ? row[8 * g + k] : 0.f;
  u4 hi, lo;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t h = pk_bf16(v[2 * k], v[2 * k + 1]);
    hi[k] = h;
    lo[k] = pk_bf16(v[2 * k] - bf16_low_f(h), v[2 * k + 1] - bf16_high_f(h));
  }
  u4 *dst = reinterpret_cast<u4 *>(As + i * kSplitWordsWide + 8 * g);
  dst[0] = hi;
  dst[1] = lo;
}

// The shared-row tables.  Kind kHotOuter: for slot k (outer row h =
// rows[k]) the core contracted with that row, M[jL][r] = sum_o A[h][o]
// G[o][jL][r]; kind kHotLast: for row l of A (= A_L) N[o][r] = sum_jL
// A[l][jL] G[o][jL][r].  FP64 sums of FP32 products, rounded to FP32, as
// split BF16 operand words in the core table's layout for one outer index:
// [k][mt][lane][hi 8, lo 8], lane (q, g) holding T[8g..8g+7][16mt + q], zero
// at and beyond R.  One thread per two words (a pair of K indices).
template <int R, int KIND>
__global__ void __launch_bounds__(256) k_hot_tables(const float *__restrict__ A,
                                                    const uint32_t *__restrict__ rows, uint64_t K,
                                                    const float *__restrict__ Gt,
                                                    uint32_t *__restrict__ out) {
  constexpr int MT = wide_tiles(R), RS = row_stride(R);
  constexpr uint64_t per_slot = (uint64_t)MT * kWave * 4;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K * per_slot) return;
  const uint64_t slot = t / per_slot;
  const int rem = (int)(t % per_slot);
  const int k = rem & 3, lane = (rem >> 2) & (kWave - 1), mt = rem >> 8;
  const int r = 16 * mt + (lane & 15), j0 = 8 * (lane >> 4) + 2 * k;
  const float *arow = A + (size_t)(KIND == kHotOuter ? rows[slot] : slot) * RS;
  float v[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int kk = j0 + i; // jL (kHotOuter) or o (kHotLast)
    double sum = 0.0;
    if (kk < R && r < R)
      for (int u = 0; u < R; ++u) {
        const size_t g = KIND == kHotOuter ? ((size_t)u * R + kk) * R + r
                                           : ((size_t)kk * R + u) * R + r;
        sum += (double)arow[u] * (double)Gt[g];
      }
    v[i] = (float)sum;
  }
  const uint32_t h = pk_bf16(v[0], v[1]);
  uint32_t *dst = out + ((slot * MT + mt) * kWave + lane) * 8;
  dst[k] = h;
  dst[4 + k] = pk_bf16(v[0] - bf16_low_f(h), v[1] - bf16_high_f(h));
}

template <int R> void wide_impl(const GramArgs &a, hipStream_t s) {
  const size_t per_wave = (size_t)WideLayout<R>::PER_WAVE * sizeof(float);
  const int wpb = 4;
  if (per_wave * wpb > lds_limit())
    throw std::invalid_argument("wide BF16 Gram: LDS per workgroup exceeds the device's");
  if (a.chunk_end <= a.chunk_begin) return;
  const unsigned grid = (unsigned)((a.chunk_end - a.chunk_begin + wpb - 1) / wpb);
  if constexpr (wide_pack(R) == 1) {
    if (a.hot_tab != nullptr && a.hot_kind == kHotOuter) {
      hipLaunchKernelGGL((k_gram_bf16w<R, kHotOuter>), dim3(grid), dim3(64 * wpb), per_wave * wpb,
                         s, a, wpb);
      return;
    }
    if (a.hot_tab != nullptr && a.hot_kind == kHotLast) {
      hipLaunchKernelGGL((k_gram_bf16w<R, kHotLast>), dim3(grid), dim3(64 * wpb), per_wave * wpb,
                         s, a, wpb);
      return;
    }
  }
  hipLaunchKernelGGL((k_gram_bf16w<R, 0>), dim3(grid), dim3(64 * wpb), per_wave * wpb, s, a, wpb);
}

} // namespace

bool gram_wide_supported(int R, int order) { return order == 3 && R >= 16 && R <= 32; }

void launch_gram_bf16_wide(int R, const GramArgs &a, hipStream_t s) {
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    wide_impl<r>(a, s);                                                                          \
    break;
    GT_CASE(16) GT_CASE(20) GT_CASE(24) GT_CASE(28) GT_CASE(30) GT_CASE(32)
#undef GT_CASE
  default:
    throw std::invalid_argument("the wide BF16 delta is built for R = 16, 20, 24, 28, 30, 32");
  }
  GT_HIP(hipGetLastError());
}

bool wide_hot_supported(int R, int order) {
  return gram_wide_supported(R, order) && wide_pack(R) == 1;
}

uint64_t hot_table_words(int R, uint64_t K) { return K * wide_tiles(R) * kWave * 16; }

void launch_hot_tables(int R, int kind, const float *A, const uint32_t *rows, uint64_t K,
                       const float *Gt, uint16_t *out, hipStream_t s) {
  if (K == 0) return;
  const uint64_t blocks = (K * wide_tiles(R) * kWave * 4 + 255) / 256;
  if (blocks >= (1ull << 31)) throw std::invalid_argument("too many shared-row tables");
  const unsigned grid = (unsigned)blocks;
  uint32_t *o32 = reinterpret_cast<uint32_t *>(out);
  switch (R * 4 + kind) {
#define GT_CASE(r)                                                                               \
  case r * 4 + kHotOuter:                                                                        \
    hipLaunchKernelGGL((k_hot_tables<r, kHotOuter>), dim3(grid), dim3(256), 0, s, A, rows, K, Gt, \
                       o32);                                                                     \
    break;                                                                                       \
  case r * 4 + kHotLast:                                                                         \
    hipLaunchKernelGGL((k_hot_tables<r, kHotLast>), dim3(grid), dim3(256), 0, s, A, rows, K, Gt,  \
                       o32);                                                                     \
    break;
    GT_CASE(16) GT_CASE(28) GT_CASE(30) GT_CASE(32)
#undef GT_CASE
  default:
    throw std::invalid_argument("the shared-row tables are built for R = 16, 28, 30, 32");
  }
  GT_HIP(hipGetLastError());
}

void launch_split_wide(int R, const float *A, uint64_t rows, uint32_t *As, hipStream_t s) {
  if (rows == 0) return;
  const unsigned grid = (unsigned)((rows * 4 + 255) / 256);
  switch (R) {
#define GT_CASE(r)                                                                               \
  case r:                                                                                        \
    hipLaunchKernelGGL(k_split_wide<r>, dim3(grid), dim3(256), 0, s, A, As, rows);              \
    break;
    GT_CASE(16) GT_CASE(20) GT_CASE(24) GT_CASE(28) GT_CASE(30) GT_CASE(32)
#undef GT_CASE
  default:
    throw std::invalid_argument("the wide split copy is built for R = 16, 20, 24, 28, 30, 32");
  }
  GT_HIP(hipGetLastError());
}

} // namespace gt
