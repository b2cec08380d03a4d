// gt_internal.hpp -- shared declarations of the gfx950 engine: launch
// wrappers exported by each kernel file, argument blocks, the rank list the
// kernels are built for, and HIP error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/gputucker_hip.h"

namespace gt {

constexpr int kMaxOrder = GT_MAX_ORDER;
constexpr int kWave = 64;      // CDNA wavefront width
constexpr int kTile = 64;      // nonzeros per tile: one per lane

// Packed Gram row: upper triangle of B (R(R+1)/2 doubles, row-major p <= q)
// followed by C (R doubles): gram_entries(R) values.  Rows are stored with
// the stride gram_width(R), rounded up to an even count so every row starts
// on a 16-byte boundary and is read with 16-byte loads; the padding entry is
// never read.
__host__ __device__ constexpr int gram_entries(int R) { return R * (R + 1) / 2 + R; }
__host__ __device__ constexpr int gram_width(int R) { return (gram_entries(R) + 1) / 2 * 2; }

// Factor rows are stored with a stride of 4 or 8 floats (R <= 8), else R
// rounded up to 16 floats, padding zero: a row is read with 16-byte loads and,
// from a 64-byte-aligned start, a random row gather of up to 16 floats touches
// one 64-byte memory segment (a 48-byte stride would straddle two for three
// rows in four).
__host__ __device__ constexpr int row_stride(int R) {
  return R <= 4 ? 4 : (R <= 8 ? 8 : (R + 15) & ~15);
}

// Ranks the templated kernels are instantiated for.  Any other rank up to
// GT_MAX_RANK runs at the next one in this list (kernel_rank): the factors'
// extra columns and the core's extra slices are zero, so the extra delta
// columns are zero, the extra rows and columns of B are lambda I and of C
// zero, and the row solves give zero there and leave the real block as it
// would be at the rank itself.
#define GT_FOR_EACH_RANK(X) \
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
  X(20) X(24) X(28) X(30) X(32) X(36) X(40) X(44) X(50)

// The instantiated rank the kernels run a rank-R model at (0: above GT_MAX_RANK).
inline int kernel_rank(int R) {
  int k = 0;
#define GT_RANK_CASE(r) if (k == 0 && R <= r) k = r;
  GT_FOR_EACH_RANK(GT_RANK_CASE)
#undef GT_RANK_CASE
  return R >= 1 ? k : 0;
}

// LDS bytes one workgroup may allocate on the current device (set at context
// creation from hipDeviceAttributeMaxSharedMemoryPerBlock; 64 KiB until then).
size_t lds_limit();
void set_lds_limit(size_t bytes);

struct HipError : std::runtime_error {
  explicit HipError(const std::string &m) : std::runtime_error(m) {}
};

#define GT_HIP(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::gt::HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                           __FILE__ + ":" + std::to_string(__LINE__));                   \
  } while (0)

// ---------------------------------------------------------------- sort/scan
// Scratch bytes the radix sort of n pairs needs.
size_t radix_sort_temp_bytes(uint64_t n);
// Stable LSD radix sort of (key, value) pairs on bits [0, bits).  keys/vals are
// sorted in place; tmp_keys/tmp_vals are n-element ping-pong buffers; temp is
// radix_sort_temp_bytes(n) of scratch.
void radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys, uint32_t *tmp_vals,
                      uint64_t n, int bits, void *temp, hipStream_t s);
size_t scan_temp_bytes(uint64_t n);
// Exclusive scan of n uint32 counts into n+1 uint64 offsets (out[n] = total).
void exclusive_scan_u32_to_u64(const uint32_t *in, uint64_t *out, uint64_t n, void *temp,
                               hipStream_t s);
void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
void fill_f64(double *p, double v, uint64_t n, hipStream_t s);
void iota_u32(uint32_t *p, uint64_t n, hipStream_t s);

// ---------------------------------------------------------------- partition
struct PartArgs {
  int order;
  uint64_t nnz;
  const uint32_t *idx[kMaxOrder]; // global 0-based indices, input order
  uint32_t bdims[kMaxOrder];
  uint32_t part[kMaxOrder];
};
// keys[e] = block id of nonzero e (tensor_manager.tpp:232-242), vals[e] = e.
void block_ids(const PartArgs &a, uint32_t blocks, uint32_t *keys, uint32_t *vals,
               uint32_t *block_counts, hipStream_t s);
// Row pointer (rows+1 entries) of n keys sorted ascending, all < rows.
void row_ptr_from_sorted(const uint32_t *keys, uint64_t n, uint64_t rows, uint64_t *row_ptr,
                         hipStream_t s);
struct U32Ptrs {
  const uint32_t *p[kMaxOrder];
};
struct U32MutPtrs {
  uint32_t *p[kMaxOrder];
};
// dst.p[k][j] = src.p[k][perm[j]] for k < order, dst_val[j] = src_val[perm[j]]
// (dst_val may be null).
void gather_coo(int order, uint64_t nnz, const uint32_t *perm, U32Ptrs src, const float *src_val,
                U32MutPtrs dst, float *dst_val, hipStream_t s);
// flags[e] = 1 for the nonzeros of a.idx (a.nnz of them) whose row of `mode`
// is in [r0, r1) and, unless all_blocks, whose block is in [b0, b1).
void flag_piece(const PartArgs &a, int mode, uint32_t r0, uint32_t r1, uint32_t b0, uint32_t b1,
                bool all_blocks, uint32_t *flags, hipStream_t s);
// hist[block] += 1 for each nonzero of a.idx whose row of `mode` is r.
void row_block_hist(const PartArgs &a, int mode, uint32_t r, uint32_t *hist, hipStream_t s);
// Stable compaction with a window: each flagged e has rank base + offs[e]
// (offs the exclusive scan of flags); those with rank in [w0, w1) go to
// dst.p[k][rank - w0] (and the value).
void scatter_flagged(int order, uint64_t n, const uint32_t *flags, const uint64_t *offs,
                     uint64_t base, uint64_t w0, uint64_t w1, U32Ptrs src, const float *src_val,
                     U32MutPtrs dst, float *dst_val, hipStream_t s);
// counts[key[e]]++ for e < n (counts must be zeroed).
void histogram_u32(const uint32_t *key, uint64_t n, uint32_t *counts, hipStream_t s);
// Per-block partial sums of x^2 in FP64 (deterministic order); returns the
// number of partials written to out (<= 1024).
int sum_squares(const float *x, uint64_t n, double *out, hipStream_t s);
// keys[e] = block id of block-ordered position perm[e] (via blk[]), vals[e] = perm[e].
void keys_from_perm(const uint32_t *perm, const uint32_t *blk, uint64_t n, uint32_t *keys,
                    hipStream_t s);
// keys[e] = row[e] << shift | slot, slot = slot_of[outer[e]] capped at
// 2^shift - 1 (the cold rows): sorting by these keys groups each row's
// nonzeros by the hot outer row they share.
void hot_keys(const uint32_t *row, const uint32_t *outer, const uint32_t *slot_of, int shift,
              uint64_t n, uint32_t *keys, hipStream_t s);
// where[j] = pos[j] - block_off[blk[pos[j]]]
void local_positions(const uint32_t *pos, const uint32_t *blk, const uint64_t *block_off,
                     uint64_t n, uint32_t *where, hipStream_t s);
// out[e] = idx[e] - base  (block-local index of one block)
void sub_u32(const uint32_t *idx, uint32_t base, uint64_t n, uint32_t *out, hipStream_t s);
// Out-of-core host streams: index arrays packed to 16 bits (two per word)
// or 24 bits (four in three words), unpacked on the device per staged
// segment; the row index of the mode-n stream is not shipped but rebuilt
// from row_ptr.  The unpackers write 4*ceil(n/4) entries.
void pack16(const uint32_t *in, uint64_t n, uint32_t *out, hipStream_t s);
void unpack16(const uint32_t *in, uint64_t n, uint32_t *out, hipStream_t s);
void pack24(const uint32_t *in, uint64_t n, uint32_t *out, hipStream_t s);
void unpack24(const uint32_t *in, uint64_t n, uint32_t *out, hipStream_t s);
void rows_from_ptr(const uint64_t *row_ptr, uint32_t rows, uint64_t lo, uint64_t len,
                   uint32_t *out, hipStream_t s);

// ---------------------------------------------------------------- ALS kernels
// Mode-n stream: the nonzeros sorted by (global row of mode n, block, position).
struct StreamView {
  const uint32_t *idx[kMaxOrder]; // global indices per mode (idx[n] is the row)
  const float *val;
  uint64_t nnz;
};

struct GramArgs {
  int order;
  int mode;
  StreamView st;
  const float *A[kMaxOrder];  // factors, (P_k*bd_k) x R row-major, FP32
  const float *Gt;            // unfolded core for this mode: [R^(N-2)][R][R] ([o][jL][r])
  const uint64_t *row_ptr;    // global CSR of this mode: rows+1
  double *gram;               // rows x W packed Gram
  double *bpart0;             // chunks x W: first incomplete run of each chunk
  double *bpart1;             // chunks x W: last incomplete run of each chunk
  uint32_t *chunk_tail;       // chunks: row whose run continues into the next chunk, else ~0u
  uint64_t chunk;             // nonzeros per chunk (multiple of 64)
  uint64_t nchunks;           // chunks of the whole stream
  uint64_t chunk_begin;       // this launch covers chunks [chunk_begin, chunk_end); the
  uint64_t chunk_end;         // stream pointers must be valid over those positions
  uint64_t rows;              // rows of this mode (row_ptr has rows + 1 entries)
  const void *Gb;             // BF16 core operand tables: K16 hi, K32 refined, K32 lo (uint16)
  // the same operands in slot order, so the kernel never indexes by mode:
  // L = last mode other than n, outer t = the other modes ascending
  int slots;                             // N - 2
  const uint32_t *idx_row;               // st.idx[n]
  const uint32_t *idx_L;                 // st.idx[L]
  const uint32_t *idx_O[kMaxOrder - 2];  // st.idx[outer t]
  const float *A_L;
  const float *A_O[kMaxOrder - 2];
  const uint32_t *As_L;                  // split BF16 copy of A_L (16 words per row) or null
  float *delta_out;                      // test hook: null, or nnz x R deltas in stream order
  int alt_source;                        // 1: refined BF16 takes its core operand from the other
                                         // source than the default (LDS table or L2), where
                                         // both are built (the autotuner's choice)
  // Shared factor rows of the wide refined BF16 delta (order 3), hot_kind:
  // kHotOuter: hot_slot gives the slot of each row of the outer factor (~0u:
  //   none) and slot k's table is M[jL][r] = sum_o A_O[h_k][o] G[o][jL][r];
  // kHotLast: every row l of A_L has a table, N[o][r] = sum_jL A_L[l][jL]
  //   G[o][jL][r], and the blocks that share l gather As_O (the split copy
  //   of the outer factor) instead of As_L.
  // Tables are split BF16 operand words in the core table's lane order.
  // hot_tab null: neither.
  int hot_kind;
  const uint32_t *hot_slot;
  const void *hot_tab;
  const uint32_t *As_O;
};
constexpr int kHotOuter = 1, kHotLast = 2;
void launch_gram_fp32(int R, const GramArgs &a, hipStream_t s);
void launch_gram_fixup(int R, const GramArgs &a, hipStream_t s);
// delta + Gram with the Gram on MFMA (R <= 15); precision selects the delta
// stage (GT_PREC_FP32: VALU, GT_PREC_BF16/BF16R: MFMA).  O = R^(N-2).
bool gram_mfma_supported(int R);

// Operand plan of the refined BF16 delta (split BF16: a = a_hi + a_lo,
// G = G_hi + G_lo, a.G ~ a_hi.G_hi + a_lo.G_hi + a_hi.G_lo) on one
// v_mfma_f32_16x16x32_bf16 per 16 nonzeros and outer index.  Lane group g
// (lanes 16g..16g+15) supplies 8 of the 32 K slots; its factor-row words and
// core words per slot are:
//   kPlanHL   [hi(a[4g..4g+3]), lo(a[4g..4g+3])]   x [Ghi[4g..], Ghi[4g..]]
//   kPlanHH   [hi(a[0..3]), hi(a[4..7])]           x [Glo[0..3], Glo[4..7]]
//   kPlanR10  [hi(a[8..11]), lo(a[8]), lo(a[9]), hi(a[8]), hi(a[9])]
//                                   x [Ghi[8..11], Ghi[8], Ghi[9], Glo[8], Glo[9]]
//   kPlanZero zeros
// R <= 10 fits all three terms in the 32 slots; above that the K=32 MFMA
// carries the first two terms (kPlanHL everywhere) and a second K=32 MFMA,
// accumulating into the first one's result, adds a_hi.G_lo: its core words
// are [Glo[4g..4g+3], 0, 0, 0, 0], so the lo words of the factor row meet
// zeros.  Both MFMAs are the same instruction, so the second may depend on
// the first back to back.
enum Bf16rPlan { kPlanZero = 0, kPlanHL = 1, kPlanHH = 2, kPlanR10 = 3 };
__host__ __device__ constexpr bool bf16r_single_mfma(int R) { return R <= 10; }
__host__ __device__ constexpr int bf16r_plan(int R, int g) {
  return R >= 11 ? kPlanHL
       : R >= 9  ? (g < 2 ? kPlanHL : (g == 2 ? kPlanR10 : kPlanHH))
       : R >= 5  ? (g < 2 ? kPlanHL : (g == 2 ? kPlanHH : kPlanZero))
                 : (g == 0 ? kPlanHL : (g == 1 ? kPlanHH : kPlanZero));
}
// BF16 core operand tables, in uint16 words: K16 hi table ([o][lane][4]),
// then the K32 refined table ([o][lane][8]), then the K32 lo table
// ([o][lane][8]; used when R > 10), then (R <= 8) the K32 pair table.
__host__ __device__ constexpr size_t gb_k16_words(int O) { return (size_t)O * kWave * 4; }
__host__ __device__ constexpr size_t gb_k32_words(int O) { return (size_t)O * kWave * 8; }
// For R <= 8 a fourth table follows, the K32 pair table ([p][lane][8]): the
// refined words of outer indices 2p and 2p + 1 in output rows 0..7 and 8..15,
// so one MFMA per 16 nonzeros covers two outer indices.
__host__ __device__ constexpr bool bf16r_pair_table(int R) { return R <= 8; }
// pairs padded to a multiple of four (zero words), so the kernel's loop over
// them unrolls by four with a four-deep ring of core operands
__host__ __device__ constexpr int gb_pairs(int O) { return ((O + 1) / 2 + 3) / 4 * 4; }
__host__ __device__ constexpr size_t gb_k32_pair_words(int O) {
  return (size_t)gb_pairs(O) * kWave * 8;
}
// Orders 4 and 5 with at most kObMaxO outer indices (R <= 10 on the single
// refined MFMA) get a last table, the o-block table ([ob][r][lane][8]): row q
// of block (ob, r) holds the K32 refined words of G[16 ob + q][.][r] (zero
// rows for 16 ob + q >= O), so one MFMA per 16 nonzeros covers 16 outer
// indices of output column r.
constexpr int kObMaxO = 128;
__host__ __device__ constexpr int gb_oblocks(int O) { return (O + 15) / 16; }
__host__ __device__ constexpr bool bf16r_ob_table(int R, int N, int O) {
  return bf16r_single_mfma(R) && (N == 4 || N == 5) && O <= kObMaxO;
}
__host__ __device__ constexpr size_t gb_ob_offset(int R, int O) {
  return gb_k16_words(O) + 2 * gb_k32_words(O) + (bf16r_pair_table(R) ? gb_k32_pair_words(O) : 0);
}
__host__ __device__ constexpr size_t gb_ob_words(int R, int O) {
  return (size_t)gb_oblocks(O) * R * kWave * 8;
}
// Order 4 at R = 4 and 8 gets a last table for the two-mode refined BF16
// delta, the Kronecker table ([kb][mu][lane][8]).  With outer slots 0 and 1
// (values o0, o1) and e = o1 * R + jL, stage 1 contracts (o1, jL) on the
// matrix cores: output row rho = 16 kb + l % 16 of block kb is (o0, r) =
// (rho / R, rho % R), and the K slots of lane group g take the entries
// e = 16 mu + 4g + i (i < 4) of the factor-side operand k[e] = A_O1[o1] *
// A_L[jL].  MFMAs mu < kr_nhl carry k_hi.G_hi + k_lo.G_hi (factor words
// [lo(k[4 entries]), hi(...)] for even mu, [hi, lo] for odd mu; core words
// G_hi twice), MFMAs kr_nhl + nu carry k_hi.G_lo over the entries of
// mu = 2 nu and 2 nu + 1 (factor words [hi(k[2 nu]), hi(k[2 nu + 1])],
// core words G_lo).  Even/odd orders let one 8-word register block per pair
// of mu hold both HL operands and the HH operand between them.  R divides
// 16, so a lane's jL = 4g % R + i and o1 = (16 mu + 4g) / R separate.
// Order 5 at R = 5 takes the same table form with (o2, jL) contracted:
// e = o2 * R + jL (dense, R^2 = 25 entries; a lane group's (o2, jL) do not
// separate, so each lane keeps their LDS offsets), and output rows (o01, r)
// with o01 = o0 * R + o1 and r padded to 8 (rho = 8 o01 + r, rows with
// r >= R zero), so a lane's o01 = 2 kb + g / 2 and r = 4 (g % 2) + i.
__host__ __device__ constexpr bool bf16r_kr_table(int R, int N) {
  return (N == 4 && (R == 4 || R == 8)) || (N == 5 && R == 5);
}
__host__ __device__ constexpr int kr_nhl(int R) { return (R * R + 15) / 16; }
__host__ __device__ constexpr int kr_nhh(int R) { return (R * R + 31) / 32; }
__host__ __device__ constexpr int kr_nkb(int R, int N) {
  return N == 4 ? R * R / 16 : (8 * R * R + 15) / 16;
}
__host__ __device__ constexpr size_t gb_kr_offset(int R, int N, int O) {
  return gb_ob_offset(R, O) + (bf16r_ob_table(R, N, O) ? gb_ob_words(R, O) : 0);
}
__host__ __device__ constexpr size_t gb_kr_words(int R, int N) {
  return (size_t)kr_nkb(R, N) * (kr_nhl(R) + kr_nhh(R)) * kWave * 8;
}
// Returns the gt_delta_path it took.
int launch_gram_mfma(int R, int precision, const GramArgs &a, int O, hipStream_t s);
// Split BF16 copy of a factor for the MFMA delta (R <= 15): 16 uint32 words
// per row, the K=32 operand words of each lane group (bf16r_plan).
constexpr int kSplitWords = 16;
void launch_split_factor(int R, const float *A, uint64_t rows, uint32_t *As, hipStream_t s);

// The refined BF16 MFMA delta of the wide ranks (order 3, 16 <= R <= 32,
// built for R = 16, 20, 24, 28, 30, 32) with the block-per-lane FP64 Gram.
// It gathers a wide split copy of A_L (kSplitWordsWide words per row: per
// lane group g, hi then lo of columns 8g..8g+7) and reads its core operand,
// [o][mt][lane][hi 8, lo 8] BF16 words for the tiles formed per outer index,
// then [ob][lane][hi 8, lo 8] for the packed last tile, from global memory.
constexpr int kSplitWordsWide = 32;
bool gram_wide_supported(int R, int order);
void launch_gram_bf16_wide(int R, const GramArgs &a, hipStream_t s);
void launch_split_wide(int R, const float *A, uint64_t rows, uint32_t *As, hipStream_t s);
// Shared factor rows of the wide kernel (ranks whose last tile is not
// packed: R = 16, 28, 30, 32): GramArgs.hot_kind / hot_slot / hot_tab.  A
// slot table entry of kNoHotSlot marks a cold row; at most kMaxHotRows slots.
constexpr uint32_t kNoHotSlot = 0xFFFFFFFFu;
constexpr int kMaxHotRows = 63;
bool wide_hot_supported(int R, int order);
uint64_t hot_table_words(int R, uint64_t K); // uint16 words of K tables
// K tables into out: kind kHotOuter contracts the outer index o with row
// rows[k] of A (A_O), kind kHotLast contracts jL with row k of A (A_L; rows
// null).
void launch_hot_tables(int R, int kind, const float *A, const uint32_t *rows, uint64_t K,
                       const float *Gt, uint16_t *out, hipStream_t s);
// The last 16-column output tile of the wide kernel holds RR = R - 16 (MT - 1)
// columns; when RR <= 8 it packs wide_pack(R) outer indices (rows (o - PO ob,
// r) with r padded to 16 / PO), one MFMA set per PO outer indices.
__host__ __device__ constexpr int wide_tiles(int R) { return (R + 15) / 16; }
__host__ __device__ constexpr int wide_pack(int R) {
  return R - 16 * (wide_tiles(R) - 1) <= 4 ? 4 : R - 16 * (wide_tiles(R) - 1) <= 8 ? 2 : 1;
}
__host__ __device__ constexpr int wide_tiles_per_o(int R) { // tiles formed per outer index
  return wide_pack(R) == 1 ? wide_tiles(R) : wide_tiles(R) - 1;
}
__host__ __device__ constexpr int wide_packed_blocks(int R) { // packed last-tile blocks
  return wide_pack(R) == 1 ? 0 : (R + wide_pack(R) - 1) / wide_pack(R);
}
__host__ __device__ constexpr size_t gb_wide_words(int R) { // uint16 words of the core table
  return ((size_t)R * wide_tiles_per_o(R) + wide_packed_blocks(R)) * kWave * 16;
}

struct SolveArgs {
  const double *gram;
  const uint64_t *row_ptr;
  uint64_t rows;
  double lambda;
  float *A; // rows x R, written
  double *fit_partials; // null, or solve_partials(R, rows) residual partials (FP64)
  uint32_t *As;         // null, or the split BF16 copy of A to refresh (kSplitWords per row)
};
// Number of fit partials launch_solve writes for `rows` rows.
uint64_t solve_partials(int R, uint64_t rows);
uint64_t launch_solve(int R, const SolveArgs &a, hipStream_t s); // returns partials written

// Contraction over block-ordered COO for one mode n, operands in slot order
// (as GramArgs): delta of every nonzero (delta_out), xhat = A_n[i_n] . delta
// (xhat_out) and per-workgroup FP64 partial sums of (x - xhat)^2.
struct CooArgs {
  int slots;                             // N - 2
  uint64_t nnz;
  const uint32_t *idx_n;                 // global indices, block order
  const uint32_t *idx_L;
  const uint32_t *idx_O[kMaxOrder - 2];
  const float *val;
  const float *A_n;
  const float *A_L;
  const float *A_O[kMaxOrder - 2];
  const float *Gt;                       // unfolded core of mode n
  float *delta_out;                      // nnz x R or null
  float *xhat_out;                       // nnz or null
  double *err_partials;                  // gridDim.x or null
};
int launch_coo_contract(int R, const CooArgs &a, hipStream_t s); // returns grid size
void reduce_f64(const double *in, int n, double *out, hipStream_t s);

} // namespace gt
