// gt_partition.hip -- the integer kernels of the tensor-block partitioner and
// the per-mode row grouping (tensor_manager.tpp:201-302, block.tpp:114-186).
// All of them are HBM-bound byte work: coalesced SoA reads, one pass each.
#include "gt_internal.hpp"

namespace gt {
namespace {

unsigned grid_for(uint64_t n, unsigned threads, unsigned cap = 1u << 20) {
  uint64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

constexpr uint32_t kLdsBins = 4096;

// block id = sum_n floor(i_n / bd_n) * prod_{m>n} P_m (tensor_manager.tpp:232-242).
// Counts go to a per-workgroup LDS histogram first (skewed tensors put most
// nonzeros in one block, and one global counter would serialise them).
__global__ void k_block_ids(PartArgs a, uint32_t blocks, uint32_t *__restrict__ keys,
                            uint32_t *__restrict__ vals, uint32_t *__restrict__ counts) {
  __shared__ uint32_t h[kLdsBins];
  const bool lds = blocks <= kLdsBins;
  if (lds)
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) h[b] = 0;
  __syncthreads();
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t id = 0;
#pragma unroll
    for (int k = 0; k < kMaxOrder; ++k)
      if (k < a.order) id = id * a.part[k] + a.idx[k][e] / a.bdims[k];
    keys[e] = id;
    vals[e] = (uint32_t)e;
    if (lds)
      atomicAdd(&h[id], 1u);
    else
      atomicAdd(&counts[id], 1u);
  }
  __syncthreads();
  if (lds)
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x)
      if (h[b]) atomicAdd(&counts[b], h[b]);
}

// row_ptr[r] = first position whose sorted key is >= r, for r in [0, rows];
// every row pointer is written by exactly one thread.
__global__ void k_row_ptr_from_sorted(const uint32_t *__restrict__ keys, uint64_t n,
                                      uint64_t rows, uint64_t *__restrict__ row_ptr) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e <= n;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = (e == 0) ? 0 : (uint64_t)keys[e - 1] + 1;
    const uint64_t hi = (e == n) ? rows : (uint64_t)keys[e];
    for (uint64_t r = lo; r <= hi; ++r) row_ptr[r] = e;
  }
}

__global__ void k_gather_coo(int order, uint64_t nnz, const uint32_t *__restrict__ perm,
                             U32Ptrs src, const float *__restrict__ src_val, U32MutPtrs dst,
                             float *__restrict__ dst_val) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t p = perm[j];
    if (p >= nnz) continue; // perm is a permutation; never read outside the source
#pragma unroll
    for (int k = 0; k < kMaxOrder; ++k)
      if (k < order) dst.p[k][j] = src.p[k][p];
    if (dst_val) dst_val[j] = src_val[p];
  }
}

__device__ inline uint32_t block_id_of(const PartArgs &a, uint64_t e) {
  uint32_t id = 0;
#pragma unroll
  for (int k = 0; k < kMaxOrder; ++k)
    if (k < a.order) id = id * a.part[k] + a.idx[k][e] / a.bdims[k];
  return id;
}

// flags[e] = 1 for the nonzeros of a partition piece: row of `mode` in
// [r0, r1) and block in [b0, b1) (the block test skipped when it is every block)
__global__ void k_flag_piece(PartArgs a, int mode, uint32_t r0, uint32_t r1, uint32_t b0,
                             uint32_t b1, bool all_blocks, uint32_t *__restrict__ flags) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.nnz;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = a.idx[mode][e];
    bool in = v >= r0 && v < r1;
    if (in && !all_blocks) {
      const uint32_t b = block_id_of(a, e);
      in = b >= b0 && b < b1;
    }
    flags[e] = in ? 1u : 0u;
  }
}

// hist[block] += 1 for the nonzeros whose row of `mode` is r
__global__ void k_row_block_hist(PartArgs a, int mode, uint32_t r, uint32_t *__restrict__ hist) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.nnz;
       e += (uint64_t)gridDim.x * blockDim.x)
    if (a.idx[mode][e] == r) atomicAdd(&hist[block_id_of(a, e)], 1u);
}

// The flagged elements whose rank (base + offs[e], offs the exclusive scan of
// flags) lies in [w0, w1) go to dst[rank - w0], in order.
__global__ void k_scatter_flagged(int order, uint64_t n, const uint32_t *__restrict__ flags,
                                  const uint64_t *__restrict__ offs, uint64_t base, uint64_t w0,
                                  uint64_t w1, U32Ptrs src, const float *__restrict__ src_val,
                                  U32MutPtrs dst, float *__restrict__ dst_val) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x) {
    if (!flags[e]) continue;
    const uint64_t rank = base + offs[e];
    if (rank < w0 || rank >= w1) continue;
    const uint64_t j = rank - w0;
#pragma unroll
    for (int k = 0; k < kMaxOrder; ++k)
      if (k < order) dst.p[k][j] = src.p[k][e];
    dst_val[j] = src_val[e];
  }
}

__global__ void k_histogram(const uint32_t *__restrict__ key, uint64_t n,
                            uint32_t *__restrict__ counts) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&counts[key[e]], 1u);
}

__global__ void k_sum_squares(const float *__restrict__ x, uint64_t n, double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const double v = x[e];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void k_keys_from_perm(const uint32_t *__restrict__ perm, const uint32_t *__restrict__ blk,
                                 uint64_t n, uint32_t *__restrict__ keys) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x)
    keys[e] = blk[perm[e]];
}

__global__ void k_hot_keys(const uint32_t *__restrict__ row, const uint32_t *__restrict__ outer,
                           const uint32_t *__restrict__ slot_of, int shift, uint64_t n,
                           uint32_t *__restrict__ keys) {
  const uint32_t cold = (1u << shift) - 1;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t sl = slot_of[outer[e]];
    keys[e] = (row[e] << shift) | (sl < cold ? sl : cold);
  }
}

__global__ void k_local_positions(const uint32_t *__restrict__ pos, const uint32_t *__restrict__ blk,
                                  const uint64_t *__restrict__ block_off, uint64_t n,
                                  uint32_t *__restrict__ where) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t p = pos[j];
    where[j] = (uint32_t)(p - block_off[blk[p]]);
  }
}

__global__ void k_sub_u32(const uint32_t *__restrict__ idx, uint32_t base, uint64_t n,
                          uint32_t *__restrict__ out) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x)
    out[e] = idx[e] - base;
}

// 24-bit packing of an index array for the out-of-core host streams: every
// group of four indices (each < 2^24) becomes three words, little-endian
// bytes in order; a short last group is padded with zeros.
__global__ void k_pack24(const uint32_t *__restrict__ in, uint64_t n, uint32_t *__restrict__ out) {
  const uint64_t groups = (n + 3) / 4;
  for (uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < groups;
       gi += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = 4 * gi + j < n ? in[4 * gi + j] : 0u;
    out[3 * gi + 0] = v[0] | (v[1] << 24);
    out[3 * gi + 1] = (v[1] >> 8) | (v[2] << 16);
    out[3 * gi + 2] = (v[2] >> 16) | (v[3] << 8);
  }
}

// 16-bit packing (indices < 2^16): two indices per word, the first in the
// low half; a short last pair is padded with zero.
__global__ void k_pack16(const uint32_t *__restrict__ in, uint64_t n, uint32_t *__restrict__ out) {
  const uint64_t pairs = (n + 1) / 2;
  for (uint64_t pi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; pi < pairs;
       pi += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t lo = in[2 * pi], hi = 2 * pi + 1 < n ? in[2 * pi + 1] : 0u;
    out[pi] = lo | (hi << 16);
  }
}

__global__ void k_unpack16(const uint32_t *__restrict__ in, uint64_t quads,
                           uint32_t *__restrict__ out) {
  for (uint64_t qi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; qi < quads;
       qi += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 w = reinterpret_cast<const uint2 *>(in)[qi];
    reinterpret_cast<uint4 *>(out)[qi] = uint4{w.x & 0xFFFFu, w.x >> 16, w.y & 0xFFFFu, w.y >> 16};
  }
}

__global__ void k_unpack24(const uint32_t *__restrict__ in, uint64_t groups,
                           uint32_t *__restrict__ out) {
  for (uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < groups;
       gi += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t w0 = in[3 * gi], w1 = in[3 * gi + 1], w2 = in[3 * gi + 2];
    uint4 v;
    v.x = w0 & 0xFFFFFFu;
    v.y = (w0 >> 24) | ((w1 & 0xFFFFu) << 8);
    v.z = (w1 >> 16) | ((w2 & 0xFFu) << 16);
    v.w = w2 >> 8;
    reinterpret_cast<uint4 *>(out)[gi] = v;
  }
}

// Row of each stream position e in [lo, lo + len) from the CSR row pointer
// (rows + 1 entries): the last r with row_ptr[r] <= e, which skips empty
// rows.  A wave finds the rows of its first and last positions, then each
// lane searches between them.
__device__ inline uint32_t row_of(const uint64_t *__restrict__ rp, uint32_t a, uint32_t b,
                                  uint64_t e) {
  while (a < b) { // invariant: rp[a] <= e, and rp[b + 1] > e
    const uint32_t m = a + (b - a + 1) / 2;
    if (rp[m] <= e)
      a = m;
    else
      b = m - 1;
  }
  return a;
}

__global__ void k_rows_from_ptr(const uint64_t *__restrict__ rp, uint32_t rows, uint64_t lo,
                                uint64_t len, uint32_t *__restrict__ out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t w0 = t - (threadIdx.x & 63); // first position of this wave
  if (w0 >= len) return;
  const uint64_t w1 = (w0 + 63 < len ? w0 + 63 : len - 1);
  const uint32_t r0 = row_of(rp, 0, rows - 1, lo + w0);
  const uint32_t r1 = row_of(rp, r0, rows - 1, lo + w1);
  if (t < len) out[t] = row_of(rp, r0, r1, lo + t);
}

} // namespace

void block_ids(const PartArgs &a, uint32_t blocks, uint32_t *keys, uint32_t *vals,
               uint32_t *block_counts, hipStream_t s) {
  if (!a.nnz) return;
  hipLaunchKernelGGL(k_block_ids, dim3(grid_for(a.nnz, 256, 4096)), dim3(256), 0, s, a, blocks,
                     keys, vals, block_counts);
  GT_HIP(hipGetLastError());
}

void row_ptr_from_sorted(const uint32_t *keys, uint64_t n, uint64_t rows, uint64_t *row_ptr,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_row_ptr_from_sorted, dim3(grid_for(n + 1, 256)), dim3(256), 0, s, keys, n,
                     rows, row_ptr);
  GT_HIP(hipGetLastError());
}

void gather_coo(int order, uint64_t nnz, const uint32_t *perm, U32Ptrs src, const float *src_val,
                U32MutPtrs dst, float *dst_val, hipStream_t s) {
  if (!nnz) return;
  hipLaunchKernelGGL(k_gather_coo, dim3(grid_for(nnz, 256)), dim3(256), 0, s, order, nnz, perm,
                     src, src_val, dst, dst_val);
  GT_HIP(hipGetLastError());
}

void flag_piece(const PartArgs &a, int mode, uint32_t r0, uint32_t r1, uint32_t b0, uint32_t b1,
                bool all_blocks, uint32_t *flags, hipStream_t s) {
  if (!a.nnz) return;
  hipLaunchKernelGGL(k_flag_piece, dim3(grid_for(a.nnz, 256)), dim3(256), 0, s, a, mode, r0, r1,
                     b0, b1, all_blocks, flags);
  GT_HIP(hipGetLastError());
}

void row_block_hist(const PartArgs &a, int mode, uint32_t r, uint32_t *hist, hipStream_t s) {
  if (!a.nnz) return;
  hipLaunchKernelGGL(k_row_block_hist, dim3(grid_for(a.nnz, 256)), dim3(256), 0, s, a, mode, r,
                     hist);
  GT_HIP(hipGetLastError());
}

void scatter_flagged(int order, uint64_t n, const uint32_t *flags, const uint64_t *offs,
                     uint64_t base, uint64_t w0, uint64_t w1, U32Ptrs src, const float *src_val,
                     U32MutPtrs dst, float *dst_val, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_scatter_flagged, dim3(grid_for(n, 256)), dim3(256), 0, s, order, n, flags,
                     offs, base, w0, w1, src, src_val, dst, dst_val);
  GT_HIP(hipGetLastError());
}

void histogram_u32(const uint32_t *key, uint64_t n, uint32_t *counts, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_histogram, dim3(grid_for(n, 256)), dim3(256), 0, s, key, n, counts);
  GT_HIP(hipGetLastError());
}

int sum_squares(const float *x, uint64_t n, double *out, hipStream_t s) {
  const int g = (int)grid_for(n, 256, 1024);
  hipLaunchKernelGGL(k_sum_squares, dim3(g), dim3(256), 0, s, x, n, out);
  GT_HIP(hipGetLastError());
  return g;
}

void keys_from_perm(const uint32_t *perm, const uint32_t *blk, uint64_t n, uint32_t *keys,
                    hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_keys_from_perm, dim3(grid_for(n, 256)), dim3(256), 0, s, perm, blk, n,
                     keys);
  GT_HIP(hipGetLastError());
}

void hot_keys(const uint32_t *row, const uint32_t *outer, const uint32_t *slot_of, int shift,
              uint64_t n, uint32_t *keys, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_hot_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, row, outer, slot_of,
                     shift, n, keys);
  GT_HIP(hipGetLastError());
}

void local_positions(const uint32_t *pos, const uint32_t *blk, const uint64_t *block_off,
                     uint64_t n, uint32_t *where, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_local_positions, dim3(grid_for(n, 256)), dim3(256), 0, s, pos, blk,
                     block_off, n, where);
  GT_HIP(hipGetLastError());
}

void sub_u32(const uint32_t *idx, uint32_t base, uint64_t n, uint32_t *out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_sub_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, idx, base, n, out);
  GT_HIP(hipGetLastError());
}

void pack24(const uint32_t *in, uint64_t n, uint32_t *out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_pack24, dim3(grid_for((n + 3) / 4, 256)), dim3(256), 0, s, in, n, out);
  GT_HIP(hipGetLastError());
}

void pack16(const uint32_t *in, uint64_t n, uint32_t *out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_pack16, dim3(grid_for((n + 1) / 2, 256)), dim3(256), 0, s, in, n, out);
  GT_HIP(hipGetLastError());
}

void unpack16(const uint32_t *in, uint64_t n, uint32_t *out, hipStream_t s) {
  if (!n) return;
  const uint64_t quads = (n + 3) / 4;
  hipLaunchKernelGGL(k_unpack16, dim3(grid_for(quads, 256)), dim3(256), 0, s, in, quads, out);
  GT_HIP(hipGetLastError());
}

void unpack24(const uint32_t *in, uint64_t n, uint32_t *out, hipStream_t s) {
  if (!n) return;
  const uint64_t groups = (n + 3) / 4;
  hipLaunchKernelGGL(k_unpack24, dim3(grid_for(groups, 256)), dim3(256), 0, s, in, groups, out);
  GT_HIP(hipGetLastError());
}

void rows_from_ptr(const uint64_t *row_ptr, uint32_t rows, uint64_t lo, uint64_t len,
                   uint32_t *out, hipStream_t s) {
  if (!len) return;
  if (!rows) throw std::invalid_argument("rows_from_ptr: no rows");
  const uint64_t g = (len + 255) / 256;
  if (g > 0xFFFFFFFFull) throw std::invalid_argument("rows_from_ptr: segment too long");
  hipLaunchKernelGGL(k_rows_from_ptr, dim3((unsigned)g), dim3(256), 0, s, row_ptr, rows, lo, len,
                     out);
  GT_HIP(hipGetLastError());
}

} // namespace gt
