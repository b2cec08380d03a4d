// gt_sort.hip -- device-wide building blocks for the partitioner: exclusive
// scan, stable LSD radix sort of (key, value) pairs, fills.
//
// The radix sort is what makes the GPU partition bit-exact with the
// reference's serial scatter (tensor_manager.tpp:267-290) and its stable
// where_nnz (block.tpp:170-181): every pass is stable, so equal keys keep
// their input order.  Stable ranking inside a pass uses wave-wide ballots to
// find the lanes holding the same digit, so no per-thread digit counters are
// needed.
#include "gt_internal.hpp"

namespace gt {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint64_t kScanTile = kScanThreads * kScanItems;

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr uint64_t kSortTile = kSortThreads * kSortItems;
constexpr int kDigitBits = 8;
constexpr int kBins = 1 << kDigitBits;

__global__ void k_scan_reduce(const uint32_t *__restrict__ in, uint64_t n,
                              uint64_t *__restrict__ partials) {
  __shared__ uint64_t red[kScanThreads];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint64_t s = 0;
  for (int it = 0; it < kScanItems; ++it) {
    uint64_t i = base + (uint64_t)it * kScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = kScanThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Exclusive scan of the per-tile partials, one workgroup, any length.
__global__ void k_scan_partials(uint64_t *__restrict__ partials, uint64_t nb,
                                uint64_t *__restrict__ total_out) {
  __shared__ uint64_t buf[1024];
  __shared__ uint64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint64_t seg = 0; seg < nb; seg += 1024) {
    uint64_t i = seg + threadIdx.x;
    uint64_t v = (i < nb) ? partials[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) { // Hillis-Steele inclusive
      uint64_t add = (threadIdx.x >= (unsigned)off) ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < nb) partials[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ void k_scan_down(const uint32_t *__restrict__ in, uint64_t n,
                            const uint64_t *__restrict__ partials, uint64_t *__restrict__ out) {
  __shared__ uint64_t tsum[kScanThreads];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (int it = 0; it < kScanItems; ++it) {
    uint64_t i = base + it;
    v[it] = (i < n) ? in[i] : 0u;
    s += v[it];
  }
  tsum[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    uint64_t add = (threadIdx.x >= (unsigned)off) ? tsum[threadIdx.x - off] : 0;
    __syncthreads();
    tsum[threadIdx.x] += add;
    __syncthreads();
  }
  uint64_t run = partials[blockIdx.x] + tsum[threadIdx.x] - s;
#pragma unroll
  for (int it = 0; it < kScanItems; ++it) {
    uint64_t i = base + it;
    if (i < n) out[i] = run;
    run += v[it];
  }
}

__global__ void k_radix_hist(const uint32_t *__restrict__ keys, uint64_t n, int shift,
                             uint32_t *__restrict__ hist, uint32_t tiles) {
  __shared__ uint32_t h[kBins];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
  for (int it = 0; it < kSortItems; ++it) {
    uint64_t i = base + (uint64_t)it * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kBins - 1)], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile.  Elements are taken in rounds of 256
// consecutive positions; inside a round wave w holds positions 64w..64w+63.
// rank = (digit's start for this tile) + (earlier rounds) + (earlier waves
// of this round) + (earlier lanes of this wave with the same digit).
__global__ void __launch_bounds__(kSortThreads)
    k_radix_scatter(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, uint64_t n, int shift,
                    const uint64_t *__restrict__ offs, uint32_t tiles) {
  __shared__ uint32_t run[kBins];
  __shared__ uint32_t wcnt[kSortThreads / kWave][kBins];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
  run[tid] = (uint32_t)offs[(uint64_t)tid * tiles + blockIdx.x];
#pragma unroll
  for (int q = 0; q < kSortThreads / kWave; ++q) wcnt[q][tid] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int it = 0; it < kSortItems; ++it) {
    const uint64_t i = base + (uint64_t)it * kSortThreads + tid;
    const bool valid = i < n;
    const uint32_t k = valid ? kin[i] : 0u;
    const uint32_t v = valid ? vin[i] : 0u;
    const uint32_t d = (k >> shift) & (kBins - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kDigitBits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int q = 0; q < w; ++q) pos += wcnt[q][d];
      if (pos < n) { // ranks are < n by construction; the guard keeps a bad offset in bounds
        kout[pos] = k;
        vout[pos] = v;
      }
    }
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int q = 0; q < kSortThreads / kWave; ++q) {
      tot += wcnt[q][tid];
      wcnt[q][tid] = 0;
    }
    run[tid] += tot;
    __syncthreads();
  }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_f64(double *p, double v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_iota_u32(uint32_t *p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

unsigned grid_for(uint64_t n, unsigned threads, unsigned cap = 65536) {
  uint64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

} // namespace

size_t scan_temp_bytes(uint64_t n) {
  uint64_t nb = (n + kScanTile - 1) / kScanTile;
  return (size_t)(nb + 1) * sizeof(uint64_t) + 256;
}

void exclusive_scan_u32_to_u64(const uint32_t *in, uint64_t *out, uint64_t n, void *temp,
                               hipStream_t s) {
  uint64_t nb = (n + kScanTile - 1) / kScanTile;
  uint64_t *partials = static_cast<uint64_t *>(temp);
  if (nb == 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t *>(out),
                       0u, (uint64_t)2);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n,
                     partials);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, partials, nb, out + n);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n,
                     partials, out);
  GT_HIP(hipGetLastError());
}

size_t radix_sort_temp_bytes(uint64_t n) {
  uint64_t tiles = (n + kSortTile - 1) / kSortTile;
  if (tiles == 0) tiles = 1;
  uint64_t hist = tiles * kBins;
  size_t b = hist * sizeof(uint32_t);
  b = (b + 255) & ~size_t(255);
  b += (hist + 1) * sizeof(uint64_t);
  b = (b + 255) & ~size_t(255);
  b += scan_temp_bytes(hist);
  return b;
}

void radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys, uint32_t *tmp_vals,
                      uint64_t n, int bits, void *temp, hipStream_t s) {
  if (n <= 1 || bits <= 0) return;
  const uint64_t tiles = (n + kSortTile - 1) / kSortTile;
  const uint64_t hist_n = tiles * kBins;
  char *t = static_cast<char *>(temp);
  uint32_t *hist = reinterpret_cast<uint32_t *>(t);
  size_t off = (hist_n * sizeof(uint32_t) + 255) & ~size_t(255);
  uint64_t *offs = reinterpret_cast<uint64_t *>(t + off);
  off += ((hist_n + 1) * sizeof(uint64_t) + 255) & ~size_t(255);
  void *scan_tmp = t + off;
  uint32_t *ka = keys, *va = vals, *kb = tmp_keys, *vb = tmp_vals;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += kDigitBits, ++passes) {
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)tiles), dim3(kSortThreads), 0, s, ka, n,
                       shift, hist, (uint32_t)tiles);
    exclusive_scan_u32_to_u64(hist, offs, hist_n, scan_tmp, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)tiles), dim3(kSortThreads), 0, s, ka, va,
                       kb, vb, n, shift, offs, (uint32_t)tiles);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  GT_HIP(hipGetLastError());
  if (ka != keys) {
    GT_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    GT_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
}

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
  GT_HIP(hipGetLastError());
}
void fill_f64(double *p, double v, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
  GT_HIP(hipGetLastError());
}
void iota_u32(uint32_t *p, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
  GT_HIP(hipGetLastError());
}

} // namespace gt
