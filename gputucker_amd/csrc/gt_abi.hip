// gt_abi.hip -- the C ABI (include/gputucker_hip.h): device context, the
// partition planner and scheduler (host integer logic, as in the
// reference), orchestration of the partition kernels, the mode update, the
// fit, and the test hooks.
//
// HBM layout per context (one device, everything resident):
//   coo_idx[k][nnz]  u32   global indices, block order (blocks ascending,
//                          input order inside a block: tensor_manager.tpp:267-290)
//   coo_val[nnz]     f32
//   coo_blk[nnz]     u32   block id of each block-ordered nonzero
//   st_idx[n][k][nnz] u32  mode-n stream: nonzeros sorted by (global row of
//   st_val[n][nnz]   f32   mode n, block, position) -- the order in which
//                          ComputingBC accumulates a row (update.cuh:73-99)
//   row_ptr[n][rows_n+1] u64  global CSR of the stream (rows_n = P_n * bd_n)
//   where[n][nnz]    u32   per-block where_nnz of mode n, blocks concatenated
//                          (built on first request: only the test hooks read it)
//   A[n]             f32   (P_n * bd_n) x RS factor (RS = row_stride(R)),
//                          padding rows and columns zero
//   As[n]            u32   split BF16 copy of A[n], 16 words per row (BF16 paths)
//   Gt[n]            f32   mode-n unfolded core [o][jL][r]
//   gram             f64   rows x (R(R+1)/2 + R) packed Gram of the current mode
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "gt_internal.hpp"

using namespace gt;

// Scratch of gt_fit_noalloc (the error partials of the reconstruction pass
// and their sum): a file-scope device array, so that entry point allocates
// nothing.  Each device the library runs on holds its own copy until the
// process ends.  On the host `scratch` is only a stand-in for the symbol:
// gt_create looks up the device's address once per device with
// hipGetSymbolAddress, and host code uses that address, never the name.
constexpr size_t kScratchDoubles = 16384; // 128 KiB
__device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_scratch_mu;
std::vector<double *> g_scratch_at; // per device ordinal, null until looked up

// The device address of `scratch` on device `dev`, which must be current.
double *scratch_address(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_at.size() <= dev) g_scratch_at.resize(dev + 1, nullptr);
  if (!g_scratch_at[dev]) {
    void *p = nullptr;
    GT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_at[dev] = static_cast<double *>(p);
  }
  return g_scratch_at[dev];
}

thread_local std::string g_create_error;

template <typename T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void swap(DevBuf &other) {
    std::swap(p, other.p);
    std::swap(n, other.n);
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    reset();
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      throw HipError(std::string("hipMalloc(") + std::to_string(bytes) + " bytes): " +
                     hipGetErrorString(e));
    }
    n = count;
  }
};

struct GtError : std::runtime_error {
  int code;
  GtError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

constexpr uint64_t kErrPartialsPerSegment = 2048; // the COO contraction's grid cap

int ceil_log2(uint64_t v) { // bits needed to hold values < v
  int b = 0;
  while (b < 64 && (1ull << b) < v) ++b;
  return b;
}

// ------------------------------------------------ planner (optimizer.tpp)
struct Plan {
  uint32_t part[GT_MAX_ORDER], bd[GT_MAX_ORDER];
  uint64_t blocks = 1, avg_nnz = 0;
  int streams = 1, type = 1;
  uint64_t avail = 0;
};

bool refresh_block_dims(int N, const uint32_t *dims, uint64_t nnz, Plan &p) {
  p.blocks = 1;
  for (int a = 0; a < N; ++a) { // optimizer.tpp:359-376
    if (p.part[a] == 0) return false;
    uint32_t bd = (uint32_t)(((uint64_t)dims[a] + p.part[a] - 1) / p.part[a]);
    if (bd == 0) return false;
    if ((uint32_t)(((uint64_t)dims[a] + bd - 1) / bd) != p.part[a]) return false;
    p.bd[a] = bd;
    p.blocks *= p.part[a];
  }
  p.avg_nnz = (nnz + p.blocks - 1) / p.blocks;
  return true;
}

struct SizeModel { // optimizer.tpp:202-335
  int N;
  const uint32_t *dims;
  uint32_t R;
  uint64_t vs;
  uint64_t core() const {
    return (uint64_t)std::pow((double)R, (double)N) * (uint64_t)(N * 4 + vs);
  }
  uint64_t all_factors() const {
    uint64_t s = 0;
    for (int a = 0; a < N; ++a) s += (uint32_t)(dims[a] * R); // index_t product
    return s * vs;
  }
  uint64_t sub_factors(const Plan &p) const {
    uint64_t s = 0;
    for (int a = 0; a < N; ++a) s += (uint32_t)(p.bd[a] * R);
    return s * vs;
  }
  uint64_t all_data(const Plan &p) const {
    return p.avg_nnz * (uint64_t)(N * 4 + vs) + core() + sub_factors(p) + p.avg_nnz * R * vs;
  }
};

// FindPartitionParms (optimizer.tpp:70-164) for one GPU; forced: P given.
bool plan_partition(int N, const uint32_t *dims, uint64_t nnz, int rank, uint64_t mem, int vs,
                    const uint32_t *forced, Plan &p) {
  SizeModel m{N, dims, (uint32_t)rank, (uint64_t)vs};
  bool is_forced = false;
  for (int a = 0; a < N; ++a) is_forced |= forced && forced[a] != 0;
  for (int a = 0; a < N; ++a) p.part[a] = is_forced ? forced[a] : 1;
  if (!refresh_block_dims(N, dims, nnz, p)) return false;
  if (is_forced) {
    p.type = (p.blocks == 1) ? 1 : 0;
    p.streams = (p.type == 1) ? 1 : 4;
  } else {
    uint64_t required = nnz * (uint64_t)(N * 4 + vs) + m.core() + m.all_factors() +
                        nnz * (uint64_t)rank * (uint64_t)vs;
    if (required <= mem) {
      p.type = 1;
      p.streams = 1;
    } else {
      p.type = 0;
      p.streams = 4;
      const uint64_t buf = mem / 4;
      do {
        int ax = 0;
        for (int a = 1; a < N; ++a)
          if (p.bd[ax] < p.bd[a]) ax = a;
        p.part[ax]++;
        if (!refresh_block_dims(N, dims, nnz, p)) return false;
      } while (!(m.all_data(p) < buf && p.blocks >= 4));
    }
  }
  const uint64_t buf = mem / (uint64_t)p.streams;
  const uint64_t avail_buf = buf - (m.core() + m.sub_factors(p));
  p.avail = avail_buf / (uint64_t)(N * 4 + vs + rank * vs);
  return true;
}

} // namespace

// ------------------------------------------------ context
struct gt_ctx {
  gt_opts o{};
  std::string err;
  int N = 0;
  int R = 0;  // the rank the kernels run at (kernel_rank of Ru)
  int Ru = 0; // the model's rank: host arrays at the boundary have Ru columns
  hipStream_t s = nullptr;
  // input
  uint64_t nnz = 0;
  uint32_t dims[GT_MAX_ORDER] = {};
  bool have_coo = false;
  DevBuf<uint32_t> in_idx[GT_MAX_ORDER];
  DevBuf<float> in_val;
  // a device budget below the in-core partition's footprint keeps the input
  // in pinned host memory, and gt_partition builds the per-mode streams from
  // it bucket by bucket, straight into the host streams
  bool host_input = false;
  bool host_partitioned = false; // the last partition ran bucketed from host memory
  uint64_t partition_pieces = 0; // pieces it built (over all modes)
  uint32_t *h_in_idx[GT_MAX_ORDER] = {};
  float *h_in_val = nullptr;
  // Host memory of the out-of-core data: page-locked where the system grants
  // it, so the copies run asynchronously at the host link's rate; once a
  // pinned allocation fails, ordinary pageable memory, whose copies the
  // runtime stages (more slowly, and the host waits on them).  The pageable
  // blocks are recorded so each is freed the way it was allocated.
  std::unordered_map<void *, size_t> pageable, pinned;
  uint64_t pageable_bytes = 0, pinned_bytes = 0;
  template <typename T> void host_alloc(T **p, uint64_t count) {
    const size_t bytes = (size_t)std::max<uint64_t>(count, 1) * sizeof(T);
    if (o.host_pinned_limit == 0 || pinned_bytes + bytes <= o.host_pinned_limit) {
      if (hipHostMalloc(reinterpret_cast<void **>(p), bytes, hipHostMallocDefault) ==
          hipSuccess) {
        pinned[*p] = bytes;
        pinned_bytes += bytes;
        return;
      }
      (void)hipGetLastError(); // a failed pinned allocation leaves no state behind
    }
    void *q = std::malloc(bytes);
    if (q == nullptr)
      throw GtError(GT_ERR_MEMORY, "host memory exhausted allocating " + std::to_string(bytes) +
                                       " bytes for the out-of-core data");
    pageable[q] = bytes;
    pageable_bytes += bytes;
    *p = static_cast<T *>(q);
  }
  template <typename T> void host_free(T *&p) {
    if (p == nullptr) return;
    const auto it = pageable.find(static_cast<void *>(p));
    if (it != pageable.end()) {
      pageable_bytes -= it->second;
      pageable.erase(it);
      std::free(static_cast<void *>(p));
    } else {
      const auto jt = pinned.find(static_cast<void *>(p));
      if (jt != pinned.end()) {
        pinned_bytes -= jt->second;
        pinned.erase(jt);
      }
      (void)hipHostFree(static_cast<void *>(p));
    }
    p = nullptr;
  }
  // Moves a pageable block into page-locked memory where the system now
  // grants it (after the partition's input copy is freed).
  template <typename T> void pin_if_possible(T *&p) {
    if (p == nullptr) return;
    const auto it = pageable.find(static_cast<void *>(p));
    if (it == pageable.end()) return;
    const size_t bytes = it->second;
    if (o.host_pinned_limit != 0 && pinned_bytes + bytes > o.host_pinned_limit) return;
    void *q = nullptr;
    if (hipHostMalloc(&q, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    std::memcpy(q, static_cast<void *>(p), bytes);
    pageable_bytes -= bytes;
    pageable.erase(it);
    std::free(static_cast<void *>(p));
    pinned[q] = bytes;
    pinned_bytes += bytes;
    p = static_cast<T *>(q);
  }
  void pin_host_streams() {
    for (int n = 0; n < GT_MAX_ORDER; ++n) {
      for (int k = 0; k < GT_MAX_ORDER; ++k) {
        pin_if_possible(h_pack[n][k]);
        pin_if_possible(h_idx[n][k]);
      }
      pin_if_possible(h_val[n]);
    }
  }
  void release_host_input() {
    for (auto &p : h_in_idx) host_free(p);
    host_free(h_in_val);
    host_input = false;
  }
  // partition
  bool partitioned = false;
  Plan plan;
  uint64_t B = 0;
  std::vector<uint64_t> block_nnz, block_off;
  double norm = 0.0;
  float norm_ref = 0.f; // opts.ref_fit: serial float norm in input order
  double ms_partition = 0.0;
  DevBuf<uint64_t> d_block_off;
  DevBuf<uint32_t> coo_idx[GT_MAX_ORDER];
  DevBuf<float> coo_val;
  DevBuf<uint32_t> coo_blk;
  DevBuf<uint32_t> st_idx[GT_MAX_ORDER][GT_MAX_ORDER];
  DevBuf<float> st_val[GT_MAX_ORDER];
  DevBuf<uint64_t> row_ptr[GT_MAX_ORDER];
  DevBuf<uint32_t> where[GT_MAX_ORDER];
  uint64_t rows[GT_MAX_ORDER] = {};
  // shared factor rows of mode n's wide refined BF16 delta (in-core
  // partitions, group_hot_rows): the kind, the number of contracted-core
  // tables (rebuilt at each mode update), and for kHotOuter the slot of each
  // outer row and the hot rows
  int hot_kind[GT_MAX_ORDER] = {};   // kHotOuter, kHotLast or 0
  uint64_t hot_k[GT_MAX_ORDER] = {}; // tables
  DevBuf<uint32_t> hot_slot[GT_MAX_ORDER];
  DevBuf<uint32_t> hot_rows[GT_MAX_ORDER];
  DevBuf<uint16_t> hot_tab[GT_MAX_ORDER];
  // model
  bool initialized = false;
  DevBuf<float> A[GT_MAX_ORDER];
  DevBuf<uint32_t> As[GT_MAX_ORDER]; // split BF16 copies of the factors (MFMA delta)
  DevBuf<float> Gt[GT_MAX_ORDER];
  DevBuf<uint16_t> Gb[GT_MAX_ORDER]; // BF16 core operand tables (MFMA delta)
  int O = 1;                         // R^(N-2)
  // work
  DevBuf<double> gram, bpart0, bpart1, err_part, scalar;
  DevBuf<uint32_t> chunk_tail; // per chunk: row continuing into the next chunk, or ~0u
  DevBuf<double> fit_part;     // residual partials of the row solves
  DevBuf<float> xhat;          // opts.ref_fit: the reconstruction of every nonzero, block order
  float *dbg_delta = nullptr;  // test hook: the Gram kernels' deltas go here (stream order)
  uint64_t chunk = 0, nchunks = 0;
  // per-mode chunk size and delta variant picked by the autotuner (in core;
  // 0 = the default chunk, the default variant)
  uint64_t mchunk[GT_MAX_ORDER] = {};
  int mvariant[GT_MAX_ORDER] = {};
  // out-of-core pipeline: the per-mode streams live in pinned host memory and
  // are pumped through two device staging buffers on a copy stream
  bool streaming = false;
  uint64_t seg_nnz = 0;
  uint32_t *h_idx[GT_MAX_ORDER][GT_MAX_ORDER] = {};
  float *h_val[GT_MAX_ORDER] = {};
  // packed host streams (every factor has at most 2^24 rows): mode n's
  // stream keeps each other mode k's indices in pack_bits[k] bits (16 when
  // factor k has at most 2^16 rows, else 24) and drops its row index, which
  // row_ptr[n] rebuilds on the device
  bool packed = false;
  int pack_bits[GT_MAX_ORDER] = {};
  uint32_t stream_bytes = 0; // host bytes per nonzero, the largest over the modes
  uint32_t *h_pack[GT_MAX_ORDER][GT_MAX_ORDER] = {};
  DevBuf<uint32_t> stage_pack[2][GT_MAX_ORDER];
  DevBuf<uint32_t> stage_idx[2][GT_MAX_ORDER];
  DevBuf<float> stage_val[2];
  hipStream_t s_copy = nullptr;
  hipEvent_t ev_copy[2] = {}, ev_comp[2] = {};
  // timing
  double *scratch = nullptr; // this device's address of the static scratch array
  // opts.graph: one in-core ALS iteration (N mode updates and the reduction
  // of the fit partials) captured once, replayed by later iterations
  bool capturing = false;
  hipGraphExec_t it_graph = nullptr;
  uint64_t it_graph_parts = 0;
  void drop_graph() {
    if (it_graph) (void)hipGraphExecDestroy(it_graph);
    it_graph = nullptr;
  }
  hipEvent_t ev[8] = {};
  hipEvent_t ev_mode[GT_MAX_ORDER][3] = {}; // per mode: start, Gram done, solve done
  bool mode_pending[GT_MAX_ORDER] = {};     // recorded, not yet read into tm
  gt_timings tm{};
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_pending;
  std::vector<hipEvent_t> ev_pool;

  hipEvent_t get_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    GT_HIP(hipEventCreate(&e));
    return e;
  }
  void release_host_streams() {
    if (s) (void)hipStreamSynchronize(s); // queued copies may still read the buffers
    if (s_copy) (void)hipStreamSynchronize(s_copy);
    for (int n = 0; n < GT_MAX_ORDER; ++n) {
      for (int k = 0; k < GT_MAX_ORDER; ++k) {
        host_free(h_idx[n][k]);
        host_free(h_pack[n][k]);
      }
      host_free(h_val[n]);
    }
    streaming = false;
    packed = false;
  }
  // Reads the event times of mode updates whose events have completed (the
  // caller has synchronised the stream).
  void collect_mode_timings() {
    for (int n = 0; n < GT_MAX_ORDER; ++n) {
      if (!mode_pending[n]) continue;
      float g = 0.f, sv = 0.f;
      GT_HIP(hipEventElapsedTime(&g, ev_mode[n][0], ev_mode[n][1]));
      GT_HIP(hipEventElapsedTime(&sv, ev_mode[n][1], ev_mode[n][2]));
      tm.ms_gram[n] = g;
      tm.ms_solve[n] = sv;
      tm.ms_mode[n] = g + sv;
      mode_pending[n] = false;
    }
  }
  void drain_profile() {
    for (auto &pr : prof_pending) {
      GT_HIP(hipEventSynchronize(pr.second));
      float ms = 0.f;
      GT_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
      tm.gram_kernel_ms_total += ms;
      tm.gram_kernel_launches += 1;
      ev_pool.push_back(pr.first);
      ev_pool.push_back(pr.second);
    }
    prof_pending.clear();
  }
  ~gt_ctx() {
    if (s) {
      (void)hipStreamSynchronize(s);
      drop_graph();
      for (auto &pr : prof_pending) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
      }
      for (auto e : ev_pool) (void)hipEventDestroy(e);
      for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
      for (auto &em : ev_mode)
        for (auto e : em)
          if (e) (void)hipEventDestroy(e);
      release_host_streams();
      release_host_input();
      if (s_copy) {
        (void)hipStreamSynchronize(s_copy);
        (void)hipStreamDestroy(s_copy);
      }
      for (int b = 0; b < 2; ++b) {
        if (ev_copy[b]) (void)hipEventDestroy(ev_copy[b]);
        if (ev_comp[b]) (void)hipEventDestroy(ev_comp[b]);
      }
      (void)hipStreamDestroy(s);
    }
  }
};

namespace {

template <typename F> int guarded(gt_ctx *ctx, F &&f) {
  if (!ctx) return GT_ERR_ARG;
  try {
    GT_HIP(hipSetDevice(ctx->o.device));
    f();
    ctx->err.clear();
    return GT_OK;
  } catch (const GtError &e) {
    ctx->err = e.what();
    return e.code;
  } catch (const HipError &e) {
    ctx->err = e.what();
    return GT_ERR_HIP;
  } catch (const std::invalid_argument &e) {
    ctx->err = e.what();
    return GT_ERR_UNSUPPORTED;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    return GT_ERR_HIP;
  }
}

void require(bool cond, int code, const char *msg) {
  if (!cond) throw GtError(code, msg);
}

uint64_t ipow(uint64_t b, int e) {
  uint64_t v = 1;
  for (int i = 0; i < e; ++i) v *= b;
  return v;
}

// The Ru^N core (linear order, last axis fastest) embedded in the RK^N core
// of the kernel rank, zero where any coordinate is Ru or above.
std::vector<float> pad_core(const float *core, int N, int Ru, int RK) {
  const uint64_t total = ipow(Ru, N);
  std::vector<float> out(ipow(RK, N), 0.f);
  for (uint64_t i = 0; i < total; ++i) {
    uint64_t rem = i, at = 0, mult = 1;
    for (int a = N - 1; a >= 0; --a) {
      at += (rem % Ru) * mult;
      rem /= Ru;
      mult *= RK;
    }
    out[at] = core[i];
  }
  return out;
}

// Core unfolding for mode n in the [o][jL][r] order the kernels read.
std::vector<float> unfold_core(const float *core, int N, int R, int n) {
  const uint64_t total = ipow(R, N);
  std::vector<float> gt(total);
  const int L = (n == N - 1) ? N - 2 : N - 1;
  int d[GT_MAX_ORDER];
  for (uint64_t i = 0; i < total; ++i) {
    uint64_t rem = i;
    for (int a = N - 1; a >= 0; --a) { // tucker.cuh:120-132, last axis fastest
      d[a] = (int)(rem % R);
      rem /= R;
    }
    uint64_t o = 0;
    for (int a = 0; a < N; ++a)
      if (a != n && a != L) o = o * R + d[a];
    gt[(o * R + d[L]) * R + d[n]] = core[i];
  }
  return gt;
}

uint16_t to_bf16(float x) { // round to nearest even
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float from_bf16(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

// BF16 core operand tables of the MFMA delta (layout and plans in
// gt_internal.hpp).  Entry (o, lane, t) of a table is the core operand word
// of K slot t for output r = lane % 16 of outer index o; G[o][jL][r] is zero
// outside jL, r < R.
std::vector<uint16_t> bf16_core_table(const std::vector<float> &gt, int N, int R) {
  const int O = (int)ipow(R, N > 2 ? N - 2 : 0);
  const size_t n16 = gb_k16_words(O), n32 = gb_k32_words(O);
  const size_t npair = bf16r_pair_table(R) ? gb_k32_pair_words(O) : 0;
  const size_t nob = bf16r_ob_table(R, N, O) ? gb_ob_words(R, O) : 0;
  const size_t nkr = bf16r_kr_table(R, N) ? gb_kr_words(R, N) : 0;
  std::vector<uint16_t> tab(n16 + 2 * n32 + npair + nob + nkr, 0);
  auto g_at = [&](int o, int jL, int r) -> float {
    return (o < O && jL < R && r < R) ? gt[((size_t)o * R + jL) * R + r] : 0.f;
  };
  auto hi = [&](float g) { return to_bf16(g); };
  auto lo = [&](float g) { return to_bf16(g - from_bf16(to_bf16(g))); };
  // the K32 refined words of lane group g for output r of outer index o
  auto k32_words = [&](uint16_t *k32, int o, int g, int r) {
    switch (bf16r_plan(R, g)) {
    case kPlanHL:
      for (int t = 0; t < 8; ++t) k32[t] = hi(g_at(o, 4 * g + (t & 3), r));
      break;
    case kPlanHH:
      for (int t = 0; t < 8; ++t) k32[t] = lo(g_at(o, t, r));
      break;
    case kPlanR10:
      for (int t = 0; t < 4; ++t) k32[t] = hi(g_at(o, 8 + t, r));
      k32[4] = hi(g_at(o, 8, r));
      k32[5] = hi(g_at(o, 9, r));
      k32[6] = lo(g_at(o, 8, r));
      k32[7] = lo(g_at(o, 9, r));
      break;
    default:
      break;
    }
  };
  for (int o = 0; o < O; ++o)
    for (int l = 0; l < kWave; ++l) {
      const int g = l / 16, r = l % 16;
      for (int t = 0; t < 4; ++t) {
        const size_t at = ((size_t)o * kWave + l) * 4 + t;
        tab[at] = hi(g_at(o, 4 * g + t, r));                 // K16 hi
        tab[n16 + n32 + ((size_t)o * kWave + l) * 8 + t] = lo(g_at(o, 4 * g + t, r)); // K32 lo
      }
      k32_words(tab.data() + n16 + ((size_t)o * kWave + l) * 8, o, g, r);
    }
  // K32 pair table (R <= 8): entry (p, lane) holds output row l % 16 =
  // 8 * (o - 2p) + r of the outer pair p, so one MFMA covers two outer indices
  for (size_t p = 0; npair && p < (size_t)(O + 1) / 2; ++p)
    for (int l = 0; l < kWave; ++l) {
      const int g = l / 16, q = l % 16;
      k32_words(tab.data() + n16 + 2 * n32 + (p * kWave + l) * 8, (int)(2 * p) + q / 8, g, q % 8);
    }
  // o-block table (orders 4-5, O <= kObMaxO): entry (ob, r, lane) holds
  // output r of outer index 16 ob + l % 16
  const size_t ob_at = gb_ob_offset(R, O);
  for (int ob = 0; nob && ob < gb_oblocks(O); ++ob)
    for (int r = 0; r < R; ++r)
      for (int l = 0; l < kWave; ++l)
        k32_words(tab.data() + ob_at + (((size_t)ob * R + r) * kWave + l) * 8, 16 * ob + l % 16,
                  l / 16, r);
  // Kronecker table (order 4, R = 4 or 8; layout in gt_internal.hpp): entry
  // (kb, mu, lane (q, g)) holds, for output row rho = 16 kb + q = (o0, r), the
  // core words of the K slots of lane group g: G[o0][o1][jL][r] for the
  // entries e = o1 * R + jL the group takes in MFMA mu
  const size_t kr_at = gb_kr_offset(R, N, O);
  const int nhl = kr_nhl(R), nmf = kr_nhl(R) + kr_nhh(R);
  for (int kb = 0; nkr && kb < kr_nkb(R, N); ++kb)
    for (int mu = 0; mu < nmf; ++mu)
      for (int l = 0; l < kWave; ++l) {
        const int q = l % 16, g = l / 16, rho = 16 * kb + q;
        // output row: (o0, r) at order 4; (o01, r) with r padded to 8 at order 5
        const int oh = N == 4 ? rho / R : rho / 8, r = N == 4 ? rho % R : rho % 8;
        const bool row_ok = r < R && oh < (N == 4 ? R : R * R);
        auto g_e = [&](int e) { // G[oh][o_last][jL][r] of entry e, zero past R^2
          return (row_ok && e < R * R) ? gt[(((size_t)oh * R + e / R) * R + e % R) * R + r]
                                       : 0.f;
        };
        uint16_t *w = tab.data() + kr_at + (((size_t)kb * nmf + mu) * kWave + l) * 8;
        if (mu < nhl) { // k_lo.G_hi + k_hi.G_hi over entries 16 mu + 4g + i
          for (int t = 0; t < 8; ++t) w[t] = hi(g_e(16 * mu + 4 * g + (t & 3)));
        } else { // k_hi.G_lo over the entries of MFMAs 2 nu and 2 nu + 1
          const int nu = mu - nhl;
          for (int t = 0; t < 8; ++t) {
            const int m = 2 * nu + t / 4;
            w[t] = m < nhl ? lo(g_e(16 * m + 4 * g + (t & 3))) : 0;
          }
        }
      }
  return tab;
}

// The core operand table of the wide refined BF16 delta (order 3): entry
// (o, mt, lane (q, g)) holds the BF16 hi words of G[o][8g + t][16 mt + q],
// t < 8, then the lo words bf16(G - hi), zero at jL or r >= R, for the
// wide_tiles_per_o(R) tiles formed per outer index.  With the last tile
// packed (PO = wide_pack(R) > 1 outer indices per MFMA), entry (ob, lane
// (q, g)) then holds the words of G[PO ob + q / (16 / PO)][8g + t][16 (MT - 1)
// + q % (16 / PO)], zero past R.
std::vector<uint16_t> bf16_core_table_wide(const std::vector<float> &gt, int R) {
  const int MT = wide_tiles(R), MTO = wide_tiles_per_o(R), PO = wide_pack(R);
  std::vector<uint16_t> tab(gb_wide_words(R), 0);
  auto put = [&](uint16_t *dst, int o, int g, int r) {
    for (int t = 0; t < 8; ++t) {
      const int jL = 8 * g + t;
      const float v = (o < R && jL < R && r < R) ? gt[((size_t)o * R + jL) * R + r] : 0.f;
      const uint16_t h = to_bf16(v);
      dst[t] = h;
      dst[8 + t] = to_bf16(v - from_bf16(h));
    }
  };
  for (int o = 0; o < R; ++o)
    for (int mt = 0; mt < MTO; ++mt)
      for (int l = 0; l < kWave; ++l)
        put(tab.data() + (((size_t)o * MTO + mt) * kWave + l) * 16, o, l / 16, 16 * mt + l % 16);
  if (PO > 1) {
    const size_t base = (size_t)R * MTO * kWave * 16;
    const int RPr = 16 / PO;
    for (int ob = 0; ob < wide_packed_blocks(R); ++ob)
      for (int l = 0; l < kWave; ++l) {
        const int q = l % 16;
        put(tab.data() + base + ((size_t)ob * kWave + l) * 16, PO * ob + q / RPr, l / 16,
            16 * (MT - 1) + q % RPr);
      }
  }
  return tab;
}

// Bounds of the default chunk size, in 64-nonzero tiles.  Past about 16384
// chunks the partial tables and the fix-up get cheaper with longer chunks,
// up to 32 tiles; at 64 the order-3 kernel slows (too few waves).  At least
// two: the FP32 delta forms two tiles per pass (order 3 and R^N <= 4096),
// and on a 1e6-nonzero tensor one-tile chunks cost it 30 % (the refined
// BF16 delta 4 %).
constexpr uint64_t kMinChunkTiles = 2, kMaxChunkTiles = 32;

void ensure_work(gt_ctx *c) {
  uint64_t max_rows = 0;
  for (int n = 0; n < c->N; ++n) max_rows = std::max(max_rows, c->rows[n]);
  const int W = gram_width(c->R);
  c->gram.alloc(max_rows * W);
  uint64_t tiles = (uint64_t)c->o.chunk_tiles;
  if (tiles == 0) { // enough chunks to give every SIMD several waves
    const uint64_t target_waves = 16384;
    tiles = c->nnz / (kTile * target_waves);
    tiles = std::min<uint64_t>(std::max<uint64_t>(tiles, kMinChunkTiles), kMaxChunkTiles);
  }
  c->chunk = tiles * kTile;
  c->nchunks = (c->nnz + c->chunk - 1) / c->chunk;
  c->bpart0.alloc(std::max<uint64_t>(c->nchunks, 1) * W);
  c->bpart1.alloc(std::max<uint64_t>(c->nchunks, 1) * W);
  c->chunk_tail.alloc(std::max<uint64_t>(c->nchunks, 1));
  uint64_t parts = 1;
  for (int n = 0; n < c->N; ++n) parts = std::max(parts, solve_partials(c->R, c->rows[n]));
  c->fit_part.alloc(parts);
  c->err_part.alloc(4096);
  c->scalar.alloc(16);
}

// With a device budget smaller than the resident layout, the per-mode streams
// move to pinned host memory after partitioning and each mode update pumps
// them through two staging buffers (the reference's async block pipeline,
// delta.cuh:165-213, with the tasks of scheduler.tpp:85-159 replaced by
// fixed-size segments of the row-ordered stream).  The partition itself runs
// with the tensor resident.
// Words holding n packed indices of `bits` bits (n a multiple of four when
// it is an offset).
uint64_t pack_words(int bits, uint64_t n) { return bits == 16 ? (n + 1) / 2 : 3 * ((n + 3) / 4); }

// Device bytes the ALS keeps whatever the layout: Gram tables, chunk
// partials, factors and their split copies, row pointers, cores, and 1 MiB
// of slack for small buffers.
uint64_t resident_fixed_bytes(const gt_ctx *c) {
  const int N = c->N, R = c->R;
  uint64_t fixed = (c->gram.n + c->bpart0.n + c->bpart1.n) * sizeof(double) + c->chunk_tail.n * 4;
  const uint64_t split = c->o.precision == GT_PREC_FP32 ? 0
                         : (R <= 15 ? kSplitWords : kSplitWordsWide);
  for (int n = 0; n < N; ++n)
    fixed += c->rows[n] * ((uint64_t)row_stride(R) + split) * 4 + (c->rows[n] + 1) * 8;
  fixed += (uint64_t)N * ipow(R, N) * 4 * 2 + (1u << 20);
  return fixed;
}

// Whether the host streams hold packed indices (every factor at most 2^24
// rows; GT_STREAM_PACK=0 turns it off), and the bits of each mode's indices.
bool choose_packing(gt_ctx *c) {
  const int N = c->N;
  const char *pack_env = std::getenv("GT_STREAM_PACK");
  bool packed = !(pack_env && pack_env[0] == '0');
  for (int k = 0; k < N; ++k) packed = packed && c->rows[k] <= (1ull << 24);
  for (int k = 0; k < N; ++k) c->pack_bits[k] = !packed ? 0 : c->rows[k] <= (1u << 16) ? 16 : 24;
  c->packed = packed;
  c->stream_bytes = 4u * (N + 1);
  if (packed) {
    uint32_t most = 0;
    for (int n = 0; n < N; ++n) {
      uint32_t b = 4;
      for (int k = 0; k < N; ++k)
        if (k != n) b += (uint32_t)c->pack_bits[k] / 8;
      most = std::max(most, b);
    }
    c->stream_bytes = most;
  }
  return packed;
}

// The ALS side of the out-of-core pipeline once the host streams exist: the
// staged segment size (a multiple of the chunk) from the device bytes left,
// two staging buffers, the copy stream and its events.
void setup_staging(gt_ctx *c, uint64_t fixed) {
  const int N = c->N;
  const uint64_t budget = c->o.device_budget, nnz = c->nnz;
  // per nonzero and staging buffer: the unpacked indices and the value, plus
  // the packed words when packing
  const uint64_t stage_bytes = 4ull * (N + 1) + (c->packed ? 3ull * N : 0);
  uint64_t seg = (budget - fixed) / (2ull * stage_bytes);
  if (c->plan.avail) seg = std::min(seg, c->plan.avail);
  // at least four segments, so the copies of later ones run under the
  // kernels of earlier ones (one segment would serialise copy and kernel)
  seg = std::min(seg, std::max<uint64_t>((nnz + 3) / 4, c->chunk));
  seg = (seg / c->chunk) * c->chunk;
  if (seg == 0) throw GtError(GT_ERR_MEMORY, "device budget too small for one staged chunk");
  for (int b = 0; b < 2; ++b) {
    for (int k = 0; k < N; ++k) {
      c->stage_idx[b][k].alloc(seg);
      if (c->packed) c->stage_pack[b][k].alloc(3 * (seg / 4));
    }
    c->stage_val[b].alloc(seg);
    if (!c->ev_copy[b]) GT_HIP(hipEventCreateWithFlags(&c->ev_copy[b], hipEventDisableTiming));
    if (!c->ev_comp[b]) GT_HIP(hipEventCreateWithFlags(&c->ev_comp[b], hipEventDisableTiming));
  }
  if (!c->s_copy) GT_HIP(hipStreamCreateWithFlags(&c->s_copy, hipStreamNonBlocking));
  const uint64_t nseg = (nnz + seg - 1) / seg;
  c->err_part.alloc(std::max<uint64_t>(4096, nseg * kErrPartialsPerSegment));
  c->seg_nnz = seg;
  c->streaming = true;
}

// Device bytes per nonzero at the peak of the in-core partition: the input
// copy, the sort's keys, values and ping-pong buffers, the block-ordered COO
// with its block ids, and every mode's stream.
uint64_t incore_partition_bytes_per_nnz(int N) {
  return (4ull * N + 4) + 16 + (4ull * N + 8) + (uint64_t)N * (4ull * N + 4);
}

// resident_fixed_bytes before the partition, from the mode sizes (the rows
// of a partition exceed them by less than P_n) and the default chunk size.
uint64_t fixed_estimate(const gt_ctx *c, uint64_t nnz, const uint32_t *dims) {
  const int N = c->N, R = c->R, W = gram_width(R);
  uint64_t tiles = (uint64_t)c->o.chunk_tiles;
  if (tiles == 0)
    tiles = std::min<uint64_t>(std::max<uint64_t>(nnz / (kTile * 16384), kMinChunkTiles),
                               kMaxChunkTiles);
  const uint64_t nchunks = (nnz + tiles * kTile - 1) / (tiles * kTile) + 1;
  const uint64_t split = c->o.precision == GT_PREC_FP32 ? 0
                         : (R <= 15 ? kSplitWords : kSplitWordsWide);
  uint64_t max_rows = 1, fixed = 0;
  for (int n = 0; n < N; ++n) {
    max_rows = std::max<uint64_t>(max_rows, dims[n]);
    fixed += (uint64_t)dims[n] * ((uint64_t)row_stride(R) + split) * 4 + ((uint64_t)dims[n] + 1) * 8;
  }
  fixed += max_rows * W * 8 + nchunks * (2ull * W * 8 + 4);
  fixed += (uint64_t)N * ipow(R, N) * 4 * 2 + (1u << 20);
  return fixed;
}

void maybe_stream(gt_ctx *c) {
  const uint64_t budget = c->o.device_budget;
  if (budget == 0) return;
  const int N = c->N;
  const uint64_t nnz = c->nnz;
  const uint64_t movable = nnz * 4ull * (uint64_t)(N * (N + 1) + (N + 2) + N);
  const uint64_t fixed = resident_fixed_bytes(c);
  if (movable + fixed <= budget) return; // fits: stay resident
  if (fixed >= budget) throw GtError(GT_ERR_MEMORY, "device budget is below the per-row state");
  const bool packed = choose_packing(c);
  DevBuf<uint32_t> tmp;
  if (packed) tmp.alloc(3 * ((nnz + 3) / 4));
  for (int n = 0; n < N; ++n) {
    for (int k = 0; k < N; ++k) {
      if (packed) {
        if (k != n) {
          const uint64_t words = pack_words(c->pack_bits[k], nnz);
          if (c->pack_bits[k] == 16)
            pack16(c->st_idx[n][k].p, nnz, tmp.p, c->s);
          else
            pack24(c->st_idx[n][k].p, nnz, tmp.p, c->s);
          c->host_alloc(&c->h_pack[n][k], words);
          GT_HIP(hipMemcpyAsync(c->h_pack[n][k], tmp.p, words * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, c->s));
          GT_HIP(hipStreamSynchronize(c->s));
        }
      } else {
        c->host_alloc(&c->h_idx[n][k], nnz);
        GT_HIP(hipMemcpy(c->h_idx[n][k], c->st_idx[n][k].p, nnz * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
      }
      c->st_idx[n][k].reset();
    }
    c->host_alloc(&c->h_val[n], nnz);
    GT_HIP(hipMemcpy(c->h_val[n], c->st_val[n].p, nnz * sizeof(float), hipMemcpyDeviceToHost));
    c->st_val[n].reset();
    c->where[n].reset();
  }
  for (int k = 0; k < N; ++k) c->coo_idx[k].reset();
  c->coo_val.reset();
  c->coo_blk.reset();
  tmp.reset();
  setup_staging(c, fixed);
}

// The refined BF16 delta of the wide ranks (order 3, 16 <= R <= 32) runs.
bool wide_bf16(const gt_ctx *c) {
  return c->o.precision == GT_PREC_BF16R && gram_wide_supported(c->R, c->N);
}

GramArgs gram_args(gt_ctx *c, int n) {
  GramArgs a{};
  a.order = c->N;
  a.mode = n;
  for (int k = 0; k < c->N; ++k) {
    a.st.idx[k] = c->st_idx[n][k].p;
    a.A[k] = c->A[k].p;
  }
  a.st.val = c->st_val[n].p;
  a.st.nnz = c->nnz;
  a.Gt = c->Gt[n].p;
  a.row_ptr = c->row_ptr[n].p;
  a.gram = c->gram.p;
  a.bpart0 = c->bpart0.p;
  a.bpart1 = c->bpart1.p;
  a.chunk_tail = c->chunk_tail.p;
  a.chunk = (c->mchunk[n] && !c->streaming) ? c->mchunk[n] : c->chunk;
  a.nchunks = (c->nnz + a.chunk - 1) / a.chunk;
  a.chunk_begin = 0;
  a.chunk_end = a.nchunks;
  a.alt_source = c->streaming ? 0 : c->mvariant[n];
  a.rows = c->rows[n];
  a.Gb = c->Gb[n].p;
  // slot order: L = last mode other than n, outer slots = the rest ascending
  const int L = (n == c->N - 1) ? c->N - 2 : c->N - 1;
  a.slots = c->N - 2;
  a.idx_row = c->st_idx[n][n].p;
  a.idx_L = c->st_idx[n][L].p;
  a.A_L = c->A[L].p;
  a.As_L = c->As[L].p; // null unless a BF16 precision is selected
  a.delta_out = c->dbg_delta;
  int t = 0;
  for (int k = 0; k < c->N; ++k) {
    if (k == n || k == L) continue;
    a.idx_O[t] = c->st_idx[n][k].p;
    a.A_O[t] = c->A[k].p;
    if (t == 0) a.As_O = c->As[k].p;
    ++t;
  }
  if (c->hot_k[n] > 0 && wide_bf16(c)) {
    a.hot_kind = c->hot_kind[n];
    a.hot_slot = c->hot_kind[n] == kHotOuter ? c->hot_slot[n].p : nullptr;
    a.hot_tab = c->hot_tab[n].p;
  }
  return a;
}

// Points the stream operands of `a` (mode n) at idx[k] / val, indexed by
// global stream position.
void set_stream_ptrs(GramArgs &a, const gt_ctx *c, int n, uint32_t *const *idx,
                     const float *val) {
  const int L = (n == c->N - 1) ? c->N - 2 : c->N - 1;
  for (int k = 0; k < c->N; ++k) a.st.idx[k] = idx[k];
  a.st.val = val;
  a.idx_row = idx[n];
  a.idx_L = idx[L];
  int t = 0;
  for (int k = 0; k < c->N; ++k)
    if (k != n && k != L) a.idx_O[t++] = idx[k];
}

void launch_gram_window(gt_ctx *c, const GramArgs &a) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool prof = c->o.profile && !c->capturing;
  if (prof) {
    e0 = c->get_event();
    e1 = c->get_event();
    GT_HIP(hipEventRecord(e0, c->s));
  }
  int path;
  if (wide_bf16(c)) {
    launch_gram_bf16_wide(c->R, a, c->s);
    path = GT_DELTA_BF16R_WIDE;
  } else if (gram_mfma_supported(c->R)) {
    path = launch_gram_mfma(c->R, c->o.precision, a, c->O, c->s);
  } else {
    launch_gram_fp32(c->R, a, c->s);
    path = GT_DELTA_FP32_WIDE;
  }
  c->tm.delta_path[a.mode] = path;
  if (prof) {
    GT_HIP(hipEventRecord(e1, c->s));
    c->prof_pending.emplace_back(e0, e1);
  }
}

// Out-of-core: copies segment k of mode n's stream into staging buffer k % 2
// on the copy stream; the compute stream waits for it.  Returns [lo, hi).
std::pair<uint64_t, uint64_t> stage_segment(gt_ctx *c, int n, uint64_t k) {
  const int b = (int)(k & 1);
  const uint64_t lo = k * c->seg_nnz, hi = std::min(lo + c->seg_nnz, c->nnz), len = hi - lo;
  // the last kernel that read buffer b (this pass or, since mode updates are
  // queued without a host wait, the previous one) must be done with it
  GT_HIP(hipStreamWaitEvent(c->s_copy, c->ev_comp[b], 0));
  if (c->packed) { // lo is a multiple of the chunk size, so of four
    for (int m = 0; m < c->N; ++m) {
      if (m == n) continue;
      const int bits = c->pack_bits[m];
      GT_HIP(hipMemcpyAsync(c->stage_pack[b][m].p, c->h_pack[n][m] + pack_words(bits, lo),
                            pack_words(bits, len) * sizeof(uint32_t), hipMemcpyHostToDevice,
                            c->s_copy));
      if (bits == 16)
        unpack16(c->stage_pack[b][m].p, len, c->stage_idx[b][m].p, c->s_copy);
      else
        unpack24(c->stage_pack[b][m].p, len, c->stage_idx[b][m].p, c->s_copy);
    }
    rows_from_ptr(c->row_ptr[n].p, (uint32_t)c->rows[n], lo, len, c->stage_idx[b][n].p,
                  c->s_copy);
  } else {
    for (int m = 0; m < c->N; ++m)
      GT_HIP(hipMemcpyAsync(c->stage_idx[b][m].p, c->h_idx[n][m] + lo, len * sizeof(uint32_t),
                            hipMemcpyHostToDevice, c->s_copy));
  }
  GT_HIP(hipMemcpyAsync(c->stage_val[b].p, c->h_val[n] + lo, len * sizeof(float),
                        hipMemcpyHostToDevice, c->s_copy));
  GT_HIP(hipEventRecord(c->ev_copy[b], c->s_copy));
  GT_HIP(hipStreamWaitEvent(c->s, c->ev_copy[b], 0));
  return {lo, hi};
}

void run_gram(gt_ctx *c, int n) {
  GramArgs a = gram_args(c, n);
  if (a.hot_tab != nullptr) { // the factors changed since the last mode update
    const bool outer = a.hot_kind == kHotOuter;
    launch_hot_tables(c->R, a.hot_kind, outer ? a.A_O[0] : a.A_L,
                      outer ? c->hot_rows[n].p : nullptr, c->hot_k[n], a.Gt, c->hot_tab[n].p,
                      c->s);
  }
  if (!c->streaming) {
    launch_gram_window(c, a);
  } else {
    const uint64_t nseg = (c->nnz + c->seg_nnz - 1) / c->seg_nnz;
    for (uint64_t k = 0; k < nseg; ++k) {
      const int b = (int)(k & 1);
      const auto [lo, hi] = stage_segment(c, n, k);
      uint32_t *ptrs[GT_MAX_ORDER];
      for (int m = 0; m < c->N; ++m) ptrs[m] = c->stage_idx[b][m].p - lo; // global positions
      set_stream_ptrs(a, c, n, ptrs, c->stage_val[b].p - lo);
      a.chunk_begin = lo / c->chunk;
      a.chunk_end = (hi + c->chunk - 1) / c->chunk;
      launch_gram_window(c, a);
      GT_HIP(hipEventRecord(c->ev_comp[b], c->s));
    }
  }
  launch_gram_fixup(c->R, a, c->s);
}

// The project's autotuner (in core): for each mode, times the delta + Gram
// pass (with its fix-up) at chunk sizes of 1 to 32 tiles and, where both
// sources of the core operand apply (refined BF16: a table in LDS or L2),
// with each; keeps the fastest.  The
// passes only overwrite the Gram scratch, not the factors.  The chunking
// changes the order of the FP64 sums inside a row, so results move at FP64
// rounding.
void autotune(gt_ctx *c) {
  require(c->partitioned && c->initialized, GT_ERR_STATE, "partition and init first");
  if (c->streaming) return; // the staged segments fix the chunk size
  c->drop_graph();
  const int W = gram_width(c->R);
  const uint64_t max_chunks = (c->nnz + kTile - 1) / kTile;
  c->bpart0.alloc(std::max<uint64_t>(max_chunks, 1) * W);
  c->bpart1.alloc(std::max<uint64_t>(max_chunks, 1) * W);
  c->chunk_tail.alloc(std::max<uint64_t>(max_chunks, 1));
  const bool prof = c->o.profile != 0;
  c->o.profile = 0;
  hipEvent_t e0 = c->get_event(), e1 = c->get_event();
  for (int n = 0; n < c->N; ++n) {
    double best = 1e30;
    uint64_t best_chunk = c->chunk;
    int best_var = 0;
    for (int var = 0; var < 2; ++var) {
      if (var == 1) { // the other core-operand source: only where it is another kernel
        c->mvariant[n] = 1;
        GramArgs probe = gram_args(c, n);
        probe.chunk_end = 0; // pick the kernel, launch nothing
        const int p1 = gram_mfma_supported(c->R) ? launch_gram_mfma(c->R, c->o.precision, probe,
                                                                    c->O, c->s) : 0;
        c->mvariant[n] = 0;
        probe = gram_args(c, n);
        probe.chunk_end = 0;
        const int p0 = gram_mfma_supported(c->R) ? launch_gram_mfma(c->R, c->o.precision, probe,
                                                                    c->O, c->s) : 0;
        if (p1 == p0) continue;
      }
      for (uint64_t tiles = 1; tiles <= 32; tiles *= 2) {
        const uint64_t chunk = tiles * kTile;
        if (tiles > 1 && (c->nnz + chunk - 1) / chunk < 256) break; // too few waves
        c->mchunk[n] = chunk;
        c->mvariant[n] = var;
        float ms_best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) { // the first run warms the caches
          GT_HIP(hipEventRecord(e0, c->s));
          run_gram(c, n);
          GT_HIP(hipEventRecord(e1, c->s));
          GT_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          GT_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (rep > 0 || ms > 20.f) ms_best = std::min(ms_best, ms);
          if (ms > 20.f) break; // a long pass is its own warm-up: time it once
        }
        if (ms_best < best) {
          best = ms_best;
          best_chunk = chunk;
          best_var = var;
        }
      }
    }
    c->mchunk[n] = best_chunk;
    c->mvariant[n] = best_var;
    c->tm.tuned_chunk_tiles[n] = (int)(best_chunk / kTile);
    c->tm.tuned_variant[n] = best_var;
    c->tm.tuned_ms[n] = best;
  }
  c->ev_pool.push_back(e0);
  c->ev_pool.push_back(e1);
  c->o.profile = prof;
}

// Refreshes the split BF16 copy of factor n that the MFMA delta gathers
// (BF16 precisions only).
void refresh_split(gt_ctx *c, int n) {
  if (wide_bf16(c)) {
    c->As[n].alloc(c->rows[n] * kSplitWordsWide);
    launch_split_wide(c->R, c->A[n].p, c->rows[n], c->As[n].p, c->s);
    return;
  }
  if (c->o.precision == GT_PREC_FP32 || !gram_mfma_supported(c->R)) return;
  c->As[n].alloc(c->rows[n] * kSplitWords);
  launch_split_factor(c->R, c->A[n].p, c->rows[n], c->As[n].p, c->s);
}

// One mode update, queued on the context's stream (sync: wait for it and
// read its timings).  With fit_partials, the row solves also leave the
// residual partials of the updated factor (valid as a fit only for the last
// mode of an iteration); returns how many were written.
uint64_t mode_update(gt_ctx *c, int n, bool fit_partials, bool sync) {
  require(c->partitioned && c->initialized, GT_ERR_STATE, "partition and init first");
  require(n >= 0 && n < c->N, GT_ERR_ARG, "mode out of range");
  if (c->prof_pending.size() > 64) c->drain_profile();
  const bool timed = !c->capturing; // no event nodes in a captured iteration
  if (timed) GT_HIP(hipEventRecord(c->ev_mode[n][0], c->s));
  run_gram(c, n);
  if (timed) GT_HIP(hipEventRecord(c->ev_mode[n][1], c->s));
  SolveArgs sa{};
  sa.gram = c->gram.p;
  sa.row_ptr = c->row_ptr[n].p;
  sa.rows = c->rows[n];
  sa.lambda = c->o.lambda;
  sa.A = c->A[n].p;
  sa.fit_partials = fit_partials ? c->fit_part.p : nullptr;
  // the solve refreshes the split copy of the narrow BF16 paths; the wide
  // one is rebuilt from the new factor after it
  sa.As = wide_bf16(c) ? nullptr : c->As[n].p;
  const uint64_t parts = launch_solve(c->R, sa, c->s);
  if (wide_bf16(c)) launch_split_wide(c->R, c->A[n].p, c->rows[n], c->As[n].p, c->s);
  if (!timed) return parts;
  GT_HIP(hipEventRecord(c->ev_mode[n][2], c->s));
  c->mode_pending[n] = true;
  if (sync) {
    GT_HIP(hipEventSynchronize(c->ev_mode[n][2]));
    c->collect_mode_timings();
  }
  return parts;
}

// I_n x Ru host rows into the padded device layout (one contiguous copy);
// columns Ru..RS-1 stay zero.
void upload_factor(gt_ctx *c, int n, const float *src) {
  const int Ru = c->Ru, RS = row_stride(c->R);
  std::vector<float> pad((size_t)c->dims[n] * RS, 0.f);
  for (uint64_t i = 0; i < c->dims[n]; ++i) std::memcpy(pad.data() + i * RS, src + i * Ru, Ru * 4);
  GT_HIP(hipMemcpyAsync(c->A[n].p, pad.data(), pad.size() * sizeof(float), hipMemcpyHostToDevice,
                        c->s));
  refresh_split(c, n);
  GT_HIP(hipStreamSynchronize(c->s));
}

CooArgs coo_args(gt_ctx *c, int mode) {
  CooArgs a{};
  const int L = (mode == c->N - 1) ? c->N - 2 : c->N - 1;
  a.slots = c->N - 2;
  a.nnz = c->nnz;
  a.idx_n = c->coo_idx[mode].p;
  a.idx_L = c->coo_idx[L].p;
  a.A_n = c->A[mode].p;
  a.A_L = c->A[L].p;
  int t = 0;
  for (int k = 0; k < c->N; ++k) {
    if (k == mode || k == L) continue;
    a.idx_O[t] = c->coo_idx[k].p;
    a.A_O[t] = c->A[k].p;
    ++t;
  }
  a.val = c->coo_val.p;
  a.Gt = c->Gt[mode].p;
  return a;
}

// nnz x R deltas of the kernel rank (device) to nnz x Ru on the host.
void copy_delta_out(gt_ctx *c, const float *d, float *out) {
  const uint64_t nnz = c->nnz;
  if (c->Ru == c->R) {
    GT_HIP(hipMemcpyAsync(out, d, nnz * c->R * sizeof(float), hipMemcpyDeviceToHost, c->s));
    GT_HIP(hipStreamSynchronize(c->s));
    return;
  }
  std::vector<float> h(nnz * c->R);
  GT_HIP(hipMemcpyAsync(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->s));
  GT_HIP(hipStreamSynchronize(c->s));
  for (uint64_t e = 0; e < nnz; ++e)
    std::memcpy(out + e * c->Ru, h.data() + e * c->R, c->Ru * sizeof(float));
}

// FP64 partials the reconstruction pass of compute_fit writes: one grid's
// worth in core, one per staged segment out of core.
uint64_t fit_partials_needed(const gt_ctx *c) {
  if (!c->streaming) return kErrPartialsPerSegment;
  return ((c->nnz + c->seg_nnz - 1) / c->seg_nnz) * kErrPartialsPerSegment;
}

// The reconstruction pass (reconstruction.cuh:58-81, 259-276): the error
// partials go to `part` (fit_partials_needed(c) doubles) and their sum to
// `scal` (one double), both device memory.
double compute_fit(gt_ctx *c, double *part, double *scal) {
  require(c->partitioned && c->initialized, GT_ERR_STATE, "partition and init first");
  GT_HIP(hipEventRecord(c->ev[3], c->s));
  if (!c->streaming) {
    CooArgs a = coo_args(c, c->N - 1);
    a.err_partials = part;
    int g = launch_coo_contract(c->R, a, c->s);
    reduce_f64(part, g, scal, c->s);
  } else { // any order of the nonzeros gives the same sum: stream mode N-1's copy
    const int n = c->N - 1, L = c->N - 2;
    const uint64_t nseg = (c->nnz + c->seg_nnz - 1) / c->seg_nnz;
    fill_f64(part, 0.0, nseg * kErrPartialsPerSegment, c->s);
    for (uint64_t k = 0; k < nseg; ++k) {
      const int b = (int)(k & 1);
      const auto [lo, hi] = stage_segment(c, n, k);
      CooArgs a{};
      a.slots = c->N - 2;
      a.nnz = hi - lo;
      a.idx_n = c->stage_idx[b][n].p;
      a.idx_L = c->stage_idx[b][L].p;
      a.A_n = c->A[n].p;
      a.A_L = c->A[L].p;
      for (int t = 0; t < c->N - 2; ++t) {
        a.idx_O[t] = c->stage_idx[b][t].p; // outer slots of mode N-1: modes 0..N-3
        a.A_O[t] = c->A[t].p;
      }
      a.val = c->stage_val[b].p;
      a.Gt = c->Gt[n].p;
      a.err_partials = part + k * kErrPartialsPerSegment;
      launch_coo_contract(c->R, a, c->s);
      GT_HIP(hipEventRecord(c->ev_comp[b], c->s));
    }
    reduce_f64(part, (int)(nseg * kErrPartialsPerSegment), scal, c->s);
  }
  double err = 0.0;
  GT_HIP(hipMemcpyAsync(&err, scal, sizeof(double), hipMemcpyDeviceToHost, c->s));
  GT_HIP(hipEventRecord(c->ev[4], c->s));
  GT_HIP(hipStreamSynchronize(c->s));
  float ms = 0.f;
  GT_HIP(hipEventElapsedTime(&ms, c->ev[3], c->ev[4]));
  c->tm.ms_fit = ms;
  if (c->norm == 0.0) return 1.0; // reconstruction.cuh:272-273
  return 1.0 - std::sqrt(err) / c->norm;
}

// The fit in the reference's value_t arithmetic (reconstruction.cuh:259-276):
// xhat of every nonzero in block order from the reconstruction pass, then on
// the host a serial float sum of (x - xhat)^2 over blocks and nonzeros in
// order, and fit = 1.0f - sqrtf(err) / norm in float, the norm being the
// serial float norm of gt_set_coo (tensor_manager.tpp:263-293).
double compute_fit_ref(gt_ctx *c) {
  require(c->partitioned && c->initialized, GT_ERR_STATE, "partition and init first");
  require(!c->streaming, GT_ERR_STATE,
          "the reference-arithmetic fit needs the block COO resident (in-core layout)");
  GT_HIP(hipEventRecord(c->ev[3], c->s));
  c->xhat.alloc(std::max<uint64_t>(c->nnz, 1));
  CooArgs a = coo_args(c, c->N - 1);
  a.xhat_out = c->xhat.p;
  launch_coo_contract(c->R, a, c->s);
  std::vector<float> xh(c->nnz), x(c->nnz);
  if (c->nnz) {
    GT_HIP(hipMemcpyAsync(xh.data(), c->xhat.p, c->nnz * sizeof(float), hipMemcpyDeviceToHost,
                          c->s));
    GT_HIP(hipMemcpyAsync(x.data(), c->coo_val.p, c->nnz * sizeof(float), hipMemcpyDeviceToHost,
                          c->s));
  }
  GT_HIP(hipStreamSynchronize(c->s));
  float err = 0.0f;
  for (uint64_t e = 0; e < c->nnz; ++e) {
    const float d = x[e] - xh[e];
    err += d * d;
  }
  GT_HIP(hipEventRecord(c->ev[4], c->s));
  GT_HIP(hipStreamSynchronize(c->s));
  float ms = 0.f;
  GT_HIP(hipEventElapsedTime(&ms, c->ev[3], c->ev[4]));
  c->tm.ms_fit = ms;
  if (c->norm_ref == 0.f) return 1.0;
  return (double)(1.0f - std::sqrt(err) / c->norm_ref);
}

// One ALS iteration (update.cuh:146-198 then reconstruction.cuh:228-277).
// The fit comes from the last mode's row solves: after that update,
// sum_e (x_e - xhat_e)^2 = |x|^2 + sum_i (a_i^T G_i a_i - 2 a_i^T C_i), so no
// reconstruction pass over the nonzeros is needed.
double iteration(gt_ctx *c) {
  if (c->o.ref_fit) {
    for (int n = 0; n < c->N; ++n) mode_update(c, n, false, false);
    return compute_fit_ref(c);
  }
  if (c->o.graph && !c->streaming) {
    // the N updates and the reduction of the fit partials as one graph:
    // kernels only (no event, copy or memset nodes), captured on first use
    // and replayed; every buffer it names stays put until the graph is
    // dropped (gt_set_coo, gt_partition, gt_set_init)
    if (!c->it_graph) {
      require(c->partitioned && c->initialized, GT_ERR_STATE, "partition and init first");
      hipGraph_t g = nullptr;
      c->drain_profile(); // no event waits while the stream is captured
      GT_HIP(hipStreamSynchronize(c->s));
      GT_HIP(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
      c->capturing = true;
      uint64_t parts = 0;
      try {
        for (int n = 0; n + 1 < c->N; ++n) mode_update(c, n, false, false);
        parts = mode_update(c, c->N - 1, true, false);
        reduce_f64(c->fit_part.p, (int)parts, c->scalar.p, c->s);
      } catch (...) {
        c->capturing = false;
        (void)hipStreamEndCapture(c->s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      c->capturing = false;
      GT_HIP(hipStreamEndCapture(c->s, &g));
      const hipError_t e = hipGraphInstantiate(&c->it_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      GT_HIP(e);
      c->it_graph_parts = parts;
    }
    GT_HIP(hipEventRecord(c->ev[5], c->s));
    GT_HIP(hipGraphLaunch(c->it_graph, c->s));
    GT_HIP(hipEventRecord(c->ev[3], c->s));
  } else {
    // the N updates are queued back to back; the host waits once, for the fit
    for (int n = 0; n + 1 < c->N; ++n) mode_update(c, n, false, false);
    const uint64_t parts = mode_update(c, c->N - 1, true, false);
    GT_HIP(hipEventRecord(c->ev[3], c->s));
    reduce_f64(c->fit_part.p, (int)parts, c->scalar.p, c->s);
  }
  double d = 0.0;
  GT_HIP(hipMemcpyAsync(&d, c->scalar.p, sizeof(double), hipMemcpyDeviceToHost, c->s));
  GT_HIP(hipEventRecord(c->ev[4], c->s));
  GT_HIP(hipStreamSynchronize(c->s));
  c->collect_mode_timings();
  float ms = 0.f;
  GT_HIP(hipEventElapsedTime(&ms, c->ev[3], c->ev[4]));
  c->tm.ms_fit = ms;
  if (c->o.graph && !c->streaming) {
    GT_HIP(hipEventElapsedTime(&ms, c->ev[5], c->ev[3]));
    c->tm.ms_iteration_graph = ms;
  }
  if (c->norm == 0.0) return 1.0; // reconstruction.cuh:272-273
  const double err = std::max(0.0, c->norm * c->norm + d);
  return 1.0 - std::sqrt(err) / c->norm;
}

// Per-block where_nnz of mode n (block.tpp:114-186), blocks concatenated:
// the stream order of mode n (a stable sort of the block-ordered nonzeros by
// row) partitioned stably by block, as positions inside each block.  Built
// only for the test hooks and the driver's host blocks.
void build_where(gt_ctx *c, int n) {
  const uint64_t nnz = c->nnz;
  hipStream_t s = c->s;
  DevBuf<uint32_t> keys, vals, tk, tv;
  keys.alloc(nnz);
  vals.alloc(nnz);
  tk.alloc(nnz);
  tv.alloc(nnz);
  DevBuf<char> temp;
  temp.alloc(radix_sort_temp_bytes(nnz) + 4096);
  GT_HIP(hipMemcpyAsync(keys.p, c->coo_idx[n].p, nnz * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                        s));
  iota_u32(vals.p, nnz, s);
  radix_sort_pairs(keys.p, vals.p, tk.p, tv.p, nnz, ceil_log2(c->rows[n]), temp.p, s);
  c->where[n].alloc(nnz);
  if (c->B == 1) {
    GT_HIP(hipMemcpyAsync(c->where[n].p, vals.p, nnz * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                          s));
  } else {
    keys_from_perm(vals.p, c->coo_blk.p, nnz, keys.p, s);
    radix_sort_pairs(keys.p, vals.p, tk.p, tv.p, nnz, ceil_log2(c->B), temp.p, s);
    local_positions(vals.p, c->coo_blk.p, c->d_block_off.p, nnz, c->where[n].p, s);
  }
  GT_HIP(hipStreamSynchronize(s));
}

// Shared factor rows (wide refined BF16 ranks without a packed tile, order
// 3).  Mode n, with L the last other mode and O the remaining one, takes one
// of two groupings of each row's nonzeros in its stream (stable, so inside a
// group the order stays that of the reference's block loop; the rows, their
// pointers and the Gram are unchanged, only the order of terms inside a
// row's sum moves):
// - by row of A_L (kind kHotLast), when at least half the nonzeros lie in
//   rows of mode n holding kHotRunNnz or more nonzeros per row of A_L (such
//   a row carries runs of that many nonzeros per A_L row on average) and A_L
//   has at most kHotLastMaxRows rows (a table each, 4 KiB at R = 32; not
//   under a device budget);
// - else by hot outer row (kind kHotOuter): a row h of O is hot when it
//   holds at least kHotRunNnz nonzeros per row of mode n; the kMaxHotRows
//   heaviest (ties: lower row) get slots, cold rows last.
// Runs long enough cover whole 16-nonzero MFMA blocks (on the order-3
// Zipf(1.5) workload, 16 beat 32 and 64 by 1-5 %).  Runs after the input
// copy is freed, with the partition's sort buffers.
constexpr uint64_t kHotRunNnz = 16;
constexpr int kHotShift = 6; // slot bits in the sort key: slots 0..62, 63 cold
constexpr uint64_t kHotLastMaxRows = 1ull << 20;

void clear_hot_rows(gt_ctx *c) {
  for (int n = 0; n < GT_MAX_ORDER; ++n) {
    c->hot_kind[n] = 0;
    c->hot_k[n] = 0;
    c->hot_slot[n].reset();
    c->hot_rows[n].reset();
    c->hot_tab[n].reset();
  }
}

// Reorders mode n's stream (every index array and the values) by perm,
// one array at a time through tmp.
void permute_stream(gt_ctx *c, int n, const uint32_t *perm, DevBuf<uint32_t> &tmp) {
  hipStream_t s = c->s;
  const uint64_t nnz = c->nnz;
  tmp.alloc(nnz);
  for (int k = 0; k < c->N; ++k) {
    U32Ptrs src{};
    U32MutPtrs dst{};
    src.p[0] = c->st_idx[n][k].p;
    dst.p[0] = tmp.p;
    gather_coo(1, nnz, perm, src, nullptr, dst, nullptr, s);
    c->st_idx[n][k].swap(tmp);
  }
  DevBuf<float> tmpv;
  GT_HIP(hipStreamSynchronize(s));
  tmp.reset(); // the value copy takes its place
  tmpv.alloc(nnz);
  gather_coo(0, nnz, perm, U32Ptrs{}, c->st_val[n].p, U32MutPtrs{}, tmpv.p, s);
  c->st_val[n].swap(tmpv);
  GT_HIP(hipStreamSynchronize(s)); // before the swapped-out copies are freed
}

std::vector<uint64_t> row_degrees(gt_ctx *c, int m) {
  std::vector<uint64_t> rp(c->rows[m] + 1);
  GT_HIP(hipMemcpyAsync(rp.data(), c->row_ptr[m].p, rp.size() * sizeof(uint64_t),
                        hipMemcpyDeviceToHost, c->s));
  GT_HIP(hipStreamSynchronize(c->s));
  for (uint64_t i = 0; i + 1 < rp.size(); ++i) rp[i] = rp[i + 1] - rp[i];
  rp.pop_back();
  return rp;
}

void group_hot_rows(gt_ctx *c, DevBuf<uint32_t> &keys, DevBuf<uint32_t> &vals,
                    DevBuf<uint32_t> &tk, DevBuf<uint32_t> &tv, DevBuf<char> &temp) {
  const int N = c->N;
  clear_hot_rows(c);
  if (c->o.precision != GT_PREC_BF16R || !wide_hot_supported(c->R, N) || c->nnz == 0) return;
  static_assert((1 << kHotShift) - 1 >= kMaxHotRows, "slot bits");
  hipStream_t s = c->s;
  const uint64_t nnz = c->nnz;
  DevBuf<uint32_t> tmp;
  for (int n = 0; n < N; ++n) {
    const int L = (n == N - 1) ? N - 2 : N - 1;
    int O = 0;
    while (O == n || O == L) ++O;
    const uint64_t rows = c->rows[n], rows_o = c->rows[O], rows_l = c->rows[L];
    if (c->o.device_budget == 0 && rows_l <= kHotLastMaxRows) {
      uint64_t long_nnz = 0;
      for (uint64_t d : row_degrees(c, n))
        if (d >= kHotRunNnz * rows_l) long_nnz += d;
      if (2 * long_nnz >= nnz) {
        // stable sort of the stream by (row, A_L row): by the A_L row, then
        // stably by row
        GT_HIP(hipMemcpyAsync(keys.p, c->st_idx[n][L].p, nnz * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s));
        iota_u32(vals.p, nnz, s);
        radix_sort_pairs(keys.p, vals.p, tk.p, tv.p, nnz, ceil_log2(rows_l), temp.p, s);
        U32Ptrs src{};
        U32MutPtrs dst{};
        src.p[0] = c->st_idx[n][n].p;
        dst.p[0] = keys.p;
        gather_coo(1, nnz, vals.p, src, nullptr, dst, nullptr, s);
        radix_sort_pairs(keys.p, vals.p, tk.p, tv.p, nnz, ceil_log2(rows), temp.p, s);
        permute_stream(c, n, vals.p, tmp);
        c->hot_tab[n].alloc(hot_table_words(c->R, rows_l));
        c->hot_kind[n] = kHotLast;
        c->hot_k[n] = rows_l;
        continue;
      }
    }
    if (ceil_log2(rows) + kHotShift > 32) continue;
    const std::vector<uint64_t> deg = row_degrees(c, O);
    std::vector<std::pair<uint64_t, uint32_t>> cand; // (degree, row)
    for (uint64_t h = 0; h < rows_o; ++h)
      if (deg[h] >= kHotRunNnz * rows) cand.emplace_back(deg[h], (uint32_t)h);
    if (cand.empty()) continue;
    std::stable_sort(cand.begin(), cand.end(),
                     [](const auto &x, const auto &y) { return x.first > y.first; });
    const int K = (int)std::min<size_t>(cand.size(), kMaxHotRows);
    std::vector<uint32_t> slot(rows_o, kNoHotSlot), hrows(K);
    for (int k = 0; k < K; ++k) {
      hrows[k] = cand[k].second;
      slot[cand[k].second] = (uint32_t)k;
    }
    c->hot_slot[n].alloc(rows_o);
    c->hot_rows[n].alloc(K);
    c->hot_tab[n].alloc(hot_table_words(c->R, K));
    GT_HIP(hipMemcpyAsync(c->hot_slot[n].p, slot.data(), rows_o * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
    GT_HIP(hipMemcpyAsync(c->hot_rows[n].p, hrows.data(), K * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
    // stable sort of the stream by (row, slot)
    hot_keys(c->st_idx[n][n].p, c->st_idx[n][O].p, c->hot_slot[n].p, kHotShift, nnz, keys.p, s);
    iota_u32(vals.p, nnz, s);
    radix_sort_pairs(keys.p, vals.p, tk.p, tv.p, nnz, ceil_log2(rows) + kHotShift, temp.p, s);
    permute_stream(c, n, vals.p, tmp);
    c->hot_kind[n] = kHotOuter;
    c->hot_k[n] = (uint64_t)K;
  }
}

void do_partition_host(gt_ctx *c);

void do_partition(gt_ctx *c) {
  require(c->have_coo, GT_ERR_STATE, "gt_set_coo first");
  if (c->host_input) {
    do_partition_host(c);
    return;
  }
  const int N = c->N;
  const uint64_t nnz = c->nnz;
  auto t0 = std::chrono::steady_clock::now();
  c->host_partitioned = false;
  c->partition_pieces = 0;
  uint64_t mem = c->o.device_budget;
  if (mem == 0) {
    size_t fr = 0, tot = 0;
    GT_HIP(hipMemGetInfo(&fr, &tot));
    mem = fr;
  }
  Plan p;
  if (!plan_partition(N, c->dims, nnz, c->Ru, mem, (int)sizeof(float), c->o.partition, p))
    throw GtError(GT_ERR_PARTITION, "Block dimension is larger than the tensor dimension.");
  require(p.blocks < (1ull << 32), GT_ERR_ARG, "too many blocks");
  c->plan = p;
  c->B = p.blocks;
  hipStream_t s = c->s;
  // 1. block id per nonzero, block histogram
  DevBuf<uint32_t> keys, vals, tk, tv, counts;
  keys.alloc(nnz);
  vals.alloc(nnz);
  tk.alloc(nnz);
  tv.alloc(nnz);
  counts.alloc(std::max<uint64_t>(c->B, 1));
  DevBuf<char> temp;
  uint64_t max_rows = 1;
  for (int n = 0; n < N; ++n) max_rows = std::max<uint64_t>(max_rows, (uint64_t)p.part[n] * p.bd[n]);
  temp.alloc(std::max(radix_sort_temp_bytes(nnz), scan_temp_bytes(std::max(nnz, max_rows) + 1)) +
             4096);
  fill_u32(counts.p, 0u, c->B, s);
  PartArgs pa{};
  pa.order = N;
  pa.nnz = nnz;
  for (int k = 0; k < N; ++k) {
    pa.idx[k] = c->in_idx[k].p;
    pa.bdims[k] = p.bd[k];
    pa.part[k] = p.part[k];
  }
  block_ids(pa, (uint32_t)c->B, keys.p, vals.p, counts.p, s);
  std::vector<uint32_t> hcounts(c->B);
  GT_HIP(hipMemcpyAsync(hcounts.data(), counts.p, c->B * sizeof(uint32_t), hipMemcpyDeviceToHost,
                        s));
  // 2. stable sort by block id: the serial in-order scatter of the reference
  radix_sort_pairs(keys.p, vals.p, tk.p, tv.p, nnz, ceil_log2(c->B), temp.p, s);
  GT_HIP(hipStreamSynchronize(s));
  c->block_nnz.assign(c->B, 0);
  c->block_off.assign(c->B + 1, 0);
  for (uint64_t b = 0; b < c->B; ++b) {
    c->block_nnz[b] = hcounts[b];
    c->block_off[b + 1] = c->block_off[b] + hcounts[b];
  }
  c->d_block_off.alloc(c->B + 1);
  GT_HIP(hipMemcpyAsync(c->d_block_off.p, c->block_off.data(), (c->B + 1) * sizeof(uint64_t),
                        hipMemcpyHostToDevice, s));
  // 3. block-ordered COO
  U32Ptrs src{};
  U32MutPtrs dst{};
  for (int k = 0; k < N; ++k) {
    c->coo_idx[k].alloc(nnz);
    src.p[k] = c->in_idx[k].p;
    dst.p[k] = c->coo_idx[k].p;
  }
  c->coo_val.alloc(nnz);
  c->coo_blk.alloc(nnz);
  gather_coo(N, nnz, vals.p, src, c->in_val.p, dst, c->coo_val.p, s);
  GT_HIP(hipMemcpyAsync(c->coo_blk.p, keys.p, nnz * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  // 4. norm = sqrt(sum x^2) (tensor_manager.tpp:289,293), FP64
  DevBuf<double> sq;
  sq.alloc(1024);
  int g = sum_squares(c->in_val.p, nnz, sq.p, s);
  std::vector<double> hsq(g);
  GT_HIP(hipMemcpyAsync(hsq.data(), sq.p, g * sizeof(double), hipMemcpyDeviceToHost, s));
  // 5. per-mode row grouping
  for (int n = 0; n < N; ++n) {
    const uint64_t rows = (uint64_t)p.part[n] * p.bd[n];
    c->rows[n] = rows;
    GT_HIP(hipMemcpyAsync(keys.p, c->coo_idx[n].p, nnz * sizeof(uint32_t),
                          hipMemcpyDeviceToDevice, s));
    iota_u32(vals.p, nnz, s);
    radix_sort_pairs(keys.p, vals.p, tk.p, tv.p, nnz, ceil_log2(rows), temp.p, s);
    // vals = stream order -> block-ordered position
    U32MutPtrs sd{};
    for (int k = 0; k < N; ++k) {
      c->st_idx[n][k].alloc(nnz);
      sd.p[k] = c->st_idx[n][k].p;
    }
    c->st_val[n].alloc(nnz);
    U32Ptrs cs{};
    for (int k = 0; k < N; ++k) cs.p[k] = c->coo_idx[k].p;
    gather_coo(N, nnz, vals.p, cs, c->coo_val.p, sd, c->st_val[n].p, s);
    // global row pointer of the stream, from the sorted rows (keys)
    c->row_ptr[n].alloc(rows + 1);
    row_ptr_from_sorted(keys.p, nnz, rows, c->row_ptr[n].p, s);
    c->where[n].reset(); // built on demand (build_where): the ALS path never reads it
    GT_HIP(hipStreamSynchronize(s)); // keys, vals and temp are reused
  }
  double ss = 0.0;
  for (double v : hsq) ss += v;
  c->norm = std::sqrt(ss);
  // the raw input copy is no longer needed
  for (int k = 0; k < N; ++k) c->in_idx[k].reset();
  c->in_val.reset();
  group_hot_rows(c, keys, vals, tk, tv, temp);
  c->have_coo = false;
  c->partitioned = true;
  ensure_work(c);
  maybe_stream(c);
  c->ms_partition =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->tm.ms_partition = c->ms_partition;
}

// The partition under a device budget (tensor_manager.tpp:201-302 and
// block.tpp:114-186 with the data out of core): the input waits in pinned
// host memory and is read in segments.  A first pass counts the nonzeros per
// block and per row of every mode (the row pointers of the streams) and sums
// the norm.  Then, for each mode n, consecutive ranges of rows ("buckets")
// whose nonzeros fit in the device bytes left are built one at a time: a
// pass over the input picks the bucket's nonzeros in input order, they are
// sorted stably by block and then by row (the stream order: row, block,
// position, exactly the in-core order), and the sorted piece goes to the
// host stream at the row pointer of its first row, packed as in core (a
// piece of a packed 16- or 24-bit stream must start at a multiple of four
// nonzeros, so up to three trailing nonzeros wait for the next piece).  The
// block COO is never resident; the ALS then stages the host streams as in
// the out-of-core mode of an in-core partition.
void do_partition_host(gt_ctx *c) {
  const int N = c->N;
  const uint64_t nnz = c->nnz, budget = c->o.device_budget;
  auto t0 = std::chrono::steady_clock::now();
  auto say = [&](const std::string &m) { // opts.verbose: one line per step
    if (!c->o.verbose) return;
    const double sec =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "[gt] partition under the device budget: %s (%.1f s)\n", m.c_str(), sec);
    std::fflush(stderr);
  };
  c->host_partitioned = true;
  c->partition_pieces = 0;
  clear_hot_rows(c); // the bucketed partition keeps each row in block order
  Plan p;
  if (!plan_partition(N, c->dims, nnz, c->Ru, budget, (int)sizeof(float), c->o.partition, p))
    throw GtError(GT_ERR_PARTITION, "Block dimension is larger than the tensor dimension.");
  require(p.blocks < (1ull << 32), GT_ERR_ARG, "too many blocks");
  c->plan = p;
  c->B = p.blocks;
  for (int n = 0; n < N; ++n) c->rows[n] = (uint64_t)p.part[n] * p.bd[n];
  hipStream_t s = c->s;
  ensure_work(c);
  const uint64_t fixed = resident_fixed_bytes(c);
  if (fixed + (64ull << 10) >= budget)
    throw GtError(GT_ERR_MEMORY, "device budget is below the per-row state");
  const uint64_t avail = budget - fixed;
  // device bytes per element: a scanned input segment (indices, value,
  // flags, offsets, block keys and values) and a bucket (its nonzeros, the
  // sort's four arrays and scratch, the sorted stream, the packing buffer)
  const uint64_t per_scan = 4ull * N + 4 + 4 + 8 + 8;
  const uint64_t per_bucket = (4ull * N + 4) + 16 + 8 + (4ull * N + 4) + 4;
  uint64_t S = std::max<uint64_t>(std::min<uint64_t>(nnz, avail / 8 / per_scan), 1);
  if (S < nnz) S = std::max<uint64_t>(S / 256 * 256, 256);
  const uint64_t scan_bytes = S * per_scan + scan_temp_bytes(S + 1) + 4096;
  require(scan_bytes < avail, GT_ERR_MEMORY, "device budget too small for a partition segment");
  uint64_t cap = (avail - scan_bytes) / per_bucket;
  while (cap > 1024 && cap * per_bucket + radix_sort_temp_bytes(cap) + 8192 > avail - scan_bytes)
    cap -= cap / 16;
  require(cap >= 1024, GT_ERR_MEMORY, "device budget too small for a partition bucket");

  { // the partition's buffers are freed before the ALS staging is allocated
  // ---- pass 1: block and row histograms, the norm
  DevBuf<uint32_t> sidx[GT_MAX_ORDER], skeys, svals, flags, counts, rowcnt[GT_MAX_ORDER];
  DevBuf<float> sval;
  DevBuf<uint64_t> offs;
  DevBuf<char> stemp;
  for (int k = 0; k < N; ++k) sidx[k].alloc(S);
  sval.alloc(S);
  skeys.alloc(S);
  svals.alloc(S);
  flags.alloc(S);
  offs.alloc(S + 1);
  stemp.alloc(scan_temp_bytes(S + 1) + 4096);
  counts.alloc(std::max<uint64_t>(c->B, 1));
  fill_u32(counts.p, 0u, c->B, s);
  for (int n = 0; n < N; ++n) {
    rowcnt[n].alloc(c->rows[n]);
    fill_u32(rowcnt[n].p, 0u, c->rows[n], s);
  }
  DevBuf<double> sq;
  sq.alloc(1024);
  std::vector<double> hsq(1024);
  double ss = 0.0;
  auto load_segment = [&](uint64_t lo, uint64_t len) {
    for (int k = 0; k < N; ++k)
      GT_HIP(hipMemcpyAsync(sidx[k].p, c->h_in_idx[k] + lo, len * sizeof(uint32_t),
                            hipMemcpyHostToDevice, s));
    GT_HIP(hipMemcpyAsync(sval.p, c->h_in_val + lo, len * sizeof(float), hipMemcpyHostToDevice,
                          s));
  };
  PartArgs pa{};
  pa.order = N;
  for (int k = 0; k < N; ++k) {
    pa.bdims[k] = p.bd[k];
    pa.part[k] = p.part[k];
  }
  for (uint64_t lo = 0; lo < nnz; lo += S) {
    const uint64_t len = std::min(S, nnz - lo);
    load_segment(lo, len);
    pa.nnz = len;
    for (int k = 0; k < N; ++k) pa.idx[k] = sidx[k].p;
    block_ids(pa, (uint32_t)c->B, skeys.p, svals.p, counts.p, s);
    for (int n = 0; n < N; ++n) histogram_u32(sidx[n].p, len, rowcnt[n].p, s);
    const int g = sum_squares(sval.p, len, sq.p, s);
    GT_HIP(hipMemcpyAsync(hsq.data(), sq.p, g * sizeof(double), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < g; ++i) ss += hsq[i];
  }
  c->norm = std::sqrt(ss);
  say("counted blocks and rows over " + std::to_string((nnz + S - 1) / S) + " segments");
  std::vector<uint32_t> hcounts(c->B);
  GT_HIP(hipMemcpyAsync(hcounts.data(), counts.p, c->B * sizeof(uint32_t), hipMemcpyDeviceToHost,
                        s));
  std::vector<std::vector<uint64_t>> hrp(N);
  for (int n = 0; n < N; ++n) {
    c->row_ptr[n].alloc(c->rows[n] + 1);
    DevBuf<char> t;
    t.alloc(scan_temp_bytes(c->rows[n]) + 4096);
    exclusive_scan_u32_to_u64(rowcnt[n].p, c->row_ptr[n].p, c->rows[n], t.p, s);
    hrp[n].resize(c->rows[n] + 1);
    GT_HIP(hipMemcpyAsync(hrp[n].data(), c->row_ptr[n].p, (c->rows[n] + 1) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    rowcnt[n].reset();
  }
  c->block_nnz.assign(c->B, 0);
  c->block_off.assign(c->B + 1, 0);
  for (uint64_t b = 0; b < c->B; ++b) {
    c->block_nnz[b] = hcounts[b];
    c->block_off[b + 1] = c->block_off[b] + hcounts[b];
  }

  // ---- the host streams
  const bool packed = choose_packing(c);
  for (int n = 0; n < N; ++n) {
    for (int k = 0; k < N; ++k) {
      if (packed && k == n) continue;
      if (packed)
        c->host_alloc(&c->h_pack[n][k], pack_words(c->pack_bits[k], nnz));
      else
        c->host_alloc(&c->h_idx[n][k], nnz);
    }
    c->host_alloc(&c->h_val[n], nnz);
  }

  // ---- pass 2: per mode, bucket by bucket
  DevBuf<uint32_t> bidx[GT_MAX_ORDER], bkeys, bvals, btk, btv, sidx_out[GT_MAX_ORDER], ptmp,
      carry[GT_MAX_ORDER];
  DevBuf<float> bval, sval_out;
  DevBuf<char> btemp;
  for (int k = 0; k < N; ++k) {
    bidx[k].alloc(cap);
    sidx_out[k].alloc(cap + 4); // 4 leading slots for the carried nonzeros
    carry[k].alloc(4);
  }
  bval.alloc(cap);
  sval_out.alloc(cap);
  bkeys.alloc(cap);
  bvals.alloc(cap);
  btk.alloc(cap);
  btv.alloc(cap);
  ptmp.alloc(3 * ((cap + 8) / 4) + 4);
  btemp.alloc(radix_sort_temp_bytes(cap) + 4096);
  DevBuf<uint32_t> bhist;
  bhist.alloc(std::max<uint64_t>(c->B, 1));
  std::vector<uint32_t> hbh(c->B);
  for (int n = 0; n < N; ++n) {
    const uint64_t rows = c->rows[n];
    uint64_t cc = 0, out_off = 0; // carried nonzeros, offset of the next packed piece
    uint64_t out_pos = 0;         // stream position of the next piece
    // one piece: the nonzeros of rows [r0, r1) and blocks [b0, b1), and of
    // those the ones ranked [w0, w1) in input order; `last` flushes the carry
    auto piece = [&](uint64_t r0, uint64_t r1, uint32_t b0, uint32_t b1, bool all_blocks,
                     uint64_t w0, uint64_t w1, uint64_t nb, bool last) {
      if (nb > 0) {
        uint64_t base = 0;
        for (uint64_t lo = 0; lo < nnz; lo += S) {
          const uint64_t len = std::min(S, nnz - lo);
          load_segment(lo, len);
          pa.nnz = len;
          for (int k = 0; k < N; ++k) pa.idx[k] = sidx[k].p;
          flag_piece(pa, n, (uint32_t)r0, (uint32_t)r1, b0, b1, all_blocks, flags.p, s);
          exclusive_scan_u32_to_u64(flags.p, offs.p, len, stemp.p, s);
          uint64_t cnt = 0;
          GT_HIP(hipMemcpyAsync(&cnt, offs.p + len, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
          GT_HIP(hipStreamSynchronize(s));
          if (cnt && base + cnt > w0 && base < w1) {
            U32Ptrs src{};
            U32MutPtrs dst{};
            for (int k = 0; k < N; ++k) {
              src.p[k] = sidx[k].p;
              dst.p[k] = bidx[k].p;
            }
            scatter_flagged(N, len, flags.p, offs.p, base, w0, w1, src, sval.p, dst, bval.p, s);
          }
          base += cnt;
          if (base >= w1) break; // the window is complete
        }
        require(std::min(base, w1) - w0 == nb, GT_ERR_HIP, "partition piece count mismatch");
        // stable by block, then stable by row: the stream order
        pa.nnz = nb;
        for (int k = 0; k < N; ++k) pa.idx[k] = bidx[k].p;
        block_ids(pa, (uint32_t)c->B, bkeys.p, bvals.p, counts.p, s);
        radix_sort_pairs(bkeys.p, bvals.p, btk.p, btv.p, nb, ceil_log2(c->B), btemp.p, s);
        U32Ptrs rsrc{};
        U32MutPtrs rdst{};
        rsrc.p[0] = bidx[n].p;
        rdst.p[0] = bkeys.p;
        gather_coo(1, nb, bvals.p, rsrc, nullptr, rdst, nullptr, s);
        radix_sort_pairs(bkeys.p, bvals.p, btk.p, btv.p, nb, ceil_log2(rows), btemp.p, s);
        U32Ptrs src{};
        U32MutPtrs dst{};
        for (int k = 0; k < N; ++k) {
          src.p[k] = bidx[k].p;
          dst.p[k] = sidx_out[k].p + 4;
        }
        gather_coo(N, nb, bvals.p, src, bval.p, dst, sval_out.p, s);
        GT_HIP(hipMemcpyAsync(c->h_val[n] + out_pos, sval_out.p, nb * sizeof(float),
                              hipMemcpyDeviceToHost, s));
      }
      if (!packed) {
        for (int k = 0; k < N && nb > 0; ++k)
          GT_HIP(hipMemcpyAsync(c->h_idx[n][k] + out_pos, sidx_out[k].p + 4,
                                nb * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      } else if (nb > 0 || (last && cc > 0)) {
        const uint64_t T = cc + nb, P = last ? T : T / 4 * 4;
        for (int k = 0; k < N; ++k) {
          if (k == n) continue;
          uint32_t *work = sidx_out[k].p + 4 - cc;
          if (cc)
            GT_HIP(hipMemcpyAsync(work, carry[k].p, cc * sizeof(uint32_t),
                                  hipMemcpyDeviceToDevice, s));
          const int bits = c->pack_bits[k];
          if (P) {
            if (bits == 16)
              pack16(work, P, ptmp.p, s);
            else
              pack24(work, P, ptmp.p, s);
            GT_HIP(hipMemcpyAsync(c->h_pack[n][k] + pack_words(bits, out_off), ptmp.p,
                                  pack_words(bits, P) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  s));
          }
          if (T > P)
            GT_HIP(hipMemcpyAsync(carry[k].p, work + P, (T - P) * sizeof(uint32_t),
                                  hipMemcpyDeviceToDevice, s));
        }
        out_off += P;
        cc = T - P;
      }
      GT_HIP(hipStreamSynchronize(s)); // the buffers are reused by the next piece
      out_pos += nb;
      ++c->partition_pieces;
      say("mode " + std::to_string(n) + ": " + std::to_string(out_pos) + " of " +
          std::to_string(nnz) + " nonzeros in the stream");
    };
    uint64_t r0 = 0;
    while (r0 < rows) {
      const uint64_t rc = hrp[n][r0 + 1] - hrp[n][r0];
      if (rc <= cap) { // whole rows: as many as fit
        uint64_t r1 = r0 + 1;
        while (r1 < rows && hrp[n][r1 + 1] - hrp[n][r0] <= cap) ++r1;
        piece(r0, r1, 0, (uint32_t)c->B, true, 0, ~0ull, hrp[n][r1] - hrp[n][r0], r1 == rows);
        r0 = r1;
        continue;
      }
      // a row with more nonzeros than a piece holds: by block ranges, and a
      // block with more by windows of its input order (its stream order)
      fill_u32(bhist.p, 0u, c->B, s);
      for (uint64_t lo = 0; lo < nnz; lo += S) {
        const uint64_t len = std::min(S, nnz - lo);
        load_segment(lo, len);
        pa.nnz = len;
        for (int k = 0; k < N; ++k) pa.idx[k] = sidx[k].p;
        row_block_hist(pa, n, (uint32_t)r0, bhist.p, s);
      }
      GT_HIP(hipMemcpyAsync(hbh.data(), bhist.p, c->B * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            s));
      GT_HIP(hipStreamSynchronize(s));
      const bool last_row = r0 + 1 == rows;
      uint64_t b = 0;
      while (b < c->B) {
        if (hbh[b] > cap) {
          for (uint64_t w = 0; w < hbh[b]; w += cap) {
            const uint64_t w1 = std::min<uint64_t>(w + cap, hbh[b]);
            piece(r0, r0 + 1, (uint32_t)b, (uint32_t)b + 1, false, w, w1, w1 - w,
                  last_row && b + 1 == c->B && w1 == hbh[b]);
          }
          ++b;
          continue;
        }
        uint64_t b1 = b + 1, sum = hbh[b];
        while (b1 < c->B && sum + hbh[b1] <= cap) sum += hbh[b1++];
        piece(r0, r0 + 1, (uint32_t)b, (uint32_t)b1, false, 0, ~0ull, sum, last_row && b1 == c->B);
        b = b1;
      }
      ++r0;
    }
    require(cc == 0, GT_ERR_HIP, "partition left nonzeros unpacked");
    require(out_pos == nnz, GT_ERR_HIP, "partition stream length mismatch");
  }
  } // the partition's device buffers go here
  c->release_host_input();
  c->pin_host_streams(); // into the page-locked memory the input held
  say("done; " + std::to_string(c->pageable_bytes) + " bytes of the streams in pageable memory");
  c->have_coo = false;
  c->partitioned = true;
  setup_staging(c, fixed);
  c->ms_partition =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->tm.ms_partition = c->ms_partition;
}

} // namespace

// ================================================================ C ABI
extern "C" {

int gt_version(void) { return 200; }

int gt_kernel_rank(int rank) { return rank > GT_MAX_RANK ? 0 : kernel_rank(rank); }

int gt_default_opts(gt_opts *o) {
  if (!o) return GT_ERR_ARG;
  std::memset(o, 0, sizeof(*o));
  o->device = 0;
  o->order = 3;
  o->rank = 10;                         // cmdline_opts.cpp:30
  o->lambda = (double)0.0001f;          // constants.hpp:18 (a float literal)
  o->max_iters = 3;                     // constants.hpp:17
  o->tol = (double)0.0001f;             // tucker.cuh:183 compares with kLambda
  o->precision = GT_PREC_FP32;
  return GT_OK;
}

int gt_plan_partition(int order, const uint32_t *dims, uint64_t nnz, int rank, uint64_t mem,
                      int value_size, const uint32_t *forced, uint32_t *part_out,
                      uint32_t *bd_out, int *streams_out, uint64_t *avail_out, int *type_out) {
  if (order < 1 || order > GT_MAX_ORDER || !dims || rank < 1) return GT_ERR_ARG;
  Plan p;
  if (!plan_partition(order, dims, nnz, rank, mem, value_size, forced, p)) return GT_ERR_PARTITION;
  for (int a = 0; a < order; ++a) {
    if (part_out) part_out[a] = p.part[a];
    if (bd_out) bd_out[a] = p.bd[a];
  }
  if (streams_out) *streams_out = p.streams;
  if (avail_out) *avail_out = p.avail;
  if (type_out) *type_out = p.type;
  return GT_OK;
}

int64_t gt_schedule(int type, int stream_count, uint64_t avail_nnz, uint64_t block_count,
                    const uint64_t *block_nnz, uint64_t nnz_total, int64_t max_tasks,
                    uint64_t *t_block, uint64_t *t_nnz, uint64_t *t_off, int *t_stream) {
  // scheduler.tpp:22-159 for one GPU
  int64_t nt = 0;
  auto push = [&](uint64_t b, uint64_t n, uint64_t off, int st) {
    if (nt >= max_tasks) return false;
    t_block[nt] = b;
    t_nnz[nt] = n;
    t_off[nt] = off;
    t_stream[nt] = st;
    ++nt;
    return true;
  };
  if (avail_nnz == 0) return -1;
  if (type == 1) {
    uint64_t task_count = 1, per = nnz_total;
    if (per > avail_nnz) {
      task_count = (nnz_total + avail_nnz - 1) / avail_nnz;
      per = avail_nnz;
    }
    uint64_t n = per;
    for (uint64_t t = 0; t < task_count; ++t) {
      uint64_t ptr = t * n;
      if (t == task_count - 1) n = nnz_total - ptr;
      if (!push(0, n, ptr, 0)) return -1;
    }
    return nt;
  }
  std::vector<uint64_t> ids(block_count);
  std::iota(ids.begin(), ids.end(), 0);
  std::sort(ids.begin(), ids.end(),
            [&](uint64_t a, uint64_t b) { return block_nnz[a] < block_nnz[b]; }); // :97
  int g = 0;
  for (uint64_t i = 0; i < block_count; ++i) {
    const uint64_t b = ids[i];
    uint64_t tasks = 1, per = block_nnz[b];
    if (block_nnz[b] > avail_nnz) {
      tasks = (block_nnz[b] + avail_nnz - 1) / avail_nnz;
      per = avail_nnz;
    }
    uint64_t t;
    for (t = 0; t + 1 < tasks; ++t) {
      if (g >= stream_count) g = 0;
      if (!push(b, per, t * per, g)) return -1;
      ++g;
    }
    if (g >= stream_count) g = 0;
    if (!push(b, block_nnz[b] - t * per, t * per, g)) return -1;
    ++g;
  }
  return nt;
}

int gt_create(gt_ctx **out, const gt_opts *opts) {
  if (!out || !opts) {
    g_create_error = "null argument";
    return GT_ERR_ARG;
  }
  *out = nullptr;
  if (opts->order < 2 || opts->order > GT_MAX_ORDER) {
    g_create_error = "order must be in [2, 8]";
    return GT_ERR_UNSUPPORTED;
  }
  if (opts->rank < 1 || opts->rank > GT_MAX_RANK || kernel_rank(opts->rank) == 0) {
    g_create_error = "rank " + std::to_string(opts->rank) + " is outside 1.." +
                     std::to_string(GT_MAX_RANK);
    return GT_ERR_UNSUPPORTED;
  }
  if (!(opts->lambda > 0.0)) {
    g_create_error = "lambda must be > 0 (the row systems are solved by Cholesky)";
    return GT_ERR_ARG;
  }
  if (opts->precision < GT_PREC_FP32 || opts->precision > GT_PREC_BF16R) {
    g_create_error = "unknown precision";
    return GT_ERR_ARG;
  }
  if (opts->core_source != 0 && opts->core_source != 1) {
    g_create_error = "core_source must be 0 or 1";
    return GT_ERR_ARG;
  }
  if (opts->precision == GT_PREC_BF16 && !gram_mfma_supported(opts->rank)) {
    g_create_error = "the plain BF16 MFMA delta is built for rank <= 15 (bf16r covers the wide "
                     "ranks of order 3)";
    return GT_ERR_UNSUPPORTED;
  }
  if (ipow(kernel_rank(opts->rank), opts->order) > (1ull << 24)) {
    g_create_error = "core too large";
    return GT_ERR_UNSUPPORTED;
  }
  std::unique_ptr<gt_ctx> c(new gt_ctx);
  c->o = *opts;
  if (c->o.max_iters <= 0) c->o.max_iters = 3;
  c->N = opts->order;
  c->Ru = opts->rank;
  c->R = kernel_rank(opts->rank);
  c->O = (int)ipow(c->R, c->N - 2);
  try {
    int count = 0;
    GT_HIP(hipGetDeviceCount(&count));
    if (opts->device < 0 || opts->device >= count) throw GtError(GT_ERR_ARG, "no such device");
    GT_HIP(hipSetDevice(opts->device));
    c->scratch = scratch_address(opts->device);
    int lds = 0;
    GT_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, opts->device));
    set_lds_limit((size_t)lds);
    GT_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    for (auto &e : c->ev) GT_HIP(hipEventCreate(&e));
    for (auto &em : c->ev_mode)
      for (auto &e : em) GT_HIP(hipEventCreate(&e));
  } catch (const GtError &e) {
    g_create_error = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_create_error = e.what();
    return GT_ERR_HIP;
  }
  *out = c.release();
  return GT_OK;
}

int gt_destroy(gt_ctx *ctx) {
  if (!ctx) return GT_ERR_ARG;
  (void)hipSetDevice(ctx->o.device);
  delete ctx;
  return GT_OK;
}

const char *gt_last_error(const gt_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int gt_set_coo(gt_ctx *ctx, uint64_t nnz, const uint32_t *const *idx, const float *vals,
               const uint32_t *dims) {
  return guarded(ctx, [&] {
    require(idx && vals && dims, GT_ERR_ARG, "null argument");
    require(nnz < (1ull << 32) - 1, GT_ERR_ARG, "nnz must be < 2^32 (uint32 positions)");
    for (int k = 0; k < ctx->N; ++k) require(dims[k] > 0, GT_ERR_ARG, "zero dimension");
    for (int k = 0; k < ctx->N; ++k) require(idx[k] != nullptr || nnz == 0, GT_ERR_ARG, "null idx");
    ctx->release_host_streams();
    ctx->release_host_input();
    ctx->drop_graph();
    ctx->partitioned = false;
    ctx->initialized = false;
    ctx->nnz = nnz;
    // under a device budget the in-core partition does not fit, the input
    // stays in pinned host memory for the bucketed partition
    ctx->host_input = ctx->o.device_budget != 0 &&
                      nnz * incore_partition_bytes_per_nnz(ctx->N) + fixed_estimate(ctx, nnz, dims) >
                          ctx->o.device_budget;
    for (int k = 0; k < ctx->N; ++k) {
      ctx->dims[k] = dims[k];
      // bounds check on the host copy (tensor_manager.tpp:237 asserts it)
      for (uint64_t e = 0; e < nnz; ++e)
        if (idx[k][e] >= dims[k]) throw GtError(GT_ERR_ARG, "Coordinate is out of bounds");
      if (ctx->host_input) {
        ctx->in_idx[k].reset();
        ctx->host_alloc(&ctx->h_in_idx[k], nnz);
        std::memcpy(ctx->h_in_idx[k], idx[k], nnz * sizeof(uint32_t));
        continue;
      }
      ctx->in_idx[k].alloc(nnz);
      if (nnz)
        GT_HIP(hipMemcpyAsync(ctx->in_idx[k].p, idx[k], nnz * sizeof(uint32_t),
                              hipMemcpyHostToDevice, ctx->s));
    }
    if (ctx->o.ref_fit) { // tensor_manager.tpp:263,289,293: serial, in value_t
      float ss = 0.0f;
      for (uint64_t e = 0; e < nnz; ++e) ss += vals[e] * vals[e];
      ctx->norm_ref = std::sqrt(ss);
    }
    if (ctx->host_input) {
      ctx->in_val.reset();
      ctx->host_alloc(&ctx->h_in_val, nnz);
      std::memcpy(ctx->h_in_val, vals, nnz * sizeof(float));
    } else {
      ctx->in_val.alloc(nnz);
      if (nnz)
        GT_HIP(hipMemcpyAsync(ctx->in_val.p, vals, nnz * sizeof(float), hipMemcpyHostToDevice,
                              ctx->s));
    }
    GT_HIP(hipStreamSynchronize(ctx->s));
    ctx->have_coo = true;
  });
}

int gt_partition(gt_ctx *ctx, uint32_t *part_out, uint64_t *block_nnz_out) {
  return guarded(ctx, [&] {
    require(ctx->nnz > 0 || !ctx->have_coo, GT_ERR_ARG, "empty tensor");
    ctx->drop_graph();
    do_partition(ctx);
    if (part_out)
      for (int k = 0; k < ctx->N; ++k) part_out[k] = ctx->plan.part[k];
    if (block_nnz_out)
      for (uint64_t b = 0; b < ctx->B; ++b) block_nnz_out[b] = ctx->block_nnz[b];
  });
}

int gt_get_partition_info(gt_ctx *ctx, gt_partition_info *info) {
  return guarded(ctx, [&] {
    require(info != nullptr, GT_ERR_ARG, "null info");
    require(ctx->partitioned, GT_ERR_STATE, "not partitioned");
    std::memset(info, 0, sizeof(*info));
    info->order = ctx->N;
    for (int k = 0; k < ctx->N; ++k) {
      info->dims[k] = ctx->dims[k];
      info->partition[k] = ctx->plan.part[k];
      info->block_dims[k] = ctx->plan.bd[k];
    }
    info->block_count = ctx->B;
    info->nnz = ctx->nnz;
    for (uint64_t b = 0; b < ctx->B; ++b) {
      info->max_block_nnz = std::max(info->max_block_nnz, ctx->block_nnz[b]);
      info->empty_blocks += ctx->block_nnz[b] == 0;
    }
    info->norm = ctx->norm;
    info->partition_type = ctx->plan.type;
    info->stream_count = ctx->plan.streams;
    info->avail_nnz_per_task = ctx->plan.avail;
    info->out_of_core = ctx->streaming ? 1 : 0;
    info->segment_nnz = ctx->seg_nnz;
    info->stream_bytes_per_nnz = ctx->streaming ? ctx->stream_bytes : 0;
    info->partition_path = ctx->host_partitioned ? 1 : 0;
    info->partition_pieces = ctx->partition_pieces;
    info->host_pageable_bytes = ctx->pageable_bytes;
    for (int n = 0; n < GT_MAX_ORDER; ++n) {
      info->hot_rows[n] = (uint32_t)ctx->hot_k[n];
      info->hot_kind[n] = ctx->hot_kind[n];
    }
  });
}

int gt_get_block_nnz(gt_ctx *ctx, uint64_t *block_nnz_out) {
  return guarded(ctx, [&] {
    require(ctx->partitioned, GT_ERR_STATE, "not partitioned");
    require(block_nnz_out != nullptr, GT_ERR_ARG, "null output");
    for (uint64_t b = 0; b < ctx->B; ++b) block_nnz_out[b] = ctx->block_nnz[b];
  });
}

int gt_get_block(gt_ctx *ctx, uint64_t block, uint32_t *const *idx, float *vals) {
  return guarded(ctx, [&] {
    require(ctx->partitioned, GT_ERR_STATE, "not partitioned");
    require(!ctx->streaming, GT_ERR_STATE, "the block COO is not resident (out-of-core mode)");
    require(block < ctx->B, GT_ERR_ARG, "block out of range");
    const uint64_t off = ctx->block_off[block], n = ctx->block_nnz[block];
    if (n == 0) return;
    uint32_t coord[GT_MAX_ORDER];
    uint64_t rem = block;
    for (int a = ctx->N - 1; a >= 0; --a) { // tensor.tpp:156-174
      coord[a] = (uint32_t)(rem % ctx->plan.part[a]);
      rem /= ctx->plan.part[a];
    }
    DevBuf<uint32_t> tmp;
    tmp.alloc(n);
    for (int k = 0; k < ctx->N; ++k) {
      if (!idx || !idx[k]) continue;
      sub_u32(ctx->coo_idx[k].p + off, coord[k] * ctx->plan.bd[k], n, tmp.p, ctx->s);
      GT_HIP(hipMemcpyAsync(idx[k], tmp.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->s));
      GT_HIP(hipStreamSynchronize(ctx->s));
    }
    if (vals)
      GT_HIP(hipMemcpyAsync(vals, ctx->coo_val.p + off, n * sizeof(float), hipMemcpyDeviceToHost,
                            ctx->s));
    GT_HIP(hipStreamSynchronize(ctx->s));
  });
}

int gt_partition_dump(gt_ctx *ctx, uint64_t block, int mode, uint64_t *row_ptr, uint32_t *where) {
  return guarded(ctx, [&] {
    require(ctx->partitioned, GT_ERR_STATE, "not partitioned");
    require(!ctx->streaming, GT_ERR_STATE, "the block COO is not resident (out-of-core mode)");
    require(block < ctx->B && mode >= 0 && mode < ctx->N, GT_ERR_ARG, "block/mode out of range");
    const uint64_t off = ctx->block_off[block], n = ctx->block_nnz[block];
    const uint32_t bd = ctx->plan.bd[mode];
    uint64_t rem = block, coord = 0;
    for (int a = ctx->N - 1; a >= 0; --a) {
      if (a == mode) coord = rem % ctx->plan.part[a];
      rem /= ctx->plan.part[a];
    }
    if (where && n) {
      if (!ctx->where[mode].p) build_where(ctx, mode);
      GT_HIP(hipMemcpyAsync(where, ctx->where[mode].p + off, n * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, ctx->s));
    }
    if (row_ptr) {
      DevBuf<uint32_t> loc, cnt;
      DevBuf<uint64_t> rp;
      DevBuf<char> temp;
      loc.alloc(n);
      cnt.alloc(bd);
      rp.alloc((uint64_t)bd + 1);
      temp.alloc(scan_temp_bytes(bd) + 256);
      fill_u32(cnt.p, 0u, bd, ctx->s);
      sub_u32(ctx->coo_idx[mode].p + off, (uint32_t)(coord * bd), n, loc.p, ctx->s);
      histogram_u32(loc.p, n, cnt.p, ctx->s);
      exclusive_scan_u32_to_u64(cnt.p, rp.p, bd, temp.p, ctx->s);
      GT_HIP(hipMemcpyAsync(row_ptr, rp.p, ((uint64_t)bd + 1) * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, ctx->s));
      GT_HIP(hipStreamSynchronize(ctx->s));
    }
    GT_HIP(hipStreamSynchronize(ctx->s));
  });
}

int gt_set_init(gt_ctx *ctx, const float *const *A, const float *core) {
  return guarded(ctx, [&] {
    require(ctx->partitioned, GT_ERR_STATE, "partition first");
    require(A && core, GT_ERR_ARG, "null argument");
    GT_HIP(hipStreamSynchronize(ctx->s));
    ctx->drop_graph();
    const int N = ctx->N, R = ctx->R;
    const std::vector<float> core_k = pad_core(core, N, ctx->Ru, R);
    for (int n = 0; n < N; ++n) {
      require(A[n] != nullptr, GT_ERR_ARG, "null factor");
      const int RS = row_stride(R);
      ctx->A[n].alloc(ctx->rows[n] * RS);
      GT_HIP(hipMemsetAsync(ctx->A[n].p, 0, ctx->rows[n] * RS * sizeof(float), ctx->s));
      upload_factor(ctx, n, A[n]);
      std::vector<float> gt = unfold_core(core_k.data(), N, R, n);
      ctx->Gt[n].alloc(gt.size());
      GT_HIP(hipMemcpyAsync(ctx->Gt[n].p, gt.data(), gt.size() * sizeof(float),
                            hipMemcpyHostToDevice, ctx->s));
      if (wide_bf16(ctx)) {
        std::vector<uint16_t> gb = bf16_core_table_wide(gt, R);
        ctx->Gb[n].alloc(gb.size());
        GT_HIP(hipMemcpyAsync(ctx->Gb[n].p, gb.data(), gb.size() * sizeof(uint16_t),
                              hipMemcpyHostToDevice, ctx->s));
      } else if (gram_mfma_supported(R)) {
        std::vector<uint16_t> gb = bf16_core_table(gt, N, R);
        ctx->Gb[n].alloc(gb.size());
        GT_HIP(hipMemcpyAsync(ctx->Gb[n].p, gb.data(), gb.size() * sizeof(uint16_t),
                              hipMemcpyHostToDevice, ctx->s));
      }
      GT_HIP(hipStreamSynchronize(ctx->s));
    }
    ctx->initialized = true;
    for (int n = 0; n < GT_MAX_ORDER; ++n) {
      ctx->mchunk[n] = 0;
      ctx->mvariant[n] = ctx->o.core_source;
    }
    if (ctx->o.autotune) autotune(ctx);
  });
}

int gt_autotune(gt_ctx *ctx) {
  return guarded(ctx, [&] { autotune(ctx); });
}

int gt_mode_update(gt_ctx *ctx, int mode) {
  return guarded(ctx, [&] { mode_update(ctx, mode, false, true); });
}

int gt_fit(gt_ctx *ctx, double *fit) {
  return guarded(ctx, [&] {
    double f = ctx->o.ref_fit ? compute_fit_ref(ctx) : compute_fit(ctx, ctx->err_part.p, ctx->scalar.p);
    if (fit) *fit = f;
  });
}

int gt_fit_noalloc(gt_ctx *ctx, double *fit) {
  return guarded(ctx, [&] {
    require(ctx->partitioned && ctx->initialized, GT_ERR_STATE, "partition and init first");
    const uint64_t need = fit_partials_needed(ctx) + 1;
    if (need > kScratchDoubles)
      throw GtError(GT_ERR_MEMORY,
                    "gt_fit_noalloc needs " + std::to_string(need) +
                        " doubles of scratch (2048 per staged segment, plus one); the static "
                        "scratch holds " + std::to_string(kScratchDoubles) +
                        ": use gt_fit for this out-of-core layout");
    double f = compute_fit(ctx, ctx->scratch, ctx->scratch + (kScratchDoubles - 1));
    if (fit) *fit = f;
  });
}

int gt_iteration(gt_ctx *ctx, double *fit) {
  return guarded(ctx, [&] {
    double f = iteration(ctx);
    if (fit) *fit = f;
  });
}

int gt_run(gt_ctx *ctx, int *iters_done, double *fits) {
  return guarded(ctx, [&] {
    int iter = 0;
    double p_fit = -1.0, fit = -1.0;
    for (;;) { // tucker.cuh:166-187
      fit = iteration(ctx);
      if (fits) fits[iter] = fit;
      ++iter;
      if (iter >= ctx->o.max_iters || (p_fit != -1.0 && std::fabs(p_fit - fit) <= ctx->o.tol))
        break;
      p_fit = fit;
    }
    if (iters_done) *iters_done = iter;
  });
}

int gt_get_factor(gt_ctx *ctx, int mode, float *out) {
  return guarded(ctx, [&] {
    require(ctx->initialized, GT_ERR_STATE, "not initialized");
    require(mode >= 0 && mode < ctx->N && out, GT_ERR_ARG, "bad argument");
    const int Ru = ctx->Ru, RS = row_stride(ctx->R);
    std::vector<float> pad((size_t)ctx->dims[mode] * RS);
    GT_HIP(hipMemcpyAsync(pad.data(), ctx->A[mode].p, pad.size() * sizeof(float),
                          hipMemcpyDeviceToHost, ctx->s));
    GT_HIP(hipStreamSynchronize(ctx->s));
    for (uint64_t i = 0; i < ctx->dims[mode]; ++i)
      std::memcpy(out + i * Ru, pad.data() + i * RS, Ru * sizeof(float));
  });
}

int gt_set_factor(gt_ctx *ctx, int mode, const float *in) {
  return guarded(ctx, [&] {
    require(ctx->initialized, GT_ERR_STATE, "not initialized");
    require(mode >= 0 && mode < ctx->N && in, GT_ERR_ARG, "bad argument");
    upload_factor(ctx, mode, in);
  });
}

int gt_get_delta(gt_ctx *ctx, int mode, float *out) {
  return guarded(ctx, [&] {
    require(ctx->partitioned && ctx->initialized, GT_ERR_STATE, "partition and init first");
    require(!ctx->streaming, GT_ERR_STATE, "the block COO is not resident (out-of-core mode)");
    require(mode >= 0 && mode < ctx->N && out, GT_ERR_ARG, "bad argument");
    DevBuf<float> d;
    d.alloc(std::max<uint64_t>(ctx->nnz * ctx->R, 1));
    CooArgs a = coo_args(ctx, mode);
    a.delta_out = d.p;
    launch_coo_contract(ctx->R, a, ctx->s);
    copy_delta_out(ctx, d.p, out);
  });
}

int gt_get_stream_delta(gt_ctx *ctx, int mode, float *delta, uint32_t *const *idx) {
  return guarded(ctx, [&] {
    require(ctx->partitioned && ctx->initialized, GT_ERR_STATE, "partition and init first");
    require(!ctx->streaming, GT_ERR_STATE, "the streams are not resident (out-of-core mode)");
    require(mode >= 0 && mode < ctx->N && delta, GT_ERR_ARG, "bad argument");
    DevBuf<float> d;
    d.alloc(std::max<uint64_t>(ctx->nnz * ctx->R, 1));
    ctx->dbg_delta = d.p;
    try {
      run_gram(ctx, mode);
    } catch (...) {
      ctx->dbg_delta = nullptr;
      throw;
    }
    ctx->dbg_delta = nullptr;
    copy_delta_out(ctx, d.p, delta);
    for (int k = 0; idx && k < ctx->N; ++k)
      if (idx[k])
        GT_HIP(hipMemcpyAsync(idx[k], ctx->st_idx[mode][k].p, ctx->nnz * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, ctx->s));
    GT_HIP(hipStreamSynchronize(ctx->s));
  });
}

int gt_get_recon(gt_ctx *ctx, float *out) {
  return guarded(ctx, [&] {
    require(ctx->partitioned && ctx->initialized, GT_ERR_STATE, "partition and init first");
    require(!ctx->streaming, GT_ERR_STATE, "the block COO is not resident (out-of-core mode)");
    require(out, GT_ERR_ARG, "bad argument");
    DevBuf<float> x;
    x.alloc(ctx->nnz);
    CooArgs a = coo_args(ctx, ctx->N - 1);
    a.xhat_out = x.p;
    launch_coo_contract(ctx->R, a, ctx->s);
    GT_HIP(hipMemcpyAsync(out, x.p, ctx->nnz * sizeof(float), hipMemcpyDeviceToHost, ctx->s));
    GT_HIP(hipStreamSynchronize(ctx->s));
  });
}

int gt_get_gram(gt_ctx *ctx, int mode, double *B, double *C) {
  return guarded(ctx, [&] {
    require(ctx->partitioned && ctx->initialized, GT_ERR_STATE, "partition and init first");
    require(mode >= 0 && mode < ctx->N && B && C, GT_ERR_ARG, "bad argument");
    run_gram(ctx, mode);
    const int R = ctx->R, Ru = ctx->Ru, W = gram_width(R);
    const uint64_t rows = ctx->rows[mode];
    std::vector<double> g(rows * W);
    std::vector<uint64_t> rp(rows + 1);
    GT_HIP(hipMemcpyAsync(g.data(), ctx->gram.p, rows * W * sizeof(double),
                          hipMemcpyDeviceToHost, ctx->s));
    GT_HIP(hipMemcpyAsync(rp.data(), ctx->row_ptr[mode].p, (rows + 1) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, ctx->s));
    GT_HIP(hipStreamSynchronize(ctx->s));
    for (uint64_t r = 0; r < rows; ++r) {
      const bool empty = rp[r + 1] == rp[r];
      const double *gr = g.data() + r * W;
      int q = 0; // packed upper triangle of the kernel rank; the first Ru rows and columns out
      for (int p = 0; p < R; ++p)
        for (int p2 = p; p2 < R; ++p2, ++q) {
          if (p2 >= Ru) continue;
          const double v = empty ? 0.0 : gr[q];
          B[(r * Ru + p) * Ru + p2] = v + (p == p2 ? ctx->o.lambda : 0.0);
          B[(r * Ru + p2) * Ru + p] = v + (p == p2 ? ctx->o.lambda : 0.0);
        }
      for (int p = 0; p < Ru; ++p) C[r * Ru + p] = empty ? 0.0 : gr[R * (R + 1) / 2 + p];
    }
  });
}

int gt_get_timings(gt_ctx *ctx, gt_timings *t) {
  return guarded(ctx, [&] {
    require(t != nullptr, GT_ERR_ARG, "null");
    GT_HIP(hipStreamSynchronize(ctx->s));
    ctx->collect_mode_timings();
    ctx->drain_profile();
    *t = ctx->tm;
  });
}

int gt_synchronize(gt_ctx *ctx) {
  return guarded(ctx, [&] { GT_HIP(hipStreamSynchronize(ctx->s)); });
}

} // extern "C"
