// tensor.cpp -- Block/Tensor containers, the coordinate-file parser and
// CreateTensorBlocks on the GPU engine.
#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "gputucker.hpp"

namespace supertensor {
namespace gputucker {

template <typename T> T frand(T x, T y) {
  return ((y - x) * (static_cast<T>(rand()) / RAND_MAX)) + x; // helper.hpp:99
}
template double frand<double>(double, double);
template float frand<float>(float, float);

// ------------------------------------------------------------------ Block
template <typename I, typename V>
Block<I, V>::Block(uint64_t id, const index_t *coord, unsigned short ord, const index_t *bdims,
                   uint64_t nnz)
    : order(ord), nnz_count(nnz), _block_id(id) {
  dims = new index_t[order];
  _base_dims = new index_t[order];
  _block_coord = new index_t[order];
  for (int a = 0; a < order; ++a) {
    dims[a] = bdims[a];
    _block_coord[a] = coord[a];
    _base_dims[a] = bdims[a] * coord[a]; // block.tpp:62-66
  }
}

template <typename I, typename V> Block<I, V>::~Block() {
  delete[] dims;
  delete[] _base_dims;
  delete[] _block_coord;
  for (int a = 0; a < order; ++a) {
    delete[] indices[a];
    delete[] where_nnz[a];
    delete[] count_nnz[a];
  }
  delete[] values;
}

template <typename I, typename V> void Block<I, V>::ToString() const {
  printf("********** BLOCK[%lu] Information **********\n", (unsigned long)_block_id);
  printf("Block coord: ");
  for (int a = 0; a < order; ++a) printf("[%lu]", (unsigned long)_block_coord[a]);
  printf("\nBlock order: %d\nBlock dims: ", order);
  for (int a = 0; a < order; ++a) printf("%lu%s", (unsigned long)dims[a], a + 1 < order ? " X " : "\n");
  printf("# nnzs: %lu\n", (unsigned long)nnz_count);
}

template class Block<uint32_t, float>;

// ------------------------------------------------------------------ Tensor
template <typename B> Tensor<B>::Tensor(unsigned short ord) : order(ord) {
  if (order < 1) throw std::runtime_error("[ERROR] Tensor order should be larger than 1.");
  dims = new index_t[order]();
  partition_dims = new index_t[order]();
  block_dims = new index_t[order]();
}

template <typename B> Tensor<B>::~Tensor() {
  if (blocks) {
    for (uint64_t b = 0; b < block_count; ++b) delete blocks[b];
    delete[] blocks;
  }
  delete[] dims;
  delete[] partition_dims;
  delete[] block_dims;
  if (engine) gt_destroy(engine);
}

template <typename B> void Tensor<B>::ToString() const { // tensor.tpp:215-229
  for (int a = 0; a < order; ++a) printf("Max. dim[%d] = %lu\n", a, (unsigned long)dims[a]);
  printf("# nnzs: %lu\n", (unsigned long)nnz_count);
  for (int a = 0; a < order; ++a)
    printf("Partition dim[%d] = %lu\n", a, (unsigned long)partition_dims[a]);
  printf("# blocks: %lu\n", (unsigned long)block_count);
  for (int a = 0; a < order; ++a) printf("Block dim[%d] = %lu\n", a, (unsigned long)block_dims[a]);
}

template class Tensor<block_t>;

// ------------------------------------------------------------------ parser
namespace {
inline bool is_delim(char c) { return c == ' ' || c == '\t'; }
} // namespace

// Semantics of _ReadData (tensor_manager.tpp:46-191): a nonzero per
// '\n'-terminated line (an unterminated last line is not read, as the
// reference counts '\n'); tokens split on spaces and tabs; the first `order`
// tokens are base-1 indices read like strtoull, the next one (if any) is the
// value read like std::stod, and a line without one draws frand(0, 1); an
// index below 1 throws; dims are the largest index seen per mode.
// Differences: blank lines are skipped and lines with fewer than `order`
// indices throw (the reference reads past them); values missing from a line
// are drawn in file order after the parallel parse (the reference draws
// them inside its OpenMP region, in no fixed order).
template <> bool ParseFromFile<tensor_t>(const std::string &file_name, tensor_t **tp) {
  tensor_t *t = *tp;
  const int order = t->order;
  int fd = open(file_name.c_str(), O_RDONLY);
  if (fd < 0)
    throw std::runtime_error("[ERROR] Canot open file \"" + file_name + "\" for reading...");
  struct stat st;
  fstat(fd, &st);
  const size_t len = (size_t)st.st_size;
  printf("Input Tensor Size (COO) \t: %zu bytes\n", len);
  const char *buf = len ? (const char *)mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0) : "";
  if (buf == MAP_FAILED) {
    close(fd);
    throw std::runtime_error("[ERROR] mmap failed on \"" + file_name + "\"");
  }
  const int T = omp_get_max_threads();
  // chunk boundaries at line starts
  std::vector<size_t> start(T + 1, len);
  start[0] = 0;
  for (int i = 1; i < T; ++i) {
    size_t p = len * (size_t)i / T;
    while (p < len && p > 0 && buf[p - 1] != '\n') ++p;
    start[i] = std::max(p, start[i - 1]);
  }
  std::vector<uint64_t> lines(T + 1, 0);
#pragma omp parallel num_threads(T)
  {
    int tid = omp_get_thread_num();
    uint64_t c = 0;
    for (size_t p = start[tid]; p < start[tid + 1]; ++p) c += buf[p] == '\n';
    lines[tid + 1] = c;
  }
  for (int i = 0; i < T; ++i) lines[i + 1] += lines[i];
  const uint64_t max_lines = lines[T];
  for (int a = 0; a < order; ++a) t->coo_indices[a].assign(max_lines, 0);
  t->coo_values.assign(max_lines, 0.f);
  std::vector<uint8_t> has_val(max_lines, 1), keep(max_lines, 0);
  // per-thread results, merged after the parallel region (each thread keeps
  // its running max / min in locals: shared per-thread slots next to each
  // other would put every token's update on a cache line the others write)
  std::vector<std::vector<uint32_t>> tmax(T, std::vector<uint32_t>(order, 0));
  std::vector<uint64_t> tmin(T, UINT64_MAX);
  std::vector<int> bad(T, 0);
#pragma omp parallel num_threads(T)
  {
    const int tid = omp_get_thread_num();
    uint32_t lmax[constants::kMaxOrder] = {0};
    uint64_t lmin = UINT64_MAX;
    int lbad = 0;
    uint64_t li = lines[tid];
    size_t p = start[tid];
    const size_t end = start[tid + 1];
    while (p < end) {
      size_t q = p;
      while (q < end && buf[q] != '\n') ++q;
      if (q >= end) break; // unterminated tail
      size_t s = p;
      int axis = 0;
      bool valued = false;
      while (s < q) {
        while (s < q && is_delim(buf[s])) ++s;
        if (s >= q) break;
        size_t e = s;
        while (e < q && !is_delim(buf[e])) ++e;
        if (axis < order) {
          unsigned long long v = 0;
          std::from_chars(buf + s, buf + e, v); // strtoull: leading digits, else 0
          lmax[axis] = std::max<uint32_t>(lmax[axis], (uint32_t)v);
          lmin = std::min<uint64_t>(lmin, v);
          t->coo_indices[axis][li] = (uint32_t)v - 1;
          ++axis;
        } else {
          // std::stod, then value_t: from_chars rounds like strtod; forms it
          // does not take whole ('+', hex, inf/nan spellings) go to strtod
          double d = 0.0;
          const auto r = std::from_chars(buf + s, buf + e, d);
          if (r.ec != std::errc() || r.ptr != buf + e) {
            std::string tok(buf + s, e - s);
            d = std::strtod(tok.c_str(), nullptr);
          }
          t->coo_values[li] = (float)d;
          valued = true;
          break;
        }
        s = e;
      }
      if (axis == 0) {
        keep[li] = 0; // blank line
      } else if (axis < order) {
        lbad = 1;
      } else {
        keep[li] = 1;
        has_val[li] = valued;
      }
      ++li;
      p = q + 1;
    }
    for (int ax = 0; ax < order; ++ax) tmax[tid][ax] = lmax[ax];
    tmin[tid] = lmin;
    bad[tid] = lbad;
  }
  if (len) munmap((void *)buf, len);
  close(fd);
  for (int i = 0; i < T; ++i)
    if (bad[i]) throw std::runtime_error("[ERROR] a line has fewer than `order` indices");
  uint64_t n = 0;
  for (uint64_t li = 0; li < max_lines; ++li) {
    if (!keep[li]) continue;
    if (!has_val[li]) t->coo_values[li] = (float)frand<double>(0, 1);
    if (n != li) {
      for (int a = 0; a < order; ++a) t->coo_indices[a][n] = t->coo_indices[a][li];
      t->coo_values[n] = t->coo_values[li];
    }
    ++n;
  }
  for (int a = 0; a < order; ++a) t->coo_indices[a].resize(n);
  t->coo_values.resize(n);
  uint64_t mn = UINT64_MAX;
  for (int i = 0; i < T; ++i) mn = std::min(mn, tmin[i]);
  if (n && mn < 1)
    throw std::runtime_error("We note that input tensors must follow base-1 indexing");
  for (int a = 0; a < order; ++a) {
    uint32_t m = 0;
    for (int i = 0; i < T; ++i) m = std::max(m, tmax[i][a]);
    t->dims[a] = m;
  }
  t->nnz_count = n;
  return true;
}

// ------------------------------------------------------------ COO cache
namespace {
constexpr char kCacheMagic[8] = {'G', 'T', 'C', 'O', 'O', '1', 0, 0};
std::string cache_path(const std::string &input) { return input + ".gtcoo"; }
} // namespace

bool LoadCache(const std::string &input, tensor_t *t) {
  const std::string path = cache_path(input);
  struct stat si, sc;
  if (stat(input.c_str(), &si) != 0 || stat(path.c_str(), &sc) != 0) return false;
  if (sc.st_mtime < si.st_mtime) return false; // the input changed since
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  char magic[8];
  uint32_t order = 0;
  uint64_t nnz = 0;
  bool ok = fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kCacheMagic, 8) == 0 &&
            fread(&order, sizeof order, 1, f) == 1 && order == t->order &&
            fread(&nnz, sizeof nnz, 1, f) == 1 &&
            fread(t->dims, sizeof(uint32_t), order, f) == order;
  for (uint32_t a = 0; ok && a < order; ++a) {
    t->coo_indices[a].resize(nnz);
    ok = fread(t->coo_indices[a].data(), sizeof(uint32_t), nnz, f) == nnz;
  }
  if (ok) {
    t->coo_values.resize(nnz);
    ok = fread(t->coo_values.data(), sizeof(float), nnz, f) == nnz;
  }
  fclose(f);
  if (!ok) return false;
  t->nnz_count = nnz;
  printf("Input Tensor read from the binary cache %s\n", path.c_str());
  return true;
}

void SaveCache(const std::string &input, const tensor_t *t) {
  const std::string path = cache_path(input), tmp = path + ".tmp";
  FILE *f = fopen(tmp.c_str(), "wb");
  if (!f) return; // a cache is an optimisation: no cache, no error
  const uint32_t order = t->order;
  const uint64_t nnz = t->nnz_count;
  bool ok = fwrite(kCacheMagic, 1, 8, f) == 8 && fwrite(&order, sizeof order, 1, f) == 1 &&
            fwrite(&nnz, sizeof nnz, 1, f) == 1 &&
            fwrite(t->dims, sizeof(uint32_t), order, f) == order;
  for (uint32_t a = 0; ok && a < order; ++a)
    ok = fwrite(t->coo_indices[a].data(), sizeof(uint32_t), nnz, f) == nnz;
  if (ok) ok = fwrite(t->coo_values.data(), sizeof(float), nnz, f) == nnz;
  ok = (fclose(f) == 0) && ok;
  if (ok) ok = rename(tmp.c_str(), path.c_str()) == 0;
  if (!ok) unlink(tmp.c_str());
}

// ------------------------------------------------------- CreateTensorBlocks
static void check(gt_ctx *e, int rc, const char *what) {
  if (rc != GT_OK)
    throw std::runtime_error(std::string(what) + " failed: " + gt_last_error(e));
}

void CreateTensorBlocks(tensor_t *t, bool with_host_blocks) {
  printf("... 1) Creating tensor blocks\n");
  const int N = t->order;
  std::vector<const uint32_t *> idx(N);
  for (int a = 0; a < N; ++a) idx[a] = t->coo_indices[a].data();
  check(t->engine, gt_set_coo(t->engine, t->nnz_count, idx.data(), t->coo_values.data(), t->dims),
        "gt_set_coo");
  check(t->engine, gt_partition(t->engine, nullptr, nullptr), "gt_partition");
  gt_partition_info info;
  check(t->engine, gt_get_partition_info(t->engine, &info), "gt_get_partition_info");
  for (int a = 0; a < N; ++a) {
    t->partition_dims[a] = info.partition[a];
    t->block_dims[a] = info.block_dims[a];
  }
  t->block_count = info.block_count;
  t->norm = (float)info.norm;
  std::vector<uint64_t> bn(info.block_count);
  check(t->engine, gt_get_block_nnz(t->engine, bn.data()), "gt_get_block_nnz");
  printf("# empty blocks: %lu / %lu\n", (unsigned long)info.empty_blocks,
         (unsigned long)info.block_count);
  if (!with_host_blocks) return;
  t->blocks = new block_t *[t->block_count];
  for (uint64_t b = 0; b < t->block_count; ++b) {
    uint32_t coord[constants::kMaxOrder];
    uint64_t rem = b;
    for (int a = N - 1; a >= 0; --a) {
      coord[a] = (uint32_t)(rem % t->partition_dims[a]);
      rem /= t->partition_dims[a];
    }
    block_t *blk = new block_t(b, coord, (unsigned short)N, t->block_dims, bn[b]);
    std::vector<uint32_t *> bi(N);
    for (int a = 0; a < N; ++a) {
      blk->indices[a] = new uint32_t[std::max<uint64_t>(bn[b], 1)];
      blk->where_nnz[a] = new uint32_t[std::max<uint64_t>(bn[b], 1)];
      blk->count_nnz[a] = new uint64_t[(uint64_t)t->block_dims[a] + 1];
      bi[a] = blk->indices[a];
    }
    blk->values = new float[std::max<uint64_t>(bn[b], 1)];
    check(t->engine, gt_get_block(t->engine, b, bi.data(), blk->values), "gt_get_block");
    for (int a = 0; a < N; ++a)
      check(t->engine, gt_partition_dump(t->engine, b, a, blk->count_nnz[a], blk->where_nnz[a]),
            "gt_partition_dump");
    t->blocks[b] = blk;
  }
}

} // namespace gputucker
} // namespace supertensor
