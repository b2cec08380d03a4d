// gputucker.hpp -- the reference's host-side class API (namespace
// supertensor::gputucker) rebuilt on the C ABI of include/gputucker_hip.h.
//
//   CommandLineOptions    cmdline_opts.hpp:19-129 (same flags; no Boost)
//   Block / Tensor        block.hpp:166-180, tensor.hpp:166-182 (same members)
//   ParseFromFile         tensor_manager.hpp:16-17 (base-1, tab/space COO)
//   CreateTensorBlocks    tensor_manager.tpp:201-302 (runs on the GPU)
//   TuckerDecomposition   tucker.cuh:30-190 (same loop, same log lines)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gputucker_hip.h"

namespace supertensor {
namespace gputucker {

namespace constants {
constexpr int kMaxOrder{8};        // constants.hpp:16
constexpr int kMaxIteration{3};    // constants.hpp:17
constexpr double kLambda{0.0001f}; // constants.hpp:18 (float literal, as in the reference)
} // namespace constants

namespace enums {
enum PartitionTypes { kDimensionPartition, kNonzeroPartition, kPartitionTypeCount };
}

// helper.hpp:99: x + (y - x) * rand() / RAND_MAX, glibc rand() left unseeded.
template <typename T> T frand(T x, T y);

class CommandLineOptions {
public:
  enum ReturnStatus { OPTS_SUCCESS, OPTS_HELP, OPTS_FAILURE };
  CommandLineOptions();
  ReturnStatus Parse(int argc, char *argv[]);
  std::string get_input_path() const { return _input_path; }
  int get_order() const { return _order; }
  int get_rank() const { return _rank; }
  int get_gpu_count() const { return _gpu_count; }
  // options beyond the reference's four, defaults as the reference compiles them in
  int get_iterations() const { return _iters; }
  double get_lambda() const { return _lambda; }
  double get_tol() const { return _tol; }
  const std::vector<uint32_t> &get_partition() const { return _partition; }
  uint64_t get_device_budget() const { return _budget; }
  int get_precision() const { return _precision; }
  int get_device() const { return _device; }
  const std::string &get_output_prefix() const { return _output; }
  bool get_parse_only() const { return _parse_only; }
  bool get_cache() const { return _cache; }
  const std::string &get_resume_prefix() const { return _resume; }
  bool get_fused_fit() const { return _fused_fit; }
  bool get_ref_fit() const { return _ref_fit; }
  bool get_graph() const { return _graph; }
  bool get_autotune() const { return _autotune; }
  const std::string &get_dump_init_prefix() const { return _dump_init; }
  void PrintHelp() const;

private:
  bool ValidateFile();
  std::string _input_path;
  int _order = 0;
  int _rank = 10;
  int _gpu_count = 1;
  int _iters = constants::kMaxIteration;
  double _lambda = constants::kLambda;
  double _tol = constants::kLambda;
  std::vector<uint32_t> _partition;
  uint64_t _budget = 0;
  int _precision = GT_PREC_FP32;
  int _device = 0;
  std::string _output;
  bool _parse_only = false;
  bool _cache = false;       // read/write a binary COO copy beside the input
  std::string _resume;       // start from factors and core written by --output
  bool _fused_fit = false;   // fit from the row solves (gt_iteration)
  bool _ref_fit = false;     // fit in the reference's float arithmetic (gt_opts.ref_fit)
  bool _graph = false;       // replay in-core iterations as a HIP graph (gt_opts.graph)
  bool _autotune = false;    // the engine's autotuner at initialisation (gt_opts.autotune)
  std::string _dump_init;    // write the initial factors and core, then stop (no GPU)
};

template <typename IndexType, typename ValueType> class Block {
public:
  using index_t = IndexType;
  using value_t = ValueType;
  Block(uint64_t id, const index_t *coord, unsigned short order, const index_t *dims,
        uint64_t nnz);
  ~Block();
  Block(const Block &) = delete;
  Block &operator=(const Block &) = delete;
  uint64_t get_block_id() const { return _block_id; }
  index_t *get_block_coord() { return _block_coord; }
  void ToString() const;

  unsigned short order;
  index_t *dims;
  uint64_t nnz_count;
  value_t *values = nullptr;
  index_t *indices[constants::kMaxOrder] = {};
  index_t *where_nnz[constants::kMaxOrder] = {};
  uint64_t *count_nnz[constants::kMaxOrder] = {};

private:
  index_t *_base_dims;
  uint64_t _block_id;
  index_t *_block_coord;
};

template <typename BlockType> class Tensor {
public:
  using block_t = BlockType;
  using index_t = typename block_t::index_t;
  using value_t = typename block_t::value_t;
  explicit Tensor(unsigned short order);
  ~Tensor();
  Tensor(const Tensor &) = delete;
  Tensor &operator=(const Tensor &) = delete;
  void ToString() const;

  unsigned short order;
  index_t *dims;
  uint64_t nnz_count = 0;
  value_t norm = 0;
  index_t *partition_dims;
  index_t *block_dims;
  uint64_t block_count = 0;
  block_t **blocks = nullptr;

  // the input COO as _ReadData leaves it (tensor_manager.tpp:174-180), 0-based
  std::vector<index_t> coo_indices[constants::kMaxOrder];
  std::vector<float> coo_values;
  gt_ctx *engine = nullptr; // set by CreateTensorBlocks
};

using block_t = Block<uint32_t, float>;
using tensor_t = Tensor<block_t>;

// Reads a base-1 tab/space separated coordinate file (README.md:78-97).
// Throws std::runtime_error like tensor_manager.tpp:21-25,166-168.
template <typename TensorType> bool ParseFromFile(const std::string &file_name, TensorType **tensor);

// Binary COO cache of a parsed input: `<input>.gtcoo` holds the order, the
// nonzero count, the dims, the 0-based indices and the values as the parser
// left them.  LoadCache returns false when the cache is missing, older than
// the input or of another order; SaveCache writes it.
bool LoadCache(const std::string &input, tensor_t *tensor);
void SaveCache(const std::string &input, const tensor_t *tensor);

// Partitions on the GPU (engine already created) and, when with_host_blocks,
// copies the blocks and their per-mode row grouping back into tensor->blocks.
void CreateTensorBlocks(tensor_t *tensor, bool with_host_blocks);

// The ALS driver of tucker.cuh:30-190 on the GPU engine.
void TuckerDecomposition(tensor_t *tensor, const CommandLineOptions &opts);

} // namespace gputucker
} // namespace supertensor
