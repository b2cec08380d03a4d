// cmdline_opts.cpp -- CommandLineOptions with the reference's flags
// (cmdline_opts.cpp:27-32: -h/--help, -i/--input, -o/--order, -r/--rank
// default 10, -g/--gpus default 1; input and order mandatory, the input must
// be a regular file, cmdline_opts.cpp:60-75,102-107) parsed by hand, plus
// options for what the reference fixes at compile time.
#include <sys/stat.h>

#include <cerrno>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>

#include "gputucker.hpp"

namespace supertensor {
namespace gputucker {

CommandLineOptions::CommandLineOptions() = default;

void CommandLineOptions::PrintHelp() const {
  std::cout << "Program Options:\n"
               "  -h [ --help ]                 Display help menu.\n"
               "  -i [ --input ] arg            Input tensor path\n"
               "  -o [ --order ] arg            Order\n"
               "  -r [ --rank ] arg (=10)       Rank (1..50; bf16 and bf16r run the MFMA delta at\n"
               "                                order 3 up to R = 15 (bf16r to 32), order 4 up to\n"
               "                                10, order 5 up to 8; elsewhere the FP32 delta)\n"
               "  -g [ --gpus ] arg (=1)        The number of GPUs\n"
               "  --iters arg (=3)              Maximum ALS iterations\n"
               "  --lambda arg (=0.0001)        Ridge regularisation\n"
               "  --tol arg (=0.0001)           Stop when |fit - previous fit| <= tol\n"
               "  --partition P1,...,PN         Force the blocks per mode (default: planner)\n"
               "  --device-budget bytes         Memory the planner may assume (default: free)\n"
               "  --precision fp32|bf16|bf16r   Delta contraction precision (default fp32)\n"
               "  --device arg (=0)             HIP device ordinal\n"
               "  --output prefix               Write factors, core and fit history as text files\n"
               "  --resume prefix               Start from the factors and core of an --output run\n"
               "  --fused-fit                   Take each iteration's fit from the row solves\n"
               "  --ref-fit                     Fit in the reference's float arithmetic (serial sums)\n"
               "  --graph                       Replay each in-core iteration as a captured HIP graph\n"
               "                                (with --fused-fit)\n"
               "  --autotune                    Time each mode's chunk sizes and core sources first\n"
               "  --dump-init prefix            Write the initial factors and core, then exit (no GPU)\n"
               "  --cache                       Read/write a binary copy of the input (<input>.gtcoo)\n"
               "  --parse-only                  Parse the input, print its summary, exit (no GPU)\n";
}

namespace {
bool parse_int(const char *s, long long &out) {
  char *end = nullptr;
  errno = 0;
  long long v = std::strtoll(s, &end, 10);
  if (errno || end == s || *end) return false;
  out = v;
  return true;
}
bool parse_double(const char *s, double &out) {
  char *end = nullptr;
  errno = 0;
  double v = std::strtod(s, &end);
  if (errno || end == s || *end) return false;
  out = v;
  return true;
}
std::string trim(const std::string &s) { // boost::algorithm::trim
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}
} // namespace

CommandLineOptions::ReturnStatus CommandLineOptions::Parse(int argc, char *argv[]) {
  bool have_input = false, have_order = false, rank_set = false, gpus_set = false;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    std::string val;
    bool has_inline = false;
    if (a.rfind("--", 0) == 0) {
      size_t eq = a.find('=');
      if (eq != std::string::npos) {
        val = a.substr(eq + 1);
        a = a.substr(0, eq);
        has_inline = true;
      }
    }
    auto need = [&](std::string &v) -> bool {
      if (has_inline) {
        v = val;
        return true;
      }
      if (i + 1 >= argc) {
        std::cout << "[ERROR] Parsing error : the required argument for option '" << a
                  << "' is missing" << std::endl;
        return false;
      }
      v = argv[++i];
      return true;
    };
    long long iv;
    double dv;
    std::string v;
    if (a == "-h" || a == "--help") {
      PrintHelp();
      return OPTS_HELP;
    } else if (a == "-i" || a == "--input") {
      if (!need(v)) return OPTS_FAILURE;
      _input_path = trim(v);
      have_input = true;
    } else if (a == "-o" || a == "--order") {
      if (!need(v) || !parse_int(v.c_str(), iv)) return OPTS_FAILURE;
      _order = (int)iv;
      have_order = true;
    } else if (a == "-r" || a == "--rank") {
      if (!need(v) || !parse_int(v.c_str(), iv)) return OPTS_FAILURE;
      _rank = (int)iv;
      rank_set = true;
    } else if (a == "-g" || a == "--gpus") {
      if (!need(v) || !parse_int(v.c_str(), iv)) return OPTS_FAILURE;
      _gpu_count = (int)iv;
      gpus_set = true;
    } else if (a == "--iters") {
      if (!need(v) || !parse_int(v.c_str(), iv) || iv < 1) return OPTS_FAILURE;
      _iters = (int)iv;
    } else if (a == "--lambda") {
      if (!need(v) || !parse_double(v.c_str(), dv) || !(dv > 0)) return OPTS_FAILURE;
      _lambda = dv;
    } else if (a == "--tol") {
      if (!need(v) || !parse_double(v.c_str(), dv) || dv < 0) return OPTS_FAILURE;
      _tol = dv;
    } else if (a == "--partition") {
      if (!need(v)) return OPTS_FAILURE;
      _partition.clear();
      std::stringstream ss(v);
      std::string tok;
      while (std::getline(ss, tok, ',')) {
        if (!parse_int(tok.c_str(), iv) || iv < 1) return OPTS_FAILURE;
        _partition.push_back((uint32_t)iv);
      }
    } else if (a == "--device-budget") {
      if (!need(v) || !parse_int(v.c_str(), iv) || iv < 0) return OPTS_FAILURE;
      _budget = (uint64_t)iv;
    } else if (a == "--precision") {
      if (!need(v)) return OPTS_FAILURE;
      if (v == "fp32") _precision = GT_PREC_FP32;
      else if (v == "bf16") _precision = GT_PREC_BF16;
      else if (v == "bf16r") _precision = GT_PREC_BF16R;
      else return OPTS_FAILURE;
    } else if (a == "--device") {
      if (!need(v) || !parse_int(v.c_str(), iv) || iv < 0) return OPTS_FAILURE;
      _device = (int)iv;
    } else if (a == "--parse-only") {
      _parse_only = true;
    } else if (a == "--output") {
      if (!need(v)) return OPTS_FAILURE;
      _output = v;
    } else if (a == "--resume") {
      if (!need(v)) return OPTS_FAILURE;
      _resume = v;
    } else if (a == "--fused-fit") {
      _fused_fit = true;
    } else if (a == "--ref-fit") {
      _ref_fit = true;
    } else if (a == "--graph") {
      _graph = true;
    } else if (a == "--autotune") {
      _autotune = true;
    } else if (a == "--dump-init") {
      if (!need(v)) return OPTS_FAILURE;
      _dump_init = v;
    } else if (a == "--cache") {
      _cache = true;
    } else {
      std::cout << "[ERROR] Parsing error : unrecognised option '" << argv[i] << "'" << std::endl;
      return OPTS_FAILURE;
    }
  }
  if (!have_input) {
    std::cout << "[ERROR] Input file must be specified!!!" << std::endl;
    PrintHelp();
    return OPTS_FAILURE;
  }
  if (!ValidateFile()) return OPTS_FAILURE;
  if (!have_order) {
    std::cout << "[ERROR] Tensor order must be specified!!!" << std::endl;
    PrintHelp();
    return OPTS_FAILURE;
  }
  if (!_partition.empty() && (int)_partition.size() != _order) {
    std::cout << "[ERROR] --partition needs one value per mode" << std::endl;
    return OPTS_FAILURE;
  }
  if (rank_set) std::cout << "[WARNING] Default value for User-Value overwritten to " << _rank << std::endl;
  if (gpus_set) std::cout << "[WARNING] Default value for GPU count overwritten to " << _gpu_count << std::endl;
  return OPTS_SUCCESS;
}

bool CommandLineOptions::ValidateFile() {
  struct stat st;
  if (stat(_input_path.c_str(), &st) != 0 || !S_ISREG(st.st_mode)) {
    std::cout << "[ERROR] Input file provided does not exist [" << _input_path << "]" << std::endl;
    return false;
  }
  return true;
}

} // namespace gputucker
} // namespace supertensor
