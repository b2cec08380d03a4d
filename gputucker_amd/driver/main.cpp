// main.cpp -- the drop-in executable for main.cu:9-48:
//   GPUTucker -i <input.tns> -o <order> [-r rank] [-g gpus] [extra options]
#include <iostream>

#include "gputucker.hpp"

int main(int argc, char *argv[]) {
  using namespace supertensor::gputucker;
  CommandLineOptions options;
  CommandLineOptions::ReturnStatus ret = options.Parse(argc, argv);
  if (ret == CommandLineOptions::OPTS_HELP) return 0;
  if (ret != CommandLineOptions::OPTS_SUCCESS) {
    std::cout << "ERROR - problem with options." << std::endl;
    return 1;
  }
  std::cout << options.get_input_path() << std::endl;
  printf("Values are float type.\n");
  try {
    tensor_t *input = new tensor_t((unsigned short)options.get_order());
    if (!options.get_cache() || !LoadCache(options.get_input_path(), input)) {
      ParseFromFile<tensor_t>(options.get_input_path(), &input);
      if (options.get_cache()) SaveCache(options.get_input_path(), input);
    }
    if (options.get_parse_only()) { // summary for scripts and tests
      double ss = 0.0;
      for (float v : input->coo_values) ss += (double)v * v;
      printf("parsed order %d nnz %lu dims", input->order, (unsigned long)input->nnz_count);
      for (int a = 0; a < input->order; ++a) printf(" %u", input->dims[a]);
      printf(" sumsq %.10g\n", ss);
      delete input;
      return 0;
    }
    TuckerDecomposition(input, options);
    delete input;
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
