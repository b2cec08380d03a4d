// tucker.cpp -- TuckerDecomposition (tucker.cuh:30-190) on the GPU engine:
// the same initialisation order of the factors and the core from glibc
// rand(), the same Gauss-Seidel loop, convergence test and log lines.
#include <omp.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "gputucker.hpp"

namespace supertensor {
namespace gputucker {

namespace {
void check(gt_ctx *e, int rc, const char *what) {
  if (rc != GT_OK) throw std::runtime_error(std::string(what) + " failed: " + gt_last_error(e));
}

// Text matrices, 9 significant digits so every float reads back exactly.
void write_matrix(const std::string &path, const float *m, uint64_t rows, int cols) {
  std::ofstream f(path);
  f << std::setprecision(9);
  for (uint64_t r = 0; r < rows; ++r) {
    for (int c = 0; c < cols; ++c) f << m[r * cols + c] << (c + 1 < cols ? '\t' : '\n');
  }
  if (!f) throw std::runtime_error("[ERROR] cannot write \"" + path + "\"");
}

// Reads a rows x cols matrix written by write_matrix (whitespace separated).
void read_matrix(const std::string &path, float *m, uint64_t rows, int cols) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("[ERROR] cannot open \"" + path + "\" to resume");
  for (uint64_t i = 0; i < rows * (uint64_t)cols; ++i) {
    if (!(f >> m[i]))
      throw std::runtime_error("[ERROR] \"" + path + "\" holds fewer than " +
                               std::to_string(rows) + " x " + std::to_string(cols) + " values");
  }
  float extra;
  if (f >> extra) throw std::runtime_error("[ERROR] \"" + path + "\" holds more values than expected");
}
} // namespace

void TuckerDecomposition(tensor_t *t, const CommandLineOptions &o) {
  const int N = t->order, R = o.get_rank();
  if (o.get_gpu_count() != 1)
    printf("[WARNING] this engine runs on one GPU; using device %d\n", o.get_device());
  // The factors and the core are drawn before the GPU runtime starts: its
  // threads may call rand() too, which would shift the sequence.
  // tucker.cuh:84-97: factors by axis, then part, then row, then column; the
  // rows of one axis are visited in global order whatever the partition.
  printf("\t... Make the factor matrices\n");
  std::vector<std::vector<float>> A(N);
  std::vector<const float *> Ap(N);
  uint64_t core_n = 1;
  for (int a = 0; a < N; ++a) core_n *= (uint64_t)R;
  std::vector<float> core(core_n);
  const std::string &resume = o.get_resume_prefix();
  for (int a = 0; a < N; ++a) {
    A[a].resize((uint64_t)t->dims[a] * R);
    if (resume.empty()) {
      printf("\t\t... Fill the factor matrix [%d]\n", a);
      for (uint64_t i = 0; i < A[a].size(); ++i) A[a][i] = (float)frand<double>(0, 1);
    } else {
      printf("\t\t... Read the factor matrix [%d] from %s_factor%d.txt\n", a, resume.c_str(), a);
      read_matrix(resume + "_factor" + std::to_string(a) + ".txt", A[a].data(), t->dims[a], R);
    }
    Ap[a] = A[a].data();
  }
  if (resume.empty()) {
    printf("\t... Make the core tensor\n"); // tucker.cuh:118-119, linear order
    for (uint64_t i = 0; i < core_n; ++i) core[i] = (float)frand<double>(0, 1);
  } else {
    printf("\t... Read the core tensor from %s_core.txt\n", resume.c_str());
    read_matrix(resume + "_core.txt", core.data(), core_n / R, R);
  }
  if (!o.get_dump_init_prefix().empty()) { // the draws above, for checking; no GPU
    for (int a = 0; a < N; ++a)
      write_matrix(o.get_dump_init_prefix() + "_factor" + std::to_string(a) + ".txt", A[a].data(),
                   t->dims[a], R);
    write_matrix(o.get_dump_init_prefix() + "_core.txt", core.data(), core_n / R, R);
    printf("DONE\n");
    return;
  }
  gt_opts opts;
  gt_default_opts(&opts);
  opts.device = o.get_device();
  opts.order = N;
  opts.rank = R;
  opts.lambda = o.get_lambda();
  opts.max_iters = o.get_iterations();
  opts.tol = o.get_tol();
  opts.precision = o.get_precision();
  opts.device_budget = o.get_device_budget();
  opts.ref_fit = o.get_ref_fit() ? 1 : 0;
  opts.graph = o.get_graph() ? 1 : 0;
  opts.autotune = o.get_autotune() ? 1 : 0;
  for (size_t a = 0; a < o.get_partition().size(); ++a) opts.partition[a] = o.get_partition()[a];
  printf("\t... Initialize the GPU engine\n");
  if (gt_create(&t->engine, &opts) != GT_OK)
    throw std::runtime_error(std::string("gt_create failed: ") + gt_last_error(nullptr));
  CreateTensorBlocks(t, false);
  t->ToString();

  check(t->engine, gt_set_init(t->engine, Ap.data(), core.data()), "gt_set_init");

  int iter = 0;
  double p_fit = -1, fit = -1;
  std::ostringstream history; // iteration, fit, seconds
  history << std::setprecision(17);
  for (;;) { // tucker.cuh:166-187
    double itertime = omp_get_wtime();
    double steptime = itertime;
    if (o.get_fused_fit()) {
      // the N updates back to back, the fit from the last mode's row solves
      check(t->engine, gt_iteration(t->engine, &fit), "gt_iteration");
      gt_timings tm;
      check(t->engine, gt_get_timings(t->engine, &tm), "gt_get_timings");
      for (int n = 0; n < N; ++n) {
        printf("[ Update factor matrix %d ]\n", n);
        printf("\t- Elapsed time for Computing Delta and B, C (fused): %lf\n", tm.ms_gram[n] / 1e3);
        printf("\t- row-wise update TIME : %lf\n", tm.ms_solve[n] / 1e3);
      }
      printf("Factor Time : %lf\n", omp_get_wtime() - steptime);
      printf("Recon Time : %lf\n\n", tm.ms_fit / 1e3);
    } else {
      for (int n = 0; n < N; ++n) {
        printf("[ Update factor matrix %d ]\n", n);
        check(t->engine, gt_mode_update(t->engine, n), "gt_mode_update");
        gt_timings tm;
        check(t->engine, gt_get_timings(t->engine, &tm), "gt_get_timings");
        printf("\t- Elapsed time for Computing Delta and B, C (fused): %lf\n", tm.ms_gram[n] / 1e3);
        printf("\t- row-wise update TIME : %lf\n", tm.ms_solve[n] / 1e3);
      }
      printf("Factor Time : %lf\n", omp_get_wtime() - steptime);
      steptime = omp_get_wtime();
      check(t->engine, gt_fit(t->engine, &fit), "gt_fit");
      printf("Recon Time : %lf\n\n", omp_get_wtime() - steptime);
    }
    ++iter;
    const double elapsed = omp_get_wtime() - itertime;
    history << iter << '\t' << fit << '\t' << elapsed << '\n';
    printf("iter %d\t Fit: %g\n", iter, fit);
    printf("iter%d :      Fit : %lf\tElapsed Time : %lf\n\n", iter, fit, elapsed);
    if (iter >= o.get_iterations() || (p_fit != -1 && std::fabs(p_fit - fit) <= o.get_tol())) break;
    p_fit = fit;
  }
  if (!o.get_output_prefix().empty()) {
    for (int a = 0; a < N; ++a) {
      check(t->engine, gt_get_factor(t->engine, a, A[a].data()), "gt_get_factor");
      write_matrix(o.get_output_prefix() + "_factor" + std::to_string(a) + ".txt", A[a].data(),
                   t->dims[a], R);
    }
    write_matrix(o.get_output_prefix() + "_core.txt", core.data(), core_n / R, R);
    std::ofstream(o.get_output_prefix() + "_fit.txt") << history.str();
  }
  printf("DONE\n");
}

} // namespace gputucker
} // namespace supertensor
