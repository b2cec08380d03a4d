// eigen_ref.cpp -- exposes the reference's own matrix inverse (the vendored
// Eigen 3.3.91 under lib/Eigen, MatrixXd::inverse() as update.cuh:173-180
// calls it) through a C symbol, so the oracle's LU can be checked against it.
// Built into oracle/_ref only; test infrastructure.
#include <Eigen/Dense>

extern "C" int eigenref_inverse(int n, const double *B, double *out) {
  Eigen::MatrixXd BB(n, n);
  for (int k = 0; k < n; ++k)
    for (int l = 0; l < n; ++l) BB(k, l) = B[k * n + l];
  Eigen::MatrixXd inv = BB.inverse();
  for (int k = 0; k < n; ++k)
    for (int l = 0; l < n; ++l) out[k * n + l] = inv(k, l);
  return 0;
}
