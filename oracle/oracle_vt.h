/*
 * oracle_vt.h -- value-type-generic part of the CPU restatement, included by
 * oracle.c once per value_t (VT = double -> *_f64, VT = float -> *_f32).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Factor matrices are held as the reference holds them: factor_matrices[n][p]
 * is a block_dims[n] x R row-major slice (tucker.cuh:75-81); the slices of one
 * mode are contiguous here, so A[n] is (P_n * bd_n) x R and the row of a
 * nonzero is coord[n] * bd_n + local index.  The core is R^N values in linear
 * order, core coordinates row-major with the last axis fastest
 * (tucker.cuh:118-134).
 */

static inline void SFX(core_digits)(int order, int rank, uint64_t i, int *digits) {
  /* tucker.cuh:120-132 */
  for (int a = order - 1; a >= 0; --a) {
    digits[a] = (int)(i % (uint64_t)rank);
    i /= (uint64_t)rank;
  }
}

static inline uint64_t SFX(ipow)(int r, int n) {
  uint64_t v = 1;
  for (int i = 0; i < n; ++i) v *= (uint64_t)r;
  return v;
}

/* tensor_manager.tpp:263,289,293: NormX accumulated serially in value_t, input order. */
VT SFX(or_norm)(const VT *vals, uint64_t nnz) {
  VT s = 0;
  for (uint64_t e = 0; e < nnz; ++e) s += vals[e] * vals[e];
  return (VT)sqrt((double)s);
}

/* delta.cuh:62-90 (computing_delta_kernel body), one nonzero per iteration. */
void SFX(or_delta)(const or_tensor *t, const VT *const *A, const VT *G, int mode,
                   VT *delta_out) {
  const int N = t->order, R = t->rank;
  const uint64_t core_n = SFX(ipow)(R, N);
  int *cidx = (int *)malloc(core_n * (size_t)N * sizeof(int));
  for (uint64_t i = 0; i < core_n; ++i) SFX(core_digits)(N, R, i, cidx + i * (uint64_t)N);
  for (uint64_t b = 0; b < t->block_count; ++b) {
    uint32_t coord[OR_MAX_ORDER];
    or_block_coord(N, t->part, b, coord);
    const int64_t lo = (int64_t)t->block_off[b], hi = (int64_t)t->block_off[b + 1];
#pragma omp parallel for schedule(static)
    for (int64_t e = lo; e < hi; ++e) {
      VT tmp[64];
      uint64_t row[OR_MAX_ORDER];
      for (int r = 0; r < R; ++r) tmp[r] = 0;
      for (int a = 0; a < N; ++a)
        row[a] = (uint64_t)coord[a] * t->bdims[a] + t->lidx[(uint64_t)a * t->nnz + (uint64_t)e];
      for (uint64_t i = 0; i < core_n; ++i) {
        const int *d = cidx + i * (uint64_t)N;
        VT beta = G[i];
        for (int a = 0; a < N; ++a)
          if (a != mode) beta *= A[a][row[a] * (uint64_t)R + (uint64_t)d[a]];
        tmp[d[mode]] += beta;
      }
      for (int r = 0; r < R; ++r) delta_out[(uint64_t)e * R + r] = tmp[r];
    }
  }
  free(cidx);
}

/* update.cuh:35-101 (ComputingBC): B = lambda I + sum delta delta^T,
 * C = sum x delta; products in value_t, sums in double; blocks ascending,
 * rows ascending, nonzeros in where order.  B_out is (P_n*bd_n) x R x R,
 * C_out is (P_n*bd_n) x R. */
void SFX(or_bc)(const or_tensor *t, const VT *vals, const VT *delta, int mode, double lambda,
                double *B_out, double *C_out) {
  const int N = t->order, R = t->rank;
  const uint64_t bd = t->bdims[mode];
  const uint64_t rows = (uint64_t)t->part[mode] * bd;
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < (int64_t)rows; ++row) { /* :52-69 */
    for (int k = 0; k < R; ++k) {
      for (int l = 0; l < R; ++l)
        B_out[((uint64_t)row * R + k) * R + l] = (k == l) ? lambda : 0.0;
      C_out[(uint64_t)row * R + k] = 0.0;
    }
  }
  for (uint64_t b = 0; b < t->block_count; ++b) { /* :73-100 */
    uint32_t coord[OR_MAX_ORDER];
    or_block_coord(N, t->part, b, coord);
    const uint64_t part_base = (uint64_t)coord[mode] * bd;
    const uint64_t *rp = t->row_ptr[mode] + b * (bd + 1);
    const uint32_t *wh = t->where[mode] + t->block_off[b];
    const VT *bvals = vals + t->block_off[b];
    const VT *bdelta = delta + t->block_off[b] * (uint64_t)R;
#pragma omp parallel for schedule(dynamic)
    for (int64_t row = 0; row < (int64_t)bd; ++row) {
      double *Bp = B_out + (part_base + (uint64_t)row) * (uint64_t)R * R;
      double *Cp = C_out + (part_base + (uint64_t)row) * (uint64_t)R;
      for (uint64_t kk = rp[row]; kk < rp[row + 1]; ++kk) {
        const uint64_t pos = wh[kk];
        const VT x = bvals[pos];
        const VT *d = bdelta + pos * (uint64_t)R;
        for (int ii = 0; ii < R; ++ii) {
          const VT cach = d[ii];
          for (int jj = 0; jj < R; ++jj) Bp[ii * R + jj] += (VT)(cach * d[jj]);
          Cp[ii] += (VT)(cach * x);
        }
      }
    }
  }
}

/* update.cuh:167-194: a_i = C_i^T inv(B_i), the inverse by partial-pivot LU,
 * the dot product accumulated in value_t. */
void SFX(or_solve)(int rank, uint64_t rows, const double *B, const double *C, VT *A_out) {
  const int R = rank;
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < (int64_t)rows; ++row) {
    double inv[64 * 64];
    or_lu_inverse(R, B + (uint64_t)row * R * R, inv);
    for (int k = 0; k < R; ++k) {
      VT res = 0;
      for (int l = 0; l < R; ++l) res += C[(uint64_t)row * R + l] * inv[l * R + k];
      A_out[(uint64_t)row * R + k] = res;
    }
  }
}

/* reconstruction.cuh:58-81 (ComputingReconstructionKernel body). */
void SFX(or_recon)(const or_tensor *t, const VT *const *A, const VT *G, VT *xhat_out) {
  const int N = t->order, R = t->rank;
  const uint64_t core_n = SFX(ipow)(R, N);
  int *cidx = (int *)malloc(core_n * (size_t)N * sizeof(int));
  for (uint64_t i = 0; i < core_n; ++i) SFX(core_digits)(N, R, i, cidx + i * (uint64_t)N);
  for (uint64_t b = 0; b < t->block_count; ++b) {
    uint32_t coord[OR_MAX_ORDER];
    or_block_coord(N, t->part, b, coord);
    const int64_t lo = (int64_t)t->block_off[b], hi = (int64_t)t->block_off[b + 1];
#pragma omp parallel for schedule(static)
    for (int64_t e = lo; e < hi; ++e) {
      uint64_t row[OR_MAX_ORDER];
      for (int a = 0; a < N; ++a)
        row[a] = (uint64_t)coord[a] * t->bdims[a] + t->lidx[(uint64_t)a * t->nnz + (uint64_t)e];
      VT res = 0;
      for (uint64_t i = 0; i < core_n; ++i) {
        const int *d = cidx + i * (uint64_t)N;
        VT temp = G[i];
        for (int a = 0; a < N; ++a) temp *= A[a][row[a] * (uint64_t)R + (uint64_t)d[a]];
        res += temp;
      }
      xhat_out[e] = res;
    }
  }
  free(cidx);
}

/* reconstruction.cuh:259-276: serial value_t error sum, fit = 1 - sqrt(Err)/norm. */
double SFX(or_fit)(const VT *vals, const VT *xhat, uint64_t nnz, VT norm) {
  VT err = 0;
  for (uint64_t e = 0; e < nnz; ++e) {
    VT d = vals[e] - xhat[e];
    err += d * d;
  }
  if (norm == 0) return 1.0;
  VT s = (VT)sqrt((double)err);
  return (double)((VT)1 - s / norm);
}

/* tucker.cuh:160-187 + update.cuh:128-199: Gauss-Seidel mode loop, fit after
 * every iteration, stop at max_iters or |fit - previous fit| <= tol.  The
 * factors A[n] ((P_n*bd_n) x R) are updated in place.  vals are in block
 * order; norm is or_norm_* over the input-ordered values, as
 * tensor_manager.tpp:267-293 sums them before the blocks exist.  Returns the number of
 * iterations run; fits_out[i] is the fit after iteration i. */
int SFX(or_als)(const or_tensor *t, const VT *vals, VT norm, VT *const *A, const VT *G,
                double lambda, int max_iters, double tol, double *fits_out) {
  const int N = t->order, R = t->rank;
  uint64_t max_rows = 0;
  for (int a = 0; a < N; ++a) {
    uint64_t rows = (uint64_t)t->part[a] * t->bdims[a];
    if (rows > max_rows) max_rows = rows;
  }
  VT *delta = (VT *)malloc((t->nnz ? t->nnz : 1) * (size_t)R * sizeof(VT));
  VT *xhat = (VT *)malloc((t->nnz ? t->nnz : 1) * sizeof(VT));
  double *B = (double *)malloc(max_rows * (size_t)R * R * sizeof(double));
  double *C = (double *)malloc(max_rows * (size_t)R * sizeof(double));
  int iter = 0;
  double p_fit = -1.0, fit = -1.0;
  const VT *const *Ac = (const VT *const *)A;
  for (;;) {
    for (int n = 0; n < N; ++n) {
      SFX(or_delta)(t, Ac, G, n, delta);
      SFX(or_bc)(t, vals, delta, n, lambda, B, C);
      SFX(or_solve)(R, (uint64_t)t->part[n] * t->bdims[n], B, C, A[n]);
    }
    SFX(or_recon)(t, Ac, G, xhat);
    fit = SFX(or_fit)(vals, xhat, t->nnz, norm);
    fits_out[iter] = fit;
    ++iter;
    if (iter >= max_iters || (p_fit != -1.0 && fabs(p_fit - fit) <= tol)) break;
    p_fit = fit;
  }
  free(delta);
  free(xhat);
  free(B);
  free(C);
  return iter;
}
