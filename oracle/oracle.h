/*
 * oracle.h -- CPU restatement of GPUTucker's row-wise ALS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gputucker_amd/) links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker and as the timed CPU
 * baseline.
 *
 * Every function follows one reference routine loop for loop; the citation
 * is given at each declaration (paths relative to the reference root).
 * The floating-point routines exist twice:
 *   *_f64  value_t = double  (the FP64 oracle; the shipped main.cu:24-25 type)
 *   *_f32  value_t = float   ("reference precision": the FP32 build of the
 *                             reference, B/C sums in double as update.cuh
 *                             does, serial float norm/error sums)
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   or_block_dims, or_block_coord  closed forms + brute force
 *   or_partition, or_assign_mode   README example (golden), invariants,
 *                                  brute-force stable sort on random inputs
 *   or_plan                        closed-form hand cases + invariants
 *   or_schedule                    invariants (coverage, sizes, round robin)
 *   or_delta_*, or_recon_*         numpy einsum brute force, R=1 / N=2 /
 *                                  one-hot special cases
 *   or_bc_*                        numpy normal-equation sums
 *   or_solve_*                     numpy.linalg.solve and the vendored Eigen
 *                                  inverse built into oracle/_ref
 *   or_als_*                       exact-model fit = 1, monotone objective
 */
#ifndef GT_ORACLE_H_
#define GT_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_ORDER 8

/* Partitioned tensor as the reference holds it after CreateTensorBlocks
 * (tensor_manager.tpp:201-302) and AssignIndicesOfEachBlock (tensor.tpp:138-147).
 * All arrays are flat; blocks are concatenated in block-id order. */
typedef struct or_tensor {
  int order;
  int rank;
  uint64_t nnz;
  uint64_t block_count;
  const uint32_t *dims;      /* [order] */
  const uint32_t *part;      /* [order] P_n */
  const uint32_t *bdims;     /* [order] block_dims = ceil(I/P) */
  const uint64_t *block_off; /* [block_count+1] */
  const uint32_t *lidx;      /* [order * nnz] block-local indices, mode-major */
  const uint64_t *row_ptr[OR_MAX_ORDER]; /* per mode: block_count*(bd_n+1) */
  const uint32_t *where[OR_MAX_ORDER];   /* per mode: nnz (block-local positions) */
} or_tensor;

/* tensor.tpp:198-213 / optimizer.tpp:359-376: bd = ceil(I/P); -1 when
 * ceil(I/bd) != P (the reference throws). */
/* OpenMP threads the parallel loops use. */
int or_max_threads(void);
int or_block_dims(int order, const uint32_t *dims, const uint32_t *part, uint32_t *bd_out);

/* tensor.tpp:156-174: mixed-radix decode, last axis fastest. */
void or_block_coord(int order, const uint32_t *part, uint64_t block_id, uint32_t *coord_out);

/* tensor_manager.tpp:224-290 + block.tpp:99-108: block histogram, serial
 * stable scatter in input order, block-local indices.  perm_out[j] is the
 * input position of the nonzero placed at block-ordered position j. */
int or_partition(int order, uint64_t nnz, const uint32_t *idx, const uint32_t *dims,
                 const uint32_t *part, uint64_t *block_nnz_out, uint64_t *perm_out,
                 uint32_t *lidx_out);

/* block.tpp:114-186: per-mode row histogram, exclusive prefix, stable where. */
void or_assign_mode(uint64_t nnz_b, uint32_t dim, const uint32_t *lidx_n,
                    uint64_t *row_ptr_out, uint32_t *where_out);

/* optimizer.tpp:70-164: partition planner.  value_size is sizeof(value_t).
 * type_out: 0 = dimension-based, 1 = nonzero-based (constants.hpp:38-42).
 * Returns -1 where the reference throws. */
int or_plan(int order, const uint32_t *dims, uint64_t nnz, int rank, int gpu_count,
            uint64_t gpu_mem, int value_size, uint32_t *part_out, uint32_t *bd_out,
            int *streams_out, uint64_t *avail_nnz_out, int *type_out);

/* scheduler.tpp:22-159.  Outputs at most max_tasks tasks as parallel arrays;
 * returns the task count or -1 if max_tasks is too small.  Block ties in the
 * nnz sort are broken by block id (the reference uses std::sort, whose tie
 * order is unspecified). */
int64_t or_schedule(int type, int gpu_count, int stream_count, uint64_t avail_nnz,
                    uint64_t block_count, const uint64_t *block_nnz, uint64_t nnz_total,
                    int64_t max_tasks, uint64_t *t_block, uint64_t *t_nnz, uint64_t *t_off,
                    int *t_stream, int *t_device);

/* ---- value_t = double ---- */
double or_norm_f64(const double *vals, uint64_t nnz);
void or_delta_f64(const or_tensor *t, const double *const *A, const double *G, int mode,
                  double *delta_out);
void or_bc_f64(const or_tensor *t, const double *vals, const double *delta, int mode,
               double lambda, double *B_out, double *C_out);
void or_solve_f64(int rank, uint64_t rows, const double *B, const double *C, double *A_out);
void or_recon_f64(const or_tensor *t, const double *const *A, const double *G, double *xhat_out);
double or_fit_f64(const double *vals, const double *xhat, uint64_t nnz, double norm);
int or_als_f64(const or_tensor *t, const double *vals, double norm, double *const *A,
               const double *G, double lambda, int max_iters, double tol, double *fits_out);

/* ---- value_t = float (reference precision) ---- */
float or_norm_f32(const float *vals, uint64_t nnz);
void or_delta_f32(const or_tensor *t, const float *const *A, const float *G, int mode,
                  float *delta_out);
void or_bc_f32(const or_tensor *t, const float *vals, const float *delta, int mode,
               double lambda, double *B_out, double *C_out);
void or_solve_f32(int rank, uint64_t rows, const double *B, const double *C, float *A_out);
void or_recon_f32(const or_tensor *t, const float *const *A, const float *G, float *xhat_out);
double or_fit_f32(const float *vals, const float *xhat, uint64_t nnz, float norm);
int or_als_f32(const or_tensor *t, const float *vals, float norm, float *const *A,
               const float *G, double lambda, int max_iters, double tol, double *fits_out);

/* Partial-pivot LU inverse (the routine Eigen's MatrixXd::inverse() runs:
 * lib/Eigen/src/LU/InverseImpl.h:23-30 -> PartialPivLU.h). out = inv(B). */
int or_lu_inverse(int n, const double *B, double *out);

#ifdef __cplusplus
}
#endif
#endif /* GT_ORACLE_H_ */
