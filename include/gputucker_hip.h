/*
 * gputucker_hip.h -- C ABI of the MI355X (gfx950) sparse Tucker engine.
 *
 * Drop-in boundary for GPUTucker's row-wise ALS hot path.  The reference has
 * no C/FFI layer: its boundary is the header-template API in namespace
 * supertensor::gputucker (TuckerDecomposition, tucker.cuh:30; ParseFromFile,
 * tensor_manager.hpp:16-17; UpdateFactorMatrices, update.cuh:128-130;
 * Reconstruction, reconstruction.cuh:228-230).  Each entry point below names
 * the reference routine it replaces.  The C++ driver in gputucker_amd/driver
 * rebuilds the reference's class API and CLI on top of these calls.
 *
 * Conventions: every call returns GT_OK (0) or a gt_status code and never
 * exits the process (the reference calls exit(1) on CUDA errors,
 * cuda_helper.cpp:102-109).  gt_last_error() gives the message.  Host arrays
 * are plain pointers owned by the caller; nothing is retained after a call
 * returns.  A context is bound to one device and is not thread-safe.
 * Indices are 0-based uint32 (the reference's index_t, main.cu:24); values
 * are float (the reference's value_t in its FP32 build).
 */
#ifndef GPUTUCKER_HIP_H_
#define GPUTUCKER_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_MAX_ORDER 8 /* constants.hpp:16 kMaxOrder */
#define GT_MAX_RANK 50 /* the reference's delta kernel holds R <= 50 sums (delta.cuh:63) */

typedef struct gt_ctx gt_ctx;

typedef enum gt_status {
  GT_OK = 0,
  GT_ERR_ARG = 1,         /* bad argument (the reference asserts or throws) */
  GT_ERR_HIP = 2,         /* HIP runtime error */
  GT_ERR_STATE = 3,       /* call out of order (e.g. mode update before partition) */
  GT_ERR_UNSUPPORTED = 4, /* rank/order outside the built kernels */
  GT_ERR_MEMORY = 5,      /* device budget too small */
  GT_ERR_PARTITION = 6    /* "Block dimension is larger than the tensor dimension." */
} gt_status;

typedef enum gt_precision {
  GT_PREC_FP32 = 0,  /* plain FP32 VALU contraction (the reference's arithmetic) */
  GT_PREC_BF16 = 1,  /* BF16 MFMA contraction, FP32 accumulate */
  GT_PREC_BF16R = 2  /* BF16 MFMA + split-BF16 refinement terms (near-FP32 delta) */
} gt_precision;

/* The delta + Gram kernel a mode update ran (gt_timings.delta_path). */
typedef enum gt_delta_path {
  GT_DELTA_NONE = 0,      /* no mode update has run */
  GT_DELTA_FP32 = 1,      /* FP32 VALU delta, FP64 MFMA Gram (R <= 15) */
  GT_DELTA_FP32_WIDE = 2, /* FP32 VALU delta, block-per-lane FP64 Gram (R > 15) */
  GT_DELTA_BF16 = 3,      /* BF16 MFMA delta, core table in LDS */
  GT_DELTA_BF16R = 4,     /* refined BF16 MFMA delta, core table in LDS */
  GT_DELTA_BF16R_L2 = 5,  /* refined BF16 MFMA delta of order 5 at R = 6..8, one MFMA per two
                             outer indices, core operand from L2 */
  GT_DELTA_BF16R_WIDE = 6, /* refined BF16 MFMA delta of order 3 at 16 <= R <= 32, core operand
                             from L2, block-per-lane FP64 Gram */
  GT_DELTA_BF16R_OB = 7,   /* refined BF16 MFMA delta at orders 4-5 with R^(N-2) <= 128, one MFMA
                              per 16 outer indices, the o-block table copied to LDS per
                              workgroup of 12 waves */
  GT_DELTA_BF16R_OB_L2 = 8, /* the same, o-block table from L2 */
  GT_DELTA_BF16R_KR = 9     /* refined BF16 MFMA delta of order 4 at R = 4 and 8 and of order 5
                               at R = 5 with two modes contracted on the matrix cores (the
                               Kronecker operand of the last outer factor and A_L, split per
                               nonzero), table in LDS per workgroup of 12 waves */
} gt_delta_path;

typedef struct gt_opts {
  int device;                       /* HIP device ordinal */
  int order;                        /* N, 2..GT_MAX_ORDER */
  int rank;                         /* R per mode, 1..GT_MAX_RANK (cmdline_opts.cpp:30, default 10) */
  double lambda;                    /* ridge term (constants.hpp:18, default (double)0.0001f) */
  int max_iters;                    /* constants.hpp:17, default 3 */
  double tol;                       /* fit tolerance (tucker.cuh:183, default (double)0.0001f) */
  int precision;                    /* gt_precision */
  uint32_t partition[GT_MAX_ORDER]; /* forced P_n; all zero = planner (optimizer.tpp:70-164) */
  uint64_t device_budget;           /* bytes the planner may assume; 0 = free device memory.
                                       When the resident layout exceeds it, the per-mode
                                       streams are staged from pinned host memory */
  int chunk_tiles;                  /* 64-nonzero tiles per Gram chunk; 0 = from the nonzero
                                       count: about 16384 chunks, 2 to 32 tiles each */
  int profile;                      /* 1: time the Gram kernel with HIP events */
  int ref_fit;                      /* 0: the fit from FP64 sums.  1: the fit in the reference's
                                       value_t arithmetic: the norm as a serial float sum of x^2
                                       in input order (tensor_manager.tpp:263-293), the error as
                                       a serial float sum of (x - xhat)^2 in block order, and
                                       fit = 1.0f - sqrtf(err) / norm in float
                                       (reconstruction.cuh:259-276).  Needs the block COO
                                       resident (in core); gt_iteration then runs the
                                       reconstruction pass instead of the fused fit */
  int autotune;                     /* 1: gt_set_init runs gt_autotune (in core) */
  int graph;                        /* 1: gt_iteration / gt_run capture one in-core iteration
                                       (N mode updates and the fit reduction, kernels only)
                                       into a HIP graph on first use and replay it; dropped
                                       by gt_set_coo, gt_partition and gt_set_init.  No per-mode
                                       or per-kernel timings inside it (ms_iteration_graph) */
  int core_source;                  /* 0: each refined BF16 delta kernel takes its core operand
                                       from its default source; 1: from the other one (a table
                                       in LDS or L2) where both are built.  gt_autotune
                                       chooses per mode instead (tuned_variant) */
  uint64_t host_pinned_limit;       /* out-of-core host data: at most this many page-locked
                                       bytes (0: as many as the system grants); the rest in
                                       pageable memory (gt_partition_info.host_pageable_bytes) */
  int verbose;                      /* 1: a line on stderr per step of a partition built under
                                       the device budget (each pass over the input) */
} gt_opts;

typedef struct gt_partition_info {
  int order;
  uint32_t dims[GT_MAX_ORDER];
  uint32_t partition[GT_MAX_ORDER];  /* P_n */
  uint32_t block_dims[GT_MAX_ORDER]; /* ceil(I_n / P_n) */
  uint64_t block_count;
  uint64_t nnz;
  uint64_t max_block_nnz;
  uint64_t empty_blocks;
  double norm;               /* sqrt(sum x^2), FP64 */
  int partition_type;        /* 0 dimension-based, 1 nonzero-based (constants.hpp:38-42) */
  int stream_count;          /* optimizer.tpp:130,140 */
  uint64_t avail_nnz_per_task;
  int out_of_core;           /* 1: streams staged from pinned host memory per mode update */
  uint64_t segment_nnz;      /* nonzeros per staged segment (out-of-core) */
  uint32_t stream_bytes_per_nnz; /* out-of-core: host bytes per nonzero shipped per mode update,
                                    the largest over the modes (indices packed to 16 or 24 bits
                                    and no row index when every factor has at most 2^24 rows,
                                    else 4 bytes per index) */
  int partition_path;        /* 0: partitioned in HBM; 1: the in-core partition exceeded the device
                                budget, so the input waited in pinned host memory and each mode's
                                stream was built piece by piece (ranges of rows, of blocks inside
                                a heavy row, windows inside a heavy block) under the budget */
  uint64_t partition_pieces; /* pieces the bucketed partition built, over all modes */
  uint64_t host_pageable_bytes; /* out-of-core host data held in pageable memory because the
                                   system granted no more page-locked memory (0 normally);
                                   copies from it are slower and synchronous */
  uint32_t hot_rows[GT_MAX_ORDER]; /* per mode: factor rows the refined BF16 delta of the wide
                                      ranks (order 3, R = 16 or 25..32, in-core partition) takes
                                      through a core contracted with the row, each row's
                                      nonzeros grouped by them in the stream (0: none) */
  int hot_kind[GT_MAX_ORDER];      /* per mode: 1 the heavy rows of its outer factor, 2 every
                                      row of its last other factor (rows long enough to hold
                                      runs of each), 0 none */
} gt_partition_info;

typedef struct gt_timings {
  double ms_gram[GT_MAX_ORDER];   /* last delta+Gram phase per mode */
  double ms_solve[GT_MAX_ORDER];  /* last row solve per mode */
  double ms_mode[GT_MAX_ORDER];   /* last whole mode update */
  double ms_fit;                  /* last reconstruction + fit */
  double ms_partition;            /* partition + row grouping */
  double gram_kernel_ms_total;    /* with opts.profile: summed event time of the Gram kernel */
  uint64_t gram_kernel_launches;
  double ms_iteration_graph;      /* opts.graph: the last replayed iteration, without the fit copy */
  int tuned_chunk_tiles[GT_MAX_ORDER]; /* gt_autotune: 64-nonzero tiles per chunk, per mode (0: none) */
  int tuned_variant[GT_MAX_ORDER];     /* gt_autotune: 1 = the other core-operand source
                                          (L2 or LDS table) than the default chosen; the
                                          delta_path field names the kernel that runs */
  double tuned_ms[GT_MAX_ORDER];       /* gt_autotune: the chosen delta + Gram pass time */
  int delta_path[GT_MAX_ORDER];   /* gt_delta_path of the last update of each mode: a BF16
                                     request that no MFMA kernel covers runs GT_DELTA_FP32 */
} gt_timings;

/* Defaults as the reference compiles them in (constants.hpp:16-18,
 * cmdline_opts.cpp:27-32). */
int gt_default_opts(gt_opts *opts);

/* Host-only partition planner, no device needed: Optimizer::FindPartitionParms
 * and _AvailableNonzeroCountPerTask (optimizer.tpp:70-164) for one GPU with
 * gpu_mem = mem.  forced (may be NULL or all zero) pins P_n instead of
 * searching.  type_out: 0 dimension-based, 1 nonzero-based. */
int gt_plan_partition(int order, const uint32_t *dims, uint64_t nnz, int rank, uint64_t mem,
                      int value_size, const uint32_t *forced, uint32_t *part_out,
                      uint32_t *bd_out, int *streams_out, uint64_t *avail_out, int *type_out);

/* Host-only task list of Scheduler::Schedule (scheduler.tpp:22-159) for one
 * GPU.  Returns the task count, or -1 when max_tasks is too small. */
int64_t gt_schedule(int type, int stream_count, uint64_t avail_nnz, uint64_t block_count,
                    const uint64_t *block_nnz, uint64_t nnz_total, int64_t max_tasks,
                    uint64_t *t_block, uint64_t *t_nnz, uint64_t *t_off, int *t_stream);

/* Replaces CudaAgent construction + AllocateMaximumBuffer (cuda_agent.tpp:20-49). */
int gt_create(gt_ctx **out, const gt_opts *opts);
int gt_destroy(gt_ctx *ctx);

/* Last error message of ctx (or of the last failed gt_create when ctx is NULL). */
const char *gt_last_error(const gt_ctx *ctx);

/* Hands over the input tensor as _ReadData leaves it (tensor_manager.tpp:174-180):
 * idx[n] points at nnz 0-based indices of mode n, vals at nnz values, dims at
 * N mode sizes (the reference uses the max index seen).  Copied to HBM. */
int gt_set_coo(gt_ctx *ctx, uint64_t nnz, const uint32_t *const *idx, const float *vals,
               const uint32_t *dims);

/* Replaces Optimizer::FindPartitionParms (optimizer.tpp:70-89),
 * CreateTensorBlocks (tensor_manager.tpp:201-302) and
 * Tensor::AssignIndicesOfEachBlock (tensor.tpp:138-147), on the GPU.
 * part_out (N entries) and block_nnz_out (block_count entries) may be NULL. */
int gt_partition(gt_ctx *ctx, uint32_t *part_out, uint64_t *block_nnz_out);
int gt_get_partition_info(gt_ctx *ctx, gt_partition_info *info);
/* Nonzeros per block (block_count entries), the histogram of
 * tensor_manager.tpp:224-256. */
int gt_get_block_nnz(gt_ctx *ctx, uint64_t *block_nnz_out);

/* Block b's COO as Block::indices/values hold it (block.hpp:166-170):
 * idx[n] receives block_nnz local indices of mode n, vals the values. */
int gt_get_block(gt_ctx *ctx, uint64_t block, uint32_t *const *idx, float *vals);

/* Block b's row grouping for mode n as Block::count_nnz / where_nnz hold it
 * (block.tpp:114-186): row_ptr gets block_dims[n]+1 entries, where block_nnz. */
int gt_partition_dump(gt_ctx *ctx, uint64_t block, int mode, uint64_t *row_ptr, uint32_t *where);

/* Initial factors and core (tucker.cuh:75-134): A[n] is I_n x R row-major,
 * core is R^N values in linear order (last core axis fastest). */
int gt_set_init(gt_ctx *ctx, const float *const *A, const float *core);

/* One mode update: delta, Gram (B = lambda I + sum dd^T, C = sum x d) and the
 * R x R row solves, factor updated in place.  Replaces ComputingDelta
 * (delta.cuh:116-214), ComputingBC (update.cuh:35-101) and the row-wise
 * inverse update (update.cuh:163-194). */
int gt_mode_update(gt_ctx *ctx, int mode);

/* The project's autotuner (no counterpart in the reference, which fixes its
 * launch shape): per mode, times the delta + Gram pass at chunk sizes of 1 to
 * 32 tiles of 64 nonzeros and, for refined BF16 at orders 4-5 where both
 * apply, with the core operand from LDS and from L2, and keeps the fastest
 * for later mode updates (gt_timings.tuned_*).  In core only (a no-op out of
 * core); factors are untouched.  Needs gt_set_init first. */
int gt_autotune(gt_ctx *ctx);

/* Replaces Reconstruction (reconstruction.cuh:228-277): fit = 1 - ||x - xhat|| / ||x||. */
int gt_fit(gt_ctx *ctx, double *fit);

/* gt_fit without any device allocation.  It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under stream capture, tight loops, harnesses that forbid hipMalloc).  The
 * error partials and their sum live in a static 128 KiB device array (16384
 * doubles), one per device for the life of the process, shared by every
 * context on that device: calls that overlap on one device race.  It needs
 * 2048 doubles per staged segment plus one (in core: one segment), and
 * returns GT_ERR_MEMORY for an out-of-core layout of more than 7 segments.
 * Same kernels and summation order as gt_fit, so the same result. */
int gt_fit_noalloc(gt_ctx *ctx, double *fit);

/* One ALS iteration, modes 0..N-1 then the fit (update.cuh:146-198 + tucker.cuh:169-176). */
int gt_iteration(gt_ctx *ctx, double *fit);

/* The whole loop of tucker.cuh:166-187: stop at max_iters or |fit - prev| <= tol.
 * fits receives up to max_iters values; *iters_done the count. */
int gt_run(gt_ctx *ctx, int *iters_done, double *fits);

/* Factor matrix of mode n, I_n x R row-major (the reference's
 * factor_matrices[n][part] slices laid end to end, without padding rows). */
int gt_get_factor(gt_ctx *ctx, int mode, float *out);
int gt_set_factor(gt_ctx *ctx, int mode, const float *in);

/* Test hooks: delta of mode n for every nonzero in block order (nnz x R),
 * as ComputingDelta hands it back (delta.cuh:206-207). */
int gt_get_delta(gt_ctx *ctx, int mode, float *out);
/* Test hook: B (rows x R x R, lambda on the diagonal) and C (rows x R) of mode n
 * over the padded row space P_n * block_dims[n], as ComputingBC leaves them. */
int gt_get_gram(gt_ctx *ctx, int mode, double *B, double *C);
/* Test hook: the deltas of mode n as the mode update's own delta + Gram
 * kernel forms them (the precision selected), one row of R per nonzero in the
 * order of mode n's row-sorted stream; idx (may be NULL, or N arrays of nnz)
 * receives the stream's global indices, which name each nonzero. */
int gt_get_stream_delta(gt_ctx *ctx, int mode, float *delta, uint32_t *const *idx);
/* Test hook: xhat for every nonzero in block order. */
int gt_get_recon(gt_ctx *ctx, float *out);

int gt_get_timings(gt_ctx *ctx, gt_timings *t);
int gt_synchronize(gt_ctx *ctx);

/* The rank the kernels run a rank-R model at: R itself when it is built,
 * else the next built rank, with zero extra factor columns and core slices
 * (results as at rank R).  0 when R is outside 1..GT_MAX_RANK. */
int gt_kernel_rank(int rank);

/* Library version: major*10000 + minor*100 + patch. */
int gt_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GPUTUCKER_HIP_H_ */
