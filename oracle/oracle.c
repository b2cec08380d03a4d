/*
 * oracle.c -- integer parts of the CPU restatement of GPUTucker's hot path
 * (partitioner, row grouping, planner, scheduler) plus the two value-type
 * instantiations of oracle_vt.h.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product never links or loads it.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

int or_max_threads(void) { return omp_get_max_threads(); }

/* tensor.tpp:198-213 and optimizer.tpp:359-376 (_RefreshDims / _RefreshBlockDims) */
int or_block_dims(int order, const uint32_t *dims, const uint32_t *part, uint32_t *bd_out) {
  for (int a = 0; a < order; ++a) {
    if (part[a] == 0) return -1;
    uint32_t bd = (uint32_t)(((uint64_t)dims[a] + part[a] - 1) / part[a]);
    if (bd == 0) return -1;
    uint32_t check = (uint32_t)(((uint64_t)dims[a] + bd - 1) / bd);
    if (check != part[a]) return -1; /* "Block dimension is larger than the tensor dimension." */
    bd_out[a] = bd;
  }
  return 0;
}

/* tensor.tpp:156-174 (_BlockIDtoBlockCoord): last axis varies fastest. */
void or_block_coord(int order, const uint32_t *part, uint64_t block_id, uint32_t *coord_out) {
  uint64_t mult = 1;
  for (int it = 0; it < order; ++it) {
    int a = order - it - 1;
    coord_out[a] = (uint32_t)((block_id / mult) % part[a]);
    mult *= part[a];
  }
}

static uint64_t block_id_of(int order, const uint32_t *idx, uint64_t nnz, uint64_t e,
                            const uint32_t *bd, const uint32_t *part) {
  /* tensor_manager.tpp:232-242: block_id = sum_n floor(i_n/bd_n) * prod_{m>n} P_m */
  uint64_t id = 0, mult = 1;
  for (int it = 0; it < order; ++it) {
    int a = order - it - 1;
    id += (uint64_t)(idx[(uint64_t)a * nnz + e] / bd[a]) * mult;
    mult *= part[a];
  }
  return id;
}

/* tensor_manager.tpp:201-302 (CreateTensorBlocks) with block.tpp:99-108
 * (InsertNonzero).  The reference stores the nonzero at position
 * nnz_b - pos where pos counts down from nnz_b, so each block keeps the
 * input order of its nonzeros. */
int or_partition(int order, uint64_t nnz, const uint32_t *idx, const uint32_t *dims,
                 const uint32_t *part, uint64_t *block_nnz_out, uint64_t *perm_out,
                 uint32_t *lidx_out) {
  uint32_t bd[OR_MAX_ORDER];
  if (order < 1 || order > OR_MAX_ORDER) return -1;
  if (or_block_dims(order, dims, part, bd) != 0) return -1;
  uint64_t blocks = 1;
  for (int a = 0; a < order; ++a) blocks *= part[a];
  for (uint64_t e = 0; e < nnz; ++e)
    for (int a = 0; a < order; ++a)
      if (idx[(uint64_t)a * nnz + e] >= dims[a]) return -1; /* :237 assert */

  /* 1. histogram (:224-256) */
  uint64_t *hist = (uint64_t *)calloc(blocks, sizeof(uint64_t));
  uint64_t *start = (uint64_t *)malloc((blocks + 1) * sizeof(uint64_t));
  for (uint64_t e = 0; e < nnz; ++e) hist[block_id_of(order, idx, nnz, e, bd, part)]++;
  start[0] = 0;
  for (uint64_t b = 0; b < blocks; ++b) {
    block_nnz_out[b] = hist[b];
    start[b + 1] = start[b] + hist[b];
  }
  /* 2. serial stable scatter (:267-290) */
  for (uint64_t e = 0; e < nnz; ++e) {
    uint64_t b = block_id_of(order, idx, nnz, e, bd, part);
    uint64_t pos = hist[b]--;
    uint64_t dst = start[b] + (block_nnz_out[b] - pos);
    uint32_t coord[OR_MAX_ORDER];
    or_block_coord(order, part, b, coord);
    perm_out[dst] = e;
    for (int a = 0; a < order; ++a) /* block.tpp:105: local = global - base_dims */
      lidx_out[(uint64_t)a * nnz + dst] = idx[(uint64_t)a * nnz + e] - bd[a] * coord[a];
  }
  free(hist);
  free(start);
  return 0;
}

/* block.tpp:114-186 (AssignIndicesToEachMode) for one block and one mode. */
void or_assign_mode(uint64_t nnz_b, uint32_t dim, const uint32_t *lidx_n, uint64_t *row_ptr_out,
                    uint32_t *where_out) {
  uint64_t *tmp = (uint64_t *)calloc((size_t)dim + 1, sizeof(uint64_t));
  memset(row_ptr_out, 0, ((size_t)dim + 1) * sizeof(uint64_t));
  for (uint64_t e = 0; e < nnz_b; ++e) row_ptr_out[lidx_n[e]]++;
  uint64_t now = 0;
  uint32_t j;
  for (j = 0; j < dim; ++j) {
    uint64_t k = row_ptr_out[j];
    row_ptr_out[j] = now;
    tmp[j] = now;
    now += k;
  }
  row_ptr_out[j] = now;
  for (uint64_t e = 0; e < nnz_b; ++e) {
    uint32_t k = lidx_n[e];
    where_out[tmp[k]] = (uint32_t)e;
    tmp[k]++;
  }
  free(tmp);
}

/* ---- optimizer.tpp ---- */
typedef struct {
  int order;
  const uint32_t *dims;
  uint64_t nnz;
  uint32_t rank; /* unsigned int in optimizer.hpp:207 */
  int vs;        /* sizeof(value_t) */
  uint32_t part[OR_MAX_ORDER], bd[OR_MAX_ORDER];
  uint64_t blocks, avg_nnz;
} plan_t;

static int plan_refresh(plan_t *p) { /* optimizer.tpp:359-376 */
  if (or_block_dims(p->order, p->dims, p->part, p->bd) != 0) return -1;
  p->blocks = 1;
  for (int a = 0; a < p->order; ++a) p->blocks *= p->part[a];
  p->avg_nnz = (p->nnz + p->blocks - 1) / p->blocks;
  return 0;
}
static uint64_t core_size(const plan_t *p) { /* :228-233, std::pow(rank, order) */
  uint64_t core_nnz = (uint64_t)pow((double)p->rank, (double)p->order);
  return core_nnz * (uint64_t)(p->order * 4 + p->vs);
}
static uint64_t all_factors_size(const plan_t *p) { /* :241-247, index_t*unsigned product */
  uint64_t s = 0;
  for (int a = 0; a < p->order; ++a) s += (uint32_t)(p->dims[a] * p->rank);
  return s * (uint64_t)p->vs;
}
static uint64_t sub_factors_size(const plan_t *p) { /* :255-262 */
  uint64_t s = 0;
  for (int a = 0; a < p->order; ++a) s += (uint32_t)(p->bd[a] * p->rank);
  return s * (uint64_t)p->vs;
}
static uint64_t all_data_size(const plan_t *p) { /* :37-46 */
  uint64_t s = p->avg_nnz * (uint64_t)(p->order * 4 + p->vs); /* sub tensor :215-220 */
  s += core_size(p);
  s += sub_factors_size(p);
  s += p->avg_nnz * p->rank * (uint64_t)p->vs; /* sub delta :282-286 */
  return s;
}

/* optimizer.tpp:70-164 (FindPartitionParms, _DeterminePartitionType,
 * _DimensionBasedPartitioning, _NextPartitionAxis, _AvailableNonzeroCountPerTask) */
int or_plan(int order, const uint32_t *dims, uint64_t nnz, int rank, int gpu_count,
            uint64_t gpu_mem, int value_size, uint32_t *part_out, uint32_t *bd_out,
            int *streams_out, uint64_t *avail_nnz_out, int *type_out) {
  plan_t p;
  if (order < 1 || order > OR_MAX_ORDER) return -1;
  p.order = order;
  p.dims = dims;
  p.nnz = nnz;
  p.rank = (uint32_t)rank;
  p.vs = value_size;
  for (int a = 0; a < order; ++a) p.part[a] = 1;
  if (plan_refresh(&p) != 0) return -1;
  int streams = 1, type;
  uint64_t total_mem = (uint64_t)gpu_count * gpu_mem;
  uint64_t required = nnz * (uint64_t)(order * 4 + value_size);
  required += (uint64_t)gpu_count * core_size(&p);
  required += (uint64_t)gpu_count * all_factors_size(&p);
  required += nnz * (uint64_t)rank * (uint64_t)value_size; /* :270-273 */
  if (required <= total_mem) {
    type = 1; /* kNonzeroPartition */
    streams = 1;
  } else {
    type = 0; /* kDimensionPartition */
    streams = 4;
    uint64_t buf = gpu_mem / (uint64_t)streams;
    uint64_t total_streams = (uint64_t)gpu_count * (uint64_t)streams;
    do {
      int max_axis = 0; /* :343-351, ties go to the lowest axis */
      for (int a = 1; a < order; ++a)
        if (p.bd[max_axis] < p.bd[a]) max_axis = a;
      p.part[max_axis]++;
      if (plan_refresh(&p) != 0) return -1;
    } while (!(all_data_size(&p) < buf && p.blocks >= total_streams));
  }
  uint64_t buf = gpu_mem / (uint64_t)streams;
  uint64_t avail_buf = buf - (core_size(&p) + sub_factors_size(&p)); /* size_t, wraps like the reference */
  *avail_nnz_out = avail_buf / (uint64_t)(order * 4 + value_size + rank * value_size);
  for (int a = 0; a < order; ++a) {
    part_out[a] = p.part[a];
    bd_out[a] = p.bd[a];
  }
  *streams_out = streams;
  *type_out = type;
  return 0;
}

/* scheduler.tpp:22-159 */
int64_t or_schedule(int type, int gpu_count, int stream_count, uint64_t avail_nnz,
                    uint64_t block_count, const uint64_t *block_nnz, uint64_t nnz_total,
                    int64_t max_tasks, uint64_t *t_block, uint64_t *t_nnz, uint64_t *t_off,
                    int *t_stream, int *t_device) {
  int64_t nt = 0;
#define PUSH(b, n, o, s, d)                                                                        \
  do {                                                                                             \
    if (nt >= max_tasks) return -1;                                                                \
    t_block[nt] = (b);                                                                             \
    t_nnz[nt] = (n);                                                                               \
    t_off[nt] = (o);                                                                               \
    t_stream[nt] = (s);                                                                            \
    t_device[nt] = (d);                                                                            \
    ++nt;                                                                                          \
  } while (0)
  if (type == 1) { /* _NonzeroBasedPartitioning :41-77 */
    uint64_t task_count = (uint64_t)gpu_count;
    uint64_t per_task = (nnz_total + task_count - 1) / task_count;
    if (per_task > avail_nnz) {
      task_count = (nnz_total + avail_nnz - 1) / avail_nnz;
      per_task = avail_nnz;
    }
    uint64_t n = per_task;
    for (uint64_t task = 0; task < task_count; ++task) {
      uint64_t ptr = task * n;
      if (task == task_count - 1) n = nnz_total - ptr;
      PUSH(0, n, ptr, 0, (int)(task % (uint64_t)gpu_count));
    }
    return nt;
  }
  /* _DimensionBasedPartitioning :85-159: blocks by ascending nnz.  Ties by
   * block id (stable); the reference's std::sort leaves them unspecified. */
  uint64_t *order_ = (uint64_t *)malloc(block_count * sizeof(uint64_t));
  for (uint64_t b = 0; b < block_count; ++b) order_[b] = b;
  for (uint64_t i = 1; i < block_count; ++i) { /* insertion sort: stable */
    uint64_t v = order_[i];
    int64_t j = (int64_t)i - 1;
    while (j >= 0 && block_nnz[order_[j]] > block_nnz[v]) {
      order_[j + 1] = order_[j];
      --j;
    }
    order_[j + 1] = v;
  }
  int total_streams = gpu_count * stream_count;
  int g = 0;
  for (uint64_t i = 0; i < block_count; ++i) {
    uint64_t b = order_[i];
    uint64_t tasks_in_block = 1, per_stream = block_nnz[b];
    if (block_nnz[b] > avail_nnz) {
      tasks_in_block = (block_nnz[b] + avail_nnz - 1) / avail_nnz;
      per_stream = avail_nnz;
    }
    uint64_t t;
    for (t = 0; t + 1 < tasks_in_block; ++t) {
      if (g >= total_streams) g = 0;
      PUSH(b, per_stream, t * per_stream, g / gpu_count, g % gpu_count);
      ++g;
    }
    if (g >= total_streams) g = 0;
    PUSH(b, block_nnz[b] - t * per_stream, t * per_stream, g / gpu_count, g % gpu_count);
    ++g;
  }
  free(order_);
#undef PUSH
  return nt;
}

/* Partial-pivot LU inverse, the algorithm behind Eigen's MatrixXd::inverse()
 * (InverseImpl.h:23-30 -> PartialPivLU: largest |pivot| in the column, first
 * one on ties; column scaled by 1/pivot via division; rank-1 trailing update;
 * then unit-lower forward and upper back substitution on the permuted I). */
int or_lu_inverse(int n, const double *B, double *out) {
  double *lu = (double *)malloc((size_t)n * n * sizeof(double));
  int *perm = (int *)malloc((size_t)n * sizeof(int));
  memcpy(lu, B, (size_t)n * n * sizeof(double));
  for (int i = 0; i < n; ++i) perm[i] = i;
  int singular = 0;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    double big = fabs(lu[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(lu[i * n + k]) > big) {
        big = fabs(lu[i * n + k]);
        piv = i;
      }
    if (big != 0.0) {
      if (piv != k) {
        for (int j = 0; j < n; ++j) {
          double t = lu[k * n + j];
          lu[k * n + j] = lu[piv * n + j];
          lu[piv * n + j] = t;
        }
        int t = perm[k];
        perm[k] = perm[piv];
        perm[piv] = t;
      }
      for (int i = k + 1; i < n; ++i) lu[i * n + k] /= lu[k * n + k];
    } else {
      singular = 1;
    }
    for (int i = k + 1; i < n; ++i)
      for (int j = k + 1; j < n; ++j) lu[i * n + j] -= lu[i * n + k] * lu[k * n + j];
  }
  /* solve LU X = P I, column by column */
  for (int c = 0; c < n; ++c) {
    double *x = (double *)malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; ++i) x[i] = (perm[i] == c) ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) x[i] -= lu[i * n + j] * x[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) x[i] -= lu[i * n + j] * x[j];
      x[i] /= lu[i * n + i];
    }
    for (int i = 0; i < n; ++i) out[i * n + c] = x[i];
    free(x);
  }
  free(lu);
  free(perm);
  return singular ? -1 : 0;
}

/* value_t = double */
#define VT double
#define SFX(name) name##_f64
#include "oracle_vt.h"
#undef VT
#undef SFX

/* value_t = float: the FP32 build of the reference */
#define VT float
#define SFX(name) name##_f32
#include "oracle_vt.h"
#undef VT
#undef SFX
