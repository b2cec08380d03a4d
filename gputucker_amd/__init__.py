"""gputucker_amd -- sparse Tucker (row-wise ALS) engine for one AMD MI355X.

The compute path is the HIP library ``libgputucker_hip.so`` (C ABI declared in
``include/gputucker_hip.h``).  This module binds it with ctypes and mirrors
the reference's flow (supertensor::gputucker): set the COO tensor, partition
it into P_1 x ... x P_N blocks with per-mode row grouping, initialise
factors and core, then run Gauss-Seidel mode updates and the fit.

There is no CPU fallback: if the library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgputucker_hip.so")
MAX_ORDER = 8

PRECISIONS = {"fp32": 0, "bf16": 1, "bf16r": 2}
# gt_delta_path: the delta + Gram kernel a mode update ran
DELTA_PATHS = {0: "none", 1: "fp32", 2: "fp32-wide", 3: "bf16", 4: "bf16r", 5: "bf16r-l2",
               6: "bf16r-wide", 7: "bf16r-ob", 8: "bf16r-ob-l2", 9: "bf16r-kr"}
STATUS = {0: "GT_OK", 1: "GT_ERR_ARG", 2: "GT_ERR_HIP", 3: "GT_ERR_STATE",
          4: "GT_ERR_UNSUPPORTED", 5: "GT_ERR_MEMORY", 6: "GT_ERR_PARTITION"}

# every extern "C" symbol include/gputucker_hip.h declares
EXPORTS = [
    "gt_default_opts", "gt_plan_partition", "gt_schedule", "gt_create", "gt_destroy",
    "gt_last_error", "gt_set_coo", "gt_partition", "gt_get_partition_info", "gt_get_block_nnz", "gt_get_block",
    "gt_partition_dump", "gt_set_init", "gt_mode_update", "gt_fit", "gt_fit_noalloc",
    "gt_iteration", "gt_run",
    "gt_get_factor", "gt_set_factor", "gt_get_delta", "gt_get_stream_delta", "gt_get_gram",
    "gt_get_recon",
    "gt_get_timings", "gt_synchronize", "gt_version", "gt_kernel_rank", "gt_autotune",
]


class GtOpts(C.Structure):
    _fields_ = [
        ("device", C.c_int), ("order", C.c_int), ("rank", C.c_int), ("lambda_", C.c_double),
        ("max_iters", C.c_int), ("tol", C.c_double), ("precision", C.c_int),
        ("partition", C.c_uint32 * MAX_ORDER), ("device_budget", C.c_uint64),
        ("chunk_tiles", C.c_int), ("profile", C.c_int), ("ref_fit", C.c_int),
        ("autotune", C.c_int), ("graph", C.c_int), ("core_source", C.c_int),
        ("host_pinned_limit", C.c_uint64), ("verbose", C.c_int),
    ]


class GtPartitionInfo(C.Structure):
    _fields_ = [
        ("order", C.c_int), ("dims", C.c_uint32 * MAX_ORDER),
        ("partition", C.c_uint32 * MAX_ORDER), ("block_dims", C.c_uint32 * MAX_ORDER),
        ("block_count", C.c_uint64), ("nnz", C.c_uint64), ("max_block_nnz", C.c_uint64),
        ("empty_blocks", C.c_uint64), ("norm", C.c_double), ("partition_type", C.c_int),
        ("stream_count", C.c_int), ("avail_nnz_per_task", C.c_uint64),
        ("out_of_core", C.c_int), ("segment_nnz", C.c_uint64),
        ("stream_bytes_per_nnz", C.c_uint32), ("partition_path", C.c_int),
        ("partition_pieces", C.c_uint64),
        ("host_pageable_bytes", C.c_uint64),
        ("hot_rows", C.c_uint32 * MAX_ORDER), ("hot_kind", C.c_int * MAX_ORDER),
    ]


class GtTimings(C.Structure):
    _fields_ = [
        ("ms_gram", C.c_double * MAX_ORDER), ("ms_solve", C.c_double * MAX_ORDER),
        ("ms_mode", C.c_double * MAX_ORDER), ("ms_fit", C.c_double),
        ("ms_partition", C.c_double), ("gram_kernel_ms_total", C.c_double),
        ("gram_kernel_launches", C.c_uint64), ("ms_iteration_graph", C.c_double),
        ("tuned_chunk_tiles", C.c_int * MAX_ORDER), ("tuned_variant", C.c_int * MAX_ORDER),
        ("tuned_ms", C.c_double * MAX_ORDER),
        ("delta_path", C.c_int * MAX_ORDER),
    ]


class GtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m gputucker_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, u32p, u64p = C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        L.gt_default_opts.argtypes = [C.POINTER(GtOpts)]
        L.gt_plan_partition.argtypes = [C.c_int, u32p, C.c_uint64, C.c_int, C.c_uint64, C.c_int,
                                        u32p, u32p, u32p, C.POINTER(C.c_int), u64p,
                                        C.POINTER(C.c_int)]
        L.gt_schedule.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, u64p, C.c_uint64,
                                  C.c_int64, u64p, u64p, u64p, C.POINTER(C.c_int)]
        L.gt_schedule.restype = C.c_int64
        L.gt_create.argtypes = [C.POINTER(vp), C.POINTER(GtOpts)]
        L.gt_destroy.argtypes = [vp]
        L.gt_last_error.argtypes = [vp]
        L.gt_last_error.restype = C.c_char_p
        L.gt_set_coo.argtypes = [vp, C.c_uint64, C.POINTER(vp), vp, u32p]
        L.gt_partition.argtypes = [vp, u32p, u64p]
        L.gt_get_partition_info.argtypes = [vp, C.POINTER(GtPartitionInfo)]
        L.gt_get_block_nnz.argtypes = [vp, u64p]
        L.gt_get_block.argtypes = [vp, C.c_uint64, C.POINTER(vp), vp]
        L.gt_partition_dump.argtypes = [vp, C.c_uint64, C.c_int, vp, vp]
        L.gt_set_init.argtypes = [vp, C.POINTER(vp), vp]
        L.gt_mode_update.argtypes = [vp, C.c_int]
        L.gt_fit.argtypes = [vp, C.POINTER(C.c_double)]
        L.gt_fit_noalloc.argtypes = [vp, C.POINTER(C.c_double)]
        L.gt_iteration.argtypes = [vp, C.POINTER(C.c_double)]
        L.gt_run.argtypes = [vp, C.POINTER(C.c_int), vp]
        L.gt_get_factor.argtypes = [vp, C.c_int, vp]
        L.gt_set_factor.argtypes = [vp, C.c_int, vp]
        L.gt_get_delta.argtypes = [vp, C.c_int, vp]
        L.gt_get_gram.argtypes = [vp, C.c_int, vp, vp]
        L.gt_get_stream_delta.argtypes = [vp, C.c_int, vp, C.POINTER(vp)]
        L.gt_get_recon.argtypes = [vp, vp]
        L.gt_get_timings.argtypes = [vp, C.POINTER(GtTimings)]
        L.gt_synchronize.argtypes = [vp]
        L.gt_kernel_rank.argtypes = [C.c_int]
        L.gt_autotune.argtypes = [vp]
        _lib = L
    return _lib


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def kernel_rank(rank):
    """The rank the kernels run a rank-`rank` model at (zero-padded above it)."""
    return lib().gt_kernel_rank(int(rank))


def default_lambda():
    """constants.hpp:18 writes kLambda as the float literal 0.0001f."""
    return float(np.float32(1e-4))


def plan_partition(dims, nnz, rank, mem, value_size=4, forced=None):
    """Host-only planner (optimizer.tpp:70-164) through the C ABI."""
    dims = _u32(dims)
    N = len(dims)
    part = np.zeros(N, np.uint32)
    bd = np.zeros(N, np.uint32)
    fr = _u32(forced if forced is not None else np.zeros(N))
    streams, avail, typ = C.c_int(), C.c_uint64(), C.c_int()
    rc = lib().gt_plan_partition(N, dims.ctypes.data_as(C.POINTER(C.c_uint32)), int(nnz), int(rank),
                                 int(mem), int(value_size), fr.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 part.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 bd.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(streams),
                                 C.byref(avail), C.byref(typ))
    if rc != 0:
        raise GtError(rc, "planner rejected the partition")
    return dict(part=part, bdims=bd, streams=streams.value, avail_nnz=avail.value, type=typ.value)


def schedule(typ, stream_count, avail_nnz, block_nnz, nnz_total, max_tasks=1 << 16):
    bn = np.ascontiguousarray(block_nnz, np.uint64)
    tb, tn, to = (np.zeros(max_tasks, np.uint64) for _ in range(3))
    ts = np.zeros(max_tasks, np.int32)
    P64 = C.POINTER(C.c_uint64)
    n = lib().gt_schedule(int(typ), int(stream_count), int(avail_nnz), len(bn),
                          bn.ctypes.data_as(P64), int(nnz_total), max_tasks, tb.ctypes.data_as(P64),
                          tn.ctypes.data_as(P64), to.ctypes.data_as(P64),
                          ts.ctypes.data_as(C.POINTER(C.c_int)))
    if n < 0:
        raise GtError(1, "too many tasks")
    return dict(block=tb[:n], nnz=tn[:n], offset=to[:n], stream=ts[:n])


class Engine:
    """One device context.  Mirrors the reference's TuckerDecomposition flow."""

    def __init__(self, order, rank, device=0, lam=None, max_iters=3, tol=None, precision="fp32",
                 partition=None, device_budget=0, chunk_tiles=0, profile=False, ref_fit=False,
                 graph=False, autotune=False, core_source=0, host_pinned_limit=0, verbose=0):
        L = lib()
        o = GtOpts()
        L.gt_default_opts(C.byref(o))
        o.device, o.order, o.rank = int(device), int(order), int(rank)
        if lam is not None:
            o.lambda_ = float(lam)
        o.max_iters = int(max_iters)
        if tol is not None:
            o.tol = float(tol)
        o.precision = PRECISIONS[precision]
        if partition is not None:
            for k, p in enumerate(partition):
                o.partition[k] = int(p)
        o.device_budget = int(device_budget)
        o.chunk_tiles = int(chunk_tiles)
        o.profile = 1 if profile else 0
        o.ref_fit = 1 if ref_fit else 0
        o.graph = 1 if graph else 0
        o.autotune = 1 if autotune else 0
        o.core_source = int(core_source)
        o.host_pinned_limit = int(host_pinned_limit)
        o.verbose = int(verbose)
        self.opts = o
        self.order, self.rank = int(order), int(rank)
        self._h = C.c_void_p()
        rc = L.gt_create(C.byref(self._h), C.byref(o))
        if rc != 0:
            raise GtError(rc, L.gt_last_error(None).decode())
        self.dims = None
        self.nnz = 0

    def close(self):
        if self._h:
            lib().gt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise GtError(rc, lib().gt_last_error(self._h).decode())

    # ------------------------------------------------------------ input
    def set_coo(self, idx, vals, dims):
        idx = _u32(idx)
        vals = np.ascontiguousarray(vals, np.float32)
        self.dims = _u32(dims)
        self.nnz = idx.shape[1]
        rows = [np.ascontiguousarray(idx[k]) for k in range(self.order)]
        ptrs = (C.c_void_p * self.order)(*[r.ctypes.data for r in rows])
        self._check(lib().gt_set_coo(self._h, self.nnz, ptrs, vals.ctypes.data,
                                     self.dims.ctypes.data_as(C.POINTER(C.c_uint32))))

    def partition(self):
        """Plans P_n (or uses the forced ones) and partitions on the GPU."""
        part = np.zeros(self.order, np.uint32)
        self._check(lib().gt_partition(self._h, part.ctypes.data_as(C.POINTER(C.c_uint32)), None))
        B = self.info()["block_count"]
        bn = np.zeros(B, np.uint64)
        self._check(lib().gt_get_block_nnz(self._h, bn.ctypes.data_as(C.POINTER(C.c_uint64))))
        self.block_nnz = bn
        return part, bn

    def info(self):
        i = GtPartitionInfo()
        self._check(lib().gt_get_partition_info(self._h, C.byref(i)))
        N = self.order
        return dict(dims=np.array(i.dims[:N]), part=np.array(i.partition[:N]),
                    bdims=np.array(i.block_dims[:N]), block_count=i.block_count, nnz=i.nnz,
                    max_block_nnz=i.max_block_nnz, empty_blocks=i.empty_blocks, norm=i.norm,
                    type=i.partition_type, streams=i.stream_count, avail_nnz=i.avail_nnz_per_task,
                    out_of_core=i.out_of_core, segment_nnz=i.segment_nnz,
                    stream_bytes_per_nnz=i.stream_bytes_per_nnz, partition_path=i.partition_path,
                    partition_pieces=i.partition_pieces,
                    host_pageable_bytes=i.host_pageable_bytes,
                    hot_rows=np.array(i.hot_rows[:N]), hot_kind=np.array(i.hot_kind[:N]))

    def get_block(self, b, n):
        idx = np.zeros((self.order, max(n, 1)), np.uint32)
        vals = np.zeros(max(n, 1), np.float32)
        ptrs = (C.c_void_p * self.order)(*[idx[k].ctypes.data for k in range(self.order)])
        self._check(lib().gt_get_block(self._h, int(b), ptrs, vals.ctypes.data))
        return idx[:, :n], vals[:n]

    def partition_dump(self, b, mode, block_nnz, bd):
        rp = np.zeros(int(bd) + 1, np.uint64)
        wh = np.zeros(max(int(block_nnz), 1), np.uint32)
        self._check(lib().gt_partition_dump(self._h, int(b), int(mode), rp.ctypes.data,
                                            wh.ctypes.data))
        return rp, wh[:block_nnz]

    # ------------------------------------------------------------ model
    def set_init(self, A, core):
        A = [np.ascontiguousarray(a, np.float32) for a in A]
        core = np.ascontiguousarray(core, np.float32).reshape(-1)
        ptrs = (C.c_void_p * self.order)(*[a.ctypes.data for a in A])
        self._check(lib().gt_set_init(self._h, ptrs, core.ctypes.data))

    def autotune(self):
        """Per-mode chunk size (and core-operand source) by timing the pass."""
        self._check(lib().gt_autotune(self._h))
        t = self.timings()
        return dict(chunk_tiles=t["tuned_chunk_tiles"], variant=t["tuned_variant"],
                    ms=t["tuned_ms"])

    def mode_update(self, mode):
        self._check(lib().gt_mode_update(self._h, int(mode)))

    def fit(self):
        f = C.c_double()
        self._check(lib().gt_fit(self._h, C.byref(f)))
        return f.value

    def fit_noalloc(self):
        """gt_fit through the entry point that allocates no device memory."""
        f = C.c_double()
        self._check(lib().gt_fit_noalloc(self._h, C.byref(f)))
        return f.value

    def iteration(self):
        f = C.c_double()
        self._check(lib().gt_iteration(self._h, C.byref(f)))
        return f.value

    def run(self):
        fits = np.zeros(max(self.opts.max_iters, 1), np.float64)
        it = C.c_int()
        self._check(lib().gt_run(self._h, C.byref(it), fits.ctypes.data))
        return fits[: it.value]

    def get_factor(self, mode):
        out = np.zeros((int(self.dims[mode]), self.rank), np.float32)
        self._check(lib().gt_get_factor(self._h, int(mode), out.ctypes.data))
        return out

    def set_factor(self, mode, a):
        a = np.ascontiguousarray(a, np.float32)
        self._check(lib().gt_set_factor(self._h, int(mode), a.ctypes.data))

    def get_delta(self, mode):
        out = np.zeros((max(self.nnz, 1), self.rank), np.float32)
        self._check(lib().gt_get_delta(self._h, int(mode), out.ctypes.data))
        return out[: self.nnz]

    def get_stream_delta(self, mode):
        """The mode update kernel's own deltas of mode `mode` (stream order)
        and the stream's global indices (N, nnz)."""
        out = np.zeros((max(self.nnz, 1), self.rank), np.float32)
        idx = np.zeros((self.order, max(self.nnz, 1)), np.uint32)
        ptrs = (C.c_void_p * self.order)(*[idx[k].ctypes.data for k in range(self.order)])
        self._check(lib().gt_get_stream_delta(self._h, int(mode), out.ctypes.data, ptrs))
        return out[: self.nnz], idx[:, : self.nnz]

    def get_recon(self):
        out = np.zeros(max(self.nnz, 1), np.float32)
        self._check(lib().gt_get_recon(self._h, out.ctypes.data))
        return out[: self.nnz]

    def get_gram(self, mode, rows):
        B = np.zeros((rows, self.rank, self.rank), np.float64)
        Cv = np.zeros((rows, self.rank), np.float64)
        self._check(lib().gt_get_gram(self._h, int(mode), B.ctypes.data, Cv.ctypes.data))
        return B, Cv

    def timings(self):
        t = GtTimings()
        self._check(lib().gt_get_timings(self._h, C.byref(t)))
        N = self.order
        return dict(ms_gram=list(t.ms_gram[:N]), ms_solve=list(t.ms_solve[:N]),
                    ms_mode=list(t.ms_mode[:N]), ms_fit=t.ms_fit, ms_partition=t.ms_partition,
                    gram_kernel_ms_total=t.gram_kernel_ms_total,
                    gram_kernel_launches=t.gram_kernel_launches,
                    ms_iteration_graph=t.ms_iteration_graph,
                    tuned_chunk_tiles=list(t.tuned_chunk_tiles[:N]),
                    tuned_variant=list(t.tuned_variant[:N]), tuned_ms=list(t.tuned_ms[:N]),
                    delta_path=[DELTA_PATHS[p] for p in t.delta_path[:N]])

    def synchronize(self):
        self._check(lib().gt_synchronize(self._h))
