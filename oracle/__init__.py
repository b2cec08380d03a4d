"""CPU oracle for GPUTucker's row-wise ALS hot path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, as the checker and as the timed CPU
baseline.  The product (``gputucker_amd``) never imports it.

``liboracle.so`` (built from ``oracle.c`` / ``oracle_vt.h`` by ``oracle/Makefile``)
restates the reference loop for loop; see ``oracle.h`` for the file:line each
routine follows and for how each one is pinned.  ``oracle/_ref/libeigenref.so``
is the reference's vendored Eigen inverse, compiled from its headers, used to
check the oracle's LU (built only where the reference tree exists).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MAX_ORDER = 8

_lib = None
_eig = None

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class OrTensor(C.Structure):
    _fields_ = [
        ("order", C.c_int),
        ("rank", C.c_int),
        ("nnz", C.c_uint64),
        ("block_count", C.c_uint64),
        ("dims", C.c_void_p),
        ("part", C.c_void_p),
        ("bdims", C.c_void_p),
        ("block_off", C.c_void_p),
        ("lidx", C.c_void_p),
        ("row_ptr", C.c_void_p * MAX_ORDER),
        ("where", C.c_void_p * MAX_ORDER),
    ]


def build():
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_use_baseline = False
_blib = None


def use_baseline(on=True):
    """Routes every oracle call to libbaseline.so, the same restatement built
    with -O3 -march=native, for timing the CPU baseline (bench.py); off again
    for checking."""
    global _use_baseline
    _use_baseline = bool(on)


def lib():
    global _lib, _blib
    if _use_baseline:
        if _blib is None:
            path = os.path.join(_HERE, "libbaseline.so")
            if not os.path.exists(path):
                build()
            _blib = _bind(C.CDLL(path))
        return _blib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _lib = _bind(C.CDLL(path))
    return _lib


def max_threads():
    """OpenMP threads the oracle's parallel loops use."""
    return lib().or_max_threads()


def _bind(L):
    if True:
        L.or_block_dims.argtypes = [C.c_int, u32p, u32p, u32p]
        L.or_block_coord.argtypes = [C.c_int, u32p, C.c_uint64, u32p]
        L.or_block_coord.restype = None
        L.or_partition.argtypes = [C.c_int, C.c_uint64, u32p, u32p, u32p, u64p, u64p, u32p]
        L.or_assign_mode.argtypes = [C.c_uint64, C.c_uint32, u32p, u64p, u32p]
        L.or_assign_mode.restype = None
        L.or_plan.argtypes = [C.c_int, u32p, C.c_uint64, C.c_int, C.c_int, C.c_uint64, C.c_int,
                              u32p, u32p, C.POINTER(C.c_int), C.POINTER(C.c_uint64),
                              C.POINTER(C.c_int)]
        L.or_schedule.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64, u64p,
                                  C.c_uint64, C.c_int64, u64p, u64p, u64p, i32p, i32p]
        L.or_schedule.restype = C.c_int64
        L.or_lu_inverse.argtypes = [C.c_int, f64p, f64p]
        for sfx, fp in (("f64", f64p), ("f32", f32p)):
            ctype = C.c_double if sfx == "f64" else C.c_float
            getattr(L, f"or_norm_{sfx}").argtypes = [fp, C.c_uint64]
            getattr(L, f"or_norm_{sfx}").restype = ctype
            getattr(L, f"or_delta_{sfx}").argtypes = [C.POINTER(OrTensor), C.c_void_p, fp, C.c_int, fp]
            getattr(L, f"or_delta_{sfx}").restype = None
            getattr(L, f"or_bc_{sfx}").argtypes = [C.POINTER(OrTensor), fp, fp, C.c_int, C.c_double,
                                                  f64p, f64p]
            getattr(L, f"or_bc_{sfx}").restype = None
            getattr(L, f"or_solve_{sfx}").argtypes = [C.c_int, C.c_uint64, f64p, f64p, fp]
            getattr(L, f"or_solve_{sfx}").restype = None
            getattr(L, f"or_recon_{sfx}").argtypes = [C.POINTER(OrTensor), C.c_void_p, fp, fp]
            getattr(L, f"or_recon_{sfx}").restype = None
            getattr(L, f"or_fit_{sfx}").argtypes = [fp, fp, C.c_uint64, ctype]
            getattr(L, f"or_fit_{sfx}").restype = C.c_double
            getattr(L, f"or_als_{sfx}").argtypes = [C.POINTER(OrTensor), fp, ctype, C.c_void_p, fp,
                                                   C.c_double, C.c_int, C.c_double, f64p]
            getattr(L, f"or_als_{sfx}").restype = C.c_int
        L.or_max_threads.argtypes = []
        L.or_max_threads.restype = C.c_int
    return L


def eigen_ref():
    """The reference's vendored Eigen inverse (oracle/_ref), or None if not built."""
    global _eig
    if _eig is None:
        path = os.path.join(_HERE, "_ref", "libeigenref.so")
        if not os.path.exists(path):
            return None
        E = C.CDLL(path)
        E.eigenref_inverse.argtypes = [C.c_int, f64p, f64p]
        _eig = E
    return _eig


# ---------------------------------------------------------------- integer part
def block_dims(dims, part):
    dims = np.ascontiguousarray(dims, np.uint32)
    part = np.ascontiguousarray(part, np.uint32)
    bd = np.zeros(len(dims), np.uint32)
    if lib().or_block_dims(len(dims), dims, part, bd) != 0:
        raise ValueError("Block dimension is larger than the tensor dimension.")
    return bd


def block_coord(part, block_id):
    part = np.ascontiguousarray(part, np.uint32)
    out = np.zeros(len(part), np.uint32)
    lib().or_block_coord(len(part), part, int(block_id), out)
    return out


def partition(idx, dims, part):
    """idx: (N, nnz) uint32, 0-based.  Returns (block_nnz, perm, local_idx)."""
    idx = np.ascontiguousarray(idx, np.uint32)
    N, nnz = idx.shape
    dims = np.ascontiguousarray(dims, np.uint32)
    part = np.ascontiguousarray(part, np.uint32)
    blocks = int(np.prod(part.astype(np.uint64)))
    bn = np.zeros(blocks, np.uint64)
    perm = np.zeros(max(nnz, 1), np.uint64)
    lidx = np.zeros((N, max(nnz, 1)), np.uint32)
    if lib().or_partition(N, nnz, idx.reshape(-1) if nnz else np.zeros(N, np.uint32), dims, part,
                          bn, perm, lidx.reshape(-1)) != 0:
        raise ValueError("invalid partition or coordinate out of bounds")
    return bn, perm[:nnz], lidx[:, :nnz].copy()


def assign_mode(lidx_n, dim):
    lidx_n = np.ascontiguousarray(lidx_n, np.uint32)
    rp = np.zeros(int(dim) + 1, np.uint64)
    wh = np.zeros(max(len(lidx_n), 1), np.uint32)
    lib().or_assign_mode(len(lidx_n), int(dim), lidx_n if len(lidx_n) else np.zeros(1, np.uint32),
                         rp, wh)
    return rp, wh[: len(lidx_n)]


def plan(dims, nnz, rank, gpu_count, gpu_mem, value_size=4):
    dims = np.ascontiguousarray(dims, np.uint32)
    N = len(dims)
    part = np.zeros(N, np.uint32)
    bd = np.zeros(N, np.uint32)
    streams, avail, typ = C.c_int(), C.c_uint64(), C.c_int()
    if lib().or_plan(N, dims, int(nnz), int(rank), int(gpu_count), int(gpu_mem), int(value_size),
                     part, bd, C.byref(streams), C.byref(avail), C.byref(typ)) != 0:
        raise ValueError("Block dimension is larger than the tensor dimension.")
    return dict(part=part, bdims=bd, streams=streams.value, avail_nnz=avail.value, type=typ.value)


def schedule(typ, gpu_count, stream_count, avail_nnz, block_nnz, nnz_total, max_tasks=1 << 16):
    block_nnz = np.ascontiguousarray(block_nnz, np.uint64)
    tb = np.zeros(max_tasks, np.uint64)
    tn = np.zeros(max_tasks, np.uint64)
    to = np.zeros(max_tasks, np.uint64)
    ts = np.zeros(max_tasks, np.int32)
    td = np.zeros(max_tasks, np.int32)
    n = lib().or_schedule(int(typ), int(gpu_count), int(stream_count), int(avail_nnz),
                          len(block_nnz), block_nnz, int(nnz_total), max_tasks, tb, tn, to, ts, td)
    if n < 0:
        raise ValueError("too many tasks")
    return dict(block=tb[:n], nnz=tn[:n], offset=to[:n], stream=ts[:n], device=td[:n])


def lu_inverse(B):
    B = np.ascontiguousarray(B, np.float64)
    out = np.zeros_like(B)
    lib().or_lu_inverse(B.shape[0], B.reshape(-1), out.reshape(-1))
    return out


# -------------------------------------------------------- partitioned tensor
class Tensor:
    """The reference's partitioned tensor (blocks + per-mode CSR), built by the
    oracle's own partitioner.  ``vals`` are kept in input order and in block
    order (``bvals``)."""

    def __init__(self, idx, vals, dims, part, rank):
        idx = np.ascontiguousarray(idx, np.uint32)
        self.order, self.nnz = idx.shape
        self.rank = int(rank)
        self.dims = np.ascontiguousarray(dims, np.uint32)
        self.part = np.ascontiguousarray(part, np.uint32)
        self.bdims = block_dims(self.dims, self.part)
        self.idx = idx
        self.vals = np.asarray(vals)
        self.block_nnz, self.perm, self.lidx = partition(idx, self.dims, self.part)
        self.block_count = len(self.block_nnz)
        self.block_off = np.zeros(self.block_count + 1, np.uint64)
        self.block_off[1:] = np.cumsum(self.block_nnz)
        self.bvals = self.vals[self.perm.astype(np.int64)]
        self.row_ptr, self.where = [], []
        for n in range(self.order):
            bd = int(self.bdims[n])
            rp = np.zeros(self.block_count * (bd + 1), np.uint64)
            wh = np.zeros(max(self.nnz, 1), np.uint32)
            for b in range(self.block_count):
                lo, hi = int(self.block_off[b]), int(self.block_off[b + 1])
                r, w = assign_mode(self.lidx[n, lo:hi], bd)
                rp[b * (bd + 1):(b + 1) * (bd + 1)] = r
                wh[lo:hi] = w
            self.row_ptr.append(rp)
            self.where.append(wh)
        self._lidx_flat = np.ascontiguousarray(self.lidx.reshape(-1)) if self.nnz else np.zeros(
            self.order, np.uint32)
        s = OrTensor()
        s.order, s.rank, s.nnz, s.block_count = self.order, self.rank, self.nnz, self.block_count
        s.dims = self.dims.ctypes.data
        s.part = self.part.ctypes.data
        s.bdims = self.bdims.ctypes.data
        s.block_off = self.block_off.ctypes.data
        s.lidx = self._lidx_flat.ctypes.data
        for n in range(self.order):
            s.row_ptr[n] = self.row_ptr[n].ctypes.data
            s.where[n] = self.where[n].ctypes.data
        self.s = s

    def rows(self, mode):
        return int(self.part[mode]) * int(self.bdims[mode])

    def global_rows(self):
        """(N, nnz) global factor-row index of each block-ordered nonzero."""
        out = np.zeros_like(self.lidx)
        for b in range(self.block_count):
            lo, hi = int(self.block_off[b]), int(self.block_off[b + 1])
            c = block_coord(self.part, b)
            out[:, lo:hi] = self.lidx[:, lo:hi] + (c * self.bdims)[:, None]
        return out


def _dt(prec):
    return (np.float64, "f64") if prec == "f64" else (np.float32, "f32")


def _ptrs(arrs):
    P = C.c_void_p * len(arrs)
    return P(*[a.ctypes.data for a in arrs])


def norm(vals, prec="f64"):
    dt, s = _dt(prec)
    v = np.ascontiguousarray(vals, dt)
    return getattr(lib(), f"or_norm_{s}")(v, len(v))


def delta(t: Tensor, A, G, mode, prec="f64"):
    dt, s = _dt(prec)
    A = [np.ascontiguousarray(a, dt) for a in A]
    G = np.ascontiguousarray(G, dt).reshape(-1)
    out = np.zeros((max(t.nnz, 1), t.rank), dt)
    getattr(lib(), f"or_delta_{s}")(C.byref(t.s), _ptrs(A), G, int(mode), out.reshape(-1))
    return out[: t.nnz]


def bc(t: Tensor, delta_, mode, lam, prec="f64"):
    dt, s = _dt(prec)
    rows = t.rows(mode)
    B = np.zeros((rows, t.rank, t.rank), np.float64)
    Cv = np.zeros((rows, t.rank), np.float64)
    bv = np.ascontiguousarray(t.bvals, dt)
    d = np.ascontiguousarray(delta_, dt).reshape(-1)
    if t.nnz == 0:
        bv, d = np.zeros(1, dt), np.zeros(1, dt)
    getattr(lib(), f"or_bc_{s}")(C.byref(t.s), bv, d, int(mode), float(lam), B.reshape(-1),
                                 Cv.reshape(-1))
    return B, Cv


def solve(B, Cv, prec="f64"):
    dt, s = _dt(prec)
    rows, R = Cv.shape
    out = np.zeros((rows, R), dt)
    getattr(lib(), f"or_solve_{s}")(R, rows, np.ascontiguousarray(B, np.float64).reshape(-1),
                                    np.ascontiguousarray(Cv, np.float64).reshape(-1),
                                    out.reshape(-1))
    return out


def recon(t: Tensor, A, G, prec="f64"):
    dt, s = _dt(prec)
    A = [np.ascontiguousarray(a, dt) for a in A]
    G = np.ascontiguousarray(G, dt).reshape(-1)
    out = np.zeros(max(t.nnz, 1), dt)
    getattr(lib(), f"or_recon_{s}")(C.byref(t.s), _ptrs(A), G, out)
    return out[: t.nnz]


def fit(bvals, xhat, nrm, prec="f64"):
    dt, s = _dt(prec)
    return getattr(lib(), f"or_fit_{s}")(np.ascontiguousarray(bvals, dt),
                                        np.ascontiguousarray(xhat, dt), len(bvals), nrm)


def als(t: Tensor, A, G, lam, max_iters, tol, prec="f64"):
    """Runs the reference ALS loop; returns (fits, factors)."""
    dt, s = _dt(prec)
    A = [np.ascontiguousarray(a, dt).copy() for a in A]
    G = np.ascontiguousarray(G, dt).reshape(-1)
    fits = np.zeros(max_iters, np.float64)
    nrm = norm(t.vals, prec)
    bv = np.ascontiguousarray(t.bvals, dt)
    it = getattr(lib(), f"or_als_{s}")(C.byref(t.s), bv, nrm, _ptrs(A), G, float(lam),
                                      int(max_iters), float(tol), fits)
    return fits[:it], A


def pad_factor(A_n, rows_padded):
    """Global-padded factor layout: P_n*bd_n rows, the tail rows zero."""
    out = np.zeros((rows_padded, A_n.shape[1]), A_n.dtype)
    out[: A_n.shape[0]] = A_n
    return out
