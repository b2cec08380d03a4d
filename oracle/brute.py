"""Dense / brute-force numpy restatements used to pin the C oracle.

TEST INFRASTRUCTURE ONLY.  These are independent formulations of the same
mathematics (einsum contractions, numpy's stable sorts, numpy.linalg.solve),
so a dropped term, a wrong index or a transposed operand in oracle.c shows up
as a mismatch against them.
"""
from __future__ import annotations

import string

import numpy as np


def core_unfold_order(order):
    """einsum letters for the core axes."""
    return string.ascii_lowercase[:order]


def delta_einsum(G, rows, A, mode):
    """delta[e, r] = sum_{j: j_mode = r} G[j] prod_{k != mode} A_k[rows[k, e], j_k].

    G: core with shape (R,)*N; rows: (N, nnz) global factor rows; A: list of
    factor matrices (rows x R)."""
    N = G.ndim
    letters = core_unfold_order(N)
    ops, subs = [G], [letters]
    for k in range(N):
        if k == mode:
            continue
        ops.append(A[k][rows[k]])
        subs.append("z" + letters[k])
    spec = ",".join(subs) + "->z" + letters[mode]
    return np.einsum(spec, *ops, optimize=True)


def recon_einsum(G, rows, A):
    N = G.ndim
    letters = core_unfold_order(N)
    ops, subs = [G], [letters]
    for k in range(N):
        ops.append(A[k][rows[k]])
        subs.append("z" + letters[k])
    return np.einsum(",".join(subs) + "->z", *ops, optimize=True)


def dense_tucker(G, A):
    """Full dense reconstruction G x_1 A_1 ... x_N A_N (tiny tensors only)."""
    N = G.ndim
    letters = core_unfold_order(N)
    outl = string.ascii_uppercase[:N]
    subs = [letters] + [outl[k] + letters[k] for k in range(N)]
    return np.einsum(",".join(subs) + "->" + outl, G, *A, optimize=True)


def gram_numpy(row_of_nnz, delta, vals, rows, lam):
    """B_i = lam I + sum delta delta^T, C_i = sum x delta via np.add.at."""
    R = delta.shape[1]
    B = np.zeros((rows, R, R))
    Cv = np.zeros((rows, R))
    np.add.at(B, row_of_nnz, delta[:, :, None].astype(np.float64) * delta[:, None, :])
    np.add.at(Cv, row_of_nnz, delta.astype(np.float64) * vals[:, None])
    B += lam * np.eye(R)[None]
    return B, Cv


def solve_numpy(B, Cv):
    return np.linalg.solve(B, Cv[:, :, None])[:, :, 0]


def partition_python(idx, dims, part):
    """Independent statement of the block partition: block id per nonzero by
    the mixed-radix formula, then Python's stable sort by block id."""
    N, nnz = idx.shape
    bd = [(int(d) + int(p) - 1) // int(p) for d, p in zip(dims, part)]
    bids = []
    for e in range(nnz):
        b = 0
        for k in range(N):
            b = b * int(part[k]) + int(idx[k, e]) // bd[k]
        bids.append(b)
    order = sorted(range(nnz), key=lambda e: bids[e])
    nblocks = int(np.prod([int(p) for p in part]))
    counts = [0] * nblocks
    for b in bids:
        counts[b] += 1
    return bd, counts, order, bids


def csr_python(local_rows, dim):
    """Stable counting sort of positions by row, independent of oracle.c."""
    order = sorted(range(len(local_rows)), key=lambda e: int(local_rows[e]))
    counts = [0] * (dim + 1)
    for r in local_rows:
        counts[int(r) + 1] += 1
    rp = np.cumsum(counts).tolist()
    return rp, order
