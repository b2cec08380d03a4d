"""Shared test data: seeded sparse tensors with distinct coordinates, and the
reference-style initial factors/core."""
import numpy as np


def random_coo(seed, dims, nnz, values="uniform", skew=None):
    """Distinct coordinates (duplicates dropped, drawn until nnz are found).
    skew: None for uniform indices, or a Zipf exponent applied per mode."""
    rng = np.random.default_rng(seed)
    dims = [int(d) for d in dims]
    N = len(dims)
    total = int(np.prod([float(d) for d in dims]))
    nnz = min(int(nnz), total)
    strides = np.cumprod([1] + dims[::-1][:-1])[::-1].astype(np.int64) if N else None
    seen = np.zeros(0, np.int64)
    while len(seen) < nnz:
        m = max(2 * (nnz - len(seen)), 64)
        cols = []
        for d in dims:
            if skew is None:
                cols.append(rng.integers(0, d, m))
            else:
                w = 1.0 / np.arange(1, d + 1) ** skew
                cols.append(rng.choice(d, size=m, p=w / w.sum()))
        keys = np.zeros(m, np.int64)
        for k in range(N):
            keys += cols[k].astype(np.int64) * strides[k]
        allk = np.concatenate([seen, keys])
        _, first = np.unique(allk, return_index=True)
        seen = allk[np.sort(first)]
    seen = seen[:nnz]
    idx = np.zeros((N, nnz), np.uint32)
    rem = seen.copy()
    for k in range(N):
        idx[k] = rem // strides[k]
        rem = rem % strides[k]
    if values == "uniform":
        vals = rng.random(nnz)
    elif values == "lognormal":
        vals = rng.lognormal(0.0, 1.0, nnz)
    else:
        raise ValueError(values)
    return idx, vals.astype(np.float32), dims


def random_init(seed, dims, rank):
    rng = np.random.default_rng(seed + 1000)
    A = [rng.random((d, rank)).astype(np.float32) for d in dims]
    G = rng.random((rank,) * len(dims)).astype(np.float32)
    return A, G
