"""Seeded synthetic sparse tensors for the benchmark workloads, and the
reference's base-1 tab-separated coordinate text format (README.md:78-97).

Coordinates are drawn independently per mode (uniform or truncated Zipf(s)
on [0, I_n)), duplicates are dropped and more are drawn until the target
count of distinct nonzeros is reached; the kept set is then put in a seeded
random order, which becomes the input order.  On a GPU the draws, the dedup
sort and the shuffle run through torch (plumbing only: the engine itself
never uses torch); without one, numpy does the same work.
"""
from __future__ import annotations

import numpy as np


def _zipf_cdf(n, s):
    w = 1.0 / np.arange(1, n + 1, dtype=np.float64) ** s
    c = np.cumsum(w)
    return c / c[-1]


def generate(dims, nnz, modes=None, values="uniform", seed=0, device=None, log=None):
    """Returns (idx (N, nnz) uint32 0-based, vals float32, dims).

    modes: per-mode "uniform" or ("zipf", s); values: "uniform" (U[0,1)) or
    "lognormal" (mu=0, sigma=1); log: called with a line per step of the
    bucketed form."""
    dims = [int(d) for d in dims]
    N = len(dims)
    modes = modes or ["uniform"] * N
    total = 1.0
    for d in dims:
        total *= d
    if nnz > total:
        raise ValueError("more nonzeros than cells")
    strides = [1] * N
    for k in range(N - 2, -1, -1):
        strides[k] = strides[k + 1] * dims[k + 1]
    use_torch = device is not None and str(device).startswith("cuda")
    if use_torch and nnz > kBucketNnz:
        return _generate_buckets(dims, nnz, modes, values, seed, device, strides, total, log)
    return _generate_one(dims, nnz, modes, values, seed, device, strides, total)


# Nonzeros above which the GPU generator works in buckets of mode-0 index
# ranges: one torch sort holds at most 2^31 - 1 elements, and the candidate
# draws are 1.25 times the target.
kBucketNnz = 600_000_000


def _generate_one(dims, nnz, modes, values, seed, device, strides, total, r0=None):
    N = len(dims)
    use_torch = device is not None and str(device).startswith("cuda")
    if total >= 2.0 ** 62:
        # two 62-bit key words: the leading modes and the trailing ones
        split = next(j for j in range(1, N) if float(np.prod([float(d) for d in dims[j:]])) < 2.0 ** 62
                     and float(np.prod([float(d) for d in dims[:j]])) < 2.0 ** 62)
        if not use_torch:
            raise ValueError("index spaces beyond 2^62 cells are generated on the GPU")
        return _generate_torch_wide(dims, nnz, modes, values, seed, device, split, r0)
    if use_torch:
        return _generate_torch(dims, nnz, modes, values, seed, device, strides, r0)
    return _generate_numpy(dims, nnz, modes, values, seed, strides, r0)


def _generate_buckets(dims, nnz, modes, values, seed, device, strides, total, log=None):
    """Large targets: mode 0's range is cut into K contiguous pieces of equal
    probability mass, each piece gets a multinomial share of the nonzeros and
    its own seeded draw with mode 0 restricted to the piece (distinct pieces
    never share a coordinate, so each piece's dedup is the global one), and
    the pieces are merged in a seeded random order: slice c of every piece
    forms chunk c, which is shuffled as a whole."""
    import torch
    N = len(dims)
    K = -(-nnz // (kBucketNnz // 2))
    m0 = modes[0]
    if m0 == "uniform":
        cdf = np.arange(1, dims[0] + 1, dtype=np.float64) / dims[0]
    else:
        cdf = _zipf_cdf(dims[0], m0[1])
    cuts = [0] + [int(np.searchsorted(cdf, (b + 1) / K, side="left")) + 1 for b in range(K - 1)]
    cuts = sorted(set(min(max(c, 1), dims[0] - 1) for c in cuts[1:]))
    cuts = [0] + cuts + [dims[0]]
    K = len(cuts) - 1
    mass = np.array([cdf[cuts[b + 1] - 1] - (cdf[cuts[b] - 1] if cuts[b] > 0 else 0.0)
                     for b in range(K)])
    rng = np.random.default_rng(seed + 104729)
    counts = rng.multinomial(nnz, mass / mass.sum())
    pieces = []
    for b in range(K):
        cells = float(cuts[b + 1] - cuts[b]) * total / dims[0]
        if counts[b] > cells:
            raise ValueError("a mode-0 piece has fewer cells than its share of nonzeros")
        ib, vb, _ = _generate_one(dims, int(counts[b]), modes, values, seed * 1009 + 17 * b + 1,
                                  device, strides, total, (cuts[b], cuts[b + 1]))
        pieces.append((ib, vb))
        if log:
            log(f"piece {b + 1}/{K}: mode-0 rows [{cuts[b]}, {cuts[b + 1]}), {counts[b]} nonzeros")
    C = K
    idx = np.empty((N, nnz), np.uint32)
    vals = np.empty(nnz, np.float32)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) + 7)
    pos = 0
    for c in range(C):
        parts_i, parts_v = [], []
        for ib, vb in pieces:
            n = ib.shape[1]
            lo, hi = n * c // C, n * (c + 1) // C
            parts_i.append(ib[:, lo:hi])
            parts_v.append(vb[lo:hi])
        ci = torch.from_numpy(np.concatenate(parts_i, axis=1).view(np.int32)).to(device)
        cv = torch.from_numpy(np.concatenate(parts_v)).to(device)
        n = cv.shape[0]
        perm = torch.randperm(n, generator=g, device=device)
        idx[:, pos:pos + n] = ci[:, perm].cpu().numpy().view(np.uint32)
        vals[pos:pos + n] = cv[perm].cpu().numpy()
        pos += n
        del ci, cv, perm
        if log:
            log(f"merged chunk {c + 1}/{C}")
    torch.cuda.empty_cache()
    return idx, vals, dims


def _generate_numpy(dims, nnz, modes, values, seed, strides, r0=None):
    rng = np.random.default_rng(seed)
    N = len(dims)
    cdfs = [None if m == "uniform" else _zipf_cdf(dims[k], m[1]) for k, m in enumerate(modes)]
    keys = np.zeros(0, np.int64)
    while len(keys) < nnz:
        m = int((nnz - len(keys)) * 1.25) + 1024
        key = np.zeros(m, np.int64)
        for k in range(N):
            lo, hi = (r0 if k == 0 and r0 is not None else (0, dims[k]))
            if cdfs[k] is None:
                col = rng.integers(lo, hi, m)
            else:
                base = cdfs[k][lo - 1] if lo > 0 else 0.0
                u = base + (cdfs[k][hi - 1] - base) * rng.random(m)
                col = np.clip(np.searchsorted(cdfs[k], u, side="right"), lo, hi - 1)
            key += col.astype(np.int64) * strides[k]
        keys = np.unique(np.concatenate([keys, key]))
    keys = keys[rng.permutation(len(keys))[:nnz]]
    idx = np.zeros((N, nnz), np.uint32)
    for k in range(N):
        idx[k] = (keys // strides[k]) % dims[k]
    vals = rng.random(nnz) if values == "uniform" else rng.lognormal(0.0, 1.0, nnz)
    return idx, vals.astype(np.float32), dims


def _draw_col(torch, k, m, dims, cdfs, g, device, r0):
    """m draws of mode k; mode 0 restricted to the index range r0 = (lo, hi)."""
    lo, hi = (r0 if k == 0 and r0 is not None else (0, dims[k]))
    if cdfs[k] is None:
        return torch.randint(lo, hi, (m,), generator=g, device=device)
    c = cdfs[k]
    base = float(c[lo - 1]) if lo > 0 else 0.0
    top = float(c[hi - 1])
    u = base + (top - base) * torch.rand(m, generator=g, device=device, dtype=torch.float64)
    return torch.searchsorted(c, u, right=True).clamp_(min=lo, max=hi - 1)


def _generate_torch(dims, nnz, modes, values, seed, device, strides, r0=None):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    N = len(dims)
    cdfs = [None if m == "uniform" else torch.from_numpy(_zipf_cdf(dims[k], m[1])).to(device)
            for k, m in enumerate(modes)]
    keys = torch.zeros(0, dtype=torch.int64, device=device)
    while keys.numel() < nnz:
        m = int((nnz - keys.numel()) * 1.25) + 1024
        key = torch.zeros(m, dtype=torch.int64, device=device)
        for k in range(N):
            col = _draw_col(torch, k, m, dims, cdfs, g, device, r0)
            key += col.to(torch.int64) * strides[k]
            del col
        keys = torch.unique(torch.cat([keys, key]))
        del key
    perm = torch.randperm(keys.numel(), generator=g, device=device)[:nnz]
    keys = keys[perm]
    del perm
    idx = np.zeros((N, nnz), np.uint32)
    for k in range(N):
        idx[k] = ((keys // strides[k]) % dims[k]).to(torch.int32).cpu().numpy().view(np.uint32)
    if values == "uniform":
        v = torch.rand(nnz, generator=g, device=device, dtype=torch.float32)
    else:
        v = torch.exp(torch.randn(nnz, generator=g, device=device, dtype=torch.float32))
    vals = v.cpu().numpy()
    del keys, v
    torch.cuda.empty_cache()
    return idx, vals, dims


def _generate_torch_wide(dims, nnz, modes, values, seed, device, split, r0=None):
    """As _generate_torch, for index spaces of 2^62 cells and more: a
    coordinate is two key words (modes [0, split) and [split, N)), and
    duplicates are found after a lexicographic sort of the word pairs (stable
    sort by the second word, then by the first)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    N = len(dims)
    cdfs = [None if m == "uniform" else torch.from_numpy(_zipf_cdf(dims[k], m[1])).to(device)
            for k, m in enumerate(modes)]
    groups = [list(range(split)), list(range(split, N))]
    gstr = []
    for grp in groups:
        st = {}
        mult = 1
        for k in reversed(grp):
            st[k] = mult
            mult *= dims[k]
        gstr.append(st)
    hi = torch.zeros(0, dtype=torch.int64, device=device)
    lo = torch.zeros(0, dtype=torch.int64, device=device)
    while hi.numel() < nnz:
        m = int((nnz - hi.numel()) * 1.25) + 1024
        words = []
        for grp, st in zip(groups, gstr):
            w = torch.zeros(m, dtype=torch.int64, device=device)
            for k in grp:
                col = _draw_col(torch, k, m, dims, cdfs, g, device, r0)
                w += col.to(torch.int64) * st[k]
                del col
            words.append(w)
        hi = torch.cat([hi, words[0]])
        lo = torch.cat([lo, words[1]])
        del words
        o = torch.sort(lo, stable=True).indices
        o = o[torch.sort(hi[o], stable=True).indices]
        hi, lo = hi[o], lo[o]
        del o
        keep = torch.ones(hi.numel(), dtype=torch.bool, device=device)
        keep[1:] = (hi[1:] != hi[:-1]) | (lo[1:] != lo[:-1])
        hi, lo = hi[keep], lo[keep]
        del keep
    perm = torch.randperm(hi.numel(), generator=g, device=device)[:nnz]
    hi, lo = hi[perm], lo[perm]
    del perm
    idx = np.zeros((N, nnz), np.uint32)
    for w, grp, st in ((hi, groups[0], gstr[0]), (lo, groups[1], gstr[1])):
        for k in grp:
            idx[k] = ((w // st[k]) % dims[k]).to(torch.int32).cpu().numpy().view(np.uint32)
    if values == "uniform":
        v = torch.rand(nnz, generator=g, device=device, dtype=torch.float32)
    else:
        v = torch.exp(torch.randn(nnz, generator=g, device=device, dtype=torch.float32))
    vals = v.cpu().numpy()
    del hi, lo, v
    torch.cuda.empty_cache()
    return idx, vals, dims


def write_tns(path, idx, vals):
    """Base-1, tab-separated: i_1 ... i_N value (README.md:78-97)."""
    N, nnz = idx.shape
    cols = [idx[k].astype(np.int64) + 1 for k in range(N)]
    with open(path, "w") as f:
        for e in range(nnz):
            f.write("\t".join(str(int(c[e])) for c in cols) + f"\t{float(vals[e])!r}\n")


def init_factors(dims, rank, seed=0):
    """Factors and core drawn U[0,1) from a seed (the reference draws them
    from glibc rand(): tucker.cuh:84-134)."""
    rng = np.random.default_rng(seed + 7919)
    A = [rng.random((d, rank), dtype=np.float32) for d in dims]
    G = rng.random((rank,) * len(dims), dtype=np.float32)
    return A, G
