"""The synthetic workload generator: distinct coordinates in bounds, and the
bucketed form it takes for targets beyond one GPU sort (mode 0's range in
pieces, merged in a seeded random order), exercised here on the CPU."""
import numpy as np
import pytest

from gputucker_amd import synth


def _space(dims):
    strides = [1] * len(dims)
    for k in range(len(dims) - 2, -1, -1):
        strides[k] = strides[k + 1] * dims[k + 1]
    return strides, float(np.prod([float(d) for d in dims]))


def test_distinct_in_bounds():
    dims = [50, 40, 30]
    idx, vals, d = synth.generate(dims, 5000, [("zipf", 1.1), "uniform", "uniform"], seed=1)
    assert idx.shape == (3, 5000) and d == dims
    assert len(set(map(tuple, idx.T.tolist()))) == 5000
    assert all(idx[k].max() < dims[k] for k in range(3))
    assert (vals >= 0).all() and (vals < 1).all()


@pytest.mark.parametrize("modes", [None, [("zipf", 1.2), "uniform", "uniform"]])
def test_bucketed_generation(monkeypatch, modes):
    pytest.importorskip("torch")
    monkeypatch.setattr(synth, "kBucketNnz", 2000)  # pieces of about 1000 nonzeros
    dims, nnz = [300, 200, 100], 9000
    strides, total = _space(dims)
    modes = modes or ["uniform"] * 3
    idx, vals, d = synth._generate_buckets(dims, nnz, modes, "uniform", 3, "cpu", strides, total)
    assert idx.shape == (3, nnz) and vals.shape == (nnz,)
    assert len(set(map(tuple, idx.T.tolist()))) == nnz  # distinct across the pieces
    assert all(idx[k].max() < dims[k] for k in range(3))
    # the merged order is mixed: the pieces' mode-0 ranges are not laid out in order
    assert not (np.diff(idx[0].astype(np.int64)) >= 0).all()
    first_half = idx[0][: nnz // 2]
    assert first_half.max() > dims[0] // 2
    if modes[0] != "uniform":  # Zipf: index 0 is the most frequent in mode 0
        assert np.bincount(idx[0]).argmax() == 0
    again = synth._generate_buckets(dims, nnz, modes, "uniform", 3, "cpu", strides, total)
    assert np.array_equal(again[0], idx) and np.array_equal(again[1], vals)
