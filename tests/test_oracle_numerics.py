"""Pins the oracle's floating-point routines (delta, B/C, row solve,
reconstruction, fit, whole ALS) against numpy einsum brute force, textbook
special cases, numpy.linalg.solve, the reference's vendored Eigen inverse and
ALS invariants.  CPU only."""
import numpy as np
import pytest

import oracle
from oracle import brute


def _rand_tensor(seed, N, dims, nnz, R, part):
    rng = np.random.default_rng(seed)
    nnz = min(nnz, int(np.prod(dims)))
    keys = set()
    idx = []
    while len(idx) < nnz:
        c = tuple(int(rng.integers(0, d)) for d in dims)
        if c not in keys:
            keys.add(c)
            idx.append(c)
    idx = np.array(idx, np.uint32).T.copy()
    vals = rng.random(nnz)
    t = oracle.Tensor(idx, vals, dims, part, R)
    A = [oracle.pad_factor(rng.random((d, R)), t.rows(k)) for k, d in enumerate(dims)]
    G = rng.random((R,) * N)
    return t, A, G


@pytest.mark.parametrize("N,R,part", [(3, 3, (1, 1, 1)), (3, 2, (2, 1, 3)), (4, 2, (1, 2, 1, 2)),
                                      (2, 4, (2, 1))])
def test_delta_and_recon_vs_einsum(N, R, part):
    dims = [7, 5, 6, 4][:N]
    t, A, G = _rand_tensor(N * 10 + R, N, dims, 60, R, part)
    rows = t.global_rows()
    for mode in range(N):
        d = oracle.delta(t, A, G, mode)
        ref = brute.delta_einsum(G, rows, A, mode)
        np.testing.assert_allclose(d, ref, rtol=1e-12, atol=1e-12)
        d32 = oracle.delta(t, A, G, mode, "f32")
        np.testing.assert_allclose(d32, ref, rtol=2e-5, atol=1e-6)
    xh = oracle.recon(t, A, G)
    np.testing.assert_allclose(xh, brute.recon_einsum(G, rows, A), rtol=1e-12)
    # reconstruction equals the dense Tucker product at the sampled positions
    dense = brute.dense_tucker(G, [a[:d] for a, d in zip(A, dims)])
    np.testing.assert_allclose(xh, dense[tuple(t.idx[:, t.perm.astype(np.int64)])], rtol=1e-12)


def test_delta_rank1_closed_form():
    t, A, G = _rand_tensor(5, 3, [6, 6, 6], 40, 1, (1, 1, 1))
    rows = t.global_rows()
    for mode in range(3):
        d = oracle.delta(t, A, G, mode)[:, 0]
        ref = G.reshape(-1)[0] * np.prod([A[k][rows[k], 0] for k in range(3) if k != mode], axis=0)
        np.testing.assert_allclose(d, ref, rtol=1e-13)


def test_delta_order2_is_matvec():
    """N=2: delta for mode 0 is G @ a1[i1] (textbook mat-vec)."""
    t, A, G = _rand_tensor(6, 2, [9, 8], 30, 4, (1, 1))
    rows = t.global_rows()
    d0 = oracle.delta(t, A, G, 0)
    np.testing.assert_allclose(d0, A[1][rows[1]] @ G.T, rtol=1e-13)
    d1 = oracle.delta(t, A, G, 1)
    np.testing.assert_allclose(d1, A[0][rows[0]] @ G, rtol=1e-13)


def test_delta_one_hot_rows_pick_core_fibre():
    t, A, G = _rand_tensor(7, 3, [4, 4, 4], 20, 3, (1, 1, 1))
    rows = t.global_rows()
    for k in range(3):
        A[k][:] = 0
        A[k][np.arange(A[k].shape[0]), np.arange(A[k].shape[0]) % 3] = 1
    d = oracle.delta(t, A, G, 0)
    for e in range(t.nnz):
        np.testing.assert_allclose(d[e], G[:, rows[1, e] % 3, rows[2, e] % 3], rtol=0)


@pytest.mark.parametrize("part", [(1, 1, 1), (2, 3, 1)])
def test_bc_and_solve_vs_numpy(part):
    t, A, G = _rand_tensor(11, 3, [8, 9, 7], 150, 3, part)
    rows = t.global_rows()
    lam = 1e-3
    for mode in range(3):
        d = oracle.delta(t, A, G, mode)
        B, Cv = oracle.bc(t, d, mode, lam)
        B2, C2 = brute.gram_numpy(rows[mode].astype(np.int64), d, t.bvals, t.rows(mode), lam)
        np.testing.assert_allclose(B, B2, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(Cv, C2, rtol=1e-12, atol=1e-14)
        a = oracle.solve(B, Cv)
        np.testing.assert_allclose(a, brute.solve_numpy(B2, C2), rtol=1e-9, atol=1e-12)
        # rows with no nonzero get a = 0 (B = lam I, C = 0)
        empty = np.bincount(rows[mode].astype(np.int64), minlength=t.rows(mode)) == 0
        assert (a[empty] == 0).all()


def test_lu_inverse_vs_numpy_and_eigen():
    rng = np.random.default_rng(3)
    eig = oracle.eigen_ref()
    for n in (1, 2, 3, 5, 10, 16, 32):
        M = rng.standard_normal((n, n)) + n * np.eye(n) * rng.random()
        inv = oracle.lu_inverse(M)
        np.testing.assert_allclose(inv, np.linalg.inv(M), rtol=1e-9, atol=1e-11)
        if eig is not None:
            out = np.zeros_like(M)
            eig.eigenref_inverse(n, np.ascontiguousarray(M), out)
            np.testing.assert_allclose(inv, out, rtol=1e-11, atol=1e-13)
    # a pivoting case: zero leading entry forces a row swap
    M = np.array([[0.0, 2.0], [3.0, 1.0]])
    np.testing.assert_allclose(oracle.lu_inverse(M), np.linalg.inv(M), rtol=1e-15)


def test_fit_exact_model_is_one():
    """X sampled exactly from (G, A): starting from the truth with a tiny
    lambda, one ALS iteration keeps fit = 1 to rounding."""
    t, A, G = _rand_tensor(21, 3, [10, 9, 8], 400, 2, (2, 1, 2))
    rows = t.global_rows()
    x = brute.recon_einsum(G, rows, A)
    vals = np.zeros(t.nnz)
    vals[t.perm.astype(np.int64)] = x
    t2 = oracle.Tensor(t.idx, vals, t.dims, t.part, t.rank)
    fits, _ = oracle.als(t2, A, G, 1e-13, 1, 1e-4)
    assert abs(fits[0] - 1.0) < 1e-6


def test_als_objective_monotone():
    t, A, G = _rand_tensor(22, 3, [12, 10, 9], 300, 3, (1, 2, 1))
    lam = 1e-2
    A = [a.copy() for a in A]

    def objective():
        xh = oracle.recon(t, A, G)
        return np.sum((t.bvals - xh) ** 2) + lam * sum(np.sum(a ** 2) for a in A)

    prev = objective()
    for _ in range(3):
        for mode in range(3):
            d = oracle.delta(t, A, G, mode)
            B, Cv = oracle.bc(t, d, mode, lam)
            A[mode] = oracle.solve(B, Cv)
            cur = objective()
            assert cur <= prev * (1 + 1e-12)
            prev = cur


def test_als_loop_matches_stepwise_and_f32_close():
    t, A, G = _rand_tensor(23, 3, [15, 12, 10], 500, 3, (1, 1, 1))
    fits, A_out = oracle.als(t, A, G, 1e-4, 3, 0.0)
    assert len(fits) == 3
    A2 = [a.copy() for a in A]
    for _ in range(3):
        for mode in range(3):
            d = oracle.delta(t, A2, G, mode)
            B, Cv = oracle.bc(t, d, mode, 1e-4)
            A2[mode] = oracle.solve(B, Cv)
    for a, b in zip(A_out, A2):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)
    xh = oracle.recon(t, A2, G)
    fit = 1 - np.sqrt(np.sum((t.bvals - xh) ** 2)) / np.sqrt(np.sum(t.vals ** 2))
    assert abs(fit - fits[-1]) < 1e-12
    f32, _ = oracle.als(t, [a.astype(np.float32) for a in A], G.astype(np.float32), 1e-4, 3, 0.0,
                        "f32")
    assert np.max(np.abs(f32 - fits)) < 1e-4


def test_als_convergence_stops_on_tol():
    t, A, G = _rand_tensor(24, 3, [10, 10, 10], 300, 2, (1, 1, 1))
    fits, _ = oracle.als(t, A, G, 1e-4, 50, 1e-2)
    assert len(fits) < 50
    assert abs(fits[-1] - fits[-2]) <= 1e-2
