"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bars: integer work (block assignment, block COO, per-block row grouping) is
bit-exact; delta, Gram, factors and fit are FP32-path numerics compared with
the FP64 oracle (tolerances written at each assert)."""
import numpy as np
import pytest

import gputucker_amd as gt
import oracle
from tests.helpers import random_coo, random_init

pytestmark = pytest.mark.gpu


def make(idx, vals, dims, R, part, **kw):
    e = gt.Engine(len(dims), R, partition=part, **kw)
    e.set_coo(idx, vals, dims)
    e.partition()
    return e


def check_partition(e, t):
    info = e.info()
    assert info["part"].tolist() == t.part.tolist()
    assert info["bdims"].tolist() == t.bdims.tolist()
    assert e.block_nnz.tolist() == t.block_nnz.tolist()
    for b in range(t.block_count):
        lo, hi = int(t.block_off[b]), int(t.block_off[b + 1])
        n = hi - lo
        idx, vals = e.get_block(b, n)
        assert np.array_equal(idx, t.lidx[:, lo:hi]), f"block {b} indices"
        assert np.array_equal(vals, t.bvals[lo:hi].astype(np.float32)), f"block {b} values"
        for m in range(t.order):
            bd = int(t.bdims[m])
            rp, wh = e.partition_dump(b, m, n, bd)
            assert np.array_equal(rp, t.row_ptr[m][b * (bd + 1):(b + 1) * (bd + 1)]), (b, m)
            assert np.array_equal(wh, t.where[m][lo:hi]), (b, m)


@pytest.mark.parametrize("dims,nnz,part,skew", [
    ([3, 2, 2], 9, [2, 1, 1], None),
    ([40, 30, 20], 3000, [1, 1, 1], None),
    ([40, 30, 20], 3000, [3, 2, 4], None),
    ([500, 400, 300], 60000, [4, 4, 4], 1.1),
    ([60, 50, 40, 30], 20000, [2, 3, 1, 2], None),
    ([30, 30, 30, 30, 5], 30000, [2, 1, 2, 1, 1], None),
    ([2000, 100], 150000, [7, 3], 1.3),
])
def test_partition_bit_exact(dims, nnz, part, skew):
    idx, vals, dims = random_coo(len(dims) * 7 + nnz, dims, nnz, skew=skew)
    t = oracle.Tensor(idx, vals, dims, part, 2)
    e = make(idx, vals, dims, 2, part)
    check_partition(e, t)
    assert abs(e.info()["norm"] - np.sqrt(np.sum(vals.astype(np.float64) ** 2))) < 1e-9 * (
        1 + e.info()["norm"])


def test_partition_readme_example_and_empty_blocks():
    from tests.golden.make_golden import read_tns
    import os
    idx, vals = read_tns(os.path.join(os.path.dirname(__file__), "golden", "readme_example.tns"), 3)
    for part in ([1, 1, 1], [2, 1, 1], [3, 2, 2]):  # (3,2,2) leaves empty blocks
        t = oracle.Tensor(idx, vals, [3, 2, 2], part, 2)
        e = make(idx, vals, [3, 2, 2], 2, part)
        check_partition(e, t)
        assert abs(e.info()["norm"] ** 2 - 192.84) < 1e-4


def _setup(seed, dims, nnz, R, part, skew=None, **kw):
    idx, vals, dims = random_coo(seed, dims, nnz, skew=skew)
    t = oracle.Tensor(idx, vals.astype(np.float64), dims, part, R)
    A, G = random_init(seed, dims, R)
    e = make(idx, vals, dims, R, part, **kw)
    e.set_init(A, G)
    Ap = [oracle.pad_factor(a.astype(np.float64), t.rows(k)) for k, a in enumerate(A)]
    return e, t, Ap, G.astype(np.float64)


@pytest.mark.parametrize("N,R,part", [(3, 10, [1, 1, 1]), (3, 10, [2, 3, 2]), (4, 8, [1, 2, 1, 1]),
                                      (2, 5, [2, 1]), (5, 3, [1, 1, 2, 1, 1]), (3, 1, [1, 1, 1]),
                                      (3, 16, [1, 1, 1]), (3, 20, [1, 1, 1]), (3, 32, [1, 1, 1]),
                                      (3, 17, [1, 2, 1]), (3, 30, [1, 1, 1]), (3, 50, [1, 1, 1]),
                                      (4, 26, [1, 1, 1, 1])])
def test_delta_parity(N, R, part):
    dims = [50, 40, 30, 20, 10][:N]
    e, t, A, G = _setup(N * 100 + R, dims, 4000, R, part)
    for m in range(N):
        d = e.get_delta(m)
        ref = oracle.delta(t, A, G, m)
        scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-30
        err = np.abs(d - ref) / scale
        assert err.max() < 2e-5 * max(1, R ** (N - 1) / 100), (m, err.max())


def _stream_positions(t, idx_stream):
    """Block-order position of each nonzero of a mode's stream, matched on
    its global coordinates."""
    radix = [t.rows(k) for k in range(t.order)]
    g = t.global_rows().astype(np.int64)
    kb = np.ravel_multi_index(tuple(g), radix)
    ks = np.ravel_multi_index(tuple(idx_stream.astype(np.int64)), radix)
    ob = np.argsort(kb)
    pos = ob[np.searchsorted(kb[ob], ks)]
    assert np.array_equal(kb[pos], ks)
    return pos


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-5), ("bf16r", 1e-4), ("bf16", 2e-2)])
@pytest.mark.parametrize("N,R,part", [(3, 10, [2, 1, 2]), (3, 13, [1, 1, 1]), (4, 8, [1, 2, 1, 1]),
                                      (5, 5, [1, 1, 1, 2, 1]), (2, 6, [2, 1]), (3, 20, [1, 1, 1]),
                                      (3, 16, [2, 1, 1]), (3, 30, [1, 1, 1]), (3, 32, [1, 2, 1]),
                                      (3, 17, [1, 1, 1]), (3, 9, [1, 1, 2])])
def test_hot_path_delta_parity(prec, tol, N, R, part):
    """The deltas the mode update's own delta + Gram kernel forms (read back
    through a test hook that stores each tile before its Gram), against the
    FP64 oracle's delta (delta.cuh:62-90) of the same nonzero, relative to
    the largest entry of that nonzero's delta.  Covers the FP32 VALU delta
    (two tiles per pass, and the R > 15 kernel), the refined BF16 MFMA delta
    (order 3 from the LDS table, with the second MFMA at R = 13; orders 4
    and 5 on the o-block kernel, its table in LDS; the wide ranks 16..32 of
    order 3 with the core operand from L2, R = 17 zero-padded at 20) and
    plain BF16."""
    if R > 15 and prec == "bf16":
        pytest.skip("plain BF16 covers R <= 15")
    dims = [90, 70, 50, 30, 20][:N]
    e, t, A, G = _setup(N * 17 + R, dims, 8000, R, part, skew=1.2, precision=prec, chunk_tiles=1)
    scale_n = max(1, R ** (N - 1) / 100)
    for m in range(N):
        d, idx_s = e.get_stream_delta(m)
        ref = oracle.delta(t, A, G, m)[_stream_positions(t, idx_s)]
        scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-30
        err = (np.abs(d - ref) / scale).max()
        assert err < tol * scale_n, (m, err)


@pytest.mark.parametrize("chunk_tiles", [1, 2, 0])
def test_gram_parity_heavy_rows(chunk_tiles):
    """Zipf rows (degrees up to thousands) with tiny chunks: many rows straddle
    chunk boundaries and go through the fix-up."""
    e, t, A, G = _setup(7, [300, 200, 100], 30000, 6, [2, 1, 2], skew=1.3,
                        chunk_tiles=chunk_tiles)
    lam = gt.default_lambda()
    for m in range(3):
        d = oracle.delta(t, A, G, m)
        B, Cv = oracle.bc(t, d, m, lam)
        Bg, Cg = e.get_gram(m, t.rows(m))
        sc = np.abs(B).max(axis=(1, 2), keepdims=True)
        assert (np.abs(Bg - B) / sc).max() < 1e-5
        scc = np.abs(Cv).max(axis=1, keepdims=True) + 1e-30
        assert (np.abs(Cg - Cv) / scc).max() < 1e-5


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("N,R,part,skew", [(3, 10, [1, 1, 1], None), (3, 10, [2, 2, 2], 1.1),
                                           (4, 4, [1, 2, 1, 2], None), (3, 32, [1, 1, 1], None),
                                           (2, 7, [3, 1], None), (3, 30, [1, 1, 1], 1.1),
                                           (3, 19, [2, 1, 1], None), (3, 50, [1, 1, 1], None)])
def test_mode_update_and_fit_parity(N, R, part, skew):
    dims = [120, 100, 80, 60][:N]
    e, t, A, G = _setup(N * 31 + R, dims, 20000, R, part, skew=skew, lam=1e-3)
    lam = 1e-3
    A_ref = [a.copy() for a in A]
    for m in range(N):
        e.mode_update(m)
        d = oracle.delta(t, A_ref, G, m)
        B, Cv = oracle.bc(t, d, m, lam)
        A_ref[m] = oracle.solve(B, Cv)
        got = e.get_factor(m)
        # forward error, against what the reference's own FP32 delta
        # (delta.cuh:62-90 in float, its order of sums) does to the FP64
        # solution: the row systems have condition numbers up to 1e6, so that
        # alone moves it by 6e-6 (R = 10) to 4e-4 (R = 30) here.  At most
        # twice that, or 2e-5
        d32 = oracle.delta(t, [a.astype(np.float32) for a in A_ref], G.astype(np.float32), m,
                           "f32").astype(np.float64)
        B32, C32 = oracle.bc(t, d32, m, lam)
        floor = _rel(oracle.solve(B32, C32)[: dims[m]], A_ref[m][: dims[m]])
        fwd = _rel(got, A_ref[m][: dims[m]])
        assert fwd < max(2e-5, 2 * floor), (m, fwd, floor)
        # normwise backward error against the oracle's exact FP64 system,
        # which the conditioning does not amplify:
        # |B a - C| <= 1e-5 (|B| |a| + |C|) per row
        ag = np.zeros_like(A_ref[m])
        ag[: dims[m]] = got
        res = np.linalg.norm(np.einsum("rij,rj->ri", B, ag) - Cv, axis=1)
        den = np.linalg.norm(B, ord=2, axis=(1, 2)) * np.linalg.norm(ag, axis=1) + np.linalg.norm(
            Cv, axis=1)
        bwd = (res / den)[np.linalg.norm(Cv, axis=1) > 0]
        assert bwd.max() < 1e-5, (m, bwd.max())
        e.set_factor(m, A_ref[m][: dims[m]].astype(np.float32))  # continue from the same point
    f = e.fit()
    xh = oracle.recon(t, [a.astype(np.float32).astype(np.float64) for a in A_ref], G)
    f_ref = 1 - np.sqrt(np.sum((t.bvals - xh) ** 2)) / np.sqrt(np.sum(t.vals ** 2))
    assert abs(f - f_ref) < 1e-5, (f, f_ref)


@pytest.mark.parametrize("part", [[1, 1, 1], [4, 4, 4]])
def test_als_fit_trajectory_vs_oracle(part):
    """Whole loop (gt_run): fits per iteration vs the FP64 oracle and the
    reference-precision (FP32) oracle."""
    dims = [200, 150, 120]
    e, t, A, G = _setup(11, dims, 40000, 10, part, lam=1e-3, max_iters=5, tol=0.0)
    fits = e.run()
    f64, _ = oracle.als(t, A, G, 1e-3, 5, 0.0, "f64")
    f32, _ = oracle.als(t, [a.astype(np.float32) for a in A], G.astype(np.float32), 1e-3, 5, 0.0,
                        "f32")
    assert len(fits) == 5
    assert np.max(np.abs(fits - f64)) < 1e-5, (fits, f64)
    assert np.max(np.abs(fits - f32)) < 1e-5, (fits, f32)


@pytest.mark.parametrize("prec", ["fp32", "bf16r"])
def test_full_size_plain_config_vs_oracle(prec):
    """The plain-method workload at its own size: 2000^3, 1e6 distinct uniform
    nonzeros, U[0,1) values, R = 10, P = 1, 5 ALS iterations, lambda = 1e-3.
    The FP32 path's fit in the reference's float arithmetic (ref_fit) must be
    within 1e-5 of the reference-precision oracle (float delta, FP32 products,
    serial float norm and error sums), and its FP64 fit within 1e-5 of the
    FP64 oracle; the BF16-refined path's fit within 1e-4 relative of FP64.
    At this size the two oracles themselves differ by 1.5e-5, all of it from
    the serial float sums of the norm and the error."""
    from gputucker_amd import synth
    print(f"C1 full size: 2000^3, 1e6 nnz, R=10, P=1, 5 iterations, {prec}", flush=True)
    idx, vals, dims = synth.generate([2000] * 3, 1_000_000, None, "uniform", seed=0)
    A, G = synth.init_factors(dims, 10, seed=0)
    t = oracle.Tensor(idx, vals.astype(np.float64), dims, [1, 1, 1], 10)
    Ap = [oracle.pad_factor(a.astype(np.float64), t.rows(k)) for k, a in enumerate(A)]
    f64, _ = oracle.als(t, Ap, G.astype(np.float64), 1e-3, 5, 0.0, "f64")
    e = make(idx, vals, dims, 10, [1, 1, 1], precision=prec, lam=1e-3, max_iters=5, tol=0.0)
    e.set_init(A, G)
    fits = e.run()
    assert len(fits) == 5
    if prec == "bf16r":
        rel = np.abs(fits - f64) / np.abs(f64)
        assert rel.max() < 1e-4, (fits, f64)
        return
    assert np.max(np.abs(fits - f64)) < 1e-5, (fits, f64)
    t32 = oracle.Tensor(idx, vals, dims, [1, 1, 1], 10)
    f32, _ = oracle.als(t32, A, G, 1e-3, 5, 0.0, "f32")
    er = make(idx, vals, dims, 10, [1, 1, 1], precision="fp32", lam=1e-3, max_iters=5, tol=0.0,
              ref_fit=True)
    er.set_init(A, G)
    fits_ref = er.run()
    assert np.max(np.abs(fits_ref - f32)) < 1e-5, (fits_ref, f32)
    # the same factors as the FP64-sum run: only the fit's arithmetic differs
    for m in range(3):
        assert np.array_equal(er.get_factor(m), e.get_factor(m)), m


def test_error_paths():
    idx, vals, dims = random_coo(1, [10, 10, 10], 100)
    with pytest.raises(gt.GtError) as ex:
        make(idx, vals, dims, 4, [6, 1, 1])  # bd=2 -> ceil(10/2)=5 != 6
    assert ex.value.code == 6
    e = gt.Engine(3, 4)
    with pytest.raises(gt.GtError) as ex:
        e.mode_update(0)
    assert ex.value.code == 3
    bad = idx.copy()
    bad[0, 0] = 10
    with pytest.raises(gt.GtError) as ex:
        e.set_coo(bad, vals, dims)
    assert ex.value.code == 1


@pytest.mark.parametrize("prec,tol_gram", [("bf16", 3e-2), ("bf16r", 2e-4), ("fp32", 1e-5)])
@pytest.mark.parametrize("N,R,part", [(3, 10, [2, 1, 2]), (4, 5, [1, 2, 1, 1]), (2, 8, [1, 3]),
                                      (3, 13, [2, 1, 2]), (3, 15, [1, 1, 1]), (3, 7, [1, 2, 1]),
                                      (2, 13, [2, 1]), (4, 7, [1, 2, 1, 1]), (4, 9, [1, 1, 2, 1]),
                                      (5, 4, [1, 1, 1, 2, 1]), (5, 5, [1, 1, 1, 2, 1]),
                                      (4, 8, [1, 1, 2, 1]), (4, 4, [2, 1, 1, 1]),
                                      (3, 20, [2, 1, 2]), (3, 32, [1, 1, 2]), (3, 24, [1, 2, 1]),
                                      (3, 9, [1, 2, 1])])
def test_mfma_delta_gram_parity(prec, tol_gram, N, R, part):
    """BF16 MFMA delta (plain and with the split-BF16 refinement; above
    R = 10 the refinement takes a second MFMA per tile; orders 4 and 5 run
    the refined delta on the o-block kernel, 16 outer indices per MFMA, at
    R = 5, 7 and 9 of order 4 and R = 4, 5 of order 5 with partial blocks
    or none, and at order 4 with R = 4 and 8 and order 5 with R = 5 the
    Kronecker kernel, two modes per MFMA; plain BF16 there takes the order-3
    kernel's loop over outer
    indices) and the FP32 VALU delta (two tiles per pass) feeding the MFMA
    Gram, against the FP64 oracle's B and C."""
    if R > 15 and prec == "bf16":
        pytest.skip("plain BF16 covers R <= 15")
    dims = [90, 70, 50, 30, 20][:N]
    e, t, A, G = _setup(N * 13 + R, dims, 12000, R, part, skew=1.2, precision=prec,
                        chunk_tiles=1)
    lam = gt.default_lambda()
    for m in range(N):
        d = oracle.delta(t, A, G, m)
        B, Cv = oracle.bc(t, d, m, lam)
        Bg, Cg = e.get_gram(m, t.rows(m))
        sc = np.abs(B).max(axis=(1, 2), keepdims=True)
        assert (np.abs(Bg - B) / sc).max() < tol_gram, (m, (np.abs(Bg - B) / sc).max())
        scc = np.abs(Cv).max(axis=1, keepdims=True) + 1e-30
        assert (np.abs(Cg - Cv) / scc).max() < tol_gram


@pytest.mark.parametrize("nnz", [1, 5, 63, 64, 65, 129])
@pytest.mark.parametrize("N,R,prec", [(3, 10, "fp32"), (3, 10, "bf16r"), (4, 8, "bf16r"),
                                      (5, 5, "bf16r"), (3, 32, "bf16r"), (3, 32, "fp32")])
def test_tiny_tensors(nnz, N, R, prec):
    """Tensors of one nonzero to a few partial tiles (a lone nonzero, one
    partial tile, exactly one tile, one tile and one nonzero, two tiles and
    one) through each kernel family: B and C against the FP64 oracle, then
    a whole ALS iteration's fit within the precision's bar.  Every row here
    has one or two nonzeros, so B = lambda I + a rank-one or rank-two term."""
    print(f"tiny: N={N} R={R} nnz={nnz} {prec}", flush=True)
    dims = [70, 60, 50, 40, 30][:N]
    e, t, A, G = _setup(nnz * 7 + N + R, dims, nnz, R, [1] * N, precision=prec, lam=1e-3,
                        max_iters=1, tol=0.0)
    tol = 1e-5 if prec == "fp32" else 2e-4
    for m in range(N):
        d = oracle.delta(t, A, G, m)
        B, Cv = oracle.bc(t, d, m, 1e-3)
        Bg, Cg = e.get_gram(m, t.rows(m))
        sc = np.abs(B).max(axis=(1, 2), keepdims=True)
        assert (np.abs(Bg - B) / sc).max() < tol, (m, (np.abs(Bg - B) / sc).max())
        scc = np.abs(Cv).max(axis=1, keepdims=True) + 1e-30
        assert (np.abs(Cg - Cv) / scc).max() < tol, m
    fits = e.run()
    f64, _ = oracle.als(t, A, G, 1e-3, 1, 0.0, "f64")
    # (a lone small value is shrunk to zero by lambda: both fits are then 0)
    bar = 1e-5 if prec == "fp32" else 1e-4 * max(abs(f64[0]), 1e-2)
    assert np.abs(fits - f64).max() <= bar, (fits, f64)


@pytest.mark.parametrize("prec", ["fp32", "bf16r"])
@pytest.mark.parametrize("N,R,part", [(4, 15, [1, 2, 1, 1]), (5, 12, [1, 1, 2, 1, 1])])
def test_gram_parity_high_order_wide_rank(prec, N, R, part):
    """Orders 4 and 5 at R = 15 and 12: the BF16 core tables (R^(N-2) outer
    indices) do not fit in LDS, so both precisions take the FP32 VALU delta,
    whose outer factor rows staged in LDS leave room for three and two waves
    per workgroup."""
    dims = [40, 30, 25, 20, 15][:N]
    e, t, A, G = _setup(N * 7 + R, dims, 1500, R, part, precision=prec)
    lam = gt.default_lambda()
    for m in range(N):
        d = oracle.delta(t, A, G, m)
        B, Cv = oracle.bc(t, d, m, lam)
        Bg, Cg = e.get_gram(m, t.rows(m))
        sc = np.abs(B).max(axis=(1, 2), keepdims=True)
        assert (np.abs(Bg - B) / sc).max() < 1e-5, (m, (np.abs(Bg - B) / sc).max())
        scc = np.abs(Cv).max(axis=1, keepdims=True) + 1e-30
        assert (np.abs(Cg - Cv) / scc).max() < 1e-5


@pytest.mark.parametrize("R", [10, 13, 20, 30])
@pytest.mark.parametrize("prec,tol", [("bf16r", 1e-4), ("bf16", 2e-3), ("fp32", 1e-5)])
def test_fit_vs_fp64_oracle_by_precision(prec, tol, R):
    """Fit after 5 ALS iterations, relative to the FP64 oracle: the BF16 path
    with refinement must stay within 1e-4, the FP32 path within 1e-5.  At
    R = 13 the refinement takes two MFMAs per block and the row solves write
    the split copy the next mode's delta gathers; at R = 20 and 30 the wide
    kernel runs and a split pass after each solve rebuilds its copy."""
    if R > 15 and prec == "bf16":
        pytest.skip("plain BF16 covers R <= 15")
    dims = [200, 150, 120]
    e, t, A, G = _setup(17, dims, 40000, R, [2, 2, 2], skew=1.1, lam=1e-3, max_iters=5,
                        tol=0.0, precision=prec)
    fits = e.run()
    f64, _ = oracle.als(t, A, G, 1e-3, 5, 0.0, "f64")
    rel = np.abs(fits - f64) / np.abs(f64)
    assert rel.max() < tol, (prec, fits, f64)


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16r", 1e-4)])
def test_skewed_sample_of_benchmark_workload(prec, tol):
    """2e6 nonzeros drawn like the benchmark workload (1e6^3 index space,
    Zipf(1.1) per mode, default lambda = 0.0001f, P = 4x4x4): most rows have
    fewer than R nonzeros and a few have ~1e5, so the row systems range from
    lambda-dominated to heavily loaded.  One ALS iteration vs the FP64 oracle."""
    from gputucker_amd import synth
    idx, vals, dims = synth.generate([1_000_000] * 3, 2_000_000, [("zipf", 1.1)] * 3,
                                     "uniform", seed=3)
    A, G = synth.init_factors(dims, 10, seed=3)
    lam = gt.default_lambda()
    e = make(idx, vals, dims, 10, [4, 4, 4], precision=prec, max_iters=1, tol=0.0)
    e.set_init(A, G)
    fits = e.run()
    t = oracle.Tensor(idx, vals.astype(np.float64), dims, [4, 4, 4], 10)
    Ap = [oracle.pad_factor(a.astype(np.float64), t.rows(k)) for k, a in enumerate(A)]
    f64, _ = oracle.als(t, Ap, G.astype(np.float64), lam, 1, 0.0, "f64")
    assert np.isfinite(fits).all()
    assert abs(fits[0] - f64[0]) / abs(f64[0]) < tol, (fits, f64)


@pytest.mark.parametrize("prec,pack", [("fp32", True), ("bf16r", True), ("fp32", False)])
def test_out_of_core_matches_in_core(prec, pack, monkeypatch):
    """A device budget below the resident layout stages the per-mode streams
    from pinned host memory segment by segment (double-buffered copies).  With
    every factor at most 2^24 rows the host streams hold packed indices (16
    bits here, every factor having at most 2^16 rows) and no row index: 8
    bytes per nonzero for order 3, unpacked and rebuilt from row_ptr on the
    device; GT_STREAM_PACK=0 ships plain 32-bit indices (16 bytes).  The kernels see the same chunks in the same order, so factors
    must be bit-identical to the in-core run; the fit sums its partials in
    another grouping, so it agrees to FP64 rounding."""
    monkeypatch.setenv("GT_STREAM_PACK", "1" if pack else "0")
    idx, vals, dims = random_coo(11, [300, 200, 100], 30000, skew=1.3)
    A, G = random_init(11, dims, 6)
    kw = dict(precision=prec, chunk_tiles=1, max_iters=2, tol=0.0)
    ref = make(idx, vals, dims, 6, [2, 1, 2], **kw)
    ref.set_init(A, G)
    assert ref.info()["out_of_core"] == 0
    # about 1.4 MB of per-row state stays resident; 1.5 MB leaves room for a
    # few thousand nonzeros per staged segment, and the partition itself runs
    # from host memory in pieces of about 1500 nonzeros: the Zipf rows are
    # split by block and the heaviest blocks by windows of their input order
    oc = make(idx, vals, dims, 6, [2, 1, 2], device_budget=1_500_000, **kw)
    oc.set_init(A, G)
    info = oc.info()
    assert info["out_of_core"] == 1
    assert info["partition_path"] == 1 and info["partition_pieces"] > 3 * 10, info
    assert info["stream_bytes_per_nnz"] == (8 if pack else 16)
    assert 0 < info["segment_nnz"] < 30000 and info["segment_nnz"] % 64 == 0
    f_ref, f_oc = ref.run(), oc.run()
    for m in range(3):
        assert np.array_equal(ref.get_factor(m), oc.get_factor(m)), m
    assert np.allclose(f_ref, f_oc, rtol=0, atol=1e-12), (f_ref, f_oc)
    with pytest.raises(gt.GtError):
        oc.get_delta(0)  # the block COO is not resident out of core


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16r", 2e-5)])
@pytest.mark.parametrize("N,R,part", [(3, 10, [2, 2, 2]), (4, 6, [1, 2, 1, 1]), (3, 20, [1, 1, 1])])
def test_iteration_fit_matches_reconstruction(prec, tol, N, R, part):
    """gt_iteration takes the fit from the last mode's row solves
    (|x|^2 + sum_i a_i^T G_i a_i - 2 a_i^T C_i); gt_fit runs the
    reconstruction pass over the nonzeros.  Both describe the same factors."""
    dims = [120, 100, 80, 60][:N]
    e, t, A, G = _setup(N * 7 + R, dims, 20000, R, part, skew=1.1, precision=prec)
    for _ in range(2):
        f_iter = e.iteration()
        f_rec = e.fit()
        assert abs(f_iter - f_rec) <= tol * abs(f_rec), (f_iter, f_rec)


def _one_iteration_vs_oracle(dims, nnz, modes, values, R, part, prec, seed, tol, **kw):
    from gputucker_amd import synth
    idx, vals, dims = synth.generate(dims, nnz, modes, values, seed=seed)
    A, G = synth.init_factors(dims, R, seed=seed)
    lam = gt.default_lambda()
    e = make(idx, vals, dims, R, part, precision=prec, max_iters=1, tol=0.0, **kw)
    e.set_init(A, G)
    fits = e.run()
    t = oracle.Tensor(idx, vals.astype(np.float64), dims, part, R)
    Ap = [oracle.pad_factor(a.astype(np.float64), t.rows(k)) for k, a in enumerate(A)]
    f64, _ = oracle.als(t, Ap, G.astype(np.float64), lam, 1, 0.0, "f64")
    assert np.isfinite(fits).all()
    assert abs(fits[0] - f64[0]) / abs(f64[0]) < tol, (fits, f64)
    return e


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16r", 1e-4)])
def test_order4_rank8_sample(prec, tol):
    """Shape of the 4-order workload (three Zipf(1.2) modes, one small uniform
    mode, lognormal values, R = 8), at 4e4 nonzeros: one ALS iteration vs the
    FP64 oracle."""
    _one_iteration_vs_oracle([20_000, 20_000, 20_000, 100], 40_000,
                             [("zipf", 1.2)] * 3 + ["uniform"], "lognormal", 8, [2, 2, 1, 1],
                             prec, 5, tol)


def test_rank32_sample():
    """Shape of the rank-32 workload (Zipf(1.5), 5e5 x 5e5 x 5e4), at 1.5e4
    nonzeros: the VALU delta + FP64 Gram for R > 15 and the wave-per-row solve."""
    _one_iteration_vs_oracle([50_000, 50_000, 5_000], 15_000, [("zipf", 1.5)] * 3, "uniform",
                             32, [1, 1, 1], "fp32", 6, 1e-5)


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16r", 1e-4)])
def test_rank32_heavy_row_sample(prec, tol):
    """The rank-32 workload's row skew: 3e5 nonzeros with Zipf(1.5) rows in
    mode 0 (its largest row holds about 1.15e5 nonzeros, spread over some
    1800 chunks, so the two-level fix-up sums it), R = 32, one ALS iteration
    against the FP64 oracle: the FP32 VALU delta and the wide refined BF16
    MFMA delta."""
    print(f"rank 32 heavy row: 5e4 x 5e4 x 5e3, 3e5 nnz, Zipf(1.5) in mode 0, {prec}", flush=True)
    e = _one_iteration_vs_oracle([50_000, 50_000, 5_000], 300_000,
                                 [("zipf", 1.5), "uniform", "uniform"], "uniform", 32, [1, 1, 1],
                                 prec, 6, tol)
    assert e.info()["nnz"] == 300_000
    assert e.timings()["delta_path"] == ["fp32-wide" if prec == "fp32" else "bf16r-wide"] * 3


@pytest.mark.parametrize("prec", ["fp32", "bf16r"])
def test_order5_out_of_core_sample(prec):
    """Shape of the 5-order rank-5 workload that runs under a device budget:
    streams staged from pinned host memory, factors bit-identical to the
    in-core run, fit against the FP64 oracle."""
    from gputucker_amd import synth
    dims0 = [500, 500, 500, 500, 20]
    idx, vals, dims = synth.generate(dims0, 30_000, None, "uniform", seed=9)
    A, G = synth.init_factors(dims, 5, seed=9)
    kw = dict(precision=prec, max_iters=1, tol=0.0, chunk_tiles=1)
    ref = make(idx, vals, dims, 5, [1, 1, 1, 1, 1], **kw)
    ref.set_init(A, G)
    oc = make(idx, vals, dims, 5, [1, 1, 1, 1, 1], device_budget=2_000_000, **kw)
    oc.set_init(A, G)
    assert oc.info()["out_of_core"] == 1 and ref.info()["out_of_core"] == 0
    assert oc.info()["stream_bytes_per_nnz"] == 2 * 4 + 4  # four 16-bit indices and the value
    f_ref, f_oc = ref.run(), oc.run()
    for m in range(5):
        assert np.array_equal(ref.get_factor(m), oc.get_factor(m)), m
    assert np.allclose(f_ref, f_oc, rtol=0, atol=1e-12)
    t = oracle.Tensor(idx, vals.astype(np.float64), dims, [1, 1, 1, 1, 1], 5)
    Ap = [oracle.pad_factor(a.astype(np.float64), t.rows(k)) for k, a in enumerate(A)]
    f64, _ = oracle.als(t, Ap, G.astype(np.float64), gt.default_lambda(), 1, 0.0, "f64")
    assert abs(f_oc[0] - f64[0]) / abs(f64[0]) < (1e-5 if prec == "fp32" else 1e-4)


@pytest.mark.parametrize("prec", ["fp32", "bf16r"])
def test_out_of_core_24bit_indices(prec):
    """Mode 0 has 70000 rows (> 2^16), so the streams of modes 1 and 2 carry
    its indices packed to 24 bits (four in three words) and the other index
    in 16 bits: 3 + 2 + 4 = 9 host bytes per nonzero, the largest over the
    modes.  Two staged segments; factors bit-identical to the in-core run."""
    print(f"out of core 24-bit: 70000x60x50, 20000 nnz, R=4, {prec}", flush=True)
    idx, vals, dims = random_coo(23, [70000, 60, 50], 20000)
    A, G = random_init(23, dims, 4)
    kw = dict(precision=prec, chunk_tiles=1, max_iters=2, tol=0.0)
    ref = make(idx, vals, dims, 4, [1, 1, 1], **kw)
    ref.set_init(A, G)
    # per-row state (Gram rows, factors and for BF16 their split copies, row
    # pointers) is about 10.6 MB here (15.1 MB for bf16r); 0.56 MB more leaves
    # room for two segments of about 11000 nonzeros, and for partition pieces
    # of about 8000, which start at offsets that are not multiples of four
    budget = 11_200_000 if prec == "fp32" else 15_700_000
    oc = make(idx, vals, dims, 4, [1, 1, 1], device_budget=budget, **kw)
    oc.set_init(A, G)
    info = oc.info()
    assert info["out_of_core"] == 1 and ref.info()["out_of_core"] == 0
    assert info["partition_path"] == 1 and info["partition_pieces"] >= 3 * 2, info
    assert info["stream_bytes_per_nnz"] == 9
    assert -(-20000 // info["segment_nnz"]) >= 2
    f_ref, f_oc = ref.run(), oc.run()
    for m in range(3):
        assert np.array_equal(ref.get_factor(m), oc.get_factor(m)), m
    assert np.allclose(f_ref, f_oc, rtol=0, atol=1e-12)


@pytest.mark.parametrize("R,RK", [(17, 20), (33, 36), (45, 50)])
def test_padded_rank_is_the_embedded_model(R, RK):
    """A rank that is not built runs at the next built rank RK with zero
    extra columns and core slices.  Running the same model embedded at rank
    RK by hand (factors and core zero-padded) must give bit-identical
    factors, zero in the extra columns, and the same fits."""
    print(f"padded rank {R} -> {RK}", flush=True)
    idx, vals, dims = random_coo(R, [60, 50, 40], 6000, skew=1.1)
    A, G = random_init(R, dims, R)
    kw = dict(lam=1e-3, max_iters=2, tol=0.0)
    e = make(idx, vals, dims, R, [1, 2, 1], **kw)
    e.set_init(A, G)
    Ak = [np.zeros((d, RK), np.float32) for d in dims]
    for k in range(3):
        Ak[k][:, :R] = A[k]
    Gk = np.zeros((RK,) * 3, np.float32)
    Gk[:R, :R, :R] = G
    ek = make(idx, vals, dims, RK, [1, 2, 1], **kw)
    ek.set_init(Ak, Gk)
    f, fk = e.run(), ek.run()
    assert np.array_equal(f, fk), (f, fk)
    for m in range(3):
        a, ak = e.get_factor(m), ek.get_factor(m)
        assert a.shape == (dims[m], R)
        assert np.array_equal(a, ak[:, :R]), m
        assert not ak[:, R:].any(), m


@pytest.mark.parametrize("prec,N,R,part", [("bf16r", 3, 10, [2, 2, 2]), ("fp32", 4, 6, [1, 2, 1, 1]),
                                           ("fp32", 3, 20, [1, 1, 1])])
def test_graph_iteration_equals_eager(prec, N, R, part):
    """An iteration captured as a HIP graph (kernels only) and replayed must
    give what the eager launches give, bit for bit, replay after replay, with
    eager work (factor reads and writes, a reconstruction pass) between the
    replays; gt_set_init drops the graph and the next iteration recaptures."""
    print(f"graph vs eager: {prec} N={N} R={R}", flush=True)
    dims = [120, 100, 80, 60][:N]
    idx, vals, dims = random_coo(N * 5 + R, dims, 20000, skew=1.1)
    A, G = random_init(N * 5 + R, dims, R)
    kw = dict(precision=prec, lam=1e-3)
    eg = make(idx, vals, dims, R, part, **kw)
    gr = make(idx, vals, dims, R, part, graph=True, **kw)
    for e in (eg, gr):
        e.set_init(A, G)
    for it in range(4):
        f_e, f_g = eg.iteration(), gr.iteration()
        assert f_e == f_g, (it, f_e, f_g)
        for m in range(N):
            assert np.array_equal(eg.get_factor(m), gr.get_factor(m)), (it, m)
        assert eg.fit() == gr.fit()
        if it == 1:  # new factors through the boundary: the graph is kept (same buffers)
            a0 = eg.get_factor(0) * np.float32(0.5)
            eg.set_factor(0, a0)
            gr.set_factor(0, a0)
    assert gr.timings()["ms_iteration_graph"] > 0
    for e in (eg, gr):
        e.set_init(A, G)  # drops the graph
    assert eg.iteration() == gr.iteration()


@pytest.mark.parametrize("prec,N,R,part", [("bf16r", 3, 10, [2, 2, 2]), ("bf16r", 4, 5, [1, 2, 1, 1]),
                                           ("fp32", 3, 20, [1, 1, 1])])
def test_autotune_keeps_parity(prec, N, R, part):
    """gt_autotune times each mode's delta + Gram pass over chunk sizes (and,
    for refined BF16 at order 4, the o-block table in LDS or from L2) and
    keeps the fastest.  The factors are untouched by the tuning, and the ALS
    run that follows agrees with the untuned one to FP64 rounding of the Gram
    sums and with the FP64 oracle to the precision's bar."""
    print(f"autotune: {prec} N={N} R={R}", flush=True)
    dims = [200, 150, 120, 60][:N]
    e, t, A, G = _setup(N * 3 + R, dims, 60000, R, part, skew=1.1, precision=prec, lam=1e-3,
                        max_iters=3, tol=0.0)
    a0 = [e.get_factor(m) for m in range(N)]
    tuned = e.autotune()
    assert all(tl in (1, 2, 4, 8, 16, 32) for tl in tuned["chunk_tiles"]), tuned
    assert all(ms > 0 for ms in tuned["ms"]), tuned
    for m in range(N):
        assert np.array_equal(e.get_factor(m), a0[m])
    fits = e.run()
    ref = make(t.idx, t.vals.astype(np.float32), [int(d) for d in t.dims], R, part,
               precision=prec, lam=1e-3, max_iters=3, tol=0.0)
    ref.set_init([a[: dims[k]].astype(np.float32) for k, a in enumerate(A)], G.astype(np.float32))
    fits_ref = ref.run()
    assert np.allclose(fits, fits_ref, rtol=1e-7, atol=0), (fits, fits_ref)
    f64, _ = oracle.als(t, A, G, 1e-3, 3, 0.0, "f64")
    assert (np.abs(fits - f64) / np.abs(f64)).max() < (1e-5 if prec == "fp32" else 1e-4)


@pytest.mark.parametrize("N,R,part,paths", [
    (4, 5, [1, 2, 1, 1], ("bf16r-ob", "bf16r-ob-l2")),
    (4, 8, [1, 1, 2, 1], ("bf16r-kr", "bf16r-ob")),
    (4, 4, [2, 1, 1, 1], ("bf16r-kr", "bf16r-ob")),
    (4, 10, [1, 1, 1, 2], ("bf16r-ob", "bf16r-ob-l2")),
    (5, 5, [1, 1, 1, 2, 1], ("bf16r-kr", "bf16r-ob")),
    (5, 2, [1, 1, 1, 1, 2], ("bf16r-ob", "bf16r-ob-l2")),
    (5, 6, [1, 1, 2, 1, 1], ("bf16r-l2", "bf16r-l2"))])
def test_core_source_variants(N, R, part, paths):
    """Both sources of the refined BF16 core operand (core_source 0 and 1):
    the o-block kernel (orders 4-5 with R^(N-2) <= 128, the o-block table in
    LDS or from L2, with partial last blocks at R = 5 and 10; at order 4 with
    R = 4 and 8 and order 5 with R = 5 the Kronecker kernel and the o-block
    kernel) and, above
    that, the pair kernel (its pair table, 83 KiB at order 5 and R = 6, does
    not go in LDS beside twelve waves: L2 either way).  Each one's fused
    deltas against the FP64 oracle's (delta.cuh:62-90), and its B and C
    against the oracle's (update.cuh:35-101), at the refined BF16 bars."""
    dims = [90, 70, 50, 30, 20][:N]
    for src, want in enumerate(paths):
        print(f"core source {src}: N={N} R={R}", flush=True)
        e, t, A, G = _setup(N * 29 + R, dims, 9000, R, part, skew=1.2, precision="bf16r",
                            chunk_tiles=1, core_source=src)
        lam = gt.default_lambda()
        scale_n = max(1, R ** (N - 1) / 100)
        for m in range(N):
            d, idx_s = e.get_stream_delta(m)
            assert e.timings()["delta_path"][m] == want, (src, m, e.timings()["delta_path"])
            ref = oracle.delta(t, A, G, m)
            dr = ref[_stream_positions(t, idx_s)]
            sc = np.abs(dr).max(axis=1, keepdims=True) + 1e-30
            assert (np.abs(d - dr) / sc).max() < 1e-4 * scale_n, (src, m)
            B, Cv = oracle.bc(t, ref, m, lam)
            Bg, Cg = e.get_gram(m, t.rows(m))
            sb = np.abs(B).max(axis=(1, 2), keepdims=True)
            assert (np.abs(Bg - B) / sb).max() < 2e-4, (src, m)
            scc = np.abs(Cv).max(axis=1, keepdims=True) + 1e-30
            assert (np.abs(Cg - Cv) / scc).max() < 2e-4, (src, m)
        e.close()


@pytest.mark.parametrize("R", [16, 30, 32])
def test_wide_hot_outer_rows(R):
    """Order 3 with mode 0 Zipf(1.5) over 40 rows.  Modes 1 and 2: a few rows
    of mode 0 (their outer factor) hold at least 16 nonzeros per row of the
    mode, so each row's nonzeros are grouped by those hot outer rows (slots
    by degree, heaviest first, the cold ones last).  Mode 0: most of its
    nonzeros lie in rows holding at least 16 per row of mode 2 (its last
    other factor), so each row's nonzeros are grouped by their mode-2 row.
    Both groupings are stable sorts of the input order, which numpy pins
    (P = 1: block order is input order).  A 16-nonzero block sharing a hot
    outer row h takes M_h[jL][r] = sum_o A_0[h][o] G[o][jL][r], one sharing a
    row l of the last factor N_l[o][r] = sum_jL A_L[l][jL] G[o][jL][r]: the
    deltas (delta.cuh:62-90), B and C (update.cuh:35-101) against the FP64
    oracle at the refined BF16 bars, and the fit after one iteration within
    1e-4 of the FP64 oracle's.  The FP32 path keeps the plain stream."""
    from gputucker_amd import synth
    print(f"wide shared rows: R={R}", flush=True)
    nnz = 40000
    idx, vals, dims = synth.generate([40, 250, 150], nnz, [("zipf", 1.5), "uniform", "uniform"],
                                     "uniform", seed=R)
    part = [1, 1, 1]
    A, G = random_init(R, dims, R)
    e = make(idx, vals, dims, R, part, precision="bf16r", chunk_tiles=1)
    e.set_init(A, G)
    info = e.info()
    hot, kind = info["hot_rows"], info["hot_kind"]
    t = oracle.Tensor(idx, vals.astype(np.float64), dims, part, R)
    Ap = [oracle.pad_factor(a.astype(np.float64), t.rows(k)) for k, a in enumerate(A)]
    Gd = G.astype(np.float64)
    lam = gt.default_lambda()
    scale_n = max(1, R ** 2 / 100)
    for m in range(3):
        L = 1 if m == 2 else 2
        O = 3 - m - L
        deg_m = np.bincount(idx[m], minlength=dims[m])
        if 2 * deg_m[deg_m >= 16 * dims[L]].sum() >= nnz:
            want = (2, dims[L])
            perm = np.lexsort((np.arange(nnz), idx[L], idx[m]))
        else:
            deg = np.bincount(idx[O], minlength=dims[O])
            cand = [h for h in np.lexsort((np.arange(dims[O]), -deg))
                    if deg[h] >= 16 * dims[m]][:63]
            want = (1 if cand else 0, len(cand))
            slot = np.full(dims[O], 63, np.int64)
            slot[cand] = np.arange(len(cand))
            perm = np.argsort(idx[m].astype(np.int64) * 64 + slot[idx[O]], kind="stable")
        assert (kind[m], hot[m]) == want, (m, kind, hot, want)
        d, idx_s = e.get_stream_delta(m)
        assert np.array_equal(idx_s, idx[:, perm]), m
        ref = oracle.delta(t, Ap, Gd, m)
        dr = ref[_stream_positions(t, idx_s)]
        sc = np.abs(dr).max(axis=1, keepdims=True) + 1e-30
        assert (np.abs(d - dr) / sc).max() < 1e-4 * scale_n, m
        B, Cv = oracle.bc(t, ref, m, lam)
        Bg, Cg = e.get_gram(m, t.rows(m))
        sb = np.abs(B).max(axis=(1, 2), keepdims=True)
        assert (np.abs(Bg - B) / sb).max() < 2e-4, m
        scc = np.abs(Cv).max(axis=1, keepdims=True) + 1e-30
        assert (np.abs(Cg - Cv) / scc).max() < 2e-4, m
    assert kind.tolist() == [2, 1, 1] and hot[1] >= 2 and hot[2] >= 2, (kind, hot)
    e.close()
    f = make(idx, vals, dims, R, part, precision="fp32", chunk_tiles=1)
    assert f.info()["hot_rows"].tolist() == [0, 0, 0]
    f.close()
    e = make(idx, vals, dims, R, part, precision="bf16r", max_iters=1, tol=0.0)
    e.set_init(A, G)
    fits = e.run()
    f64, _ = oracle.als(t, Ap, Gd, lam, 1, 0.0, "f64")
    assert abs(fits[0] - f64[0]) / abs(f64[0]) < 1e-4, (fits, f64)
    e.close()


def test_host_partition_windows_inside_a_heavy_block():
    """A row holding most of the tensor in one block (P = 1): under a tight
    device budget the bucketed partition cuts it into windows of its input
    order, a few thousand nonzeros each; the streams, and so the factors,
    are those of the in-core partition, bit for bit.  Also the
    order-5 case with 16-bit packing over many pieces."""
    from gputucker_amd import synth
    print("host partition: heavy row windows", flush=True)
    idx, vals, dims = synth.generate([60, 500, 400], 20000, [("zipf", 2.0), "uniform", "uniform"],
                                     "uniform", seed=4)
    A, G = synth.init_factors(dims, 5, seed=4)
    kw = dict(precision="bf16r", chunk_tiles=1, max_iters=2, tol=0.0)
    ref = make(idx, vals, dims, 5, [1, 1, 1], **kw)
    ref.set_init(A, G)
    oc = make(idx, vals, dims, 5, [1, 1, 1], device_budget=1_600_000, **kw)
    oc.set_init(A, G)
    info = oc.info()
    assert info["partition_path"] == 1 and info["out_of_core"] == 1, info
    assert np.bincount(idx[0]).max() > 8000  # the heavy row: about 12000 of the 20000
    assert np.array_equal(ref.block_nnz, oc.block_nnz)
    assert abs(ref.info()["norm"] - info["norm"]) <= 1e-12 * info["norm"]
    f_ref, f_oc = ref.run(), oc.run()
    for m in range(3):
        assert np.array_equal(ref.get_factor(m), oc.get_factor(m)), m
    assert np.allclose(f_ref, f_oc, rtol=0, atol=1e-12)
    idx, vals, dims = synth.generate([300, 300, 300, 300, 20], 40_000, None, "uniform", seed=8)
    A, G = synth.init_factors(dims, 4, seed=8)
    kw = dict(precision="fp32", chunk_tiles=1, max_iters=1, tol=0.0)
    ref = make(idx, vals, dims, 4, [1, 2, 1, 1, 1], **kw)
    ref.set_init(A, G)
    oc = make(idx, vals, dims, 4, [1, 2, 1, 1, 1], device_budget=1_600_000, **kw)
    oc.set_init(A, G)
    assert oc.info()["partition_path"] == 1 and oc.info()["stream_bytes_per_nnz"] == 4 * 2 + 4
    f_ref, f_oc = ref.run(), oc.run()
    for m in range(5):
        assert np.array_equal(ref.get_factor(m), oc.get_factor(m)), m


@pytest.mark.parametrize("limit", [1, 700_000])
def test_out_of_core_pageable_host_memory(limit):
    """When the system grants no more page-locked memory, the out-of-core
    host data (the waiting input, the per-mode streams) lives in pageable
    memory instead; host_pinned_limit forces that here (1 byte: all of it;
    700 kB: the input and the first streams pinned, the rest pageable).  Copies from pageable
    memory go through the runtime's staging, and the factors must still be
    those of the in-core run, bit for bit."""
    print(f"out of core, pageable host memory past {limit} pinned bytes", flush=True)
    idx, vals, dims = random_coo(11, [300, 200, 100], 30000, skew=1.3)
    A, G = random_init(11, dims, 6)
    kw = dict(precision="bf16r", chunk_tiles=1, max_iters=2, tol=0.0)
    ref = make(idx, vals, dims, 6, [2, 1, 2], **kw)
    ref.set_init(A, G)
    oc = make(idx, vals, dims, 6, [2, 1, 2], device_budget=1_500_000, host_pinned_limit=limit,
              **kw)
    oc.set_init(A, G)
    info = oc.info()
    assert info["out_of_core"] == 1 and info["partition_path"] == 1, info
    # streams: 30000 nonzeros x (two 16-bit indices + a value) x 3 modes
    assert 0 < info["host_pageable_bytes"] <= 3 * 30000 * 8 + 4096, info
    f_ref, f_oc = ref.run(), oc.run()
    for m in range(3):
        assert np.array_equal(ref.get_factor(m), oc.get_factor(m)), m
    assert np.allclose(f_ref, f_oc, rtol=0, atol=1e-12), (f_ref, f_oc)
