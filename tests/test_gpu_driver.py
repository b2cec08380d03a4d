"""The drop-in executable on the GPU: a full run, resume from its --output
files, and the fused-fit iteration."""
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import random_coo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gputucker_amd", "bin", "GPUTucker")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(BIN), reason="driver not built")]


def run(*args):
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def fits(prefix):
    return np.loadtxt(f"{prefix}_fit.txt", ndmin=2)[:, 1]


@pytest.fixture(scope="module")
def tns(tmp_path_factory):
    idx, vals, dims = random_coo(21, [60, 50, 40], 5000)
    p = tmp_path_factory.mktemp("drv") / "t.tns"
    with open(p, "w") as f:
        for e in range(idx.shape[1]):
            f.write("\t".join(str(int(i) + 1) for i in idx[:, e]) + f"\t{vals[e]:.6f}\n")
    return p


def test_resume_continues_the_run(tns, tmp_path):
    a, b, c = (str(tmp_path / x) for x in "abc")
    common = ["-i", str(tns), "-o", "3", "-r", "5", "--tol", "0"]
    run(*common, "--iters", "3", "--output", a)
    run(*common, "--iters", "2", "--output", b)
    run(*common, "--iters", "1", "--resume", b, "--output", c)
    fa, fb, fc = fits(a), fits(b), fits(c)
    assert len(fa) == 3 and len(fb) == 2 and len(fc) == 1
    assert fb[1] == fa[1]
    # factors and core are written with 9 significant digits, so they read
    # back exactly and the resumed iteration is the third one of run a
    assert fc[0] == fa[2], (fa, fc)
    for m in range(3):
        assert np.array_equal(np.loadtxt(f"{a}_factor{m}.txt"), np.loadtxt(f"{c}_factor{m}.txt"))


def test_fused_fit_run(tns, tmp_path):
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    common = ["-i", str(tns), "-o", "3", "-r", "5", "--tol", "0", "--iters", "2"]
    run(*common, "--output", a)
    out = run(*common, "--fused-fit", "--output", b)
    assert "iter2 :" in out
    assert np.allclose(fits(a), fits(b), rtol=2e-6, atol=0)


@pytest.mark.parametrize("prec", ["fp32", "bf16r"])
def test_driver_fits_match_oracle(tns, tmp_path, prec):
    """The drop-in executable end to end (parse, partition, the glibc rand()
    initialisation, the ALS loop with the reference's default lambda) against
    the FP64 oracle run from the driver's own initial factors and core
    (--dump-init) on the tensor as the driver read it: the FP32 path within
    1e-5, the refined BF16 path within 1e-4 relative; with --ref-fit the FP32
    path's fit in the reference's float arithmetic within 1e-5 of the
    reference-precision oracle."""
    from oracle import Tensor, als
    print(f"driver vs oracle: 60x50x40, 5000 nnz, R=5, 3 iterations, {prec}", flush=True)
    init, out, ref = (str(tmp_path / x) for x in ("init", "out", "ref"))
    common = ["-i", str(tns), "-o", "3", "-r", "5", "--partition", "1,1,1"]
    run(*common, "--dump-init", init)
    run(*common, "--iters", "3", "--tol", "0", "--precision", prec, "--output", out)
    raw = np.loadtxt(str(tns), ndmin=2)
    idx = (raw[:, :3].astype(np.int64) - 1).T.astype(np.uint32)
    vals = raw[:, 3].astype(np.float32)
    dims = [int(idx[m].max()) + 1 for m in range(3)]
    A = [np.loadtxt(f"{init}_factor{m}.txt", ndmin=2) for m in range(3)]
    assert [a.shape for a in A] == [(d, 5) for d in dims]
    G = np.loadtxt(f"{init}_core.txt", ndmin=2).reshape(-1)
    lam = float(np.float32(1e-4))  # the reference's lambda, a float literal
    t = Tensor(idx, vals.astype(np.float64), dims, [1, 1, 1], 5)
    f64, _ = als(t, A, G, lam, 3, 0.0, "f64")
    got = fits(out)
    assert len(got) == 3
    if prec == "bf16r":
        assert np.max(np.abs(got - f64) / np.abs(f64)) < 1e-4, (got, f64)
        return
    assert np.max(np.abs(got - f64)) < 1e-5, (got, f64)
    run(*common, "--iters", "3", "--tol", "0", "--ref-fit", "--output", ref)
    f32, _ = als(Tensor(idx, vals, dims, [1, 1, 1], 5), [a.astype(np.float32) for a in A],
                 G.astype(np.float32), lam, 3, 0.0, "f32")
    assert np.max(np.abs(fits(ref) - f32)) < 1e-5, (fits(ref), f32)
