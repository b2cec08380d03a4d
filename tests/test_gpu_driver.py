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
