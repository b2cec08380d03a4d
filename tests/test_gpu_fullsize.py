"""The benchmark workload at full size (3-order 1e6^3, 1e8 nonzeros with
Zipf(1.1) coordinates, R = 10, P = 4x4x4), where the oracle is too slow to
run: properties that hold at any size.

- Each mode update minimises every row's regularised least-squares problem
  exactly, so the ALS objective sum_e (x_e - x^_e)^2 + lambda sum_n |A_n|^2
  (update.cuh:167-194 solves the normal equations of exactly this problem)
  never increases from one mode update to the next.
- The fit taken from the last mode's row solves equals the reconstruction
  pass of reconstruction.cuh:228-277.
- The BF16-refined and FP32 paths reach the same fit.
The smaller parity tests compare the same kernels with the oracle directly."""
import numpy as np
import pytest

import gputucker_amd as gt

pytestmark = pytest.mark.gpu

R = 10


def objective(e, norm, lam):
    """sum_e (x_e - x^_e)^2 + lambda sum_n |A_n|^2, with the error recovered
    from the reconstruction fit 1 - sqrt(err) / |x|."""
    err = ((1.0 - e.fit()) * norm) ** 2
    reg = sum(float(np.square(e.get_factor(n).astype(np.float64)).sum()) for n in range(3))
    return err + lam * reg


def test_benchmark_workload_properties():
    from gputucker_amd import synth
    print("generating 1e8 nonzeros", flush=True)
    idx, vals, dims = synth.generate([1_000_000] * 3, 100_000_000, [("zipf", 1.1)] * 3,
                                     "uniform", seed=0, device="cuda")
    A, G = synth.init_factors(dims, R, seed=0)
    lam = gt.default_lambda()
    norm = float(np.sqrt(np.square(vals.astype(np.float64)).sum()))
    fits = {}
    for prec in ("bf16r", "fp32"):
        print(f"precision {prec}", flush=True)
        with gt.Engine(3, R, partition=[4, 4, 4], lam=lam, precision=prec) as e:
            e.set_coo(idx, vals, dims)
            e.partition()
            e.set_init(A, G)
            prev = objective(e, norm, lam)
            for n in range(3):
                e.mode_update(n)
                obj = objective(e, norm, lam)
                # exact minimisation per row; FP32 factors and rounded deltas
                # leave at most rounding-level slack
                assert obj <= prev * (1 + 1e-6), (prec, n, prev, obj)
                prev = obj
            f_iter = e.iteration()
            f_recon = e.fit()
            assert np.isfinite(f_iter) and 0.0 < f_recon < 1.0
            assert abs(f_iter - f_recon) <= 2e-5 * abs(f_recon), (prec, f_iter, f_recon)
            fits[prec] = f_recon
    assert abs(fits["bf16r"] - fits["fp32"]) <= 1e-6 * abs(fits["fp32"]), fits
