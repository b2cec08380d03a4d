"""gt_fit_noalloc: the fit entry point that allocates no device memory.  Its
error partials live in the library's static device scratch (16384 doubles
per device); it must give what gt_fit gives, call after call, including
after a call on a smaller tensor has left its partials in the shared
scratch, and refuse layouts whose partials exceed the scratch."""
import numpy as np
import pytest

import gputucker_amd as gt
from tests.helpers import random_coo, random_init

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 16384
PARTIALS_PER_SEGMENT = 2048


def _engine(seed, dims, nnz, R, part, **kw):
    idx, vals, dims = random_coo(seed, dims, nnz, skew=1.1)
    A, G = random_init(seed, dims, R)
    e = gt.Engine(len(dims), R, partition=part, **kw)
    e.set_coo(idx, vals, dims)
    e.partition()
    e.set_init(A, G)
    return e


def test_fit_noalloc_matches_fit_on_consecutive_calls():
    print("noalloc: 3-order 300x250x200, 2e5 nnz, R=10, then 90x80x70, 5000 nnz, R=6", flush=True)
    big = _engine(3, [300, 250, 200], 200_000, 10, [2, 2, 2], lam=1e-3)
    big.iteration()
    f_big = big.fit()
    assert big.fit_noalloc() == f_big
    # a smaller tensor next: fewer partials, the first call's tail left in scratch
    small = _engine(4, [90, 80, 70], 5000, 6, [1, 1, 1], lam=1e-3)
    small.iteration()
    f_small = small.fit()
    assert small.fit_noalloc() == f_small
    assert big.fit_noalloc() == f_big
    assert np.isfinite(f_big) and np.isfinite(f_small)


@pytest.mark.parametrize("budget,fits_scratch", [(3_000_000, True), (1_500_000, False)])
def test_fit_noalloc_out_of_core(budget, fits_scratch):
    """Out of core the pass writes 2048 partials per staged segment: a layout
    of four segments fits the scratch, one of eleven is refused."""
    print(f"noalloc: out of core, budget {budget}", flush=True)
    kw = dict(chunk_tiles=1, lam=1e-3)
    e = _engine(11, [300, 200, 100], 30000, 6, [2, 1, 2], device_budget=budget, **kw)
    info = e.info()
    assert info["out_of_core"] == 1
    nseg = -(-30000 // info["segment_nnz"])
    assert (nseg * PARTIALS_PER_SEGMENT + 1 <= SCRATCH_DOUBLES) == fits_scratch, nseg
    e.iteration()
    f = e.fit()
    if fits_scratch:
        assert e.fit_noalloc() == f
    else:
        with pytest.raises(gt.GtError) as ex:
            e.fit_noalloc()
        assert ex.value.code == 5
        assert e.fit() == f  # the refused call left the context usable
