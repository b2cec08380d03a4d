"""Pins the oracle's integer routines (partitioner, row grouping, planner,
scheduler) against the README example, hand-derived values, an independent
Python statement and closed-form invariants.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import brute

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _readme():
    from tests.golden.make_golden import read_tns
    return read_tns(os.path.join(GOLD, "readme_example.tns"), 3)


def test_readme_example_counts():
    idx, vals = _readme()
    gold = json.load(open(os.path.join(GOLD, "readme_partition.json")))
    assert [int(idx[k].max()) + 1 for k in range(3)] == gold["dims"] == [3, 2, 2]
    assert idx.shape[1] == gold["nnz"] == 9
    assert abs(np.sum(vals ** 2) - 192.84) < 1e-9
    # norm in value_t (double and float), serial input order
    assert abs(oracle.norm(vals, "f64") - np.sqrt(192.84)) < 1e-12
    assert abs(oracle.norm(vals, "f32") - np.sqrt(192.84)) < 1e-5


def test_readme_partition_by_hand_P211():
    """P=(2,1,1): bd=(2,2,2); only (3,2,1) lands in block 1 (hand derived)."""
    idx, _ = _readme()
    bn, perm, lidx = oracle.partition(idx, [3, 2, 2], [2, 1, 1])
    assert bn.tolist() == [8, 1]
    assert perm.tolist() == [0, 1, 2, 3, 5, 6, 7, 8, 4]
    assert lidx[:, 8].tolist() == [0, 1, 0]  # global (2,1,0) minus base (2,0,0)
    rp, wh = oracle.assign_mode(lidx[0, :8], 2)
    assert rp.tolist() == [0, 4, 8] and wh.tolist() == [0, 1, 4, 5, 2, 3, 6, 7]
    rp, wh = oracle.assign_mode(lidx[1, :8], 2)
    assert rp.tolist() == [0, 4, 8] and wh.tolist() == [0, 2, 4, 6, 1, 3, 5, 7]


@pytest.mark.parametrize("case", range(6))
def test_readme_partition_golden(case):
    idx, _ = _readme()
    gold = json.load(open(os.path.join(GOLD, "readme_partition.json")))
    c = gold["cases"][case]
    t = oracle.Tensor(idx, np.zeros(9), gold["dims"], c["part"], 2)
    assert t.bdims.tolist() == c["block_dims"]
    assert t.block_nnz.tolist() == c["block_nnz"]
    assert t.perm.tolist() == c["perm"]
    for n in range(3):
        bd = c["block_dims"][n]
        for b, blk in enumerate(c["csr"][n]):
            lo, hi = int(t.block_off[b]), int(t.block_off[b + 1])
            assert t.row_ptr[n][b * (bd + 1):(b + 1) * (bd + 1)].tolist() == blk["row_ptr"]
            assert t.where[n][lo:hi].tolist() == blk["where"]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_partition_random_vs_python(seed):
    rng = np.random.default_rng(seed)
    N = 3 + seed % 2
    dims = rng.integers(5, 40, N)
    nnz = 500
    idx = np.stack([rng.integers(0, d, nnz) for d in dims]).astype(np.uint32)
    part = [int(rng.integers(1, min(4, d) + 1)) for d in dims]
    try:
        bd_ref = oracle.block_dims(dims, part)
    except ValueError:
        pytest.skip("invalid P for these dims")
    bd, counts, order, bids = brute.partition_python(idx, dims, part)
    bn, perm, lidx = oracle.partition(idx, dims, part)
    assert bd_ref.tolist() == bd
    assert bn.tolist() == counts
    assert perm.tolist() == order
    for j, e in enumerate(order):
        coord = oracle.block_coord(part, bids[e])
        assert (lidx[:, j] == idx[:, e] - coord * bd_ref).all()
    # row grouping against Python's stable sort
    starts = np.concatenate([[0], np.cumsum(counts)])
    for n in range(N):
        for b in range(len(counts)):
            loc = lidx[n, starts[b]:starts[b + 1]]
            rp, wh = oracle.assign_mode(loc, int(bd_ref[n]))
            rp2, wh2 = brute.csr_python(loc.tolist(), int(bd_ref[n]))
            assert rp.tolist() == rp2 and wh.tolist() == wh2


def test_block_dims_rules():
    assert oracle.block_dims([10], [3]).tolist() == [4]
    assert oracle.block_dims([2000, 2000, 2000], [4, 4, 4]).tolist() == [500, 500, 500]
    with pytest.raises(ValueError):
        oracle.block_dims([3], [4])  # P > I
    assert oracle.block_dims([10], [4]).tolist() == [3]  # ceil(10/3) = 4 = P: valid
    with pytest.raises(ValueError):
        oracle.block_dims([10], [6])  # bd = 2 but ceil(10/2) = 5 != 6


def test_block_coord_roundtrip():
    part = [3, 1, 4, 2]
    for b in range(24):
        c = oracle.block_coord(part, b)
        # row-major, last axis fastest
        assert ((c[0] * 1 + 0) * 4 + c[2]) * 2 + c[3] == b


def test_plan_small_fits_in_memory():
    p = oracle.plan([2000, 2000, 2000], 1_000_000, 10, 1, 64 << 30)
    assert p["type"] == 1 and p["streams"] == 1 and p["part"].tolist() == [1, 1, 1]
    # avail = (mem - core - subfactors) / (N*4 + 4 + R*4)
    core = 1000 * (3 * 4 + 4)
    sub = 3 * 2000 * 10 * 4
    assert p["avail_nnz"] == ((64 << 30) - (core + sub)) // (12 + 4 + 40)


# Hand-derived plans of the dimension-based branch (optimizer.tpp:137-153,
# 202-376), worked step by step below with the reference's size model:
#   all_data = avg*(4N + 4) + R^N*(4N + 4) + sum_n bd_n*R*4 + avg*R*4,
#   avg = ceil(nnz / blocks); stop at the first P with all_data < mem/4 and
#   blocks >= 4; each step increments the axis of the largest block_dims,
#   ties to the lowest axis (optimizer.tpp:343-351);
#   avail = (mem/4 - R^N*(4N + 4) - sum_n bd_n*R*4) / (4N + 4 + 4R).
PLAN_CASES = [
    # dims (1e5, 5e4, 2e3), 5e7 nnz, R = 10, mem = 256 MiB, mem/4 = 67108864.
    # Steps (axis chosen): (2,1,1) 0; (3,1,1) 0 [tie 50000/50000]; (3,2,1) 1;
    # (4,2,1) 0; (5,2,1) 0 [tie 25000]; (5,3,1) 1; (6,3,1) 0; (7,3,1) 0
    # [tie 16667]; (7,4,1) 1; (8,4,1) 0; (9,4,1) 0 [tie 12500]:
    #   36 blocks, avg 1388889: 77777784 + 16000 + (11112+12500+2000)*40
    #   = 78818264, not below mem/4;
    # (9,5,1) 1: 45 blocks, avg 1111112: 62222272 + 16000 + 924480 = 63162752 < mem/4.
    # avail = (67108864 - 16000 - 924480) / 56 = 1181578.
    dict(dims=[100000, 50000, 2000], nnz=50_000_000, R=10, mem=256 << 20,
         part=[9, 5, 1], bd=[11112, 10000, 2000], avail=1181578),
    # dims (4000, 4000, 10), 1e6 nnz, R = 4, mem = 24e6 (mem/4 = 6e6); the
    # whole tensor needs 16e6 + 1024 + 128160 + 16e6 > mem: dimension-based.
    # (2,1,1) 0 [tie 4000]: 2 blocks; (2,2,1) 1: avg 250000, 8000000 + 1024 +
    # (2000+2000+10)*16 = 8065184; (3,2,1) 0 [tie 2000]: avg 166667, 5333344 +
    # 1024 + (1334+2000+10)*16 = 5387872 < 6e6.  A tie broken to the highest
    # axis would give (2,3,1).  avail = (6e6 - 1024 - 53504) / 32 = 185796.
    dict(dims=[4000, 4000, 10], nnz=1_000_000, R=4, mem=24_000_000,
         part=[3, 2, 1], bd=[1334, 2000, 10], avail=185796),
    # order 4: dims (3000, 3000, 2000, 5), 2e6 nnz, R = 6, mem = 36e6
    # (mem/4 = 9e6); the whole tensor needs 2e6*44 + 1296*20 + 8005*24 =
    # 88218040 > mem.  (2,1,1,1) 0 [tie 3000]; (2,2,1,1) 1; (2,2,2,1) 2:
    # avg 250000, 11000000 + 25920 + (1500+1500+1000+5)*24 = 11122040;
    # (3,2,2,1) 0 [tie 1500]: avg 166667, 7333348 + 25920 + 3505*24 = 7443388
    # < 9e6.  Ties to the highest axis would give (2,3,2,1).
    # avail = (9e6 - 25920 - 84120) / 44 = 202044.
    dict(dims=[3000, 3000, 2000, 5], nnz=2_000_000, R=6, mem=36_000_000,
         part=[3, 2, 2, 1], bd=[1000, 1500, 1000, 5], avail=202044),
]


@pytest.mark.parametrize("case", PLAN_CASES)
def test_plan_dimension_branch_hand_derived(case):
    p = oracle.plan(case["dims"], case["nnz"], case["R"], 1, case["mem"])
    assert p["type"] == 0 and p["streams"] == 4
    assert p["part"].tolist() == case["part"]
    assert p["bdims"].tolist() == case["bd"]
    assert p["avail_nnz"] == case["avail"]


def test_plan_dimension_partitioning_invariants():
    dims, nnz, R, mem = [100000, 50000, 2000], 50_000_000, 10, 256 << 20
    p = oracle.plan(dims, nnz, R, 1, mem)
    assert p["type"] == 0 and p["streams"] == 4
    part = p["part"].astype(np.int64)
    blocks = int(np.prod(part))
    assert blocks >= 4
    bd = p["bdims"].astype(np.int64)
    avg = -(-nnz // blocks)
    size = avg * (3 * 4 + 4) + 1000 * 16 + int(np.sum(bd * R)) * 4 + avg * R * 4
    assert size < mem // 4
    # the step before the last must have failed the stop test (minimality):
    # undo one increment on the axis the planner touched last
    assert (bd == np.ceil(np.array(dims) / part)).all()


def test_schedule_nonzero_and_dimension():
    s = oracle.schedule(1, 1, 1, 400, [1000], 1000)
    assert s["nnz"].tolist() == [400, 400, 200]
    assert s["offset"].tolist() == [0, 400, 800]
    assert (s["block"] == 0).all() and (s["stream"] == 0).all()
    bn = [50, 10, 900, 0, 30]
    s = oracle.schedule(0, 1, 4, 400, bn, sum(bn))
    # blocks by ascending nnz: 3(0),1(10),4(30),0(50),2(900 -> 400,400,100)
    assert s["block"].tolist() == [3, 1, 4, 0, 2, 2, 2]
    assert s["nnz"].tolist() == [0, 10, 30, 50, 400, 400, 100]
    assert s["offset"].tolist() == [0, 0, 0, 0, 0, 400, 800]
    assert s["stream"].tolist() == [0, 1, 2, 3, 0, 1, 2]
    # every nonzero of every block covered exactly once
    for b, n in enumerate(bn):
        m = s["block"] == b
        assert int(s["nnz"][m].sum()) == n
