"""CPU-side checks of the HIP library: it loads, exports every symbol the
header declares, and its host logic (planner, scheduler, defaults, error
paths) matches the oracle.  No kernel is launched."""
import os
import re

import numpy as np
import pytest

import gputucker_amd as gt
import oracle

HDR = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "gputucker_hip.h")


def header_symbols():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gt_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = gt.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), f"missing export {s}"
    assert sorted(gt.EXPORTS) == syms


def test_version_and_defaults():
    import ctypes as C
    assert gt.lib().gt_version() >= 100
    o = gt.GtOpts()
    assert gt.lib().gt_default_opts(C.byref(o)) == 0
    assert o.rank == 10 and o.max_iters == 3 and o.order == 3
    assert o.lambda_ == float(np.float32(1e-4)) == o.tol  # 0.0001f, constants.hpp:18


@pytest.mark.parametrize("case", [
    ([2000, 2000, 2000], 1_000_000, 10, 64 << 30),
    ([100000, 50000, 2000], 50_000_000, 10, 256 << 20),
    ([300, 40000, 5000, 70], 20_000_000, 5, 128 << 20),
    ([1_000_000] * 3, 100_000_000, 10, 1 << 30),
    ([50000, 50000, 50000, 50000, 200], 2_000_000_000, 5, 16 << 30),
])
def test_planner_matches_oracle(case):
    dims, nnz, R, mem = case
    a = gt.plan_partition(dims, nnz, R, mem)
    b = oracle.plan(dims, nnz, R, 1, mem)
    assert a["part"].tolist() == b["part"].tolist()
    assert a["bdims"].tolist() == b["bdims"].tolist()
    assert (a["streams"], a["avail_nnz"], a["type"]) == (b["streams"], b["avail_nnz"], b["type"])


@pytest.mark.parametrize("case", __import__("tests.test_oracle_partition",
                                            fromlist=["PLAN_CASES"]).PLAN_CASES)
def test_planner_hand_derived(case):
    """The library's planner against the hand-derived plans (not against the
    oracle's copy of the same code)."""
    a = gt.plan_partition(case["dims"], case["nnz"], case["R"], case["mem"])
    assert a["type"] == 0 and a["streams"] == 4
    assert a["part"].tolist() == case["part"]
    assert a["bdims"].tolist() == case["bd"]
    assert a["avail_nnz"] == case["avail"]


def test_planner_forced_and_invalid():
    a = gt.plan_partition([1000, 1000, 1000], 10**6, 10, 1 << 30, forced=[4, 4, 4])
    assert a["part"].tolist() == [4, 4, 4] and a["bdims"].tolist() == [250, 250, 250]
    assert a["type"] == 0 and a["streams"] == 4
    with pytest.raises(gt.GtError):
        gt.plan_partition([10, 10], 5, 2, 1 << 20, forced=[6, 1])  # bd=2 -> ceil(10/2)=5 != 6


def test_scheduler_matches_oracle():
    rng = np.random.default_rng(5)
    bn = rng.permutation(np.arange(1, 41) * 1000).astype(np.uint64)  # distinct: no sort ties
    for avail in (500, 7000, 10**9):
        a = gt.schedule(0, 4, avail, bn, int(bn.sum()))
        b = oracle.schedule(0, 1, 4, avail, bn, int(bn.sum()))
        for k in ("block", "nnz", "offset", "stream"):
            assert a[k].tolist() == b[k].tolist()
    a = gt.schedule(1, 1, 300, [1000], 1000)
    b = oracle.schedule(1, 1, 1, 300, [1000], 1000)
    assert a["nnz"].tolist() == b["nnz"].tolist() == [300, 300, 300, 100]


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_fails_cleanly():
    with pytest.raises(gt.GtError) as e:
        gt.Engine(3, 10)
    assert e.value.code != 0


def test_create_rejects_bad_options():
    with pytest.raises(gt.GtError) as e:
        gt.Engine(3, 51)  # above the reference's 50 (delta.cuh:63)
    assert e.value.code == 4
    with pytest.raises(gt.GtError) as e:
        gt.Engine(1, 4)  # order 1
    assert e.value.code == 4
    with pytest.raises(gt.GtError) as e:
        gt.Engine(3, 4, lam=0.0)
    assert e.value.code == 1


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of gt_opts, gt_partition_info and gt_timings have
    the header's size and field offsets (a C program built from the header
    prints them): a field added on one side only would make the library
    write past the caller's struct."""
    import ctypes
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = {"gt_opts": gt.GtOpts, "gt_partition_info": gt.GtPartitionInfo,
               "gt_timings": gt.GtTimings}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "gputucker_hip.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in py._fields_:  # (a Python keyword gets a trailing underscore: lambda_)
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0].rstrip("_")}));')
    lines += ["  return 0;", "}"]
    src, exe = tmp_path / "layout.c", tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HDR), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True,
                               check=True).stdout.splitlines():
        c, field, v = line.split()
        got[(c, field)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
