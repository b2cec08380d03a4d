"""bench.py's host-side contract, without a GPU: the one-GPU guard, the
algorithmic bytes per nonzero behind `roofline.achieved`, the kernel each
precision runs, and the committed traffic figure behind `roofline.traffic`."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_algorithmic_bytes_per_nonzero():
    # N indices + value + row-grouping entry + (N-1) gathered rows of R elements
    assert bench.bytes_per_nnz(3, 10, 4) == 4 * 3 + 8 + 2 * 10 * 4 == 100
    assert bench.bytes_per_nnz(3, 10, 2) == 60  # the BF16 figure of the headline workload
    assert bench.bytes_per_nnz(4, 8, 4) == 16 + 8 + 3 * 8 * 4
    assert bench.flops_per_nnz(3, 10) == 2220
    # what the kernels request: the split BF16 copy is 64 bytes per row
    assert bench.kernel_bytes_per_nnz(3, 10, "bf16r") == 12 + 4 + 64 + 40
    assert bench.kernel_bytes_per_nnz(3, 10, "fp32") == 12 + 4 + 40 + 40


def test_kernel_per_delta_path():
    assert bench.kernel_name("bf16r", 10) == "k_gram_bf16<10, 2>"
    assert bench.kernel_name("bf16", 10) == "k_gram_bf16<10, 1>"
    assert bench.kernel_name("fp32", 10) == "k_gram_valu<10>"
    assert bench.kernel_name("fp32-wide", 32) == "k_gram_fp32<32>"
    assert bench.kernel_name("bf16r-l2", 6, 5) == "k_gram_bf16_o5<6>"
    assert bench.kernel_name("bf16r-ob", 5, 5) == "k_gram_bf16_ob<5, 3, true>"
    assert bench.kernel_name("bf16r-ob-l2", 8, 4) == "k_gram_bf16_ob<8, 2, false>"


def test_kernel_rank():
    import gputucker_amd as gt
    assert [gt.kernel_rank(r) for r in (1, 10, 16, 17, 30, 31, 45, 50, 51)] == [
        1, 10, 16, 20, 30, 32, 50, 50, 0]


def test_committed_counters_describe_their_kernel():
    assert bench.committed_counters("c2", "k_gram_valu<999>") is None
    t = bench.measured_traffic("c2", "k_gram_bf16<10, 2>")
    assert t is None or 5e9 < t < 2e10  # bytes per launch, 1e8 nonzeros


def test_workloads():
    c2 = bench.CONFIGS["c2"]
    assert c2["nnz"] == 100_000_000 and c2["dims"] == [1_000_000] * 3 and c2["rank"] == 10
    assert c2["part"] == [4, 4, 4] and c2["modes"] == [("zipf", 1.1)] * 3
    c1 = bench.CONFIGS["c1"]
    assert c1["nnz"] == 1_000_000 and c1["dims"] == [2000] * 3 and c1["lam"] == 1e-3
