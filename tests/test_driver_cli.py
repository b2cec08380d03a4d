"""The drop-in CLI driver: flags and input-format semantics of the reference
(cmdline_opts.cpp:27-107, tensor_manager.tpp:46-191), exercised through its
--parse-only mode, which needs no GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gputucker_amd", "bin", "GPUTucker")
README_TNS = os.path.join(ROOT, "tests", "golden", "readme_example.tns")

pytestmark = pytest.mark.skipif(not os.path.exists(BIN), reason="driver not built")


def run(*args):
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout + r.stderr


def summary(out):
    line = [l for l in out.splitlines() if l.startswith("parsed ")][0].split()
    order, nnz = int(line[2]), int(line[4])
    dims = [int(x) for x in line[6:6 + order]]
    return order, nnz, dims, float(line[-1])


def test_readme_example_tabs():
    rc, out = run("-i", README_TNS, "-o", "3", "--parse-only")
    assert rc == 0, out
    order, nnz, dims, ss = summary(out)
    assert (order, nnz, dims) == (3, 9, [3, 2, 2])
    assert abs(ss - 192.84) < 1e-4


def test_readme_example_spaces(tmp_path):
    p = tmp_path / "spaces.tns"
    p.write_text(open(README_TNS).read().replace("\t", "    "))
    rc, out = run("-i", str(p), "-o", "3", "--parse-only")
    assert rc == 0
    assert summary(out)[:3] == (3, 9, [3, 2, 2])


def test_zero_based_index_is_rejected(tmp_path):
    p = tmp_path / "zero.tns"
    p.write_text("0\t1\t1\t1.0\n1\t1\t1\t2.0\n")
    rc, out = run("-i", str(p), "-o", "3", "--parse-only")
    assert rc != 0 and "base-1" in out


def test_missing_value_draws_uniform(tmp_path):
    p = tmp_path / "noval.tns"
    p.write_text("1 1 1\n2 2 2 3.5\n")
    rc, out = run("-i", str(p), "-o", "3", "--parse-only")
    assert rc == 0
    order, nnz, dims, ss = summary(out)
    assert nnz == 2 and dims == [2, 2, 2]
    assert 12.25 <= ss <= 13.25  # 3.5^2 + u^2, u in [0, 1]


def test_unterminated_last_line_is_not_read(tmp_path):
    p = tmp_path / "tail.tns"
    p.write_text("1 1 1 1.0\n2 2 2 2.0")  # the reference counts '\n'
    rc, out = run("-i", str(p), "-o", "3", "--parse-only")
    assert rc == 0
    assert summary(out)[1] == 1


def test_option_errors(tmp_path):
    assert run("-o", "3")[0] != 0                       # input mandatory
    assert run("-i", README_TNS)[0] != 0                # order mandatory
    assert run("-i", str(tmp_path), "-o", "3")[0] != 0  # not a regular file
    rc, out = run("-h")
    assert rc == 0 and "--input" in out and "--rank" in out and "--gpus" in out
    assert run("-i", README_TNS, "-o", "3", "--partition", "1,2")[0] != 0


def test_binary_cache_round_trip(tmp_path):
    p = tmp_path / "t.tns"
    p.write_text(open(README_TNS).read())
    rc, out = run("-i", str(p), "-o", "3", "--parse-only", "--cache")
    assert rc == 0 and "binary cache" not in out
    cache = tmp_path / "t.tns.gtcoo"
    assert cache.exists()
    first = summary(out)
    rc, out = run("-i", str(p), "-o", "3", "--parse-only", "--cache")
    assert rc == 0 and "binary cache" in out
    assert summary(out) == first
    # an input newer than its cache is parsed again
    os.utime(cache, (1, 1))
    rc, out = run("-i", str(p), "-o", "3", "--parse-only", "--cache")
    assert rc == 0 and "binary cache" not in out and summary(out) == first


def test_cache_of_another_order_is_ignored(tmp_path):
    p = tmp_path / "t.tns"
    p.write_text("1 1 1 1 2.0\n2 1 1 1 3.0\n")
    rc, out = run("-i", str(p), "-o", "4", "--parse-only", "--cache")
    assert rc == 0
    rc, out = run("-i", str(p), "-o", "3", "--parse-only", "--cache")
    assert rc == 0 and "binary cache" not in out
    assert summary(out)[:3] == (3, 2, [2, 1, 1])


def test_new_flags_are_accepted():
    rc, out = run("-i", README_TNS, "-o", "3", "--parse-only", "--fused-fit", "--resume", "x")
    assert rc == 0, out
    rc, out = run("-i", README_TNS, "-o", "3", "--parse-only", "--graph", "--autotune",
                  "--ref-fit")
    assert rc == 0, out


def test_dump_init_follows_glibc_rand_order(tmp_path):
    """The initial factors and core come from unseeded glibc rand() in the
    reference's order: a value missing from the input is drawn first, while
    parsing (tensor_manager.tpp:150-155); then every factor, axis by axis, row
    by row over the whole mode (the parts are consecutive row ranges), column
    by column (tucker.cuh:84-97); then the core in linear order
    (tucker.cuh:118-119).  Each draw is frand<double>(0, 1) =
    rand() / RAND_MAX in double (helper.hpp:99), stored as float.  Restated
    here with libc's own rand() through ctypes."""
    import ctypes
    import numpy as np
    p = tmp_path / "t.tns"
    p.write_text("1\t1\t1\t1.5\n2\t3\t1\n3\t1\t2\t0.5\n")
    prefix = str(tmp_path / "init")
    rc, out = run("-i", str(p), "-o", "3", "-r", "4", "--dump-init", prefix)
    assert rc == 0, out
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)  # the state of a process that never seeded (C standard: seed 1)
    rand_max = 2147483647  # glibc RAND_MAX

    def frand():
        return np.float32(float(libc.rand()) / rand_max)

    missing = frand()
    rc, out = run("-i", str(p), "-o", "3", "--parse-only")
    assert rc == 0
    ss = summary(out)[3]
    assert abs(ss - (1.5 ** 2 + 0.5 ** 2 + float(missing) ** 2)) < 1e-6
    dims, R = [3, 3, 2], 4
    for a, d in enumerate(dims):
        want = np.array([[frand() for _ in range(R)] for _ in range(d)], np.float32)
        got = np.loadtxt(f"{prefix}_factor{a}.txt", ndmin=2).astype(np.float32)
        assert np.array_equal(got, want), a
    want = np.array([frand() for _ in range(R ** 3)], np.float32).reshape(-1, R)
    got = np.loadtxt(f"{prefix}_core.txt", ndmin=2).astype(np.float32)
    assert np.array_equal(got, want)


def test_value_spellings_read_like_stod(tmp_path):
    """Values are read as std::stod reads them (tensor_manager.tpp:150-155):
    signs, exponents, hex floats and trailing text after the number."""
    p = tmp_path / "spell.tns"
    p.write_text("1 1 1 +2.5\n2 1 1 0x1p1\n1 2 1 1e-3\n2 2 1 -3\n1 1 2 7.25abc\n")
    rc, out = run("-i", str(p), "-o", "3", "--parse-only")
    assert rc == 0, out
    order, nnz, dims, ss = summary(out)
    assert (nnz, dims) == (5, [2, 2, 2])
    want = 2.5 ** 2 + 2.0 ** 2 + 0.001 ** 2 + 9.0 + 7.25 ** 2
    assert abs(ss - want) < 1e-6 * want
