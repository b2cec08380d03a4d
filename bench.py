"""Benchmark: nnz/s per ALS mode update on one MI355X.

Workload (default): 3-order 1,000,000^3 sparse tensor, 100,000,000 distinct
nonzeros with per-mode Zipf(s=1.1) coordinates, values U[0,1) FP32, Tucker
rank 10, P_n = 4 per mode (64 tensor blocks), all resident in HBM.

A step is one ALS mode update (delta + Gram + row solves over every
nonzero, cycling modes 0..N-1).  W warmup steps, then K timed steps between
device synchronisations; t_mode is the median over the timed ALS iterations
of (iteration time without the fit) / N, and value = nnz / t_mode.  The fit
is computed after the timed region.

    python bench.py [--steps K] [--warmup W] [--config c2|c1|c3|c4|c5s] [--precision fp32]
                    [--no-graph] [--autotune] [--budget BYTES] [--no-cpu]

By default each ALS iteration replays a captured HIP graph (the N mode
updates and the fit reduction); --autotune lets the engine's autotuner pick
each mode's chunk size at initialisation.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import resource
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # the headline workload: its mode-update throughput is the reported value
    "c2": dict(name="3-order 1e6^3, 1e8 nnz Zipf(1.1), R=10, P=4x4x4", dims=[1_000_000] * 3,
               nnz=100_000_000, modes=[("zipf", 1.1)] * 3, values="uniform", rank=10,
               part=[4, 4, 4], iters=10, lam=None),
    # a small uniform case of the size the reference itself runs: 2000^3, 1e6 nonzeros
    "c1": dict(name="3-order 2000^3, 1e6 nnz uniform, R=10, P=1", dims=[2000] * 3,
               nnz=1_000_000, modes=None, values="uniform", rank=10, part=[1, 1, 1], iters=5,
               lam=1e-3),
    # order 4, three Zipf(1.2) modes and a short uniform one, lognormal values;
    # the planner picks P
    "c3": dict(name="4-order 2e5^3 x 1e3, 2e8 nnz Zipf(1.2) x3 + uniform, lognormal, R=8",
               dims=[200_000] * 3 + [1_000], nnz=200_000_000,
               modes=[("zipf", 1.2)] * 3 + ["uniform"], values="lognormal", rank=8, part=None,
               iters=3, lam=None),
    # order 5, R = 5 at a size the 64-bit coordinate keys of the generator
    # reach (the 5e4^4 x 200 index space needs 71 bits), in core
    "c5s": dict(name="5-order 1e4^4 x 200, 2e8 nnz uniform, R=5", dims=[10_000] * 4 + [200],
                nnz=200_000_000, modes=None, values="uniform", rank=5, part=None, iters=3,
                lam=None),
    # the out-of-core workload's shape: 5e4^4 x 200 (71-bit coordinates),
    # R = 5, under a forced 16 GB device budget, with 2e8 of its 2e9 nonzeros;
    # the partition and every mode update stream from pinned host memory
    "c5": dict(name="5-order 5e4^4 x 200, 2e8 nnz uniform, R=5, 16 GB device budget",
               dims=[50_000] * 4 + [200], nnz=200_000_000, modes=None, values="uniform", rank=5,
               part=None, iters=2, lam=None, budget=16_000_000_000),
    # R = 32 (the wide refined BF16 kernel, the block-per-lane FP64 Gram) with power-law rows
    # in mode 0; Zipf(1.5) in every mode would need about 2e10 draws for 3e8
    # distinct coordinates
    "c4": dict(name="3-order 5e5 x 5e5 x 5e4, 3e8 nnz Zipf(1.5) in mode 0, R=32",
               dims=[500_000, 500_000, 50_000], nnz=300_000_000,
               modes=[("zipf", 1.5), "uniform", "uniform"], values="uniform", rank=32, part=None,
               iters=2, lam=None),
}

HBM_PEAK_GBS = 8000.0

# the delta + Gram kernel behind each gt_delta_path, as rocprofv3 names it
_PATH_KERNEL = {"fp32": "k_gram_valu<{R}>", "fp32-wide": "k_gram_fp32<{R}>",
                "bf16": "k_gram_bf16<{R}, 1>", "bf16r": "k_gram_bf16<{R}, 2>",
                "bf16r-l2": "k_gram_bf16_o5<{R}>", "bf16r-wide": "k_gram_bf16w<{R}>",
                "bf16r-ob": "k_gram_bf16_ob<{R}, {NS}, true>",
                "bf16r-ob-l2": "k_gram_bf16_ob<{R}, {NS}, false>", "bf16r-kr": "k_gram_bf16_kr<{R}>", "bf16r-kr5": "k_gram_bf16_kr5<{R}>"}
# the arithmetic each path's delta runs in
_PATH_DTYPE = {"fp32": "fp32", "fp32-wide": "fp32", "bf16": "bf16", "bf16r": "bf16r",
               "bf16r-l2": "bf16r", "bf16r-wide": "bf16r", "bf16r-ob": "bf16r",
               "bf16r-ob-l2": "bf16r", "bf16r-kr": "bf16r"}


def committed_counters(workload, kernel):
    """The committed rocprofv3 counter summary of `kernel` on `workload`
    (profiles/counters_<workload>.json, written from FETCH_SIZE / WRITE_SIZE
    runs scaled by a copy kernel's factors and from SQ_ counter runs), or None
    when it does not describe this kernel."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"counters_{workload}.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return k if k else None


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary, or None."""
    k = committed_counters(workload, kernel)
    return k.get("traffic_bytes_per_launch") if k else None


def kernel_name(path, R, N=3):
    """The delta + Gram kernel of a gt_delta_path at kernel rank R."""
    if path == "bf16r-kr" and N == 5:
        path = "bf16r-kr5"
    return _PATH_KERNEL[path].format(R=R, NS=max(N - 2, 0))


def bytes_per_nnz(N, R, s_f):
    """Algorithmic HBM bytes per nonzero per mode update: N indices, the value,
    the row-grouping entry and the N-1 gathered factor rows of s_f bytes per
    element (4 FP32, 2 BF16)."""
    return 4 * N + 8 + (N - 1) * R * s_f


def kernel_bytes_per_nnz(N, R, path):
    """Bytes per nonzero the delta + Gram kernel itself requests: N stream
    indices and the value (no row-grouping entry: the stream is row-sorted),
    the split BF16 copy of the last other factor (16 words per row) on the
    MFMA paths or its FP32 row, and the other factors' FP32 rows."""
    a_l = (128 if path == "bf16r-wide" else
           (4 * R if path == "bf16r-kr" else (64 if path.startswith("bf16") else 4 * R)))
    return 4 * N + 4 + a_l + (N - 2) * 4 * R


def flops_per_nnz(N, R):
    return 2 * R ** N + 2 * R * R + 2 * R


def limiter(ctr):
    """What the committed counters say bounds the kernel: HBM when its traffic
    runs near the peak; the SIMDs' issue when MFMA busy plus VALU busy (which
    do not overlap on gfx950) fills most cycles; else latency (the waves
    waiting on memory and LDS with neither pipe nor HBM saturated)."""
    if not ctr or ctr.get("mfma_busy") is None:
        return None
    issue = ctr["mfma_busy"] + (ctr.get("valu_busy") or 0.0)
    hbm = ctr.get("hbm_util") or 0.0
    if hbm >= 0.75:
        return f"hbm ({hbm:.0%} of 8 TB/s)"
    if issue >= 0.75:
        return f"issue (MFMA {ctr['mfma_busy']:.0%} + VALU {ctr.get('valu_busy') or 0:.0%})"
    return (f"latency (issue {issue:.0%}: MFMA {ctr['mfma_busy']:.0%} + VALU "
            f"{ctr.get('valu_busy') or 0:.0%}; HBM {hbm:.0%})")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(idx, vals, dims, part, R, lam, A, G, sample=20_000_000):
    """Times the oracle's restatement of the reference host path, built with
    -O3 -march=native and OpenMP (libbaseline.so): for every mode in turn, the
    delta as the reference kernel computes it (delta.cuh:62-90, parallel over
    nonzeros), B/C (update.cuh:35-101) and the LU row solves
    (update.cuh:167-194), in FP32 as the reference's FP32 build.  It runs one
    ALS iteration's N mode updates on the first `sample` nonzeros of the
    workload (input order, full index space and partition) and reports
    sample / (time / N): measured, not extrapolated."""
    import oracle
    N = len(dims)
    S = min(int(sample), idx.shape[1])
    oracle.use_baseline(True)
    try:
        t = oracle.Tensor(idx[:, :S], vals[:S].astype(np.float32), dims, part, R)
        Ap = [oracle.pad_factor(a, t.rows(k)) for k, a in enumerate(A)]
        t0 = time.perf_counter()
        for m in range(N):
            d = oracle.delta(t, Ap, G, m, "f32")
            B, Cv = oracle.bc(t, d, m, lam, "f32")
            Ap[m] = oracle.solve(B, Cv, "f32")
            del d, B, Cv
        t_mode = (time.perf_counter() - t0) / N
        threads = oracle.max_threads()
    finally:
        oracle.use_baseline(False)
    return dict(value=S / t_mode, unit="nnz/s", cores=threads, kind="port",
                cpu=cpu_model(), ms_per_mode=t_mode * 1e3,
                sample=(f"one ALS iteration ({N} mode updates: delta, B/C, LU row solves, FP32) "
                        f"of the oracle built -O3 -march=native on the first {S} of the "
                        f"{idx.shape[1]} nonzeros (full index space, partition {list(part)}), "
                        f"{threads} OpenMP threads; value = {S} / (time / {N})"))


def run_engine(gt, idx, vals, dims, A, G, part, R, lam, precision, steps, warmup, chunk_tiles,
               budget=0, graph=False, autotune=False):
    """Partitions, initialises, runs `warmup` untimed mode updates, then times
    exactly `steps` mode updates (whole ALS iterations of N updates, then
    single updates for the remainder) between device synchronisations.  With
    graph, the iterations are replays of a captured HIP graph, which holds no
    timing events: the delta + Gram kernel's time then comes from a second,
    eager engine that runs the same steps."""
    N = len(dims)
    if graph:
        r = run_engine(gt, idx, vals, dims, A, G, part, R, lam, precision, steps, warmup,
                       chunk_tiles, budget, False, autotune)
    e = gt.Engine(N, R, partition=part, lam=lam, precision=precision, chunk_tiles=chunk_tiles,
                  profile=not graph, device_budget=budget, graph=graph, autotune=autotune,
                  verbose=1 if budget else 0)
    t0 = time.perf_counter()
    e.set_coo(idx, vals, dims)
    e.partition()
    e.set_init(A, G)
    e.synchronize()
    setup_s = time.perf_counter() - t0

    def updates(k):
        its = []
        for _ in range(k // N):
            ti = time.perf_counter()
            e.iteration()  # N mode updates and the fit from the row solves
            its.append(time.perf_counter() - ti)
        for m in range(k % N):
            e.mode_update(m)
        return its

    updates(warmup)
    e.synchronize()
    base = e.timings()
    t_start = time.perf_counter()
    it_times = updates(steps)
    e.synchronize()
    total = time.perf_counter() - t_start
    tm = e.timings()
    fit = e.fit()  # the reconstruction pass (reconstruction.cuh), after the timed region
    launches = tm["gram_kernel_launches"] - base["gram_kernel_launches"]
    k_ms = (tm["gram_kernel_ms_total"] - base["gram_kernel_ms_total"]) / max(launches, 1)
    t_mode = statistics.median(it_times) / N if it_times else total / steps
    info = e.info()
    e.close()
    if graph:  # the kernel times and delta path of the eager engine
        launches, k_ms = r["launches"], r["kernel_ms"]
        tm = dict(tm, ms_gram=[r["gram_phase_ms"]], ms_solve=[r["solve_ms"]],
                  delta_path=r["delta_path"])
    return dict(t_mode=t_mode, value=idx.shape[1] / t_mode, fit=fit,
                out_of_core=info["out_of_core"], segment_nnz=info["segment_nnz"],
                kernel_ms=k_ms, gram_phase_ms=statistics.mean(tm["ms_gram"]),
                solve_ms=statistics.mean(tm["ms_solve"]), delta_path=tm["delta_path"],
                launches=launches, total=total, setup_s=setup_s,
                partition_ms=tm["ms_partition"], part=[int(p) for p in info["part"]],
                tuned_chunk_tiles=tm["tuned_chunk_tiles"], graph=graph,
                partition_path=info["partition_path"],
                stream_bytes_per_nnz=info["stream_bytes_per_nnz"],
                host_pageable_bytes=info["host_pageable_bytes"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed mode updates per precision")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="bf16r", help="headline path")
    ap.add_argument("--compare", default="fp32", help="baseline GPU path ('' to skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk-tiles", type=int, default=0)
    ap.add_argument("--budget", type=int, default=0,
                    help="device budget in bytes; below the resident size the streams are "
                         "staged from pinned host memory every mode update (PCIe-inclusive)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--nnz", type=float, default=None,
                    help="nonzeros of the workload (default: the config's)")
    ap.add_argument("--rank", type=int, default=None,
                    help="Tucker rank per mode (default: the config's)")
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="launch each iteration's kernels eagerly instead of replaying the "
                         "captured HIP graph")
    ap.add_argument("--autotune", action="store_true",
                    help="pick each mode's chunk size (and core-operand source) by timing the "
                         "candidates at initialisation")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is "
                                   "designed but not built (run with --gpus 1)"}))
        return 2
    cfg = dict(CONFIGS[a.config])
    if a.nnz is not None:
        cfg["nnz"] = int(a.nnz)
        cfg["name"] = re.sub(r"[0-9.e]+ nnz", f"{cfg['nnz']:.3g} nnz", cfg["name"])
    if a.rank is not None:
        cfg["rank"] = int(a.rank)
        cfg["name"] = re.sub(r"R=\d+", f"R={cfg['rank']}", cfg["name"])
    N, R = len(cfg["dims"]), cfg["rank"]
    if R > 15 and a.precision == "bf16":
        # plain BF16 is built for R <= 15; the wide ranks have the refined delta
        print(f"[bench] R = {R}: plain BF16 is built for R <= 15, timing bf16r",
              file=sys.stderr, flush=True)
        a.precision = "bf16r"
    if a.compare == a.precision:
        a.compare = ""

    steps = a.steps if a.steps is not None else cfg["iters"] * N
    warmup = a.warmup if a.warmup is not None else N
    if steps < 1 or warmup < 0:
        print(json.dumps({"error": "--steps must be >= 1 and --warmup >= 0"}))
        return 2

    import gputucker_amd as gt
    from gputucker_amd import synth

    t0 = time.perf_counter()
    idx, vals, dims = synth.generate(cfg["dims"], cfg["nnz"], cfg["modes"], cfg["values"],
                                     seed=a.seed, device="cuda",
                                     log=lambda m: print(f"[bench] generating: {m}",
                                                         file=sys.stderr, flush=True))
    max_deg = [int(np.bincount(idx[k]).max()) for k in range(N)]
    print(f"[bench] generated {cfg['name']}: {idx.shape[1]} nnz in "
          f"{time.perf_counter() - t0:.1f}s, largest row degree per mode {max_deg}",
          file=sys.stderr, flush=True)
    A, G = synth.init_factors(dims, R, seed=a.seed)
    lam = cfg["lam"] if cfg["lam"] is not None else gt.default_lambda()

    runs = {}
    for prec in [p for p in (a.compare, a.precision) if p]:
        r = run_engine(gt, idx, vals, dims, A, G, cfg["part"], R, lam, prec, steps, warmup,
                       a.chunk_tiles, a.budget or cfg.get("budget", 0), a.graph, a.autotune)
        runs[prec] = r
        print(f"[bench] {prec}: t_mode {r['t_mode'] * 1e3:.3f} ms ({r['value']:.3e} nnz/s), "
              f"delta path {r['delta_path']}, gram kernel {r['kernel_ms']:.3f} ms avg over "
              f"{r['launches']} (gram phase {r['gram_phase_ms']:.3f} ms, solve "
              f"{r['solve_ms']:.3f} ms), out of core {r['out_of_core']} (segment "
              f"{r['segment_nnz']}), fit {r['fit']:.7f}, setup {r['setup_s']:.1f}s (partition "
              f"{r['partition_ms']:.0f} ms, P {r['part']}), pageable host bytes "
              f"{r['host_pageable_bytes']}, peak host RSS "
              f"{resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 2**20:.1f} GiB",
              file=sys.stderr, flush=True)
    head = runs[a.precision]
    paths = sorted(set(head["delta_path"]))
    dtype = "+".join(sorted({_PATH_DTYPE[p] for p in paths}))
    RK = gt.kernel_rank(R)
    kernel = kernel_name(paths[0], RK, N) if len(paths) == 1 else None
    s_f = 2 if dtype.startswith("bf16") and len(paths) == 1 else 4
    bpn = bytes_per_nnz(N, R, s_f)
    achieved = bpn * idx.shape[1] / (head["kernel_ms"] * 1e-3) / 1e9
    ctr = committed_counters(a.config, kernel) if kernel else None

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(idx, vals, dims, head["part"], R, lam, A, G)
    out = {
        "metric": "nnz/s per ALS mode update",
        "value": head["value"],
        "unit": "nnz/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": head["t_mode"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": cfg["name"], "nnz": int(idx.shape[1]), "order": N, "rank": R,
                   "partition": head["part"], "precision": a.precision,
                   "delta_path": head["delta_path"], "graph": a.graph,
                   "chunk_tiles": head["tuned_chunk_tiles"] if a.autotune else a.chunk_tiles,
                   "parallelism": "single GPU",
                   "out_of_core": bool(head["out_of_core"]),
                   "device_budget": a.budget or cfg.get("budget", 0),
                   "partition_from_host": bool(head["partition_path"])},
        "fit": head["fit"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": ctr.get("traffic_bytes_per_launch") if ctr else None,
                     "kernel": kernel, "kernel_ms": head["kernel_ms"],
                     "bytes_per_nnz": bpn,
                     "kernel_bytes_per_nnz": kernel_bytes_per_nnz(N, RK, paths[0]),
                     "flops_per_nnz": flops_per_nnz(N, R),
                     "mfma_util": ctr.get("mfma_busy") if ctr else None,
                     "valu_util": ctr.get("valu_busy") if ctr else None,
                     "hbm_util": ctr.get("hbm_util") if ctr else None,
                     "limiter": limiter(ctr),
                     "counters": f"profiles/counters_{a.config}.md" if ctr else None},
        "cpu_baseline": cpu,
    }
    if head["out_of_core"]:  # the host-link rate of the staged streams (never `value`)
        out["pcie"] = {"host_bytes_per_nnz": head["stream_bytes_per_nnz"],
                       "h2d_gbs": head["stream_bytes_per_nnz"] * idx.shape[1] / head["t_mode"] / 1e9,
                       "partition_s": head["partition_ms"] / 1e3,
                       "host_pageable_bytes": head["host_pageable_bytes"]}
    if a.compare and a.compare in runs:
        base = runs[a.compare]
        out["compare"] = {"precision": a.compare, "value": base["value"],
                          "ms_per_step": base["t_mode"] * 1e3, "fit": base["fit"],
                          "delta_path": base["delta_path"], "kernel_ms": base["kernel_ms"],
                          "speedup": head["value"] / base["value"],
                          "fit_rel_diff": abs(head["fit"] - base["fit"]) / abs(base["fit"])}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
