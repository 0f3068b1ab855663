"""Benchmark of the MI355X bundle-adjustment engine.

Workload (one GPU): the 100-keyframe / 20k-landmark / ~160k-observation problem
(forward random walk, 1 px noise, 5% outliers, Huber 2 px), 20 Levenberg-Marquardt
iterations.  One step = one full solve from the perturbed initial state, which is
restored on the device before every step (inputs resident in HBM).

Metric: BA observation-iterations per second = observations x outer LM iterations
executed / solve wall time.  Prints one JSON line (see DESIGN.md for the roofline
byte model and the CPU baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def algorithmic_bytes(n_obs, n_poses, n_points, n_pairs, s):
    """Per-launch algorithmic HBM bytes of each kernel class (s = bytes per scalar)."""
    n = 6 * n_poses
    return {
        # K2+K3 linearisation: obs uv + ids, W / J write, point in + H_ll/b_l out, pose in + H_pp/b_p out
        "linearisation": n_obs * (16 + 8) + n_pairs * 18 * s + n_points * (24 + 9 * s) + n_poses * (56 + 27 * s),
        "cost": n_obs * 24 + n_points * 24 + n_poses * 56,
        "schur": n_pairs * 18 * s + n_points * (9 * s + 9 * s) + 8 * n * n // 2 + 8 * n,
        "backsub": n_pairs * 18 * s + n_points * (9 * s + 24) + 8 * n,
    }


def cpu_baseline(problem, iters: int) -> dict:
    """The reference's own optimiser (compiled from its sources into oracle/_ref)
    on a bounded sample of the same workload, on this host, one thread."""
    from oracle.oracle import CpuGraph, have_reference
    kind = "reference" if have_reference() else "port"
    g = CpuGraph(problem, kind if kind == "reference" else "oracle")
    t0 = time.perf_counter()
    g.solve(iters)
    dt = time.perf_counter() - t0
    executed = iters
    if kind == "port":
        executed = g.stats()["iterations"]
    value = problem.n_obs * executed / dt
    return {"value": value, "unit": "obs*iter/s", "cores": 1, "kind": kind,
            "sample": f"solve({iters}) of the same 100KF/20k-pt workload ({problem.n_obs} obs), "
                      f"{dt:.2f} s single-threaded on {cpu_model()} (nproc {os.cpu_count()})"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp64_fast", "mixed"])
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        return 2

    from zeroslam_amd.engine import FP64, FP64_FAST, MIXED, Graph
    from zeroslam_amd.problem import ITERATIONS, synth

    problem = synth(args.config)
    iters = ITERATIONS[args.config]
    precision = {"fp64": FP64, "fp64_fast": FP64_FAST, "mixed": MIXED}[args.precision]
    g = Graph.from_problem(problem, timing=True)
    g.solve(0)          # structure analysis + upload (untimed)
    g.snapshot()
    for _ in range(args.warmup):
        g.restore()
        g.solve(iters, False, precision)
    kernel_ms = {}
    kernel_launches = {}
    obs_iters = 0
    chi2 = None
    stats = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.restore()
        g.solve(iters, False, precision)
        stats = g.stats()
        obs_iters += stats["n_obs"] * stats["iterations"]
        chi2 = stats["final_chi2"]
        for k, v in stats["kernel_ms"].items():
            kernel_ms[k] = kernel_ms.get(k, 0.0) + v
            kernel_launches[k] = kernel_launches.get(k, 0) + stats["kernel_launches"][k]
    elapsed = time.perf_counter() - t0
    value = obs_iters / elapsed

    # roofline of the dominant kernel class (live HIP-event durations)
    s = 8
    n_pairs = int(problem.n_obs)  # one W block per (pose, point) pair; the workload has no repeats
    B = algorithmic_bytes(problem.n_obs, problem.n_poses, problem.n_points, n_pairs, s)
    avg = {k: kernel_ms[k] / kernel_launches[k] for k in kernel_ms if kernel_launches[k]}
    groups = {
        "linearisation": ["linearise", "pose_reduce"],
        "cost": ["cost"],
        "schur": ["invert", "form_Y", "fill", "schur"],
        "backsub": ["backsub"],
        "cholesky": ["cholesky", "trsv"],
    }
    total = {gname: sum(kernel_ms.get(k, 0.0) for k in ks) for gname, ks in groups.items()}
    dominant = max(total, key=total.get)
    if dominant == "cholesky":
        n = 6 * problem.n_poses
        dur = avg.get("cholesky", float("nan")) + avg.get("trsv", 0.0)
        flops = n ** 3 / 3 + 2 * n * n
        achieved = flops / (dur * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP64_MFMA_PEAK_TFLOPS, "traffic": None, "kernel": "cholesky+trsv"}
    else:
        dur = sum(avg.get(k, 0.0) for k in groups[dominant])
        achieved = B[dominant] / (dur * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dominant,
                    "algorithmic_bytes": B[dominant], "avg_ms": dur}
    out = {
        "metric": "BA observation*iters/s, 100KF/20k pts: mixed vs FP64 GPU at equal final cost",
        "value": value, "unit": "obs*iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if precision == MIXED else "f64", "data": "synthetic",
        "config": {"workload": f"c{args.config}: {problem.n_poses} KF, {problem.n_points} pts, {problem.n_obs} obs, "
                               f"{iters} LM iterations", "precision": args.precision,
                   "lm_iterations": stats["iterations"], "attempts": stats["attempts"], "final_chi2": chi2},
        "roofline": roofline,
        "kernel_avg_ms": {k: round(v, 5) for k, v in avg.items()},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(problem, args.cpu_iters)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
