"""Benchmark of the MI355X bundle-adjustment engine.

Workload (one GPU): the 100-keyframe / 20k-landmark / ~160k-observation problem
(forward random walk, 1 px noise, 5% outliers, Huber 2 px), 20 Levenberg-Marquardt
iterations.  One step = one full solve from the perturbed initial state, which is
restored on the device before every step (inputs resident in HBM).

Metric: BA observation-iterations per second.  For the binary64 paths:
observations x LM iterations executed / solve wall time.  For the mixed path, at
equal final cost: observations x the LM iterations the binary64 run needs to reach
a cost level / the wall time the mixed run needs to reach the same level, so that
the ratio of the two values is the ratio of their times to equal cost.  The level
is the binary64 cost after 20 iterations, or, when the mixed run settles in a
local minimum above it (the robust problem has many, and rounding decides which
one a run ends in), the mixed run's converged cost.  Prints one JSON line, with
the roofline of the dominant kernel group and the reference CPU optimiser's
throughput on a sample of the same workload.
"""
from __future__ import annotations

import argparse
import dataclasses
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
        # back-substitution and update: dx_l = H_ll^-1 (b_l - W^T dx_p) per point (the
        # mixed path receives dx_l from its refinement round and only reads it), then
        # point and pose read, backed up and written
        "backsub": ((n_points * 24 if s == 4 else n_pairs * 18 * s + n_points * (9 * s + 24)) + 8 * n
                    + n_points * 24 * 3 + n_poses * 56 * 3),
        # mixed refinement round: the matrix-free binary64 residual (observation
        # record, pose and point in, W blocks, point rows out) and the reduced residual
        "refinement": n_obs * 32 + n_pairs * 18 * s + n_points * (24 + 9 * s + 3 * 8 * 3) + n_poses * (56 + 6 * 8 * 3)
                      + 8 * n * 3,
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


def envelope_first_column(problem) -> np.ndarray:
    """First structurally non-zero column of each row of the reduced system's lower
    triangle: pose i couples to pose j < i when both are free and observe a common
    free point; its 6 rows start at 6 * (smallest such j, or i)."""
    n_p = problem.n_poses
    first = np.arange(n_p)
    pose, point = problem.obs_pose, problem.obs_point
    keep = (~problem.pose_fixed.astype(bool)[pose]) & (~problem.point_fixed.astype(bool)[point])
    pose, point = pose[keep], point[keep]
    m = np.full(problem.n_points, n_p, dtype=np.int64)
    np.minimum.at(m, point, pose)
    np.minimum.at(first, pose, m[point])
    return np.repeat(6 * first, 6)


def cholesky_flops(problem) -> float:
    """Multiply-subtract pairs (2 flops each) plus the divisions and square roots
    that a Cholesky factorisation restricted to the envelope must perform."""
    f = envelope_first_column(problem)
    total = 0.0
    for i in range(len(f)):
        j = np.arange(f[i], i + 1)
        total += 2.0 * np.sum(j - np.maximum(f[i], f[j])) + len(j)
    return total


def time_solves(g, precision, iters, steps, warmup):
    """Restore + solve(iters) `steps` times after `warmup` untimed runs; returns
    (obs*iters/s, ms per step, last stats, per-class kernel ms, launches)."""
    for _ in range(warmup):
        g.restore()
        g.solve(iters, False, precision)
    kernel_ms, kernel_launches = {}, {}
    obs_iters = 0
    stats = None
    t0 = time.perf_counter()
    for _ in range(steps):
        g.restore()
        g.solve(iters, False, precision)
        stats = g.stats()
        obs_iters += stats["n_obs"] * stats["iterations"]
        for k, v in stats["kernel_ms"].items():
            kernel_ms[k] = kernel_ms.get(k, 0.0) + v
            kernel_launches[k] = kernel_launches.get(k, 0) + stats["kernel_launches"][k]
    elapsed = time.perf_counter() - t0
    return obs_iters / elapsed, elapsed / steps * 1e3, stats, kernel_ms, kernel_launches


def iterations_to_reach(g, precision, target, lo, hi):
    """Smallest iteration budget k in [lo, hi] whose solve ends at chi2 <= target.
    A solve with budget k runs the first k iterations of one with budget k + 1, so
    the final chi2 does not increase with k and a bisection applies."""
    def chi(k):
        g.restore()
        g.solve(k, False, precision)
        return g.stats()["final_chi2"]
    if chi(hi) > target:
        return None
    while lo < hi:
        mid = (lo + hi) // 2
        if chi(mid) <= target:
            hi = mid
        else:
            lo = mid + 1
    return lo


def state_error(P, X, Pr, Xr):
    return float(max(np.abs(P - Pr).max() / max(np.abs(Pr).max(), 1e-300),
                     np.abs(X - Xr).max() / max(np.abs(Xr).max(), 1e-300)))


def equal_cost(g, gfz, problem, iters, steps, warmup, chi64, ms64, P64, X64):
    """Mixed against the FP64 (reference order) and FP64-fast paths at equal cost.

    The level is the FP64 path's cost after `iters` iterations.  Each path is
    charged the wall time of its smallest iteration budget that ends at or below
    (1 + 1e-6) x that cost.  The mixed path runs with the behind-camera freeze
    (its default, zsba_options.freeze_behind); FP64-fast is measured with the same
    freeze (`gfz`, a handle with freeze_behind = 1), so that the two differ only in
    precision.  A mixed run without the freeze is reported beside it.  Should the
    mixed run never reach the level within 4 x `iters` iterations, the comparison
    falls back to the mixed run's converged cost and says so."""
    from zeroslam_amd.engine import FP64, FP64_FAST, MIXED
    level = chi64 * (1.0 + 1e-6)
    level_name = f"FP64 cost after {iters} iterations"
    k64, ms64_level = iters, ms64
    kmix = iterations_to_reach(g, MIXED, level, 1, 4 * iters)
    if kmix is None:
        g.restore()
        g.solve(4 * iters, False, MIXED)
        cmix = g.stats()["final_chi2"]
        level = cmix * (1.0 + 1e-6)
        level_name = "converged mixed cost (above the FP64 cost)"
        kmix = iterations_to_reach(g, MIXED, level, 1, 4 * iters)
        k64 = iterations_to_reach(g, FP64, level, 1, iters)
        ms64_level = time_solves(g, FP64, k64, steps, warmup)[1]
        print(f"bench: mixed settles at chi2 {cmix!r} (FP64 {chi64!r}); comparing at that level",
              file=sys.stderr, flush=True)
    kfast = iterations_to_reach(gfz, FP64_FAST, level, 1, 4 * iters)
    msf_level, st_fast = time_solves(gfz, FP64_FAST, kfast, steps, warmup)[1:3] if kfast else (None, None)
    vraw, ms, stats, kms, kl = time_solves(g, MIXED, kmix, steps, warmup)
    chi_mix = stats["final_chi2"]
    state_err = state_error(g.poses(), g.points(), P64, X64)
    # the same budget as the FP64 run, and the mixed run without the freeze
    g.restore()
    g.solve(iters, False, MIXED)
    chi_mix_full = g.stats()["final_chi2"]
    cmp = {"equal_cost_level": level_name, "level_chi2": level,
           "iterations_mixed": kmix, "iterations_fp64": k64, "iterations_fp64_fast": kfast,
           "iterations_by_precision": {"mixed": kmix},
           "ms_mixed": ms, "ms_fp64": ms64_level, "ms_fp64_fast": msf_level,
           "mixed_obs_iters_per_s": vraw, "final_chi2": chi_mix,
           "chi2_rel_vs_fp64": (chi_mix - chi64) / chi64,
           "chi2_rel_vs_fp64_at_equal_iterations": (chi_mix_full - chi64) / chi64,
           "state_rel_err_vs_fp64": state_err,
           "frozen_points": stats["frozen_points"],
           "speedup_vs_fp64": ms64_level / ms,
           "speedup_vs_fp64_fast": (msf_level / ms) if msf_level else None,
           # per LM attempt (independent of how many iterations each run draws)
           "attempts_mixed": stats["attempts"], "attempts_fp64_fast": st_fast["attempts"] if st_fast else None,
           "ms_per_attempt_mixed": ms / max(stats["attempts"], 1),
           "ms_per_attempt_fp64_fast": (msf_level / max(st_fast["attempts"], 1)) if st_fast else None,
           "solver_rule": "behind-camera freeze (mixed and FP64-fast); FP64 is the reference's LM unchanged"}
    gnf = getattr(g, "_no_freeze", None)
    if gnf is not None:
        gnf.restore()
        gnf.solve(iters, False, MIXED)
        c_nf = gnf.stats()["final_chi2"]
        k_nf = iterations_to_reach(gnf, MIXED, level, 1, 4 * iters)
        cmp["mixed_without_freeze"] = {"chi2_rel_vs_fp64_at_equal_iterations": (c_nf - chi64) / chi64,
                                       "iterations_to_level": k_nf}
    print(f"bench: mixed reaches the level in {kmix} iterations, {ms:.2f} ms (FP64 {k64} iterations, "
          f"{ms64_level:.2f} ms; FP64-fast {kfast} iterations, {msf_level if msf_level else float('nan'):.2f} ms)",
          file=sys.stderr, flush=True)
    return cmp, (vraw, ms, stats, kms, kl), k64


def make_graphs(problem, **kw):
    """The headline handle (mixed with its default freeze), the FP64-fast handle
    with the freeze, and a mixed handle without it; all at the initial state."""
    from zeroslam_amd.engine import Graph
    t0 = time.perf_counter()
    g = Graph.from_problem(problem, **kw)
    g.solve(0)
    g.setup = {"graph_build_and_first_solve_ms": (time.perf_counter() - t0) * 1e3,
               "structure_and_upload_ms": g.stats()["setup_ms"]}
    gfz = Graph.from_problem(problem, freeze_behind=1)
    gnf = Graph.from_problem(problem, freeze_behind=-1)
    for h in (gfz, gnf):
        h.solve(0)
    for h in (g, gfz, gnf):
        h.snapshot()
    g._no_freeze = gnf
    return g, gfz


def seed_compare(seed, config, iters):
    """The equal-cost comparison on another instance of the same workload
    (generator seed `seed`), two timed solves each."""
    from zeroslam_amd.engine import FP64
    from zeroslam_amd.problem import synth
    pr = synth(config, seed=seed)
    g, gfz = make_graphs(pr)
    _, ms64, st64, _, _ = time_solves(g, FP64, iters, 2, 1)
    cmp, _, _ = equal_cost(g, gfz, pr, iters, 2, 1, st64["final_chi2"], ms64, g.poses(), g.points())
    keep = ("equal_cost_level", "iterations_mixed", "iterations_fp64", "iterations_fp64_fast", "ms_mixed", "ms_fp64",
            "ms_fp64_fast", "chi2_rel_vs_fp64", "chi2_rel_vs_fp64_at_equal_iterations", "frozen_points",
            "speedup_vs_fp64", "speedup_vs_fp64_fast", "ms_per_attempt_mixed", "ms_per_attempt_fp64_fast",
            "mixed_without_freeze")
    out = {"seed": seed, "n_obs": pr.n_obs}
    out.update({k: cmp[k] for k in keep if k in cmp})
    return out


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "c1_mixed_hbm_traffic.md")
GROUP_KERNELS = {
    "cholesky": ["k_band_factor<80>"],
    "linearisation": ["k_lin_blocks_f", "k_lin_finish_f"],
    "schur": ["k_damp_invert_f", "k_form_Y_f", "k_fill_zero_d", "k_schur_f<2>"],
    "cost": ["k_cost"],
    "backsub": ["k_backsub_update<float>"],
}


def measured_traffic(group: str):
    """HBM bytes per launch of a kernel group, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the mixed 100-keyframe solve (calibrated
    against a copy kernel); None when the summary does not cover the group."""
    try:
        lines = open(TRAFFIC_PROFILE).read().splitlines()
    except OSError:
        return None
    table = {}
    for line in lines:
        parts = line.split()
        if len(parts) == 4 and parts[0] == "traffic":
            table[parts[1]] = float(parts[2]) + float(parts[3])
    names = GROUP_KERNELS.get(group, [])
    if not names or any(k not in table for k in names):
        return None
    return sum(table[k] for k in names)


def roofline_of(problem, precision_name, kernel_ms, kernel_launches, stats, steps, config=1):
    """Roofline of the kernel group with the largest device time over `steps`
    solves timed class by class with HIP events (each solve's counts in `stats`).
    achieved = the group's algorithmic bytes (or flops) per unit of work x units
    / the group's device time; a unit is one linearisation, cost evaluation,
    attempt (Schur, back-substitution, factorisation) or refinement round."""
    s = 4 if precision_name == "mixed" else 8
    n_pairs = int(problem.n_obs)  # one W block per (pose, point) pair; the workload has no repeats
    B = algorithmic_bytes(problem.n_obs, problem.n_poses, problem.n_points, n_pairs, s)
    avg = {k: kernel_ms[k] / kernel_launches[k] for k in kernel_ms if kernel_launches[k]}
    attempts = stats["attempts"] * steps
    refine_rounds = attempts if precision_name == "mixed" else 0  # one refinement round per attempt
    groups = {
        "linearisation": (["linearise", "pose_reduce"], stats["linearisations"] * steps),
        "cost": (["cost"], attempts + steps),
        "schur": (["invert", "form_Y", "fill", "schur"], attempts),
        "backsub": (["backsub"], attempts),
        "refinement": (["refine"], refine_rounds),
        "cholesky": (["cholesky", "trsv"], attempts),
    }
    total = {gname: sum(kernel_ms.get(k, 0.0) for k in ks) for gname, (ks, _) in groups.items()}
    dominant = max(total, key=total.get)
    traffic = measured_traffic(dominant) if (precision_name == "mixed" and config == 1) else None
    units = groups[dominant][1]
    dur = total[dominant] / max(units, 1)  # ms per unit of work
    share = {gname: round(t / max(sum(total.values()), 1e-30), 4) for gname, t in total.items()}
    if dominant == "cholesky":
        n = 6 * problem.n_poses
        fdur = avg.get("cholesky", float("nan"))
        flops = cholesky_flops(problem)
        achieved = flops / (fdur * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP64_MFMA_PEAK_TFLOPS, "traffic": traffic, "kernel": "cholesky",
                    "algorithmic_flops": flops, "avg_ms": fdur, "n": n, "group_share": share}
    else:
        achieved = B[dominant] / (dur * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dominant,
                    "algorithmic_bytes": B[dominant], "avg_ms": dur, "group_share": share}
    return roofline, {k: round(v, 5) for k, v in avg.items()}


def bench_batch(args) -> int:
    """Batched local BA: 256 independent windows (keyframes U{5..30}, points
    U{200..3000}, tracks U{2..window}, 5% outliers, Huber 2 px), mixed precision,
    10 LM iterations each, one hipGraph launch per solve."""
    from zeroslam_amd.engine import MIXED, Batch, Graph
    from zeroslam_amd.problem import synth_window

    n_windows, iters = 256, 10
    windows = [synth_window(i) for i in range(n_windows)]
    graphs = [Graph.from_problem(w) for w in windows]
    for g in graphs:
        g.solve(0)
        g.snapshot()
    batch = Batch(graphs)
    total_obs = sum(w.n_obs for w in windows)
    print(f"bench: batched {n_windows} windows, {total_obs} obs", file=sys.stderr, flush=True)

    def one_solve():
        for g in graphs:
            g.restore()
        t0 = time.perf_counter()
        batch.solve(iters, False, MIXED)
        return time.perf_counter() - t0

    for _ in range(args.warmup):
        one_solve()
    elapsed = 0.0
    obs_iters = 0
    kernel_ms = lin_ms_sum = 0.0
    attempts = accepted = 0
    for _ in range(args.steps):
        elapsed += one_solve()
        info = batch.info()
        kernel_ms += info["kernel_wall_ms"]
        lin_ms_sum += info["lin_ms"]
        for g in graphs:
            st = g.stats()
            obs_iters += st["n_obs"] * st["iterations"]
            attempts += st["attempts"]
            accepted += st["accepted"]
    value = obs_iters / elapsed
    # algorithmic HBM bytes: per window, a cost pass at the start and per applied
    # step, its linearisations, Schur + back-substitution per attempt (binary32
    # blocks).  Every phase runs as grid-wide launches over every window; the
    # linearisation's device time (lin_ms) comes from the solver's phase clock.
    s = 4
    per_step_bytes = lin_bytes = 0.0
    for w, g in zip(windows, graphs):
        st = g.stats()
        B = algorithmic_bytes(w.n_obs, w.n_poses, w.n_points, w.n_obs, s)
        lin_bytes += B["linearisation"] * st["linearisations"]
        per_step_bytes += (B["linearisation"] * st["linearisations"] + B["cost"] * (1 + st["attempts"]) +
                           (B["schur"] + B["backsub"]) * st["attempts"])
    kernel_avg_ms = kernel_ms / args.steps
    lin_ms = lin_ms_sum / args.steps
    achieved = per_step_bytes / (kernel_avg_ms * 1e-3) / 1e9
    lin_achieved = lin_bytes / (lin_ms * 1e-3) / 1e9 if lin_ms > 0 else 0.0
    out = {
        "metric": "BA observation*iters/s, batched local BA (256 windows, mixed, 10 LM iterations)",
        "value": value, "unit": "obs*iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"c2: {n_windows} windows, {total_obs} obs, {iters} LM iterations each",
                   "factor_threads": info["threads"], "schur_inflight": info["schur_inflight"],
                   "autotune_ms": info["tune_ms"], "lockstep_steps": info["steps"],
                   "graph_launches": info["graph_launches"],
                   "phase_ms": {k: round(v, 4) for k, v in info["phase_ms"].items()},
                   "attempts_per_window": attempts / (args.steps * n_windows),
                   "accepted_per_window": accepted / (args.steps * n_windows)},
        # the grid-wide linearisation (k_batch_lin_blocks, then k_batch_lin_finish, per round)
        "roofline": {"bound": "hbm", "achieved": lin_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": lin_achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "linearisation (k_batch_lin_blocks, k_batch_lin_finish)",
                     "algorithmic_bytes": lin_bytes, "avg_ms": lin_ms},
        "roofline_whole_solve": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": per_step_bytes,
                                 "avg_ms": kernel_avg_ms},
    }
    if not args.no_cpu:
        from oracle.oracle import CpuGraph, have_reference
        kind = "reference" if have_reference() else "port"
        sample = windows[:16]
        t0 = time.perf_counter()
        done = 0
        for w in sample:
            o = CpuGraph(w, kind if kind == "reference" else "oracle")
            o.solve(iters)
            done += w.n_obs * iters
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / dt, "unit": "obs*iter/s", "cores": 1, "kind": kind,
                               "sample": f"first 16 of the 256 windows, solve({iters}) each, {dt:.2f} s "
                                         f"single-threaded on {cpu_model()}"}
    print(json.dumps(out))
    return 0


def bench_match(args) -> int:
    """Projection search of a new frame against the map (zsba_match_by_projection,
    the loop of slam.hpp:341-384): 200k landmarks (the loop-closure map's size)
    against 2000 matched keypoints, one step = one search.  The metric counts the
    landmark-keypoint tests the reference's loop makes; the roofline prices the 6
    binary64 operations of each of those tests' distance (2 differences, 2 squares,
    2 sums) against the FP64 vector peak over the device span of a call.  The
    banded search executes only the tests of keypoints in the landmark's and the
    neighbouring 8-px image bands (about 5% of them at this size), so the fraction
    is that of a brute-force search reaching the same time."""
    from zeroslam_amd.engine import Matcher
    from zeroslam_amd.problem import match_scene

    L, N = 200_000, 2000
    scene = match_scene(L, N, seed=11)
    m = Matcher()
    print(f"bench: projection search {L} landmarks x {N} keypoints", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        m.match(scene)
    dev_ms, wall = [], []
    n = 0
    for _ in range(args.steps):
        t0 = time.perf_counter()
        _, _, n = m.match(scene)
        wall.append(time.perf_counter() - t0)
        dev_ms.append(m.stats()["device_ms"])
    st = m.stats()
    pairs = float(L) * N
    dev_s = float(np.mean(dev_ms)) * 1e-3
    achieved = 6.0 * pairs / dev_s / 1e12
    out = {"metric": "projection-search landmark-keypoint tests per second", "value": pairs / dev_s,
           "unit": "tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": dev_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64+u32", "data": "synthetic",
           "config": {"workload": f"projection search: {L} landmarks x {N} keypoints, 640x480 frame",
                      "matches": int(n), "candidates": int(st["candidates"])},
           "end_to_end_tests_per_s": pairs / float(np.mean(wall)),
           "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_MFMA_PEAK_TFLOPS, "traffic": None,
                        "note": "brute-force equivalent: 6 flops x every landmark-keypoint test of the reference; "
                                "the banded search executes about 5% of the tests"}}
    if not args.no_cpu:
        from oracle.oracle import match_by_projection
        sub = dataclasses.replace(scene, landmark_xyz=scene.landmark_xyz[:20000],
                                  landmark_descriptor=scene.landmark_descriptor[:20000],
                                  landmark_observed=scene.landmark_observed[:20000])
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 10.0:
            match_by_projection(sub)
            reps += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": reps * 20000.0 * N / dt, "unit": "tests/s", "cores": 1, "kind": "port",
                               "sample": f"the first 20000 landmarks x {N} keypoints, {reps} searches in {dt:.1f} s, "
                                         f"single-threaded on {cpu_model()}"}
    print(json.dumps(out))
    return 0


def bench_track(args) -> int:
    """The tracking step's BA, map::optimise(1, true, 50) (slam.hpp:316,387,469):
    the last frame of the 100-keyframe workload free, every landmark it observes
    fixed (zeroslam_amd.problem.motion_only), solve(50) with relative
    convergence.  One step = one frame's BA as a tracker calls it: a fresh handle
    is built, loaded and solved, and its state read back; the numeric solve alone
    (restore + solve on a built handle) is reported beside it, for each
    precision.  The CPU baseline runs the same graph through the reference's own
    optimiser (oracle/_ref), one thread."""
    from zeroslam_amd.engine import FP64, FP64_FAST, MIXED, Graph
    from zeroslam_amd.problem import motion_only, synth

    pr = motion_only(synth(1), 1)
    iters = 50
    print(f"bench: tracking BA, one frame, {pr.n_obs} observations of {pr.n_points} fixed landmarks",
          file=sys.stderr, flush=True)

    def frame(prec):
        g = Graph.from_problem(pr)
        g.solve(iters, True, prec)
        pose = g.poses()
        g.close()
        return pose

    def numeric(prec, steps):
        g = Graph.from_problem(pr)
        g.solve(0)
        g.snapshot()
        for _ in range(args.warmup):
            g.restore()
            g.solve(iters, True, prec)
        t0 = time.perf_counter()
        for _ in range(steps):
            g.restore()
            g.solve(iters, True, prec)
        dt = (time.perf_counter() - t0) / steps
        st = g.stats()
        g.close()
        return dt, st

    for _ in range(args.warmup):
        frame(FP64)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame(FP64)
    per_frame = (time.perf_counter() - t0) / args.steps
    paths = {}
    for name, prec in (("fp64", FP64), ("fp64_fast", FP64_FAST), ("mixed", MIXED)):
        dt, st = numeric(prec, max(args.steps, 20))
        paths[name] = {"ms_per_solve": dt * 1e3, "iterations": st["iterations"], "attempts": st["attempts"],
                       "final_chi2": st["final_chi2"]}
    out = {"metric": "tracking-step BA frames per second (map::optimise(1, true, 50), FP64 reference order)",
           "value": 1.0 / per_frame, "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": per_frame * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"tracking BA: last frame of the 100-keyframe workload, {pr.n_obs} observations, "
                                  f"{pr.n_points} fixed landmarks, solve({iters}) with relative convergence"},
           "numeric_solve": paths}
    if not args.no_cpu:
        from oracle.oracle import CpuGraph, have_reference
        kind = "reference" if have_reference() else "port"
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 10.0:
            o = CpuGraph(pr, kind if kind == "reference" else "oracle")
            o.solve(iters, True)
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "frames/s", "cores": 1, "kind": kind,
                               "sample": f"{reps} frames (graph build + solve({iters}, relative)) in "
                                         f"{dt * reps:.1f} s, single-threaded on {cpu_model()}"}
    print(json.dumps(out))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "fp64_fast", "mixed"],
                    help="mixed (default): the headline, mixed vs FP64 at equal final cost")
    ap.add_argument("--single", action="store_true",
                    help="time only --precision at the workload's iteration count (no FP64 comparison; profiling)")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--match", action="store_true",
                    help="bench the projection search (data association) instead of the solver")
    ap.add_argument("--track", action="store_true",
                    help="bench the tracking step's motion-only BA (one free frame) instead of the headline")
    ap.add_argument("--seeds", type=int, default=4,
                    help="repeat the equal-cost comparison on this many other instances of the workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is not implemented)", file=sys.stderr)
        return 2

    if args.config == 2:
        return bench_batch(args)
    if args.match:
        return bench_match(args)
    if args.track:
        return bench_track(args)

    from zeroslam_amd.engine import FP64, FP64_FAST, MIXED, Graph
    from zeroslam_amd.problem import ITERATIONS, synth

    problem = synth(args.config)
    iters = ITERATIONS[args.config]
    precision = {"fp64": FP64, "fp64_fast": FP64_FAST, "mixed": MIXED}[args.precision]
    # Set-up as a drop-in call pays it (untimed in `value`): handle creation, the
    # host-side graph, structure analysis and upload, the first (0-iteration) solve.
    g, gfz = make_graphs(problem)
    setup = g.setup
    print(f"bench: c{args.config} obs={problem.n_obs} poses={problem.n_poses} points={problem.n_points}",
          file=sys.stderr, flush=True)

    def class_times(prec, k):
        """Per-kernel-class device time of `k`-iteration solves from HIP events
        around every class, on a separate handle (outside the timed region)."""
        gt = Graph.from_problem(problem, timing=True)
        gt.solve(0)
        gt.snapshot()
        out = time_solves(gt, prec, k, 2, 1)
        gt.close()
        return out[3], out[4], out[2]

    if args.single:
        value, ms, stats, _, _ = time_solves(g, precision, iters, args.steps, args.warmup)
        kms, kl, cst = class_times(precision, iters)
        roofline, avg = roofline_of(problem, args.precision, kms, kl, cst, 2, args.config)
        print(json.dumps({"metric": "BA observation*iters/s (single precision run)", "value": value,
                          "unit": "obs*iter/s", "ms_per_step": ms, "precision": args.precision,
                          "roofline": roofline, "kernel_avg_ms": avg}))
        return 0

    # the plain FP64 GPU path (the reference's operation order): throughput and final cost
    v64, ms64, st64, k64, l64 = time_solves(g, FP64, iters, args.steps, args.warmup)
    chi64 = st64["final_chi2"]
    P64, X64 = g.poses(), g.points()
    print(f"bench: fp64 {v64:.4g} obs*iter/s, chi2 {chi64!r}", file=sys.stderr, flush=True)
    extra = {"fp64": {"value": v64, "ms_per_step": ms64, "lm_iterations": st64["iterations"],
                      "attempts": st64["attempts"], "final_chi2": chi64}}
    run_iters = iters
    if precision == FP64:
        value, ms, stats = v64, ms64, st64
    else:
        vf, msf, stf, _, _ = time_solves(gfz, FP64_FAST, iters, args.steps, args.warmup)
        extra["fp64_fast"] = {"value": vf, "ms_per_step": msf, "lm_iterations": stf["iterations"],
                              "final_chi2": stf["final_chi2"], "freeze_behind": True,
                              "chi2_rel_vs_fp64": (stf["final_chi2"] - chi64) / chi64,
                              "state_rel_err_vs_fp64": state_error(gfz.poses(), gfz.points(), P64, X64)}
        print(f"bench: fp64_fast {vf:.4g} obs*iter/s", file=sys.stderr, flush=True)
        if precision == FP64_FAST:
            value, ms, stats = vf, msf, stf
        else:
            cmp, (vraw, ms, stats, _, _), k64 = equal_cost(g, gfz, problem, iters, args.steps, args.warmup, chi64,
                                                           ms64, P64, X64)
            run_iters = cmp["iterations_mixed"]
            # observation*iterations of the FP64 run to the same cost, per second of the mixed run
            value = problem.n_obs * k64 / (ms * 1e-3)
            extra["mixed"] = cmp
            if args.seeds > 0:
                extra["other_seeds"] = [seed_compare(seed, args.config, iters) for seed in range(2, 2 + args.seeds)]
    kms, kl, cst = class_times(precision, run_iters)
    roofline, avg = roofline_of(problem, args.precision, kms, kl, cst, 2, args.config)
    out = {
        "metric": "BA observation*iters/s, 100KF/20k pts: mixed vs FP64 GPU at equal final cost",
        "value": value, "unit": "obs*iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if precision == MIXED else "f64", "data": "synthetic",
        "config": {"workload": f"c{args.config}: {problem.n_poses} KF, {problem.n_points} pts, {problem.n_obs} obs, "
                               f"FP64 reference {iters} LM iterations", "precision": args.precision,
                   "lm_iterations": stats["iterations"], "iteration_budget": run_iters,
                   "attempts": stats["attempts"], "final_chi2": stats["final_chi2"]},
        "roofline": roofline,
        "compare": extra,
        "kernel_avg_ms": avg,
        "setup": setup,
    }
    if not args.no_cpu:
        print("bench: reference CPU optimiser sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(problem, args.cpu_iters)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
