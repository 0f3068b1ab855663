"""Writes the fixtures of the full-size workloads, which are too slow for the
CPU oracle inside a test, from the compiled reference (oracle/_ref) or the
oracle (bitwise equal to it, tests/test_oracle_reference.py):

  * ``c1_seeds.npz``: the 100-keyframe / 20k-point / ~160k-observation workload
    (forward walk, Huber 2 px, 5% outliers) for generator seeds 1..5: the
    reference's get_current_chi() after solve(20) on a fresh factor graph
    (source/factor_graph.hpp:489-629), its accepted-step count and poses, and a
    digest of the problem arrays so the test can tell that the generator made
    the same problem.
  * ``c4_full.npz``: the 1000-keyframe loop-closure workload (200k points, ~2M
    observations, 6000-row reduced system) after solve(1), solve(2) and
    solve(15) on fresh graphs with the oracle: accepted steps, chi2, every pose,
    the SHA-256 of the point array and a sample of 2000 points.

Run from the repository root: ``python tests/golden/make_large.py [c1|c4]``.
The problems are regenerated by zeroslam_amd.problem.synth; nothing here
needs a GPU.
"""
from __future__ import annotations

import hashlib
import sys
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

OUT = Path(__file__).resolve().parent
C1_SEEDS = (1, 2, 3, 4, 5)
C4_SAMPLE = 2000


def problem_digest(pr) -> str:
    h = hashlib.sha256()
    for a in (pr.poses, pr.points, pr.obs_pose, pr.obs_point, pr.obs_uv, pr.pose_fixed, pr.point_fixed):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _c1(seed: int):
    from oracle.oracle import CpuGraph, have_reference
    from zeroslam_amd.problem import synth
    pr = synth(1, seed=seed)
    g = CpuGraph(pr, "reference" if have_reference() else "oracle")
    acc = g.solve(20)
    P, _ = g.state()
    return seed, problem_digest(pr), acc, g.current_chi(True), P


def make_c1() -> None:
    with ProcessPoolExecutor(max_workers=len(C1_SEEDS)) as pool:
        rows = list(pool.map(_c1, C1_SEEDS))
    out = {"seeds": np.array([r[0] for r in rows], np.int32),
           "digest": np.array([r[1] for r in rows]),
           "accepted": np.array([r[2] for r in rows], np.int32),
           "chi2": np.array([r[3] for r in rows]),
           "poses": np.stack([r[4] for r in rows])}
    np.savez_compressed(OUT / "c1_seeds.npz", **out)
    print("c1_seeds:", [(int(s), float(c)) for s, c in zip(out["seeds"], out["chi2"])])


def _c4(k: int):
    from oracle.oracle import CpuGraph
    from zeroslam_amd.problem import synth
    pr = synth(4)
    g = CpuGraph(pr, "oracle")
    acc = g.solve(k)
    P, X = g.state()
    return k, problem_digest(pr), acc, g.current_chi(True), P, X, g.stats()


def make_c4(ks=(1, 2, 15)) -> None:
    with ProcessPoolExecutor(max_workers=len(ks)) as pool:
        rows = list(pool.map(_c4, ks))
    n_points = rows[0][5].shape[0]
    sample = np.linspace(0, n_points - 1, C4_SAMPLE).astype(np.int64)
    out = {"digest": np.array([rows[0][1]]), "sample": sample}
    for k, _, acc, chi, P, X, st in rows:
        out[f"accepted_{k}"] = np.array([acc], np.int32)
        out[f"attempts_{k}"] = np.array([st["attempts"]], np.int32)
        out[f"chi2_{k}"] = np.array([chi])
        out[f"poses_{k}"] = P
        out[f"points_sample_{k}"] = X[sample]
        out[f"points_sha256_{k}"] = np.array([hashlib.sha256(np.ascontiguousarray(X).tobytes()).hexdigest()])
        print(f"c4 solve({k}): accepted {acc}, attempts {st['attempts']}, chi2 {chi!r}", flush=True)
    np.savez_compressed(OUT / "c4_full.npz", **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["c1", "c4"]
    if "c1" in which:
        make_c1()
    if "c4" in which:
        make_c4()
