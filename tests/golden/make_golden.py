"""Writes the golden fixtures of the bundle-adjustment path from the compiled
reference (oracle/_ref, built by `make -C oracle ref` from the reference checkout).

Each fixture ``tests/golden/<name>.npz`` holds a problem (arrays prefixed ``p_``),
and what the reference's own factor graph returns for it through its public API
(source/factor_graph.hpp:385-629):
  * ``k``/``accepted``/``chi2``: solve(k) on a fresh graph for each k, the
    accepted-step count and get_current_chi() afterwards;
  * ``poses_<k>``/``points_<k>``: the parameters after solve(k);
  * ``trace``: the verbose LM lines (factor_graph(true)) of the largest k;
  * ``edge_idx``/``edge_eval``: compute_residual + compute_jacobians + robust_info
    + chi2 + robust_chi2 of sampled edges at the initial state, and
    ``edge_eval_late`` at the state after solve(``late_k``).
Run from the repository root: ``python tests/golden/make_golden.py``.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.oracle import CpuGraph, have_reference  # noqa: E402
from zeroslam_amd.problem import Problem, synth  # noqa: E402

OUT = Path(__file__).resolve().parent


class pcg32_test:
    """The PCG32 stream of tests/factor_graph.test.cpp:62-83 (fixed seed)."""

    def __init__(self):
        self.state = 0x853C49E6748FEA9B
        self.inc = 0xDA3E39CB94B95BDB

    def raw(self) -> int:
        prev = self.state
        self.state = (prev * 0x5851F42D4C957F2D + self.inc) & 0xFFFFFFFFFFFFFFFF
        xs = (((prev >> 18) ^ prev) >> 27) & 0xFFFFFFFF
        rot = prev >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF


def seven_camera_problem() -> Problem:
    """The 7-camera / 3-landmark problem of tests/factor_graph.test.cpp:85-156:
    identity rotations, +-0.1 m translations, camera 0 fixed, K = (1, 1, 0, 0),
    landmark noise (rng % 10 - 5) * 0.01, observation noise (rng % 10 - 5) * 1e-4."""
    rng = pcg32_test()
    cams = [(0, 0, 0), (0.1, 0, 0), (0, 0.1, 0), (0, 0, 0.1), (-0.1, 0, 0), (0, -0.1, 0), (0, 0, -0.1)]
    lms = [(0.0, 0.0, 1.0), (0.0, 1.0, 1.0), (1.0, 0.0, 1.0)]
    poses = np.array([[t[0], t[1], t[2], 0, 0, 0, 1.0] for t in cams])
    noisy = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            noisy[i, j] = lms[i][j] + float(int(rng.raw() % 10) - 5) * 0.01
    obs_pose, obs_point, uv = [], [], []
    for ci, t in enumerate(cams):
        for li, X in enumerate(lms):
            Pc = (X[0] + t[0], X[1] + t[1], X[2] + t[2])
            u, v = Pc[0] * (1.0 / Pc[2]), Pc[1] * (1.0 / Pc[2])
            du = float(int(rng.raw() % 10) - 5) * 0.0001
            dv = float(int(rng.raw() % 10) - 5) * 0.0001
            obs_pose.append(ci); obs_point.append(li); uv.append((u + du, v + dv))
    n = len(obs_pose)
    return Problem(poses=poses, pose_fixed=np.array([1, 0, 0, 0, 0, 0, 0], np.uint8), points=noisy,
                   point_fixed=np.zeros(3, np.uint8), obs_pose=np.array(obs_pose, np.int32),
                   obs_point=np.array(obs_point, np.int32), obs_uv=np.array(uv),
                   obs_K=np.tile([1.0, 1.0, 0.0, 0.0], (n, 1)), obs_loss=np.zeros(n, np.int32),
                   obs_delta=np.zeros(n))


def negative_information_problem() -> Problem:
    """tests/factor_graph.test.cpp:182-216: information -I makes the reduced
    system indefinite until the damping grows."""
    return Problem(poses=np.array([[0, 0, 0, 0, 0, 0, 1.0], [0.1, 0, 0, 0, 0, 0, 1.0]]),
                   pose_fixed=np.array([1, 0], np.uint8), points=np.array([[0.1, 0.2, 1.0]]),
                   point_fixed=np.array([1], np.uint8), obs_pose=np.array([1], np.int32),
                   obs_point=np.array([0], np.int32), obs_uv=np.array([[0.15, 0.25]]),
                   obs_K=np.array([[1.0, 1.0, 0.0, 0.0]]), obs_loss=np.zeros(1, np.int32), obs_delta=np.zeros(1),
                   obs_info=np.array([[-1.0, -0.0, -0.0, -1.0]]))


def behind_camera_problem() -> Problem:
    """tests/factor_graph.test.cpp:226-323: a point in front (0, 0, 1) and one
    behind (0.1, 0.2, -10) an identity camera, observation (0.1, -0.2)."""
    return Problem(poses=np.array([[0, 0, 0, 0, 0, 0, 1.0]]), pose_fixed=np.array([0], np.uint8),
                   points=np.array([[0.0, 0.0, 1.0], [0.1, 0.2, -10.0]]), point_fixed=np.zeros(2, np.uint8),
                   obs_pose=np.array([0, 0], np.int32), obs_point=np.array([0, 1], np.int32),
                   obs_uv=np.array([[0.1, -0.2], [0.1, -0.2]]), obs_K=np.tile([1.0, 1.0, 0.0, 0.0], (2, 1)),
                   obs_loss=np.zeros(2, np.int32), obs_delta=np.zeros(2))


def mixed_problem() -> Problem:
    """Everything the per-edge path distinguishes in one small graph: all robust
    losses, non-identity information, two cameras, fixed points, repeated
    (pose, point) edges and points behind cameras."""
    pr = synth(1, seed=11, n_frames=8, n_points=150, track_min=2, track_max=8)
    rng = np.random.Generator(np.random.PCG64(11))
    n = pr.n_obs
    kinds = np.array([0, 1, 2, 3, 4], np.int32)[np.arange(n) % 5]
    deltas = np.where(kinds == 2, 2.0, np.where(kinds == 3, 3.0, np.where(kinds == 4, 5.0, 0.0)))
    info = np.tile([1.0, 0.0, 0.0, 1.0], (n, 1))
    sel = rng.random(n) < 0.25
    a = rng.uniform(0.5, 2.0, n); c = rng.uniform(0.5, 2.0, n); b = rng.uniform(-0.3, 0.3, n)
    info[sel] = np.stack([a, b, b, c], 1)[sel]
    K = np.tile([525.0, 525.0, 320.0, 240.0], (n, 1))
    K[pr.obs_pose % 2 == 1] = [500.0, 510.0, 322.0, 238.0]
    dup = np.arange(0, n, max(1, n // 10))
    pr.point_fixed[[0, 5]] = 1
    pr.points[9] = pr.points[9] - np.array([0.0, 0.0, 60.0])  # behind the cameras
    return Problem(poses=pr.poses, pose_fixed=pr.pose_fixed, points=pr.points, point_fixed=pr.point_fixed,
                   obs_pose=np.concatenate([pr.obs_pose, pr.obs_pose[dup]]),
                   obs_point=np.concatenate([pr.obs_point, pr.obs_point[dup]]),
                   obs_uv=np.concatenate([pr.obs_uv, pr.obs_uv[dup] + 0.5]),
                   obs_K=np.concatenate([K, K[dup]]), obs_loss=np.concatenate([kinds, kinds[dup]]),
                   obs_delta=np.concatenate([deltas, deltas[dup]]), obs_info=np.concatenate([info, info[dup]]))


CASES = {
    # name: (problem factory, ks, stored-state ks, use_relative, edge sample, late k)
    "c0": (lambda: synth(0), list(range(1, 11)), [1, 2, 4, 10], False, 512, 3),
    "c1s": (lambda: synth(1, n_frames=20, n_points=1500), [1, 2, 3, 5, 10, 20], [1, 3, 20], False, 512, 5),
    "c3s": (lambda: synth(3, n_frames=12, n_points=800), [1, 2, 5, 10], [2, 10], False, 256, 2),
    "seven": (seven_camera_problem, [1, 2, 5, 50], [1, 50], True, None, 5),
    "neginfo": (negative_information_problem, [1, 2, 10], [1, 10], False, None, 2),
    "behind": (behind_camera_problem, [1, 3], [3], False, None, 1),
    "mixed": (mixed_problem, [1, 2, 5, 10], [1, 10], False, None, 3),
}


def make(name: str) -> None:
    factory, ks, store, use_rel, sample, late_k = CASES[name]
    pr = factory()
    out = {f"p_{k}": v for k, v in vars(pr).items() if v is not None}
    accepted, chi2 = [], []
    for k in ks:
        g = CpuGraph(pr, "reference", verbose=1 if k == max(ks) else 0)
        if k == max(ks):
            acc, trace = g.solve_capture(k, use_rel)
            out["trace"] = np.frombuffer(trace.encode(), dtype=np.uint8)
        else:
            acc = g.solve(k, use_rel)
        accepted.append(acc)
        chi2.append(g.current_chi(True))
        if k in store:
            P, X = g.state()
            out[f"poses_{k}"] = P
            out[f"points_{k}"] = X
    out["k"] = np.array(ks, np.int32)
    out["accepted"] = np.array(accepted, np.int32)
    out["chi2"] = np.array(chi2)
    out["use_relative"] = np.array([int(use_rel)], np.int32)
    n = pr.n_obs
    idx = np.arange(n) if sample is None or sample >= n else np.linspace(0, n - 1, sample).astype(np.int64)
    g = CpuGraph(pr, "reference")
    out["edge_idx"] = idx.astype(np.int32)
    out["edge_eval"] = g.edge_eval(idx)
    g.solve(late_k, use_rel)
    out["late_k"] = np.array([late_k], np.int32)
    out["edge_eval_late"] = g.edge_eval(idx)
    np.savez_compressed(OUT / f"{name}.npz", **out)
    print(f"{name}: obs={n} accepted={accepted} chi2={chi2[-1]:.10g} -> {OUT / (name + '.npz')}")


def load(name: str) -> tuple[Problem, dict]:
    path = OUT / f"{name}.npz"
    pr = Problem.load(path, prefix="p_")
    with np.load(path, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files if not k.startswith("p_")}
    return pr, data


if __name__ == "__main__":
    if not have_reference():
        sys.exit("oracle/_ref/libzsba_ref.so is missing: run `make -C oracle ref` first")
    for name in (sys.argv[1:] or list(CASES)):
        make(name)
