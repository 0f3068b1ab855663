"""Bundle-adjustment problems as plain arrays, and the synthetic workloads.

A :class:`Problem` is the data a ZeroSLAM factor graph holds for bundle
adjustment (reference ``source/map.hpp:129-176``): pose vertices in the reference
parameter layout ``[tx, ty, tz, qx, qy, qz, qw]`` (``factor_graph.hpp:941``),
marginalised point vertices, and reprojection edges (pose id, point id, pixel
observation, 2x2 information, pinhole intrinsics, robust loss) in insertion order.
"""
from __future__ import annotations

import dataclasses
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import ptr

C_CONFIG_NAMES = {
    0: "c0: 10 KF circle, 500 pts, FP64, 10 LM iterations",
    1: "c1: 100 KF walk, 20k pts, ~160k obs, Huber 2 px, 5% outliers, 20 LM iterations",
    2: "c2: 256 local windows (5-30 KF, 200-3000 pts), mixed, 10 LM iterations",
    3: "c3: 50 KF, 5k pts, ~40k obs, no outliers (adjoint)",
    4: "c4: 1000 KF double loop, 200k pts, ~2M obs, Huber, n=6000",
}
ITERATIONS = {0: 10, 1: 20, 2: 10, 3: 20, 4: 15}


@dataclasses.dataclass
class Problem:
    poses: np.ndarray          # (Np, 7) float64
    pose_fixed: np.ndarray     # (Np,) uint8
    points: np.ndarray         # (Nl, 3) float64
    point_fixed: np.ndarray    # (Nl,) uint8
    obs_pose: np.ndarray       # (No,) int32
    obs_point: np.ndarray      # (No,) int32
    obs_uv: np.ndarray         # (No, 2) float64
    obs_K: np.ndarray          # (No, 4) float64 (fx, fy, cx, cy)
    obs_loss: np.ndarray       # (No,) int32 loss kind
    obs_delta: np.ndarray      # (No,) float64 loss parameter
    obs_info: np.ndarray | None = None  # (No, 4) float64 or None = identity
    true_poses: np.ndarray | None = None
    true_points: np.ndarray | None = None
    outlier: np.ndarray | None = None

    @property
    def n_poses(self) -> int:
        return int(self.poses.shape[0])

    @property
    def n_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def n_obs(self) -> int:
        return int(self.obs_pose.shape[0])

    def copy(self) -> "Problem":
        return Problem(**{f.name: (None if getattr(self, f.name) is None else getattr(self, f.name).copy())
                          for f in dataclasses.fields(self)})

    def uniform_camera(self) -> np.ndarray | None:
        if self.n_obs and np.all(self.obs_K == self.obs_K[0]):
            return self.obs_K[0].copy()
        return None

    def uniform_loss(self) -> tuple[int, float] | None:
        if self.n_obs and np.all(self.obs_loss == self.obs_loss[0]) and np.all(self.obs_delta == self.obs_delta[0]):
            return int(self.obs_loss[0]), float(self.obs_delta[0])
        return None

    def save(self, path: str | Path) -> None:
        arrays = {f.name: getattr(self, f.name) for f in dataclasses.fields(self) if getattr(self, f.name) is not None}
        np.savez_compressed(path, **arrays)

    @staticmethod
    def load(path: str | Path, prefix: str = "") -> "Problem":
        with np.load(path, allow_pickle=False) as z:
            kw = {}
            for f in dataclasses.fields(Problem):
                key = prefix + f.name
                kw[f.name] = z[key].copy() if key in z.files else None
        return Problem(**kw)


def synth_params(config: int, **overrides) -> _lib.SynthParams:
    lib = _lib.load()
    p = _lib.SynthParams()
    if lib.zsba_synth_default_params(int(config), p) != 0:
        raise ValueError(f"unknown configuration {config}")
    for k, v in overrides.items():
        if k == "K":
            for i in range(4):
                p.K[i] = float(v[i])
        elif not hasattr(p, k):
            raise KeyError(k)
        else:
            setattr(p, k, v)
    return p


def synth(config: int, seed: int | None = None, **overrides) -> Problem:
    """Deterministic synthetic problem of benchmark configuration `config`
    (seed defaults to the configuration index)."""
    lib = _lib.load()
    params = synth_params(config, **overrides)
    h = lib.zsba_synth_create(params, int(config if seed is None else seed))
    if not h:
        raise ValueError("invalid synthetic-problem parameters")
    try:
        sizes = (np.zeros(3, dtype=np.int64))
        lib.zsba_synth_sizes(h, ptr(sizes, _lib.C.c_int64))
        n_p, n_l, n_o = (int(x) for x in sizes)
        poses = np.zeros((n_p, 7)); pose_fixed = np.zeros(n_p, np.uint8)
        points = np.zeros((n_l, 3)); point_fixed = np.zeros(n_l, np.uint8)
        obs_pose = np.zeros(n_o, np.int32); obs_point = np.zeros(n_o, np.int32)
        obs_uv = np.zeros((n_o, 2)); true_poses = np.zeros((n_p, 7)); true_points = np.zeros((n_l, 3))
        outlier = np.zeros(n_o, np.uint8)
        lib.zsba_synth_copy(h, ptr(poses), ptr(pose_fixed, _lib.C.c_uint8), ptr(points),
                            ptr(point_fixed, _lib.C.c_uint8), ptr(obs_pose, _lib.C.c_int32),
                            ptr(obs_point, _lib.C.c_int32), ptr(obs_uv), ptr(true_poses), ptr(true_points),
                            ptr(outlier, _lib.C.c_uint8))
    finally:
        lib.zsba_synth_destroy(h)
    K = np.array(params.K[:], dtype=np.float64)
    return Problem(poses=poses, pose_fixed=pose_fixed, points=points, point_fixed=point_fixed,
                   obs_pose=obs_pose, obs_point=obs_point, obs_uv=obs_uv,
                   obs_K=np.tile(K, (n_o, 1)), obs_loss=np.full(n_o, params.loss_kind, np.int32),
                   obs_delta=np.full(n_o, params.loss_delta, np.float64), obs_info=None,
                   true_poses=true_poses, true_points=true_points, outlier=outlier)


def motion_only(pr: Problem, n_free: int = 1) -> Problem:
    """The graph ``map::optimise(n_free, true, ...)`` builds from a map
    (``map.hpp:79-192``): the last ``n_free`` frames free, every landmark fixed,
    edges only to the free frames (an edge to a fixed frame is skipped when the
    landmarks are fixed, ``map.hpp:150-152``), and frames and landmarks left
    without edges dropped (``map.hpp:181-192``).  With one free frame this is the
    motion-only BA of the tracking step (``slam.hpp:316,387,469``): a single 6x6
    system per iteration."""
    free = np.zeros(pr.n_poses, bool)
    free[pr.n_poses - n_free:] = True
    keep = free[pr.obs_pose]
    pose_ids = np.flatnonzero(free)
    point_ids = np.unique(pr.obs_point[keep])
    pose_map = np.full(pr.n_poses, -1, np.int32); pose_map[pose_ids] = np.arange(len(pose_ids))
    point_map = np.full(pr.n_points, -1, np.int32); point_map[point_ids] = np.arange(len(point_ids))
    sub = lambda a, m: None if a is None else a[m].copy()
    return Problem(poses=pr.poses[pose_ids].copy(), pose_fixed=np.zeros(len(pose_ids), np.uint8),
                   points=pr.points[point_ids].copy(), point_fixed=np.ones(len(point_ids), np.uint8),
                   obs_pose=pose_map[pr.obs_pose[keep]], obs_point=point_map[pr.obs_point[keep]],
                   obs_uv=sub(pr.obs_uv, keep), obs_K=sub(pr.obs_K, keep), obs_loss=sub(pr.obs_loss, keep),
                   obs_delta=sub(pr.obs_delta, keep), obs_info=sub(pr.obs_info, keep),
                   true_poses=sub(pr.true_poses, pose_ids), true_points=sub(pr.true_points, point_ids),
                   outlier=sub(pr.outlier, keep))


def synth_window(index: int, seed_base: int = 2) -> Problem:
    """One independent local-BA window of the batched workload: keyframes
    U{5..30}, points U{200..3000}, track length U{2..window}, camera and noise
    model of the 100-keyframe workload."""
    rng = np.random.Generator(np.random.PCG64(1000003 * seed_base + index))
    n_frames = int(rng.integers(5, 31))
    n_points = int(rng.integers(200, 3001))
    return synth(2, seed=1000003 * seed_base + index, n_frames=n_frames, n_points=n_points,
                 track_min=2, track_max=n_frames)
