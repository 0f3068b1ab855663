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


@dataclasses.dataclass
class MatchScene:
    """The inputs of the projection search of a new frame against the map
    (``slam.hpp:341-384``), as arrays: the frame's rotation (row-major, camera =
    R world + t), translation, intrinsics and image size; the landmarks in the
    map's iteration order with the descriptor of each one's first observation and
    whether the frame already observes it; the frame's matched keypoints with
    their descriptors and previous-frame indices; the previous-frame indices
    taken before the search."""
    rotation: np.ndarray             # (3, 3) float64
    translation: np.ndarray          # (3,) float64
    K: np.ndarray                    # (4,) float64
    cols: int
    rows: int
    landmark_xyz: np.ndarray         # (L, 3) float64
    landmark_descriptor: np.ndarray  # (L, 32) uint8
    landmark_observed: np.ndarray    # (L,) uint8
    keypoint_xy: np.ndarray          # (N, 2) float64
    keypoint_descriptor: np.ndarray  # (N, 32) uint8
    keypoint_key: np.ndarray         # (N,) int32
    key_taken: np.ndarray            # (n_keys,) uint8


def match_scene(n_landmarks: int, n_keypoints: int, seed: int = 0, cols: int = 640, rows: int = 480) -> MatchScene:
    """A synthetic frame and map for the projection search.  Most landmarks are
    in view; some are behind the camera or outside the image; some come in pairs
    that project within a pixel of each other.  About half the keypoints sit
    within 0-2.5 px of a landmark's projection with that landmark's descriptor
    with 40-90 bits flipped (either side of the 64-bit bound); the rest are
    scattered with random descriptors.  Some previous-frame indices are shared by
    two keypoints, some are taken before the search, and some landmarks are
    already observed."""
    rng = np.random.default_rng(seed)
    K = np.array([525.0, 525.0, 319.5, 239.5])
    axis = rng.standard_normal(3)
    axis /= np.linalg.norm(axis)
    a = 0.1
    Kx = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(a) * Kx + (1 - np.cos(a)) * Kx @ Kx
    t = rng.normal(0, 0.2, 3)
    # landmarks in the camera frame, then to the world
    z = rng.uniform(1.5, 25.0, n_landmarks)
    u = rng.uniform(-40, cols + 40, n_landmarks)
    v = rng.uniform(-40, rows + 40, n_landmarks)
    twin = rng.random(n_landmarks) < 0.1
    src = rng.integers(0, n_landmarks, n_landmarks)
    u[twin] = u[src[twin]] + rng.uniform(-0.7, 0.7, twin.sum())
    v[twin] = v[src[twin]] + rng.uniform(-0.7, 0.7, twin.sum())
    z[twin] = z[src[twin]] * rng.uniform(0.9, 1.1, twin.sum())
    cam = np.stack([(u - K[2]) / K[0] * z, (v - K[3]) / K[1] * z, z], axis=1)
    behind = rng.random(n_landmarks) < 0.05
    cam[behind, 2] *= -1.0
    world = (cam - t) @ R  # R^T (c - t) per row
    lm_desc = rng.integers(0, 256, (n_landmarks, 32), dtype=np.uint8)
    lm_desc[twin] = lm_desc[src[twin]]
    observed = (rng.random(n_landmarks) < 0.05).astype(np.uint8)
    # keypoints: half near landmark projections, half scattered
    n_near = n_keypoints // 2
    owner = rng.integers(0, n_landmarks, n_near)
    ang = rng.uniform(0, 2 * np.pi, n_near)
    rad = rng.uniform(0.0, 2.5, n_near)
    kp = np.empty((n_keypoints, 2))
    kp[:n_near, 0] = u[owner] + rad * np.cos(ang)
    kp[:n_near, 1] = v[owner] + rad * np.sin(ang)
    kp[n_near:, 0] = rng.uniform(0, cols, n_keypoints - n_near)
    kp[n_near:, 1] = rng.uniform(0, rows, n_keypoints - n_near)
    kp_desc = rng.integers(0, 256, (n_keypoints, 32), dtype=np.uint8)
    bits = np.unpackbits(lm_desc[owner], axis=1)
    for r in range(n_near):
        flip = rng.choice(256, size=int(rng.integers(40, 91)), replace=False)
        bits[r, flip] ^= 1
    kp_desc[:n_near] = np.packbits(bits, axis=1)
    order = rng.permutation(n_keypoints)
    kp, kp_desc = kp[order], kp_desc[order]
    n_keys = int(n_keypoints * 1.25) + 1
    key = rng.choice(n_keys, size=n_keypoints, replace=False).astype(np.int32)
    shared = rng.random(n_keypoints) < 0.1
    key[shared] = key[rng.integers(0, n_keypoints, shared.sum())]
    taken = (rng.random(n_keys) < 0.1).astype(np.uint8)
    return MatchScene(rotation=R, translation=t, K=K, cols=cols, rows=rows, landmark_xyz=np.ascontiguousarray(world),
                      landmark_descriptor=lm_desc, landmark_observed=observed, keypoint_xy=np.ascontiguousarray(kp),
                      keypoint_descriptor=np.ascontiguousarray(kp_desc), keypoint_key=key, key_taken=taken)
