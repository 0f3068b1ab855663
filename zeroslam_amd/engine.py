"""Python handle over the engine's C ABI (``include/zsba.h``).

:class:`Graph` owns one ``zsba_graph`` handle.  It is a thin layer: every call goes
to ``libzsba.so`` and every numeric result comes from the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import ptr
from .problem import MatchScene, Problem

FP64 = _lib.ZSBA_PRECISION_FP64
MIXED = _lib.ZSBA_PRECISION_MIXED
FP64_FAST = _lib.ZSBA_PRECISION_FP64_FAST


def math_probe(fn: int, x, on_device: bool) -> np.ndarray:
    """The engine's correctly rounded sin (0), cos (1), cube (2), sqrt (3)."""
    lib = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    rc = lib.zsba_debug_math(int(fn), len(x), ptr(x), ptr(out), int(bool(on_device)))
    if rc < 0:
        raise ZsbaError(rc, "math probe failed")
    return out


class ZsbaError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{_lib.ERRORS.get(code, code)}: {message}")
        self.code = code


class Graph:
    def __init__(self, device: int = 0, verbose: int = 0, refine_steps: int = 0, timing: bool = False,
                 floor_exp: int = 0, factor_path: int = 0, timing_classes=None, freeze_behind: int = 0,
                 attempt_graph: int = 0):
        """timing: time kernel classes with HIP events (stats()["kernel_ms"]);
        timing_classes: names from KERNEL_NAMES to time (None: all);
        freeze_behind: the behind-camera freeze (0 default: MIXED only, 1 also
        FP64_FAST, -1 off; never FP64);
        attempt_graph: 0 replays each LM attempt as one captured HIP graph, -1
        launches its kernels one by one."""
        self.lib = _lib.load()
        mask = 0
        if timing_classes:
            for name in timing_classes:
                mask |= 1 << _lib.KERNEL_NAMES.index(name)
        opts = _lib.Options(device=device, verbose=verbose, refine_steps=refine_steps, timing=int(bool(timing)),
                            floor_exp=floor_exp, factor_path=factor_path, timing_classes=mask,
                            freeze_behind=freeze_behind, attempt_graph=attempt_graph)
        self.h = self.lib.zsba_create(C.byref(opts))
        if not self.h:
            raise MemoryError("zsba_create failed")

    # -- lifetime ------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.zsba_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int) -> int:
        if rc < 0:
            raise ZsbaError(rc, self.lib.zsba_last_error(self.h).decode())
        return rc

    # -- construction ----------------------------------------------------------
    @classmethod
    def from_problem(cls, problem: Problem, **kw) -> "Graph":
        g = cls(**kw)
        g.load_problem(problem)
        return g

    def load_problem(self, pr: Problem) -> None:
        self.add_poses(pr.poses, pr.pose_fixed)
        self.add_points(pr.points, pr.point_fixed)
        K = pr.uniform_camera()
        loss = pr.uniform_loss()
        if K is not None and loss is not None:
            self.add_observations(pr.obs_pose, pr.obs_point, pr.obs_uv, K, loss[0], loss[1], pr.obs_info)
        else:
            for e in range(pr.n_obs):
                info = None if pr.obs_info is None else pr.obs_info[e]
                self.add_observation(int(pr.obs_pose[e]), int(pr.obs_point[e]), pr.obs_uv[e], pr.obs_K[e],
                                     int(pr.obs_loss[e]), float(pr.obs_delta[e]), info)

    def add_poses(self, params, fixed=None) -> int:
        params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 7)
        fixed = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        return self._check(self.lib.zsba_add_poses(self.h, len(params), ptr(params), ptr(fixed, C.c_uint8)))

    def add_points(self, xyz, fixed=None) -> int:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        fixed = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        return self._check(self.lib.zsba_add_points(self.h, len(xyz), ptr(xyz), ptr(fixed, C.c_uint8)))

    def add_observations(self, pose_ids, point_ids, uv, K, loss_kind=0, loss_delta=0.0, info=None) -> int:
        pose_ids = np.ascontiguousarray(pose_ids, dtype=np.int32)
        point_ids = np.ascontiguousarray(point_ids, dtype=np.int32)
        uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
        K = np.ascontiguousarray(K, dtype=np.float64)
        info = None if info is None else np.ascontiguousarray(info, dtype=np.float64).reshape(-1, 4)
        return self._check(self.lib.zsba_add_observations(
            self.h, len(pose_ids), ptr(pose_ids, C.c_int32), ptr(point_ids, C.c_int32), ptr(uv), ptr(info), ptr(K),
            int(loss_kind), float(loss_delta)))

    def add_observation(self, pose_id, point_id, uv, K, loss_kind=0, loss_delta=0.0, info=None) -> int:
        uv = np.ascontiguousarray(uv, dtype=np.float64)
        K = np.ascontiguousarray(K, dtype=np.float64)
        info = None if info is None else np.ascontiguousarray(info, dtype=np.float64)
        return self._check(self.lib.zsba_add_observation(self.h, int(pose_id), int(point_id), ptr(uv), ptr(info), ptr(K),
                                                         int(loss_kind), float(loss_delta)))

    def remove_observation(self, i: int) -> None:
        self._check(self.lib.zsba_remove_observation(self.h, int(i)))

    def remove_pose(self, i: int) -> None:
        self._check(self.lib.zsba_remove_pose(self.h, int(i)))

    def remove_point(self, i: int) -> None:
        self._check(self.lib.zsba_remove_point(self.h, int(i)))

    def set_pose(self, i: int, params) -> None:
        params = np.ascontiguousarray(params, dtype=np.float64)
        self._check(self.lib.zsba_set_pose(self.h, int(i), ptr(params)))

    def set_point(self, i: int, xyz) -> None:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self._check(self.lib.zsba_set_point(self.h, int(i), ptr(xyz)))

    def set_pose_fixed(self, i: int, fixed: bool) -> None:
        self._check(self.lib.zsba_set_pose_fixed(self.h, int(i), int(bool(fixed))))

    def set_point_fixed(self, i: int, fixed: bool) -> None:
        self._check(self.lib.zsba_set_point_fixed(self.h, int(i), int(bool(fixed))))

    def counts(self) -> tuple[int, int, int]:
        c = np.zeros(3, np.int64)
        self._check(self.lib.zsba_counts(self.h, ptr(c, C.c_int64)))
        return int(c[0]), int(c[1]), int(c[2])

    # -- optimisation -------------------------------------------------------------
    def solve(self, iterations: int = 10, use_relative_convergence: bool = False, precision: int = FP64) -> int:
        return self._check(self.lib.zsba_solve(self.h, int(iterations), int(bool(use_relative_convergence)),
                                               int(precision)))

    def snapshot(self) -> None:
        """Keep the current device parameters aside (no host transfer)."""
        self._check(self.lib.zsba_snapshot(self.h))

    def restore(self) -> None:
        """Put the snapshot back on the device (no host transfer)."""
        self._check(self.lib.zsba_restore(self.h))

    def current_chi(self, recompute: bool = True) -> float:
        v = self.lib.zsba_current_chi(self.h, int(bool(recompute)))
        if np.isnan(v) and self.lib.zsba_last_error(self.h):
            msg = self.lib.zsba_last_error(self.h).decode()
            if msg:
                raise ZsbaError(-3, msg)
        return float(v)

    def poses(self) -> np.ndarray:
        out = np.full((self.counts()[0], 7), np.nan)
        self._check(self.lib.zsba_get_poses(self.h, ptr(out)))
        return out

    def points(self) -> np.ndarray:
        out = np.full((self.counts()[1], 3), np.nan)
        self._check(self.lib.zsba_get_points(self.h, ptr(out)))
        return out

    def inlier_flags(self) -> np.ndarray:
        out = np.zeros(self.counts()[2], np.uint8)
        self._check(self.lib.zsba_get_inlier_flags(self.h, ptr(out, C.c_uint8)))
        return out

    def reprojection_cull(self, threshold: float = 5.991) -> tuple[np.ndarray, np.ndarray]:
        """map::cull's reprojection test at the current state (zsba_reprojection_cull):
        per point id, the mean reprojection error (float32, +inf without
        observations) and the erase flag."""
        n = self.counts()[1]
        mean = np.full(n, np.nan, np.float32)
        erase = np.zeros(n, np.uint8)
        self._check(self.lib.zsba_reprojection_cull(self.h, C.c_float(threshold), ptr(mean, C.c_float),
                                                    ptr(erase, C.c_uint8)))
        return mean, erase

    def residuals(self) -> np.ndarray:
        out = np.full((self.counts()[2], 2), np.nan)
        self._check(self.lib.zsba_get_residuals(self.h, ptr(out)))
        return out

    def stats(self) -> dict:
        st = _lib.Stats()
        self._check(self.lib.zsba_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    # -- diagnostics ---------------------------------------------------------------
    def eval_edges(self) -> np.ndarray:
        out = np.full((self.counts()[2], _lib.EDGE_EVAL_STRIDE), np.nan)
        self._check(self.lib.zsba_debug_eval_edges(self.h, ptr(out)))
        return out

    def reduced_system(self, lam: float, precision: int = FP64):
        n_p = int(np.count_nonzero(~np.isnan(self.poses()[:, 0])))
        n_l = int(np.count_nonzero(~np.isnan(self.points()[:, 0])))
        n = 6 * n_p
        S = np.zeros((n, n)); bs = np.zeros(n); xp = np.zeros(n); xl = np.zeros(3 * n_l)
        ok = self._check(self.lib.zsba_debug_reduced_system(self.h, float(lam), int(precision), ptr(S), ptr(bs),
                                                            ptr(xp), ptr(xl)))
        return bool(ok), S, bs, xp, xl

    def adjoint(self, dL_dpoints, dL_dposes=None, return_iterations: bool = False):
        """Gradient of a loss L(x*) of the converged state with respect to every
        observation's pixel coordinates (by observation id) and to the pinhole
        intrinsics (fx, fy, cx, cy) shared by the observations, by implicit
        differentiation (zsba_adjoint).  dL_dpoints: (n_points, 3) by point id;
        dL_dposes: (n_poses, 6) in the pose retraction's local coordinates
        (rotation, translation) or None."""
        n_p, n_l, n_o = self.counts()
        dL_dpoints = np.ascontiguousarray(dL_dpoints, dtype=np.float64).reshape(n_l, 3)
        dL_dposes = None if dL_dposes is None else np.ascontiguousarray(dL_dposes, dtype=np.float64).reshape(n_p, 6)
        d_obs = np.zeros((n_o, 2))
        d_K = np.zeros(4)
        it = self._check(self.lib.zsba_adjoint(self.h, ptr(dL_dpoints), ptr(dL_dposes), ptr(d_obs), ptr(d_K)))
        return (d_obs, d_K, it) if return_iterations else (d_obs, d_K)

    def adjoint_ops(self, v):
        """(H v, M^-1 v) for v ordered [live poses 6 each | live points 3 each]."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        Hv = np.zeros_like(v)
        Mv = np.zeros_like(v)
        self._check(self.lib.zsba_debug_adjoint_ops(self.h, ptr(v), ptr(Hv), ptr(Mv)))
        return Hv, Mv

    def gauss_newton(self, steps: int = 4) -> None:
        """Undamped Gauss-Newton steps, every one applied (zsba_gauss_newton)."""
        self._check(self.lib.zsba_gauss_newton(self.h, int(steps)))

    def hold_indefinite_points(self, rel_tol: float = 0.0) -> np.ndarray:
        """Set fixed the points without a local optimum at the current state
        (zsba_hold_indefinite_points); returns their flags by point id."""
        held = np.zeros(self.counts()[1], np.uint8)
        self._check(self.lib.zsba_hold_indefinite_points(self.h, float(rel_tol), ptr(held, C.c_uint8)))
        return held

    def solve_to_convergence(self, precision: int = FP64_FAST, max_rounds: int = 20, iterations: int = 50,
                             polish: int = 4) -> int:
        """Repeat solve(iterations) until a solve accepts no step (the LM restarts
        with a fresh damping each round), then `polish` undamped Gauss-Newton
        steps; returns the total accepted LM steps."""
        total = 0
        for _ in range(max_rounds):
            acc = self.solve(iterations, False, precision)
            total += acc
            if acc == 0:
                break
        if polish:
            self.gauss_newton(polish)
        return total


class Batch:
    """Many independent graphs (local BA windows) solved together (zsba_batch_*).

    Windows that fit one workgroup run in one hipGraph launch per solve (rounds of
    grid-wide linearisation and per-window attempts); the graphs must outlive the
    batch."""

    def __init__(self, graphs: list[Graph]):
        self.lib = _lib.load()
        self.graphs = list(graphs)
        arr = (C.c_void_p * len(self.graphs))(*[g.h for g in self.graphs])
        self.h = self.lib.zsba_batch_create(arr, len(self.graphs))
        if not self.h:
            raise ValueError("zsba_batch_create failed")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.zsba_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def solve(self, iterations: int, use_relative_convergence: bool = False, precision: int = MIXED) -> np.ndarray:
        acc = np.zeros(len(self.graphs), np.int32)
        rc = self.lib.zsba_batch_solve(self.h, int(iterations), int(bool(use_relative_convergence)), int(precision),
                                       ptr(acc, C.c_int32))
        if rc < 0:
            raise ZsbaError(rc, self.lib.zsba_batch_last_error(self.h).decode())
        return acc

    def info(self) -> dict:
        inf = _lib.BatchInfo()
        rc = self.lib.zsba_batch_get_info(self.h, C.byref(inf))
        if rc < 0:
            raise ZsbaError(rc, "batch info")
        return inf.as_dict()


def solve_batch(graphs: list[Graph], iterations: int, use_relative_convergence: bool = False,
                precision: int = MIXED) -> np.ndarray:
    lib = _lib.load()
    arr = (C.c_void_p * len(graphs))(*[g.h for g in graphs])
    acc = np.zeros(len(graphs), np.int32)
    rc = lib.zsba_solve_batch(arr, len(graphs), int(iterations), int(bool(use_relative_convergence)), int(precision),
                              ptr(acc, C.c_int32))
    if rc < 0:
        msgs = [lib.zsba_last_error(g.h).decode() for g in graphs if lib.zsba_last_error(g.h)]
        raise ZsbaError(rc, "; ".join(m for m in msgs if m))
    return acc


class Matcher:
    """Projection search of a new frame against the map on the device
    (zsba_match_by_projection, the loop of slam.hpp:341-384).  Keeps its device
    buffers between calls."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.h = self.lib.zsba_matcher_create(int(device))
        if not self.h:
            raise ZsbaError(-3, "zsba_matcher_create failed")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.zsba_matcher_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def match(self, scene: MatchScene, max_matches: int | None = None) -> tuple[np.ndarray, np.ndarray, int]:
        """(landmark, keypoint) index pairs in the order the reference adds the
        observations, and the number of matches."""
        f = _lib.MatchFrame()
        f.rotation[:] = [float(x) for x in np.asarray(scene.rotation, np.float64).reshape(9)]
        f.translation[:] = [float(x) for x in np.asarray(scene.translation, np.float64)]
        f.K[:] = [float(x) for x in np.asarray(scene.K, np.float64)]
        f.cols, f.rows = int(scene.cols), int(scene.rows)
        xyz = np.ascontiguousarray(scene.landmark_xyz, np.float64)
        lm_desc = np.ascontiguousarray(scene.landmark_descriptor, np.uint8)
        observed = np.ascontiguousarray(scene.landmark_observed, np.uint8)
        kp_xy = np.ascontiguousarray(scene.keypoint_xy, np.float64)
        kp_desc = np.ascontiguousarray(scene.keypoint_descriptor, np.uint8)
        key = np.ascontiguousarray(scene.keypoint_key, np.int32)
        taken = np.ascontiguousarray(scene.key_taken, np.uint8)
        cap = max(1, len(kp_xy) if max_matches is None else int(max_matches))
        lm = np.zeros(cap, np.int32)
        kp = np.zeros(cap, np.int32)
        n = self.lib.zsba_match_by_projection(
            self.h, C.byref(f), len(xyz), ptr(xyz), ptr(lm_desc, C.c_uint8), ptr(observed, C.c_uint8), len(kp_xy),
            ptr(kp_xy), ptr(kp_desc, C.c_uint8), ptr(key, C.c_int32), len(taken), ptr(taken, C.c_uint8), cap,
            ptr(lm, C.c_int32), ptr(kp, C.c_int32))
        if n < 0:
            raise ZsbaError(n, self.lib.zsba_matcher_last_error(self.h).decode())
        k = min(n, cap)
        return lm[:k].copy(), kp[:k].copy(), n

    def stats(self) -> dict:
        ms = C.c_double()
        nc = C.c_int32()
        self.lib.zsba_matcher_stats(self.h, C.byref(ms), C.byref(nc))
        return {"device_ms": ms.value, "candidates": nc.value}
