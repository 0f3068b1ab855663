"""ctypes binding of ``libzsba.so`` (the C ABI declared in ``include/zsba.h``).

The library must have been built in-tree (``python -m zeroslam_amd.build``); there
is no fallback implementation, so a missing or stale library is an error.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libzsba.so"

ZSBA_LOSS_NONE, ZSBA_LOSS_TRIVIAL, ZSBA_LOSS_HUBER, ZSBA_LOSS_CAUCHY, ZSBA_LOSS_TUKEY = range(5)
ZSBA_PRECISION_FP64, ZSBA_PRECISION_MIXED, ZSBA_PRECISION_FP64_FAST = 0, 1, 2
EDGE_EVAL_STRIDE = 27

KERNEL_NAMES = ("linearise", "pose_reduce", "cost", "invert", "form_Y", "fill", "schur", "cholesky", "trsv",
                "backsub", "control", "refine")

ERRORS = {-1: "EINVAL", -2: "ESTATE", -3: "EDEVICE", -4: "ENOMEM", -5: "EUNSUP"}


class Options(C.Structure):
    _fields_ = [("device", C.c_int32), ("verbose", C.c_int32), ("refine_steps", C.c_int32),
                ("timing", C.c_int32), ("floor_exp", C.c_int32), ("factor_path", C.c_int32),
                ("timing_classes", C.c_int32), ("freeze_behind", C.c_int32), ("attempt_graph", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("attempts", C.c_int32), ("accepted", C.c_int32),
                ("rejected", C.c_int32), ("solve_failures", C.c_int32), ("precision", C.c_int32),
                ("n_obs", C.c_int64), ("n_poses", C.c_int64), ("n_points", C.c_int64),
                ("initial_chi2", C.c_double), ("final_chi2", C.c_double), ("lambda_", C.c_double),
                ("solve_ms", C.c_double), ("setup_ms", C.c_double), ("kernel_ms", C.c_double * 16),
                ("kernel_launches", C.c_int64 * 16), ("linearisations", C.c_int32), ("frozen_points", C.c_int32),
                ("reserved", C.c_double * 7)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_
             if k not in ("reserved", "pad", "kernel_ms", "kernel_launches")}
        d["kernel_ms"] = {name: self.kernel_ms[i] for i, name in enumerate(KERNEL_NAMES)}
        d["kernel_launches"] = {name: int(self.kernel_launches[i]) for i, name in enumerate(KERNEL_NAMES)}
        return d


class Ate(C.Structure):
    _fields_ = [("scale", C.c_double), ("rotation", C.c_double * 9), ("translation", C.c_double * 3),
                ("max", C.c_double), ("min", C.c_double), ("mean", C.c_double), ("median", C.c_double),
                ("rmse", C.c_double), ("sse", C.c_double), ("std", C.c_double),
                ("distance_ground_truth", C.c_double), ("distance_estimated", C.c_double),
                ("poses", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pad", "rotation", "translation")}
        d["rotation"] = list(self.rotation)
        d["translation"] = list(self.translation)
        return d


class SynthParams(C.Structure):
    _fields_ = [("layout", C.c_int32), ("n_frames", C.c_int32), ("n_points", C.c_int32),
                ("track_min", C.c_int32), ("track_max", C.c_int32), ("track_geometric", C.c_int32),
                ("track_mean", C.c_double), ("revisit_rate", C.c_double), ("noise_px", C.c_double),
                ("outlier_rate", C.c_double), ("loss_kind", C.c_int32), ("loss_delta", C.c_double),
                ("rot_perturb", C.c_double), ("trans_perturb", C.c_double), ("point_perturb", C.c_double),
                ("n_fixed", C.c_int32), ("K", C.c_double * 4), ("width", C.c_int32), ("height", C.c_int32),
                ("min_parallax", C.c_double), ("reserved", C.c_int32 * 2)]


class BatchInfo(C.Structure):
    _fields_ = [("windows", C.c_int32), ("individual", C.c_int32), ("threads", C.c_int32),
                ("graph_launches", C.c_int32), ("tune_ms", C.c_float * 3), ("pad", C.c_float),
                ("autotune_wall_ms", C.c_double), ("kernel_wall_ms", C.c_double), ("solve_ms", C.c_double),
                ("phase_ms", C.c_double * 8), ("lin_ms", C.c_double), ("steps", C.c_int32),
                ("schur_inflight", C.c_int32), ("reserved", C.c_double * 2)]

    PHASES = ("cost", "linearise", "invert", "schur", "factor", "solve", "refine", "update")

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("reserved", "pad", "tune_ms", "phase_ms")}
        d["tune_ms"] = list(self.tune_ms)
        d["phase_ms"] = {name: self.phase_ms[i] for i, name in enumerate(self.PHASES)}
        return d


class MatchFrame(C.Structure):
    _fields_ = [("rotation", C.c_double * 9), ("translation", C.c_double * 3), ("K", C.c_double * 4),
                ("cols", C.c_int64), ("rows", C.c_int64)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)
_I64 = C.POINTER(C.c_int64)

_SIGNATURES = {
    "zsba_abi_version": (C.c_int32, []),
    "zsba_create": (_P, [C.POINTER(Options)]),
    "zsba_destroy": (None, [_P]),
    "zsba_last_error": (C.c_char_p, [_P]),
    "zsba_add_pose": (C.c_int32, [_P, _D, C.c_int32]),
    "zsba_add_point": (C.c_int32, [_P, _D, C.c_int32]),
    "zsba_add_poses": (C.c_int32, [_P, C.c_int32, _D, _U8]),
    "zsba_add_points": (C.c_int32, [_P, C.c_int32, _D, _U8]),
    "zsba_add_observation": (C.c_int32, [_P, C.c_int32, C.c_int32, _D, _D, _D, C.c_int32, C.c_double]),
    "zsba_add_observations": (C.c_int32, [_P, C.c_int32, _I32, _I32, _D, _D, _D, C.c_int32, C.c_double]),
    "zsba_remove_observation": (C.c_int32, [_P, C.c_int32]),
    "zsba_remove_pose": (C.c_int32, [_P, C.c_int32]),
    "zsba_remove_point": (C.c_int32, [_P, C.c_int32]),
    "zsba_set_pose": (C.c_int32, [_P, C.c_int32, _D]),
    "zsba_set_point": (C.c_int32, [_P, C.c_int32, _D]),
    "zsba_set_pose_fixed": (C.c_int32, [_P, C.c_int32, C.c_int32]),
    "zsba_set_point_fixed": (C.c_int32, [_P, C.c_int32, C.c_int32]),
    "zsba_counts": (C.c_int32, [_P, _I64]),
    "zsba_get_poses": (C.c_int32, [_P, _D]),
    "zsba_get_points": (C.c_int32, [_P, _D]),
    "zsba_solve": (C.c_int32, [_P, C.c_int32, C.c_int32, C.c_int32]),
    "zsba_current_chi": (C.c_double, [_P, C.c_int32]),
    "zsba_get_inlier_flags": (C.c_int32, [_P, _U8]),
    "zsba_reprojection_cull": (C.c_int32, [_P, C.c_float, C.POINTER(C.c_float), _U8]),
    "zsba_get_residuals": (C.c_int32, [_P, _D]),
    "zsba_get_stats": (C.c_int32, [_P, C.POINTER(Stats)]),
    "zsba_snapshot": (C.c_int32, [_P]),
    "zsba_restore": (C.c_int32, [_P]),
    "zsba_solve_batch": (C.c_int32, [C.POINTER(_P), C.c_int32, C.c_int32, C.c_int32, C.c_int32, _I32]),
    "zsba_batch_create": (_P, [C.POINTER(_P), C.c_int32]),
    "zsba_batch_destroy": (None, [_P]),
    "zsba_batch_last_error": (C.c_char_p, [_P]),
    "zsba_batch_solve": (C.c_int32, [_P, C.c_int32, C.c_int32, C.c_int32, _I32]),
    "zsba_batch_get_info": (C.c_int32, [_P, C.POINTER(BatchInfo)]),
    "zsba_adjoint": (C.c_int32, [_P, _D, _D, _D, _D]),
    "zsba_matcher_create": (_P, [C.c_int32]),
    "zsba_matcher_destroy": (None, [_P]),
    "zsba_matcher_last_error": (C.c_char_p, [_P]),
    "zsba_matcher_stats": (C.c_int32, [_P, _D, _I32]),
    "zsba_match_by_projection": (C.c_int32, [_P, C.POINTER(MatchFrame), C.c_int32, _D, _U8, _U8, C.c_int32, _D, _U8,
                                             _I32, C.c_int32, _U8, C.c_int32, _I32, _I32]),
    # device pointers (as integers) for every array, then the stream
    "zsba_match_by_projection_noalloc": (C.c_int32, [C.POINTER(MatchFrame), C.c_int32, _P, _P, _P, C.c_int32, _P, _P,
                                                     _P, C.c_int32, _P, C.c_int32, _P, _P, _P, _P]),
    "zsba_gauss_newton": (C.c_int32, [_P, C.c_int32]),
    "zsba_hold_indefinite_points": (C.c_int32, [_P, C.c_double, C.POINTER(C.c_uint8)]),
    "zsba_debug_adjoint_ops": (C.c_int32, [_P, _D, _D, _D]),
    "zsba_debug_eval_edges": (C.c_int32, [_P, _D]),
    "zsba_debug_reduced_system": (C.c_int32, [_P, C.c_double, C.c_int32, _D, _D, _D, _D]),
    "zsba_debug_math": (C.c_int32, [C.c_int32, C.c_int32, _D, _D, C.c_int32]),
    "zsba_pose_to_frame": (C.c_int32, [_D, _D, _D]),
    "zsba_write_trajectory_tum": (C.c_int32, [C.c_char_p, C.c_int32, _I32, _D, _D]),
    "zsba_write_map_ply": (C.c_int32, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _D, _D, C.c_int32, _I32, _D, _D,
                                       _D]),
    "zsba_trajectory_ate": (C.c_int32, [C.c_int32, _D, _D, C.c_int32, C.POINTER(Ate)]),
    "zsba_synth_default_params": (C.c_int32, [C.c_int32, C.POINTER(SynthParams)]),
    "zsba_synth_create": (_P, [C.POINTER(SynthParams), C.c_uint64]),
    "zsba_synth_destroy": (None, [_P]),
    "zsba_synth_sizes": (None, [_P, _I64]),
    "zsba_synth_copy": (None, [_P, _D, _U8, _D, _U8, _I32, _I32, _D, _D, _D, _U8]),
}

EXPORTED = tuple(_SIGNATURES)

_lib = None


def load() -> C.CDLL:
    """Load the engine library, binding every exported symbol; raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m zeroslam_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (restype, argtypes) in _SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is not exported
        fn.restype = restype
        fn.argtypes = argtypes
    if lib.zsba_abi_version() != 2:
        raise RuntimeError("libzsba.so ABI version mismatch")
    _lib = lib
    return lib


def ptr(a, ctype=C.c_double):
    """ctypes pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
