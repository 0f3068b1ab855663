"""Output formats over the engine's results against the reference's own code,
compiled from its sources by oracle/Makefile (target `output`):

  * zsba_write_trajectory_tum against save_trajectory_as_txt
    (source/main.cpp:132-156): byte-identical files;
  * zsba_write_map_ply against save_trajectory_and_map_as_ply
    (source/main.cpp:158-289): byte-identical files, landmark colours clamped;
  * zsba_trajectory_ate against tools/trajectory_alignment (main.cpp:195-276):
    the printed alignment and statistics agree to their printed digits;
  * zsba_pose_to_frame against the rotation matrix of the quaternion.
No GPU is used.
"""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from tests.map_scenario import REF_DIR, Scenario, from_problem, walk_scenario
from zeroslam_amd import _lib
from zeroslam_amd.problem import synth

OUTPUT_DRIVER = REF_DIR / "output_driver"
ALIGN_TOOL = REF_DIR / "trajectory_alignment"
need_ref = pytest.mark.skipif(not (OUTPUT_DRIVER.exists() and ALIGN_TOOL.exists()),
                              reason="reference output harnesses not built")


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def frames_of(sc: Scenario, ids=None):
    by_id = {fid: (R, t) for fid, R, t, _ in sc.frames}
    ids = [fid for fid, _, _, _ in sc.frames] if ids is None else ids
    R = np.ascontiguousarray([np.asarray(by_id[i][0], np.float64).reshape(9) for i in ids])
    t = np.ascontiguousarray([np.asarray(by_id[i][1], np.float64) for i in ids])
    return np.ascontiguousarray(ids, np.int32), R, t


def reference_outputs(sc, tmp, width=640, height=480):
    src = tmp / "s.scn"
    sc.write(src)
    tum, ply, order = tmp / "ref.tum", tmp / "ref.ply", tmp / "order.bin"
    proc = subprocess.run([str(OUTPUT_DRIVER), str(src), str(tum), str(ply), str(order), str(width), str(height)],
                          capture_output=True, text=True, timeout=60)
    assert proc.returncode == 0, proc.stderr
    raw = np.fromfile(order, np.int32)
    nl = int(raw[0])
    lm_order, nf = raw[1:1 + nl], int(raw[1 + nl])
    fr_order = raw[2 + nl:2 + nl + nf]
    return tum.read_bytes(), ply.read_bytes(), lm_order, fr_order


def scenario_with_pose_frames(seed=7):
    """Frames whose rotation is zsba_pose_to_frame of the problem's poses: what
    map::optimise writes back from the solved vertices."""
    lib = _lib.load()
    pr = synth(1, seed=seed, n_frames=16, n_points=700, track_min=3, track_max=8)
    sc = from_problem(pr, [])
    frames = []
    for (fid, _, _, kps), p in zip(sc.frames, pr.poses):
        R, t = np.zeros(9), np.zeros(3)
        assert lib.zsba_pose_to_frame(dptr(np.ascontiguousarray(p)), dptr(R), dptr(t)) == 0
        frames.append((fid, R.reshape(3, 3), t, kps))
    sc.frames = frames
    return sc, pr


def test_pose_to_frame_is_the_rotation_of_the_quaternion():
    lib = _lib.load()
    rng = np.random.default_rng(1)
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        p = np.concatenate([rng.normal(size=3), q])
        R, t = np.zeros(9), np.zeros(3)
        assert lib.zsba_pose_to_frame(dptr(p), dptr(R), dptr(t)) == 0
        assert np.allclose(R.reshape(3, 3), Rotation.from_quat(q).as_matrix(), atol=1e-15)
        assert np.array_equal(t, p[:3])


@need_ref
@pytest.mark.parametrize("case", ["synthetic", "pose_frames"])
def test_tum_and_ply_bytes_equal_reference(tmp_path, case):
    if case == "synthetic":
        sc = walk_scenario(7)
    else:
        sc, _ = scenario_with_pose_frames(11)
    # a frame with a rotation of trace below zero exercises the other quaternion branch
    fid, R, t, kps = sc.frames[3]
    sc.frames[3] = (fid, Rotation.from_euler("y", 170, degrees=True).as_matrix() @ np.asarray(R), t, kps)
    tum_ref, ply_ref, lm_order, fr_order = reference_outputs(sc, tmp_path)
    lib = _lib.load()
    ids, R, t = frames_of(sc)
    out = tmp_path / "zsba.tum"
    assert lib.zsba_write_trajectory_tum(str(out).encode(), len(ids), iptr(ids), dptr(R), dptr(t)) == 0
    assert out.read_bytes() == tum_ref
    # PLY in the map's iteration order, with the driver's colours
    xyz = {lid: np.asarray(x, np.float64) for lid, x in sc.landmarks}
    lm = np.ascontiguousarray([xyz[i] for i in lm_order])
    col = np.ascontiguousarray([[1.2 * (i % 7) / 6.0 - 0.1, 1.2 * (i % 11) / 10.0 - 0.1, 1.2 * (i % 13) / 12.0 - 0.1]
                                for i in lm_order])
    fids, fR, ft = frames_of(sc, list(fr_order))
    K = np.ascontiguousarray(sc.K, np.float64)
    out = tmp_path / "zsba.ply"
    assert lib.zsba_write_map_ply(str(out).encode(), 640, 480, len(lm), dptr(lm), dptr(col), len(fids), iptr(fids),
                                  dptr(fR), dptr(ft), dptr(K)) == 0
    assert out.read_bytes() == ply_ref


def read_tum_xyz(path):
    return np.loadtxt(path)[:, 1:4]


def ate_of(gt, est, first):
    lib = _lib.load()
    out = _lib.Ate()
    n = min(len(gt), len(est))
    g, e = np.ascontiguousarray(gt[:n]), np.ascontiguousarray(est[:n])
    assert lib.zsba_trajectory_ate(n, dptr(g), dptr(e), int(first), C.byref(out)) == 0
    return out.as_dict()


def parse_tool(text):
    nums = lambda line: [float(x) for x in re.findall(r"[-+]?\d+\.\d+", line)]  # noqa: E731
    lines = text.splitlines()
    get = lambda key: next(nums(l) for l in lines if l.strip().startswith(key))  # noqa: E731
    rot = [nums(l) for l in lines if l.strip().startswith("Rotation:")]
    return {"scale": get("Scale:")[0], "rotation": sum(rot, []), "translation": get("Translation:"),
            "max": get("max:")[0], "min": get("min:")[0], "mean": get("mean:")[0], "median": get("median:")[0],
            "rmse": get("rmse:")[0], "sse": get("sse:")[0], "std": get("std:")[0],
            "distance_ground_truth": get("Ground Truth:")[0], "distance_estimated": get("Estimated:")[0]}


@need_ref
@pytest.mark.parametrize("first", [False, True])
def test_ate_matches_reference_tool(tmp_path, first):
    """Ground truth: the unperturbed poses; estimate: the perturbed ones, rotated,
    scaled and shifted so the alignment has work to do; both written as TUM."""
    lib = _lib.load()
    pr = synth(1, seed=3, n_frames=60, n_points=2000)
    files = []
    for name, poses in (("gt", pr.true_poses), ("est", pr.poses)):
        ids = np.arange(len(poses), dtype=np.int32)
        R, t = np.zeros((len(poses), 9)), np.zeros((len(poses), 3))
        for i, p in enumerate(poses):
            p = np.ascontiguousarray(p, np.float64)
            lib.zsba_pose_to_frame(dptr(p), dptr(R[i]), dptr(t[i]))
        if name == "est":  # world moved: centre c -> 0.7 Q c + d, so t -> 0.7 t - R Q^T d ... via R' = R Q^T
            Q = Rotation.from_euler("xyz", [10, -20, 35], degrees=True).as_matrix()
            d = np.array([1.0, -2.0, 0.5])
            for i in range(len(poses)):
                Ri = R[i].reshape(3, 3)
                c = -Ri.T @ t[i]
                c2 = 0.7 * Q @ c + d
                R2 = Ri @ Q.T
                R[i] = R2.reshape(9)
                t[i] = -R2 @ c2
        path = tmp_path / f"{name}.tum"
        assert lib.zsba_write_trajectory_tum(str(path).encode(), len(ids), iptr(ids), dptr(np.ascontiguousarray(R)),
                                             dptr(np.ascontiguousarray(t))) == 0
        files.append(path)
    cmd = [str(ALIGN_TOOL), str(files[0]), str(files[1])] + (["--first"] if first else [])
    proc = subprocess.run(cmd, capture_output=True, text=True, timeout=60)
    assert proc.returncode == 0, proc.stderr
    ref = parse_tool(proc.stdout)
    ours = ate_of(read_tum_xyz(files[0]), read_tum_xyz(files[1]), first)
    print(proc.stdout)
    for k, v in ref.items():
        a = np.atleast_1d(np.asarray(ours[k], np.float64))
        b = np.atleast_1d(np.asarray(v, np.float64))
        # the tool prints 5 or 6 decimals
        assert np.allclose(a, b, rtol=0, atol=1.2e-5 if k in ("scale", "rotation", "translation") else 1.2e-6), (k, a, b)
    assert 1.0 / 0.7 - 0.01 < ours["scale"] < 1.0 / 0.7 + 0.01
