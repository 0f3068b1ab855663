"""Scenario files for oracle/map_driver.cpp: ZeroSLAM's own map (map.hpp) filled
with frames, keypoints, landmarks and observations, and a list of
map::optimise / map::cull calls.  oracle/_ref/map_ref runs them with the
reference's factor graph, oracle/_ref/map_zsba with the engine's drop-in header
in its place; both write the map's frames and landmarks after every call.
"""
from __future__ import annotations

import struct
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
REF_DIR = ROOT / "oracle" / "_ref"
MAP_REF = REF_DIR / "map_ref"
MAP_ZSBA = REF_DIR / "map_zsba"

OPTIMISE, CULL = 0, 1


@dataclass
class Scenario:
    K: tuple = (525.0, 525.0, 320.0, 240.0)
    frames: list = field(default_factory=list)        # (id, R 3x3, t 3, keypoints n x 2 float32)
    landmarks: list = field(default_factory=list)     # (id, xyz 3)
    observations: list = field(default_factory=list)  # (frame id, landmark id, keypoint index)
    calls: list = field(default_factory=list)         # (op, local_window, fix_landmarks, rounds, use_relative)

    def write(self, path: Path) -> None:
        with open(path, "wb") as f:
            f.write(struct.pack("<4i", len(self.frames), len(self.landmarks), len(self.observations), len(self.calls)))
            f.write(struct.pack("<4d", *self.K))
            for fid, R, t, kps in self.frames:
                kps = np.ascontiguousarray(kps, dtype="<f4").reshape(-1, 2)
                f.write(struct.pack("<i", fid))
                f.write(np.ascontiguousarray(R, dtype="<f8").reshape(9).tobytes())
                f.write(np.ascontiguousarray(t, dtype="<f8").reshape(3).tobytes())
                f.write(struct.pack("<i", len(kps)))
                f.write(kps.tobytes())
            for lid, xyz in self.landmarks:
                f.write(struct.pack("<i", lid))
                f.write(np.ascontiguousarray(xyz, dtype="<f8").reshape(3).tobytes())
            for o in self.observations:
                f.write(struct.pack("<3i", *o))
            for c in self.calls:
                f.write(struct.pack("<5i", *c))


def quat_matrix(q) -> np.ndarray:
    """Rotation matrix of the unit quaternion (x, y, z, w)."""
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def from_problem(pr, calls) -> Scenario:
    """A scenario from a synthetic problem: frame i at pose i (world to camera),
    landmark j at point j, one keypoint (the observed pixel, rounded to binary32
    as feature::point stores it) per observation, observations in problem order."""
    sc = Scenario(K=tuple(float(k) for k in pr.obs_K[0]), calls=list(calls))
    kps = [[] for _ in range(pr.n_poses)]
    for e in range(pr.n_obs):
        p, l = int(pr.obs_pose[e]), int(pr.obs_point[e])
        sc.observations.append((p, l, len(kps[p])))
        kps[p].append(pr.obs_uv[e])
    for i in range(pr.n_poses):
        t, q = pr.poses[i, :3], pr.poses[i, 3:]
        sc.frames.append((i, quat_matrix(q / np.linalg.norm(q)), t, np.array(kps[i], np.float64).reshape(-1, 2)))
    for j in range(pr.n_points):
        sc.landmarks.append((j, pr.points[j]))
    return sc


def walk_scenario(seed: int = 7, n_frames: int = 16, n_points: int = 700, calls=None) -> Scenario:
    """A forward walk (1 px noise, 5% outliers) as a tracker leaves it: the
    local-window BA of slam.hpp:472, culling, the tracking BA of
    slam.hpp:316/387/469, a global BA."""
    from zeroslam_amd.problem import synth
    pr = synth(1, seed=seed, n_frames=n_frames, n_points=n_points, track_min=3, track_max=8)
    if calls is None:
        calls = [(OPTIMISE, 10, 0, 50, 1), (CULL, 0, 0, 0, 0), (OPTIMISE, 1, 1, 50, 0), (OPTIMISE, 0, 0, 10, 0),
                 (CULL, 0, 0, 0, 0)]
    return from_problem(pr, calls)


def tiny_scenario() -> Scenario:
    """tests/map.test.cpp:109-123: two frames, one landmark, optimise(2, false, 3)."""
    zero = np.zeros((1, 2))
    sc = Scenario(K=(1.0, 1.0, 0.5, 0.5), calls=[(OPTIMISE, 2, 0, 3, 0)])
    sc.frames = [(0, np.zeros((3, 3)), np.zeros(3), zero), (1, np.zeros((3, 3)), np.zeros(3), zero)]
    sc.landmarks = [(0, np.zeros(3))]
    sc.observations = [(0, 0, 0), (1, 0, 0)]
    return sc


def run(exe: Path, sc: Scenario, tmp: Path, tag: str, timeout: float = 120.0):
    """Runs a driver on a scenario; returns (stdout, raw output bytes)."""
    src, out = tmp / f"{tag}.scn", tmp / f"{tag}.out"
    sc.write(src)
    proc = subprocess.run([str(exe), str(src), str(out)], capture_output=True, text=True, timeout=timeout)
    assert proc.returncode == 0, f"{exe.name} exit {proc.returncode}: {proc.stdout[-2000:]} {proc.stderr[-2000:]}"
    return proc.stdout, out.read_bytes()


def parse(raw: bytes, n_calls: int):
    """The map after each call: [(frames {id: (R, t)}, landmarks {id: xyz})]."""
    states, off = [], 0
    for _ in range(n_calls):
        (nf,) = struct.unpack_from("<i", raw, off); off += 4
        frames = {}
        for _ in range(nf):
            (fid,) = struct.unpack_from("<i", raw, off); off += 4
            v = np.frombuffer(raw, "<f8", 12, off); off += 96
            frames[fid] = (v[:9].reshape(3, 3), v[9:])
        (nl,) = struct.unpack_from("<i", raw, off); off += 4
        lms = {}
        for _ in range(nl):
            (lid,) = struct.unpack_from("<i", raw, off); off += 4
            lms[lid] = np.frombuffer(raw, "<f8", 3, off); off += 24
        states.append((frames, lms))
    assert off == len(raw)
    return states
