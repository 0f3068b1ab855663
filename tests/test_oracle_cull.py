"""The oracle's restatement of map::cull's reprojection test (map.hpp:287-312),
pinned by answers the arithmetic fixes: exact 3-4-5 pixel offsets, a point
behind the camera (5.991 counted, and the mean 5.991 erases: the comparison
includes equality), a point without observations, binary32 accumulation (a sum
whose binary64 value differs) and the rotation taken as the matrix of a
non-unit quaternion (lie::so3::get_matrix), which projects differently from the
quaternion sandwich the solver uses."""
import numpy as np

from oracle.oracle import CpuGraph
from zeroslam_amd.problem import Problem

K = [525.0, 525.0, 320.0, 240.0]
IDENTITY = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0]  # [t, q_xyz, q_w]


def problem(poses, points, obs):
    """obs: (pose, point, u, v) in insertion order."""
    n = len(obs)
    return Problem(poses=np.array(poses, dtype=np.float64), pose_fixed=np.zeros(len(poses), np.uint8),
                   points=np.array(points, dtype=np.float64), point_fixed=np.zeros(len(points), np.uint8),
                   obs_pose=np.array([o[0] for o in obs], np.int32), obs_point=np.array([o[1] for o in obs], np.int32),
                   obs_uv=np.array([[o[2], o[3]] for o in obs], dtype=np.float64),
                   obs_K=np.tile(np.array(K), (n, 1)), obs_loss=np.zeros(n, np.int32), obs_delta=np.zeros(n))


def cull(pr, threshold=5.991):
    return CpuGraph(pr, "oracle").reprojection_cull(threshold)


def test_exact_offsets_and_threshold():
    # point (0, 0, 2) projects to (320, 240) from the identity pose
    pr = problem([IDENTITY], [[0, 0, 2], [0, 0, 2], [0, 0, 2]],
                 [(0, 0, 323, 244), (0, 0, 316, 237),      # errors 5 and 5: mean 5, kept
                  (0, 1, 326, 240),                        # error 6: erased
                  (0, 2, 320, 240), (0, 2, 320, 240)])     # error 0: kept
    mean, erase = cull(pr)
    assert mean.dtype == np.float32
    assert mean.tolist() == [5.0, 6.0, 0.0]
    assert erase.tolist() == [0, 1, 0]


def test_behind_camera_counts_the_threshold_and_equality_erases():
    pr = problem([IDENTITY], [[0, 0, -1], [0, 0, 2]],
                 [(0, 0, 320, 240),                        # behind: 5.991, mean 5.991 >= 5.991
                  (0, 1, 320, 240), (0, 1, 320, 240), (0, 1, 320, 240), (0, 1, 320, 240)])
    mean, erase = cull(pr)
    assert mean[0] == np.float32(5.991) and erase[0] == 1
    assert mean[1] == 0.0 and erase[1] == 0
    # a different threshold is both the behind-camera count and the bound
    mean, erase = cull(pr, threshold=7.5)
    assert mean[0] == np.float32(7.5) and erase[0] == 1


def test_point_without_observations():
    pr = problem([IDENTITY], [[0, 0, 2], [1, 1, 3]], [(0, 0, 320, 240)])
    mean, erase = cull(pr)
    assert mean[0] == 0.0 and erase[0] == 0
    assert np.isinf(mean[1]) and erase[1] == 1


def test_binary32_accumulation():
    # errors 2^24, 1, 1 in insertion order: binary32 sums 2^24 + 1 -> 2^24, twice
    big = float(2 ** 24)
    pr = problem([IDENTITY], [[0, 0, 2]], [(0, 0, 320 + big, 240), (0, 0, 321, 240), (0, 0, 319, 240)])
    mean, _ = cull(pr)
    expect = np.float32(np.float32(np.float32(big) + np.float32(1)) + np.float32(1)) / np.float32(3)
    assert mean[0] == expect
    assert float(mean[0]) != np.float32((big + 2.0) / 3.0)


def test_rotation_is_the_quaternion_matrix():
    # q = (w 2, x 0, y 0, z 0) is not unit: so3::get_matrix gives I, so the point
    # (1, 0, 2) with t = (0, 0, 1) maps to (1, 0, 3) -> u = 525 / 3 + 320 = 495; the
    # sandwich q X q* + t would map it to (4, 0, 9) -> u = 553.3
    pose = [0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 2.0]
    pr = problem([pose], [[1, 0, 2]], [(0, 0, 495, 240)])
    mean, erase = cull(pr)
    assert mean[0] == 0.0 and erase[0] == 0
