"""The synthetic workloads: deterministic, sized as configured, observations
visible and noisy as specified."""
import numpy as np
import pytest

from zeroslam_amd.problem import synth


def test_deterministic():
    a, b = synth(1, n_frames=20, n_points=300), synth(1, n_frames=20, n_points=300)
    for f in ("poses", "points", "obs_pose", "obs_point", "obs_uv"):
        assert np.array_equal(getattr(a, f), getattr(b, f))
    c = synth(1, seed=99, n_frames=20, n_points=300)
    assert not np.array_equal(a.obs_uv, c.obs_uv)


@pytest.mark.parametrize("cfg,lo,hi", [(0, 3500, 4800), (1, 140_000, 170_000), (3, 35_000, 45_000)])
def test_sizes(cfg, lo, hi):
    pr = synth(cfg)
    assert lo <= pr.n_obs <= hi
    assert pr.pose_fixed[:2].all() and not pr.pose_fixed[2:].any()


def test_tracks_and_noise():
    pr = synth(1)
    counts = np.bincount(pr.obs_point, minlength=pr.n_points)
    assert counts.min() >= 4 and counts.max() <= 12
    # consecutive keyframes per landmark
    for l in range(0, pr.n_points, 997):
        f = pr.obs_pose[pr.obs_point == l]
        assert np.array_equal(f, np.arange(f[0], f[0] + len(f)))
    assert 0.04 < pr.outlier.mean() < 0.06
    # inlier observations sit ~1 px from the true projection
    q = pr.true_poses[pr.obs_pose][:, 3:7]
    from scipy.spatial.transform import Rotation
    Pc = Rotation.from_quat(q).apply(pr.true_points[pr.obs_point]) + pr.true_poses[pr.obs_pose][:, :3]
    uv = 525.0 * Pc[:, :2] / Pc[:, 2:3] + [320.0, 240.0]
    err = (pr.obs_uv - uv)[pr.outlier == 0]
    assert 0.9 < err.std() < 1.1
    assert (uv[:, 0] >= 0).all() and (uv[:, 0] < 640).all() and (uv[:, 1] >= 0).all() and (uv[:, 1] < 480).all()
