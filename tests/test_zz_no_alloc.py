"""zsba_match_by_projection_noalloc, the projection search with device inputs
and outputs that allocates nothing (its intermediate buffers are a static
device array per GPU), called through the C ABI against
zsba_match_by_projection on the same scenes: identical matches and counts, on
two calls in a row (the second with a smaller input, so stale scratch from the
first would show), and the error codes for inputs it does not take."""
import ctypes as C

import numpy as np
import pytest

from zeroslam_amd import _lib
from zeroslam_amd.engine import Matcher
from zeroslam_amd.problem import match_scene

pytestmark = pytest.mark.gpu

EUNSUP, EINVAL = -5, -1


def frame_of(scene):
    f = _lib.MatchFrame()
    f.rotation[:] = [float(x) for x in np.asarray(scene.rotation, np.float64).reshape(9)]
    f.translation[:] = [float(x) for x in np.asarray(scene.translation, np.float64)]
    f.K[:] = [float(x) for x in np.asarray(scene.K, np.float64)]
    f.cols, f.rows = int(scene.cols), int(scene.rows)
    return f


def noalloc(scene, max_matches=None, keys=None):
    """Device copies of the scene (torch tensors as plain device memory), the
    no-alloc search on the null stream, the outputs read back."""
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")

    def d(a, dtype):
        return torch.from_numpy(np.ascontiguousarray(a, dtype)).to(dev)

    key = scene.keypoint_key if keys is None else keys
    xyz, lmd, obs = d(scene.landmark_xyz, np.float64), d(scene.landmark_descriptor, np.uint8), \
        d(scene.landmark_observed, np.uint8)
    kxy, kd, kk, tk = d(scene.keypoint_xy, np.float64), d(scene.keypoint_descriptor, np.uint8), d(key, np.int32), \
        d(scene.key_taken, np.uint8)
    cap = max(1, len(scene.keypoint_xy) if max_matches is None else max_matches)
    out_lm = torch.full((cap,), -7, dtype=torch.int32, device=dev)
    out_kp = torch.full((cap,), -7, dtype=torch.int32, device=dev)
    count = torch.full((1,), -99, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    lib = _lib.load()
    f = frame_of(scene)
    rc = lib.zsba_match_by_projection_noalloc(
        C.byref(f), len(scene.landmark_xyz), xyz.data_ptr(), lmd.data_ptr(), obs.data_ptr(), len(scene.keypoint_xy),
        kxy.data_ptr(), kd.data_ptr(), kk.data_ptr(), len(scene.key_taken), tk.data_ptr(), cap, out_lm.data_ptr(),
        out_kp.data_ptr(), count.data_ptr(), None)
    torch.cuda.synchronize()
    n = int(count.cpu()[0])
    k = max(0, min(n, cap))
    return rc, n, out_lm.cpu().numpy()[:k], out_kp.cpu().numpy()[:k]


@pytest.mark.parametrize("sizes", [((20000, 1500), (3000, 400)), ((200000, 2000), (50000, 900))])
def test_noalloc_equals_matcher_twice(sizes):
    m = Matcher()
    for (L, N), seed in zip(sizes, (31, 32)):
        scene = match_scene(L, N, seed=seed)
        print(f"no-alloc search: {L} landmarks x {N} keypoints, seed {seed}", flush=True)
        lm, kp, n = m.match(scene)
        rc, n2, lm2, kp2 = noalloc(scene)
        assert rc == 0
        assert n2 == n and n > 0
        assert np.array_equal(lm2, lm) and np.array_equal(kp2, kp)
        # fewer outputs than matches: the first ones, the full count
        rc, n3, lm3, kp3 = noalloc(scene, max_matches=max(1, n // 3))
        assert rc == 0 and n3 == n
        assert np.array_equal(lm3, lm[:len(lm3)]) and np.array_equal(kp3, kp[:len(kp3)])


def test_noalloc_limits_and_errors():
    scene = match_scene(2000, 300, seed=5)
    # a previous-frame index out of range: reported on the device, no read outside the arrays
    bad = np.array(scene.keypoint_key, np.int32).copy()
    bad[7] = len(scene.key_taken) + 3
    rc, n, _, _ = noalloc(scene, keys=bad)
    assert rc == 0 and n == EINVAL
    # sizes beyond the scratch: rejected before anything is enqueued
    lib = _lib.load()
    f = frame_of(scene)
    rc = lib.zsba_match_by_projection_noalloc(C.byref(f), 1 << 19, 1, 1, 1, 10, 1, 1, 1, 10, 1, 1, 1, 1, 1, None)
    assert rc == EUNSUP
    rc = lib.zsba_match_by_projection_noalloc(C.byref(f), 10, 1, 1, 1, 10, 1, 1, 1, 0, 1, 1, 1, 1, 1, None)
    assert rc == EINVAL
    # the same scene again afterwards still matches the handle-based search
    lm, kp, n = Matcher().match(scene)
    rc, n2, lm2, kp2 = noalloc(scene)
    assert rc == 0 and n2 == n and np.array_equal(lm2, lm) and np.array_equal(kp2, kp)
