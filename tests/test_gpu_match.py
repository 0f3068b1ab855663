"""The projection search on the device (zsba_match_by_projection) against the
oracle: the same matches in the same order, on the known-answer scenes and on
synthetic frames from a few hundred to 200k landmarks."""
import numpy as np
import pytest

from oracle.oracle import match_by_projection
from tests import test_oracle_match as kat
from zeroslam_amd.engine import Matcher, ZsbaError
from zeroslam_amd.problem import match_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def matcher():
    m = Matcher()
    yield m
    m.close()


def same(m, sc, **kw):
    lm, kp, n = m.match(sc, **kw)
    lm_o, kp_o, n_o = match_by_projection(sc, **kw)
    assert n == n_o
    assert np.array_equal(lm, lm_o) and np.array_equal(kp, kp_o)
    return n


def test_known_answer_scenes(matcher):
    rng = np.random.default_rng(7)
    d = rng.integers(0, 256, 32, dtype=np.uint8)
    scenes = [
        kat.scene([kat.at_pixel(100, 100)], [[102, 100], [100, 102.125], [98, 100], [100, 98]],
                  [d], [d, d, kat.desc_with_distance(d, 63, rng), kat.desc_with_distance(d, 64, rng)], [0, 1, 2, 3]),
        kat.scene([[0, 0, -1.0], [0, 0, 0.0], kat.at_pixel(kat.CX, kat.CY), kat.at_pixel(kat.CX, kat.CY)],
                  [[kat.CX, kat.CY]], [d] * 4, [d], [0], observed=[2]),
        kat.scene([kat.at_pixel(200, 200), kat.at_pixel(203.5, 200)], [[200, 200], [202, 200]], [d, d], [d, d], [3, 3]),
        kat.scene([[0, 0, 1]], [[512, kat.CY], [kat.CX + 64, kat.CY]], [d], [d, d], [0, 1],
                  R=[[1, 0, 0.5], [0, 1, 0], [0, 0, 1]], t=[0.25, 0, 1]),
    ]
    for u, v in [(10, 100), (9.875, 100), (629, 100), (629.125, 100), (100, 10), (100, 9.875), (100, 469), (100, 469.125)]:
        scenes.append(kat.scene([kat.at_pixel(u, v)], [[u, v]], [d], [d], [0]))
    expect = [2, 1, 1, 1, 1, 0, 1, 0, 1, 0, 1, 0]
    got = [same(matcher, sc) for sc in scenes]
    assert got == expect


@pytest.mark.parametrize("L,N,seed", [(1, 1, 0), (300, 200, 1), (3000, 1500, 2), (20000, 2000, 3), (200000, 2000, 4)])
def test_synthetic_frames(matcher, L, N, seed):
    sc = match_scene(L, N, seed=seed)
    print(f"projection search L={L} N={N} seed={seed}", flush=True)
    n = same(matcher, sc)
    st = matcher.stats()
    print(f"  matches {n} candidates {st['candidates']} device {st['device_ms']:.3f} ms", flush=True)
    assert st["candidates"] >= n
    if L >= 3000:
        assert n > 0 and st["candidates"] > n  # taken indices remove candidates


def test_truncation_reuse_and_errors(matcher):
    big = match_scene(5000, 1500, seed=5)
    small = match_scene(50, 40, seed=6)
    n = same(matcher, big)
    assert same(matcher, big, max_matches=7) == n
    same(matcher, small)  # smaller after larger: buffers reused
    same(matcher, big)
    empty = match_scene(10, 10, seed=7)
    empty.landmark_xyz = empty.landmark_xyz[:0]
    empty.landmark_descriptor = empty.landmark_descriptor[:0]
    empty.landmark_observed = empty.landmark_observed[:0]
    assert matcher.match(empty)[2] == 0
    bad = match_scene(10, 10, seed=8)
    bad.keypoint_key = bad.keypoint_key.copy()
    bad.keypoint_key[3] = len(bad.key_taken)
    with pytest.raises(ZsbaError):
        matcher.match(bad)


def test_banded_search_edge_inputs(matcher):
    """Keypoints off the image (clamped to the first or last band), with infinite
    or NaN coordinates (tested against every landmark), and landmarks at the
    border rows."""
    sc = match_scene(20000, 2000, seed=9)
    rng = np.random.default_rng(9)
    xy = sc.keypoint_xy.copy()
    idx = rng.choice(len(xy), 60, replace=False)
    xy[idx[:10], 1] = -rng.uniform(0, 50, 10)
    xy[idx[10:20], 1] = sc.rows + rng.uniform(0, 50, 10)
    xy[idx[20:25], 0] = np.inf
    xy[idx[25:30], 1] = -np.inf
    xy[idx[30:35], 0] = np.nan
    xy[idx[35:40], 1] = np.nan
    # keypoints 1 px outside the border rows next to landmarks on them
    lm = sc.landmark_xyz.copy()
    for r, (v_lm, v_kp) in enumerate([(10.0, 9.0), (sc.rows - 11.0, sc.rows - 10.0)]):
        lm[r] = kat.at_pixel(300.0, v_lm, 5.0)
        xy[idx[40 + r]] = [300.0, v_kp]
        sc.keypoint_descriptor[idx[40 + r]] = sc.landmark_descriptor[r]
    sc.keypoint_xy, sc.landmark_xyz = xy, lm
    sc.landmark_observed = sc.landmark_observed.copy()
    sc.landmark_observed[:2] = 0
    sc.key_taken = sc.key_taken.copy()
    sc.key_taken[sc.keypoint_key[idx[40:42]]] = 0
    sc.rotation, sc.translation, sc.K = np.eye(3), np.zeros(3), np.array([kat.FX, kat.FX, kat.CX, kat.CY])
    n = same(matcher, sc)
    print(f"edge inputs: {n} matches", flush=True)
    assert n > 0


def test_two_pass_fallbacks(matcher):
    """Inputs the banded search leaves to the ordered two-pass search: a landmark
    whose projection is NaN, more candidates than one workgroup sorts, a frame
    taller than the band table."""
    sc = match_scene(3000, 1500, seed=10)
    sc.landmark_xyz = sc.landmark_xyz.copy()
    sc.landmark_xyz[5] = [np.nan, 0.0, 1.0]
    sc.landmark_observed = sc.landmark_observed.copy()
    sc.landmark_observed[5] = 0
    assert same(matcher, sc) > 0
    # 5000 landmarks on a 5-px grid, each with its own keypoint and descriptor
    rng = np.random.default_rng(11)
    gu, gv = np.meshgrid(np.arange(20.0, 620.0, 5.0), np.arange(20.0, 460.0, 5.0))
    uv = np.stack([gu.ravel(), gv.ravel()], 1)[:5000]
    desc = rng.integers(0, 256, (5000, 32), dtype=np.uint8)
    dense = kat.scene([kat.at_pixel(u, v) for u, v in uv], uv, desc, desc, list(range(5000)))
    assert same(matcher, dense) == 5000
    assert matcher.stats()["candidates"] == 5000
    tall = match_scene(3000, 1500, seed=12, rows=40000)
    same(matcher, tall)
