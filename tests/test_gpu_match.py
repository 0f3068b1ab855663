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
