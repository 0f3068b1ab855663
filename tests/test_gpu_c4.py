"""The 1000-keyframe loop-closure workload at its full size (200k points, ~2M
observations, Huber 2 px with 5% outliers, a 6000-row reduced system whose
envelope is dense) against the oracle, which is bitwise the reference
(tests/golden/make_large.py writes the fixture on the CPU):

  * FP64 (the reference's operation order) after solve(1) and solve(2) on a
    fresh graph: accepted steps, attempts, chi2, every pose and every point
    (by SHA-256) identical;
  * MIXED after solve(15): at or below (1 + 1e-6) x the reference's cost after
    solve(15), with the pose error printed.
"""
import hashlib
from pathlib import Path

import numpy as np
import pytest

from tests.golden.make_large import problem_digest
from zeroslam_amd.engine import FP64, MIXED, Graph
from zeroslam_amd.problem import synth

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "c4_full.npz"


@pytest.fixture(scope="module")
def c4():
    with np.load(FIXTURE, allow_pickle=False) as z:
        fx = {k: z[k] for k in z.files}
    pr = synth(4)
    assert problem_digest(pr) == str(fx["digest"][0]), "the workload generator made a different problem"
    print(f"c4: poses {pr.n_poses} points {pr.n_points} obs {pr.n_obs}", flush=True)
    return pr, fx


@pytest.mark.parametrize("k", [1, 2])
def test_exact_full_c4(c4, k):
    pr, fx = c4
    g = Graph.from_problem(pr)
    acc = g.solve(k, False, FP64)
    st = g.stats()
    P, X = g.poses(), g.points()
    chi = g.current_chi(True)
    print(f"c4 FP64 solve({k}): accepted {acc} attempts {st['attempts']} chi2 {chi!r} "
          f"(oracle {float(fx[f'chi2_{k}'][0])!r}) solve {st['solve_ms']:.1f} ms", flush=True)
    assert acc == int(fx[f"accepted_{k}"][0]) and st["attempts"] == int(fx[f"attempts_{k}"][0])
    assert np.array_equal(np.float64(chi).view(np.uint64), fx[f"chi2_{k}"].view(np.uint64)[0])
    assert np.array_equal(P.view(np.uint64), fx[f"poses_{k}"].view(np.uint64))
    assert np.array_equal(X[fx["sample"]].view(np.uint64), fx[f"points_sample_{k}"].view(np.uint64))
    assert hashlib.sha256(np.ascontiguousarray(X).tobytes()).hexdigest() == str(fx[f"points_sha256_{k}"][0])


def test_mixed_full_c4_reaches_reference_cost(c4):
    pr, fx = c4
    if "chi2_15" not in fx:
        pytest.skip("fixture without solve(15)")
    g = Graph.from_problem(pr)
    g.solve(15, False, MIXED)
    st = g.stats()
    chi = g.current_chi(True)
    ref = float(fx["chi2_15"][0])
    P = g.poses()
    err = float(np.abs(P - fx["poses_15"]).max() / np.abs(fx["poses_15"]).max())
    print(f"c4 mixed solve(15): it {st['iterations']} att {st['attempts']} frozen {st['frozen_points']} chi2 {chi!r} "
          f"reference {ref!r} rel {(chi - ref) / ref:+.3e} pose rel err {err:.3e} solve {st['solve_ms']:.1f} ms",
          flush=True)
    assert chi <= ref * (1 + 1e-6)
