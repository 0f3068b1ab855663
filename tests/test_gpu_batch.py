"""Batched local BA: many independent windows solved in lockstep (each phase of an
LM attempt one grid-wide launch over every window, the factorisation one
workgroup per window), against per-graph solves and the CPU oracle."""
import dataclasses

import numpy as np
import pytest

from oracle.oracle import CpuGraph
from tests.helpers import rel_err, state_rel_err
from zeroslam_amd.engine import FP64, FP64_FAST, MIXED, Batch, Graph, solve_batch
from zeroslam_amd.problem import motion_only, synth, synth_window

pytestmark = pytest.mark.gpu


def clean_window(i, frames=None, points=None):
    """A window like the batched workload's, without outliers and robust loss
    (where the LM path does not depend on rounding)."""
    rng = np.random.default_rng(100 + i)
    nf = frames or int(rng.integers(5, 31))
    npt = points or int(rng.integers(200, 1500))
    return synth(2, seed=500 + i, n_frames=nf, n_points=npt, track_min=2, track_max=nf, outlier_rate=0.0,
                 loss_kind=1)


@pytest.mark.parametrize("precision,tol", [(FP64_FAST, 1e-9), (MIXED, 1e-6)])
def test_batch_matches_individual_solves(precision, tol):
    probs = [clean_window(i) for i in range(6)]
    batch_graphs = [Graph.from_problem(p) for p in probs]
    acc = solve_batch(batch_graphs, 8, False, precision)
    for k, p in enumerate(probs):
        g = Graph.from_problem(p)
        a = g.solve(8, False, precision)
        st_b, st = batch_graphs[k].stats(), g.stats()
        err = state_rel_err(batch_graphs[k].poses(), batch_graphs[k].points(), g.poses(), g.points())
        print(f"window {k}: obs {p.n_obs} batch acc {acc[k]} it {st_b['iterations']} chi2 {st_b['final_chi2']!r} | "
              f"single acc {a} chi2 {st['final_chi2']!r} | state {err:.2e}", flush=True)
        assert acc[k] == a
        assert st_b["linearisations"] == st["linearisations"] >= 1
        assert abs(st_b["final_chi2"] - st["final_chi2"]) <= tol * st["final_chi2"]
        assert err <= tol


def with_repeated_pairs(p, seed, frac=0.05):
    """The window with a fraction of its observations repeated (same pose and
    point, the pixel moved by up to 0.5 px), appended after the originals."""
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(p.n_obs, size=max(1, int(p.n_obs * frac)), replace=False))
    extra = {"obs_pose": p.obs_pose[idx], "obs_point": p.obs_point[idx],
             "obs_uv": p.obs_uv[idx] + rng.uniform(-0.5, 0.5, (len(idx), 2)), "obs_K": p.obs_K[idx],
             "obs_loss": p.obs_loss[idx], "obs_delta": p.obs_delta[idx]}
    fields = {k: np.concatenate([getattr(p, k), v]) for k, v in extra.items()}
    if p.outlier is not None:
        fields["outlier"] = np.concatenate([p.outlier, p.outlier[idx]])
    return dataclasses.replace(p, **fields)


def test_batch_repeated_pairs_match_individual_solves():
    """Windows in which some (pose, point) pairs are observed twice: their W blocks
    accumulate in observation order, in the batched linearisation as in the
    per-graph solve, bit for bit."""
    probs = [with_repeated_pairs(clean_window(i), 40 + i) for i in range(4)]
    batch_graphs = [Graph.from_problem(p) for p in probs]
    acc = solve_batch(batch_graphs, 6, False, MIXED)
    for k, p in enumerate(probs):
        g = Graph.from_problem(p)
        a = g.solve(6, False, MIXED)
        print(f"window {k}: obs {p.n_obs} batch acc {acc[k]} single acc {a}", flush=True)
        assert acc[k] == a
        assert batch_graphs[k].stats()["final_chi2"] == g.stats()["final_chi2"]
        assert np.array_equal(batch_graphs[k].poses(), g.poses())
        assert np.array_equal(batch_graphs[k].points(), g.points())


def with_long_track(p, seed, count=140):
    """The window with one point observed `count` more times by the poses that
    already see it (pixels moved by up to 0.5 px): more observations than a
    linearisation block holds, so the window is linearised per point."""
    rng = np.random.default_rng(seed)
    first = np.flatnonzero(p.obs_point == p.obs_point[0])
    idx = rng.choice(first, size=count, replace=True)
    extra = {"obs_pose": p.obs_pose[idx], "obs_point": p.obs_point[idx],
             "obs_uv": p.obs_uv[idx] + rng.uniform(-0.5, 0.5, (count, 2)), "obs_K": p.obs_K[idx],
             "obs_loss": p.obs_loss[idx], "obs_delta": p.obs_delta[idx]}
    fields = {k: np.concatenate([getattr(p, k), v]) for k, v in extra.items()}
    if p.outlier is not None:
        fields["outlier"] = np.concatenate([p.outlier, p.outlier[idx]])
    return dataclasses.replace(p, **fields)


def test_batch_block_and_per_point_windows_match_individual_solves():
    """One batch holding a window linearised per point (a track longer than a
    linearisation block) and windows linearised by blocks: every window equals its
    per-graph mixed solve bit for bit."""
    probs = [with_long_track(clean_window(0), 7), clean_window(1), with_repeated_pairs(clean_window(2), 9)]
    batch_graphs = [Graph.from_problem(p) for p in probs]
    acc = solve_batch(batch_graphs, 6, False, MIXED)
    for k, p in enumerate(probs):
        g = Graph.from_problem(p)
        a = g.solve(6, False, MIXED)
        print(f"window {k}: obs {p.n_obs} batch acc {acc[k]} single acc {a}", flush=True)
        assert acc[k] == a
        assert batch_graphs[k].stats()["final_chi2"] == g.stats()["final_chi2"]
        assert np.array_equal(batch_graphs[k].poses(), g.poses())
        assert np.array_equal(batch_graphs[k].points(), g.points())


def with_closing_edges(p, a=2, b=None, count=20):
    """The window with pose a also observing `count` of the points pose b sees
    (same pixels), so the reduced system couples the first and last poses."""
    b = p.n_poses - 1 if b is None else b
    idx = np.flatnonzero(p.obs_pose == b)[:count]
    extra = {"obs_pose": np.full(len(idx), a, np.int32), "obs_point": p.obs_point[idx], "obs_uv": p.obs_uv[idx],
             "obs_K": p.obs_K[idx], "obs_loss": p.obs_loss[idx], "obs_delta": p.obs_delta[idx]}
    fields = {k: np.concatenate([getattr(p, k), v]) for k, v in extra.items()}
    if p.outlier is not None:
        fields["outlier"] = np.concatenate([p.outlier, np.zeros(len(idx), p.outlier.dtype)])
    return dataclasses.replace(p, **fields)


def test_batch_tall_panels_skip_small_workgroups():
    """Windows of 40 poses whose first free pose shares points with the last one
    (8-tile panels, 224 rows below the first diagonal tile) need more than 256
    threads in the factorisation: the autotuner does not consider 256, and the
    results equal the per-graph solves."""
    probs = [with_closing_edges(synth(4, seed=60 + i, n_frames=40, n_points=500)) for i in range(3)]
    graphs = [Graph.from_problem(p) for p in probs]
    b = Batch(graphs)
    acc = b.solve(5, False, MIXED)
    info = b.info()
    print(f"tall panels: {info}", flush=True)
    assert info["windows"] == 3 and info["threads"] >= 512 and info["tune_ms"][0] == 0.0
    for k, p in enumerate(probs):
        g = Graph.from_problem(p)
        assert acc[k] == g.solve(5, False, MIXED)
        assert np.array_equal(graphs[k].poses(), g.poses())


def test_batch_fp64_fast_matches_oracle():
    probs = [clean_window(i) for i in range(4)]
    graphs = [Graph.from_problem(p) for p in probs]
    acc = solve_batch(graphs, 10, False, FP64_FAST)
    for k, p in enumerate(probs):
        o = CpuGraph(p, "oracle")
        a = o.solve(10)
        Po, Xo = o.state()
        chi, chi_o = graphs[k].current_chi(True), o.current_chi(True)
        assert acc[k] == a
        assert abs(chi - chi_o) <= 1e-9 * chi_o
        assert state_rel_err(graphs[k].poses(), graphs[k].points(), Po, Xo) <= 1e-9


def test_batch_mixed_on_batched_workload_windows():
    """Windows of the batched workload (5% outliers, Huber): the mixed batch
    reaches the oracle's 10-iteration cost within 1e-6 on every window, after its
    own 10 iterations or one more mixed solve of 10 on the same handle.  The
    windows are far from converged at 10 iterations (the oracle's 30-iteration
    cost is 2e-3 to 2e-2 lower), so the cost at equal iterations depends on the
    rounding of the LM path and is printed, not bounded."""
    probs = [synth_window(i) for i in range(6)]
    graphs = [Graph.from_problem(p) for p in probs]
    solve_batch(graphs, 10, False, MIXED)
    for k, p in enumerate(probs):
        o = CpuGraph(p, "oracle")
        o.solve(10)
        chi_o = o.current_chi(True)
        rel10 = rel = (graphs[k].current_chi(True) - chi_o) / chi_o
        extra = 0
        if rel > 1e-6:
            graphs[k].solve(10, False, MIXED)
            extra = graphs[k].stats()["iterations"]
            rel = (graphs[k].current_chi(True) - chi_o) / chi_o
        print(f"window {k}: {p.n_poses} KF {p.n_points} pts {p.n_obs} obs: chi2 rel vs oracle {rel10:+.3e} at 10 "
              f"iterations, {rel:+.3e} after {extra} more", flush=True)
        assert rel <= 1e-6


def test_batch_replays_and_tunes():
    probs = [clean_window(i, frames=8, points=300) for i in range(5)]
    graphs = [Graph.from_problem(p) for p in probs]
    for g in graphs:
        g.solve(0)
        g.snapshot()
    b = Batch(graphs)
    acc1 = b.solve(6, False, MIXED)
    info1 = b.info()
    P1 = [g.poses() for g in graphs]
    for g in graphs:
        g.restore()
    acc2 = b.solve(6, False, MIXED)
    info2 = b.info()
    print(f"batch info first {info1} second {info2}", flush=True)
    assert np.array_equal(acc1, acc2)
    for g, P in zip(graphs, P1):
        assert np.array_equal(g.poses(), P)
    assert info1["threads"] in (256, 512, 1024) and info2["threads"] == info1["threads"]
    assert info1["schur_inflight"] in (1, 2) and info2["schur_inflight"] == info1["schur_inflight"]
    # the first launch holds 6 steps; each later one 2, checked one launch behind
    assert info1["windows"] == 5 and 2 <= info1["graph_launches"] == (info1["steps"] - 6) // 2 + 1
    assert info1["steps"] >= 6 and info2["steps"] == info1["steps"]
    assert info2["autotune_wall_ms"] == 0.0


def test_batch_fallbacks():
    """A window too large for one workgroup, an empty graph and the reference-order
    FP64 precision are solved one by one, with the same results as zsba_solve."""
    big = synth(1, n_frames=50, n_points=800)  # 300-row reduced system
    small = clean_window(1, frames=6, points=250)
    graphs = [Graph.from_problem(big), Graph(), Graph.from_problem(small)]
    b = Batch(graphs)
    acc = b.solve(5, False, MIXED)
    info = b.info()
    assert info["windows"] == 1 and info["individual"] == 1 and acc[1] == 0
    ref = Graph.from_problem(big)
    assert acc[0] == ref.solve(5, False, MIXED)
    assert np.array_equal(graphs[0].poses(), ref.poses())
    # reference order: every graph individually, bit for bit as zsba_solve
    graphs = [Graph.from_problem(small), Graph.from_problem(big)]
    acc = solve_batch(graphs, 4, False, FP64)
    for g, p in zip(graphs, (small, big)):
        r = Graph.from_problem(p)
        r.solve(4, False, FP64)
        assert np.array_equal(g.points(), r.points())


@pytest.mark.parametrize("precision,tol", [(FP64_FAST, 1e-9), (MIXED, 1e-6)])
def test_batch_motion_only_frames(precision, tol):
    """Tracking-step BAs of many frames at once: each window one free frame and
    its fixed landmarks (a 6x6 system), without outliers or robust loss; the
    batch against the per-graph solves, and FP64_FAST against the oracle."""
    probs = [motion_only(synth(1, seed=70 + i, n_frames=12, n_points=600, outlier_rate=0.0, loss_kind=1), 1)
             for i in range(16)]
    graphs = [Graph.from_problem(p) for p in probs]
    acc = solve_batch(graphs, 50, False, precision)
    for k, p in enumerate(probs):
        g = Graph.from_problem(p)
        a = g.solve(50, False, precision)
        chi_b, chi = graphs[k].stats()["final_chi2"], g.stats()["final_chi2"]
        print(f"frame {k}: obs {p.n_obs} batch acc {acc[k]} chi2 {chi_b!r} | single acc {a} chi2 {chi!r}", flush=True)
        assert acc[k] == a
        assert abs(chi_b - chi) <= tol * chi
        assert rel_err(graphs[k].poses(), g.poses()) <= tol
        if precision == FP64_FAST:
            # run to convergence, the last accepted steps change the cost at the
            # rounding floor: whether one more is taken may differ by one
            o = CpuGraph(p, "oracle")
            assert abs(int(acc[k]) - o.solve(50)) <= 1
            assert abs(chi_b - o.current_chi(True)) <= tol * chi_b
            assert rel_err(graphs[k].poses(), o.state()[0]) <= tol
