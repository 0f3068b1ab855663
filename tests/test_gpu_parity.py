"""Parity of the HIP engine (through the C ABI) against the CPU oracle and the
golden fixtures of the compiled reference.

FP64 (exact) solves must reproduce the reference bit for bit: accepted-step
count, LM path (iterations, attempts), final chi2, every pose and point
parameter, and the inlier flags.  FP64_FAST differs from it only by the rounding
of parallel reductions; on problems without robust losses (where the LM path is
not sensitive to rounding) it must match within 1e-9 relative.
"""
import numpy as np
import pytest

from oracle.oracle import CpuGraph
from tests.golden.make_golden import CASES, load
from tests.helpers import rel_err, state_rel_err
from zeroslam_amd.engine import FP64, FP64_FAST, Graph, math_probe
from zeroslam_amd.problem import motion_only, synth

pytestmark = pytest.mark.gpu

TOL = 1e-9


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_math_probe_device_equals_host():
    rng = np.random.default_rng(0)
    x = np.concatenate([np.sign(rng.standard_normal(50000)) * 10 ** rng.uniform(-20, 1.3, 50000),
                        rng.uniform(-1, 1, 50000)])
    for fn in (0, 1, 2, 3):
        arg = np.abs(x) if fn == 3 else x
        assert np.array_equal(bits(math_probe(fn, arg, True)), bits(math_probe(fn, arg, False))), fn


@pytest.mark.parametrize("cfg,kw", [(0, {}), (1, dict(n_frames=20, n_points=1500))])
def test_edge_eval_bitwise(cfg, kw):
    pr = synth(cfg, **kw)
    print(f"edge eval cfg={cfg} {kw} obs={pr.n_obs}", flush=True)
    g = Graph.from_problem(pr)
    gpu = g.eval_edges()
    ref = CpuGraph(pr, "oracle").edge_eval()
    mism = np.argwhere(bits(gpu) != bits(ref))
    assert mism.size == 0, f"{len(mism)} mismatching values, first {mism[:5]}"


def test_reduced_system_matches_oracle():
    pr = synth(1, n_frames=20, n_points=1500)
    g = Graph.from_problem(pr)
    lam = 1e-3
    ok, S, bs, xp, xl = g.reduced_system(lam)
    ok_o, S_o, bs_o, xp_o, xl_o = CpuGraph(pr, "oracle").reduced_system(lam)
    assert ok and ok_o
    low = np.tril_indices(S.shape[0])
    assert rel_err(S[low], S_o[low]) < 1e-12
    assert rel_err(bs, bs_o) < 1e-12
    assert rel_err(xp, xp_o) < 1e-9
    assert rel_err(xl, xl_o) < 1e-9


def exact_compare(pr, iters, use_rel=False, **graph_kw):
    g = Graph.from_problem(pr, **graph_kw)
    acc = g.solve(iters, use_rel, FP64)
    st = g.stats()
    o = CpuGraph(pr, "oracle")
    acc_o = o.solve(iters, use_rel)
    so = o.stats()
    P, X = g.poses(), g.points()
    Po, Xo = o.state()
    chi, chi_o = g.current_chi(True), o.current_chi(True)
    print(f"gpu acc={acc} it={st['iterations']} att={st['attempts']} chi={chi!r} | oracle acc={acc_o} "
          f"it={so['iterations']} att={so['attempts']} chi={chi_o!r} | max rel state diff "
          f"{state_rel_err(P, X, Po, Xo):.3e} solve {st['solve_ms']:.2f} ms", flush=True)
    assert (acc, st["iterations"], st["attempts"], st["rejected"], st["solve_failures"]) == \
        (acc_o, so["iterations"], so["attempts"], so["rejected"], so["solve_failures"])
    assert bits(chi) == bits(chi_o)
    assert np.array_equal(bits(P), bits(Po)) and np.array_equal(bits(X), bits(Xo))
    assert bits(g.current_chi(False)) == bits(o.current_chi(False))
    return g, o


@pytest.mark.parametrize("name", list(CASES))
def test_exact_solve_matches_reference_fixture(name):
    """GPU exact solve vs what the compiled reference returned (golden fixture)."""
    pr, ref = load(name)
    use_rel = bool(ref["use_relative"][0])
    for i, k in enumerate(ref["k"]):
        print(f"fixture {name} solve({k})", flush=True)
        g = Graph.from_problem(pr)
        assert g.solve(int(k), use_rel, FP64) == ref["accepted"][i]
        assert bits(g.current_chi(True)) == bits(ref["chi2"][i])
        if f"poses_{k}" in ref:
            assert np.array_equal(bits(g.poses()), bits(ref[f"poses_{k}"]))
            assert np.array_equal(bits(g.points()), bits(ref[f"points_{k}"]))


@pytest.mark.parametrize("cfg,kw,iters", [
    (0, {}, 10),
    (1, dict(n_frames=20, n_points=1500), 20),
    (3, dict(n_frames=12, n_points=800), 10),
    (1, dict(n_frames=30, n_points=2000, track_max=30), 15),
])
def test_exact_solve_matches_oracle(cfg, kw, iters):
    pr = synth(cfg, **kw)
    print(f"exact solve cfg={cfg} {kw} obs={pr.n_obs} iters={iters}", flush=True)
    g, o = exact_compare(pr, iters)
    flags = g.inlier_flags()
    r_o = o.edge_eval()
    flags_o = (r_o[:, 25] <= pr.obs_delta ** 2) | ~np.isin(pr.obs_loss, [2, 4])
    assert np.array_equal(flags.astype(bool), flags_o)


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("cfg,kw,iters", [(1, dict(n_frames=20, n_points=1500), 12), (3, dict(n_frames=12, n_points=800), 8)])
def test_exact_solve_both_factor_paths(cfg, kw, iters, path):
    """The one-workgroup factorisation and the per-panel launches (both skipping
    the tiles outside the envelope) give the reference's factor bit for bit."""
    pr = synth(cfg, **kw)
    print(f"factor path {path} cfg={cfg} {kw}", flush=True)
    exact_compare(pr, iters, factor_path=path)


@pytest.mark.parametrize("frames", [80, 85])
def test_exact_one_workgroup_panel_limit(frames):
    """Loop closures of 15 and 16 tiles with the one-workgroup factorisation
    requested: 15-tile panels (448 rows below the diagonal tile, one row thread
    each) stay in the one workgroup, 16-tile panels go to the per-panel launches;
    both reproduce the reference bit for bit."""
    pr = synth(4, n_frames=frames, n_points=3000)
    print(f"one-workgroup limit frames={frames} n={6 * (frames - 2)}", flush=True)
    exact_compare(pr, 3, factor_path=1)


def test_exact_loop_closure_dense_envelope():
    """A loop with revisits: the envelope is dense and the per-panel launches run."""
    pr = synth(4, n_frames=120, n_points=6000)
    print(f"exact solve loop closure obs={pr.n_obs}", flush=True)
    exact_compare(pr, 5)


def test_exact_full_config1():
    """The 100-keyframe / 20k-point / 160k-observation workload, 20 iterations."""
    pr = synth(1)
    print(f"exact solve c1 obs={pr.n_obs}", flush=True)
    exact_compare(pr, 20)


def test_exact_repeated_solves_share_hash_state():
    """A second solve on the same graph continues from the first one's state and
    from its hash-map bucket state, as the reference's factor graph object does."""
    pr = synth(1, n_frames=12, n_points=600)
    g = Graph.from_problem(pr)
    o = CpuGraph(pr, "oracle")
    for it in (3, 4):
        assert g.solve(it, False, FP64) == o.solve(it)
        P, X = g.poses(), g.points()
        Po, Xo = o.state()
        assert np.array_equal(bits(P), bits(Po)) and np.array_equal(bits(X), bits(Xo))


@pytest.mark.parametrize("frames", [60, 130])
def test_fast_blocked_factorisation(frames):
    """The per-panel launches in blocks of 8 panels, each block's update of the
    tiles right of it on FP64 matrix cores (FP64_FAST and MIXED), against the
    oracle and against the one-workgroup factorisation."""
    pr = synth(1, n_frames=frames, n_points=40 * frames, outlier_rate=0.0, loss_kind=1)
    print(f"blocked factorisation frames={frames} obs={pr.n_obs}", flush=True)
    o = CpuGraph(pr, "oracle")
    acc_o = o.solve(6)
    Po, Xo = o.state()
    chi_o = o.current_chi(True)
    for path in (2, 1):
        g = Graph.from_problem(pr, factor_path=path)
        assert g.solve(6, False, FP64_FAST) == acc_o
        chi = g.current_chi(True)
        err = state_rel_err(g.poses(), g.points(), Po, Xo)
        print(f"  path {path}: chi2 rel {(chi - chi_o) / chi_o:+.2e} state {err:.2e}", flush=True)
        assert abs(chi - chi_o) <= TOL * abs(chi_o)
        assert err <= TOL


@pytest.mark.parametrize("cfg,kw,iters", [
    (0, {}, 10),
    (3, dict(n_frames=12, n_points=800), 10),
    # loop closure, 2160-row reduced system: the blocked factorisation and the
    # blocked substitution; no outliers or robust loss, so the LM path does not
    # hang on rounding
    (4, dict(n_frames=360, n_points=12000, outlier_rate=0.0, loss_kind=1), 4),
])
def test_fast_solve_within_tolerance(cfg, kw, iters):
    pr = synth(cfg, **kw)
    g = Graph.from_problem(pr)
    acc = g.solve(iters, False, FP64_FAST)
    o = CpuGraph(pr, "oracle")
    assert acc == o.solve(iters)
    P, X = g.poses(), g.points()
    Po, Xo = o.state()
    chi, chi_o = g.current_chi(True), o.current_chi(True)
    assert abs(chi - chi_o) <= TOL * abs(chi_o)
    assert state_rel_err(P, X, Po, Xo) <= TOL


@pytest.mark.parametrize("cfg,kw,iters", [(1, dict(n_frames=20, n_points=1500), 0), (1, dict(n_frames=20, n_points=1500), 5),
                                          (4, dict(n_frames=60, n_points=3000), 3)])
def test_reprojection_cull_matches_oracle(cfg, kw, iters):
    """map::cull's reprojection test on the device against the oracle, at the
    initial state and after exact solves: the float32 means bit for bit, the same
    erase flags; outliers and points behind cameras give flagged points."""
    pr = synth(cfg, **kw)
    print(f"reprojection cull cfg={cfg} {kw} iters={iters}", flush=True)
    g = Graph.from_problem(pr)
    o = CpuGraph(pr, "oracle")
    if iters:
        assert g.solve(iters, False, FP64) == o.solve(iters)
    for thr in (5.991, 50.0):
        mean, erase = g.reprojection_cull(thr)
        mean_o, erase_o = o.reprojection_cull(thr)
        print(f"  threshold {thr}: flagged {int(erase.sum())} of {len(erase)}", flush=True)
        assert np.array_equal(mean.view(np.uint32), mean_o.view(np.uint32))
        assert np.array_equal(erase, erase_o)
        assert 0 < int(erase.sum()) < len(erase)


def test_reprojection_cull_after_edits():
    """Ids after removals and additions: a point whose observations were all
    removed and a point added without any are flagged with an infinite mean, a
    removed point's entries are left untouched, the rest match the oracle."""
    pr = synth(1, n_frames=10, n_points=400)
    g = Graph.from_problem(pr)
    for e in np.flatnonzero(pr.obs_point == 0):
        g.remove_observation(int(e))
    g.remove_point(1)
    extra = g.add_points(np.array([[0.5, 0.5, 4.0]]))
    mean, erase = g.reprojection_cull()
    mean_o, erase_o = CpuGraph(pr, "oracle").reprojection_cull()
    assert np.isinf(mean[0]) and erase[0] == 1
    assert np.isnan(mean[1]) and erase[1] == 0
    assert np.isinf(mean[extra]) and erase[extra] == 1
    assert np.array_equal(mean[2:pr.n_points].view(np.uint32), mean_o[2:].view(np.uint32))
    assert np.array_equal(erase[2:pr.n_points], erase_o[2:])


@pytest.mark.parametrize("n_free", [1, 3])
def test_exact_motion_only(n_free):
    """map::optimise(n, true, 50), the tracking step's BA for one free frame:
    every landmark fixed, edges only to the free frames, 50 iterations."""
    pr = motion_only(synth(1, seed=30 + n_free, n_frames=20, n_points=1500), n_free)
    print(f"motion-only n_free={n_free} obs={pr.n_obs}", flush=True)
    exact_compare(pr, 50)
