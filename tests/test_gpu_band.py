"""The banded factorisation and substitutions of the reduced camera system
(kernels_band.hip, factor path 3: FP64-fast and mixed when every row of the
reduced system spans at most 120 columns) against the CPU oracle's
dense Cholesky (matrix_decomposition_cholesky.hpp:26-95) and against the
one-workgroup tiled factorisation (factor path 1)."""
import numpy as np
import pytest

from oracle.oracle import CpuGraph
from tests.helpers import rel_err, state_rel_err
from zeroslam_amd.engine import FP64, FP64_FAST, MIXED, Graph
from zeroslam_amd.problem import synth

pytestmark = pytest.mark.gpu

CASES = [
    # (config, generator arguments): band width in columns, reduced-system rows
    (0, {}),                                             # every pose coupled: band = n = 60
    (1, dict(n_frames=21, n_points=1500)),               # n = 126, not a multiple of the panel width
    (1, dict(n_frames=100)),                             # the headline workload: band 72, n = 600
    (3, dict(n_frames=50, n_points=3000)),
    # longer tracks widen the band: windows of 96, 112 and 128 positions (the
    # four-wave factorisation; the cases above take windows of 64 and 80)
    (1, dict(n_frames=60, track_max=13)),                # band 78
    (1, dict(n_frames=60, track_max=16)),                # band 96
    (1, dict(n_frames=60, track_max=18)),                # band 108 (window 128)
    (1, dict(n_frames=60, track_max=20)),                # band 120: past the limit, the tiled path
]


@pytest.mark.parametrize("cfg,kw", CASES)
def test_band_reduced_solve(cfg, kw):
    """The damped reduced system solved through the banded path equals the
    oracle's binary64 solve to the system's conditioning, and the tiled path's
    (the debug entry point factorises with the reassociating paths)."""
    pr = synth(cfg, **kw)
    o = CpuGraph(pr, "oracle")
    _, S_o, _, _, _ = o.reduced_system(1e-3)
    lam = 1e-5 * float(np.max(np.diag(S_o)))
    ok_o, S_o, bs_o, xp_o, xl_o = o.reduced_system(lam)
    gb = Graph.from_problem(pr, factor_path=3)
    gt = Graph.from_problem(pr, factor_path=1)
    ok_b, S_b, bs_b, xp_b, xl_b = gb.reduced_system(lam, FP64)
    ok_t, _, _, xp_t, xl_t = gt.reduced_system(lam, FP64)
    assert ok_o and ok_b and ok_t
    # the system itself is formed before the factorisation, identically
    low = np.tril_indices(S_b.shape[0])
    assert rel_err(S_b[low], S_o[low]) < 1e-12
    eb, et = rel_err(xp_b, xp_o), rel_err(xp_t, xp_o)
    el = rel_err(xl_b, xl_o)
    print(f"band solve cfg={cfg} {kw}: n {S_b.shape[0]} dx_p band {eb:.2e} tiled {et:.2e} dx_l band {el:.2e}",
          flush=True)
    assert eb < 1e-9 and el < 1e-9
    assert rel_err(xp_b, xp_t) < 1e-9


def test_band_failed_pivot():
    """A damping that makes the reduced system indefinite fails the decomposition
    on both paths (the pivot rule of matrix_decomposition_cholesky.hpp:26-59)."""
    pr = synth(1, n_frames=20, n_points=1500)
    o = CpuGraph(pr, "oracle")
    _, S_o, _, _, _ = o.reduced_system(1e-3)
    lam = -0.5 * float(np.max(np.diag(S_o)))
    ok_o = o.reduced_system(lam)[0]
    ok_b = Graph.from_problem(pr, factor_path=3).reduced_system(lam, FP64)[0]
    ok_t = Graph.from_problem(pr, factor_path=1).reduced_system(lam, FP64)[0]
    assert not ok_o and not ok_b and not ok_t


@pytest.mark.parametrize("precision", [FP64_FAST, MIXED])
def test_band_solve_path_equals_tiled(precision):
    """Whole solves through the banded path reach the tiled path's cost (the
    paths differ by rounding only) on a loss-free forward walk."""
    pr = synth(1, n_frames=40, n_points=3000, outlier_rate=0.0, loss_kind=1)
    out = {}
    for path in (3, 1):
        g = Graph.from_problem(pr, factor_path=path)
        acc = g.solve(8, False, precision)
        out[path] = (acc, g.current_chi(True), g.poses(), g.points())
    (a3, c3, P3, X3), (a1, c1, P1, X1) = out[3], out[1]
    err = state_rel_err(P3, X3, P1, X1)
    print(f"band vs tiled precision {precision}: accepted {a3}/{a1} chi2 rel {(c3 - c1) / c1:+.2e} state {err:.2e}",
          flush=True)
    assert a3 == a1
    assert abs(c3 - c1) <= 1e-9 * c1
    assert err < 1e-6
