"""Each LM attempt's launch sequence captured once as a HIP graph and replayed
(zsba_options.attempt_graph = 0, the default) against the same solves launched
kernel by kernel (attempt_graph = -1): every replay must give the eager result
bit for bit, across repeated solves on one handle, precisions alternating (one
captured graph per precision and freeze state), and a structure edit between
solves (the graphs are captured again)."""
import numpy as np
import pytest

from zeroslam_amd.engine import FP64, FP64_FAST, MIXED, Graph
from zeroslam_amd.problem import synth

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def run_sequence(pr, attempt_graph):
    g = Graph.from_problem(pr, attempt_graph=attempt_graph)
    out = []
    for precision, iters in ((MIXED, 4), (FP64, 3), (MIXED, 3), (FP64_FAST, 4), (MIXED, 2)):
        acc = g.solve(iters, False, precision)
        st = g.stats()
        out.append((acc, st["iterations"], st["attempts"], bits(g.poses()), bits(g.points())))
    # a structure edit: the next solve captures its graphs again
    g.remove_observation(0)
    acc = g.solve(3, False, MIXED)
    st = g.stats()
    out.append((acc, st["iterations"], st["attempts"], bits(g.poses()), bits(g.points())))
    return out


def test_attempt_graph_replays_equal_eager_launches():
    pr = synth(1, n_frames=30, n_points=2500)
    replay, eager = run_sequence(pr, 0), run_sequence(pr, -1)
    for k, (a, b) in enumerate(zip(replay, eager)):
        print(f"solve {k}: accepted {a[0]} iterations {a[1]} attempts {a[2]}", flush=True)
        assert a[:3] == b[:3]
        assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
