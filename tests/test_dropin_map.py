"""ZeroSLAM's own map::optimise and map::cull (source/map.hpp:65-323), compiled
unchanged twice by oracle/Makefile (target `map`): with the reference's factor
graph (oracle/_ref/map_ref) and with the engine's drop-in header
include/dropin/factor_graph.hpp in its place (oracle/_ref/map_zsba, linked to
libzsba.so).  The scenarios run the calls a tracker makes (the local-window BA
of slam.hpp:472, culling, the tracking BA of slam.hpp:316/387/469, a global BA)
and the two binaries must leave the map's frames and landmarks identical bit
for bit, with identical printed lines.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.map_scenario import MAP_REF, MAP_ZSBA, ROOT, parse, run, tiny_scenario, walk_scenario

REFERENCE = Path("/root/reference/source/map.hpp")


@pytest.mark.skipif(not REFERENCE.exists(), reason="reference checkout not present")
def test_reference_map_compiles_against_the_dropin():
    """The reference's map.hpp, unmodified, builds against the drop-in; the
    result links libzsba.so and solves through its C ABI."""
    proc = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "map"], capture_output=True, text=True)
    assert proc.returncode == 0, proc.stdout[-3000:] + proc.stderr[-3000:]
    assert MAP_REF.exists() and MAP_ZSBA.exists()
    ldd = subprocess.run(["ldd", str(MAP_ZSBA)], capture_output=True, text=True).stdout
    assert "libzsba.so" in ldd
    undefined = subprocess.run(["nm", "-u", str(MAP_ZSBA)], capture_output=True, text=True).stdout
    assert "zsba_solve" in undefined and "zsba_current_chi" in undefined
    assert "zsba_" not in subprocess.run(["nm", "-u", str(MAP_REF)], capture_output=True, text=True).stdout


@pytest.mark.skipif(not MAP_REF.exists(), reason="oracle/_ref/map_ref not built")
def test_reference_map_harness_runs(tmp_path):
    sc = walk_scenario()
    out, raw = run(MAP_REF, sc, tmp_path, "walk")
    states = parse(raw, len(sc.calls))
    lines = out.splitlines()
    assert lines[0].startswith("Optimised:") and lines[1].startswith("Culled:")
    # the local-window BA lowers the error and the culling removes landmarks
    a, b = (float(x) for x in lines[0].split()[1:4:2])
    assert b < a
    assert len(states[1][1]) < len(states[0][1])


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.gpu
@pytest.mark.skipif(not (MAP_REF.exists() and MAP_ZSBA.exists()), reason="map drivers not built")
@pytest.mark.parametrize("case", ["walk7", "walk21", "tiny"])
def test_dropin_map_bitwise(tmp_path, case):
    sc = {"walk7": lambda: walk_scenario(7), "walk21": lambda: walk_scenario(21, n_frames=24, n_points=1200),
          "tiny": tiny_scenario}[case]()
    print(f"map scenario {case}: frames {len(sc.frames)} landmarks {len(sc.landmarks)} "
          f"observations {len(sc.observations)} calls {len(sc.calls)}", flush=True)
    out_ref, raw_ref = run(MAP_REF, sc, tmp_path, "ref")
    out_gpu, raw_gpu = run(MAP_ZSBA, sc, tmp_path, "zsba")
    print(out_gpu, flush=True)
    assert out_gpu == out_ref
    ref, gpu = parse(raw_ref, len(sc.calls)), parse(raw_gpu, len(sc.calls))
    for k, ((fr, lr), (fg, lg)) in enumerate(zip(ref, gpu)):
        assert sorted(fr) == sorted(fg) and sorted(lr) == sorted(lg), f"call {k}: different frames or landmarks"
        for i in fr:
            assert same(fr[i][0], fg[i][0]) and same(fr[i][1], fg[i][1]), f"call {k}: frame {i}"
        for j in lr:
            assert same(lr[j], lg[j]), f"call {k}: landmark {j}"
    assert raw_gpu == raw_ref
