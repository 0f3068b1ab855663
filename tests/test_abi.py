"""The C-ABI library: it loads, exports every function include/zsba.h declares,
and its host-side graph bookkeeping validates arguments without a device."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from zeroslam_amd import _lib
from zeroslam_amd.engine import Graph, ZsbaError

HEADER = Path(__file__).resolve().parents[1] / "include" / "zsba.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zsba_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_lib.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared surface
    assert sorted(_lib.EXPORTED) == names


def test_abi_version_and_handles():
    lib = _lib.load()
    assert lib.zsba_abi_version() == 2
    g = Graph()
    assert g.add_poses(np.tile([0, 0, 0, 0, 0, 0, 1.0], (3, 1)), [1, 0, 0]) == 0
    assert g.add_points(np.zeros((2, 3))) == 0
    assert g.add_point([1.0, 2.0, 3.0]) if hasattr(g, "add_point") else True
    K = [525.0, 525.0, 320.0, 240.0]
    assert g.add_observations([0, 1], [0, 1], [[1, 2], [3, 4]], K, 2, 2.0) == 0
    assert g.counts() == (3, 2, 2)
    with pytest.raises(ZsbaError):
        g.add_observations([5], [0], [[0, 0]], K)          # unknown pose id
    with pytest.raises(ZsbaError):
        g.add_observations([0], [9], [[0, 0]], K)          # unknown point id
    with pytest.raises(ZsbaError):
        g.add_observations([0], [0], [[0, 0]], [0.0, 1.0, 0, 0])  # fx must be > 0
    with pytest.raises(ZsbaError):
        g.add_observations([0], [0], [[0, 0]], K, 7, 1.0)  # unknown loss
    g.remove_observation(1)
    with pytest.raises(ZsbaError):
        g.remove_observation(1)                            # already removed (reference returns false)
    g.remove_pose(0)
    with pytest.raises(ZsbaError):
        g.remove_pose(0)
    P = g.poses()
    assert np.isnan(P[0]).all() and np.array_equal(P[1], [0, 0, 0, 0, 0, 0, 1.0])
    g.close()


def test_empty_graph_solve_returns_zero():
    # factor_graph.hpp:494-499: no edges or no vertices -> 0, without touching a device
    g = Graph()
    assert g.solve(10) == 0
    g.add_poses([[0, 0, 0, 0, 0, 0, 1.0]])
    assert g.solve(10) == 0


def test_package_star_import():
    # every name in zeroslam_amd.__all__ is importable
    ns = {}
    exec("from zeroslam_amd import *", ns)
    import zeroslam_amd
    for name in zeroslam_amd.__all__:
        assert name in ns, name
