"""Shared helpers for the parity tests."""
import numpy as np


def rel_err(a, b) -> float:
    """max |a - b| / max(|b|, tiny) over all entries."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def state_rel_err(poses_a, points_a, poses_b, points_b) -> float:
    return max(rel_err(poses_a, poses_b), rel_err(points_a, points_b))
