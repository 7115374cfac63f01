"""Shared helpers for parity checks against the CPU oracle."""

import numpy as np


def rel_dev(test, ref, floor=1e-30):
    """Max relative deviation over bins above floor * peak of the reference."""
    ref = np.asarray(ref)
    test = np.asarray(test)
    peak = np.abs(ref).max()
    mask = np.abs(ref) > floor * peak
    if not mask.any():
        return 0.0, 0
    rel = np.abs(test[mask] - ref[mask]) / np.abs(ref[mask])
    return float(rel.max()), int(mask.sum())
