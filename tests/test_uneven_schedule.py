"""Host schedule of the uneven batch: the group size for folding windows ahead
into two value slots, checked against brute force over random window starts."""

import numpy as np

from prince_amd.solvers.uneven import UnevenBatchSolverETD2


def _brute_ok(starts, S):
    """Direct statement of the slot rule: with window w of a column in slot
    w & 1, the fold of every window starting in [g, g + S) happens before step
    g, so the window it overwrites (two windows back) must have ended before g
    and the window before it must have started before g."""
    T, K = starts.shape
    for k in range(K):
        st = np.flatnonzero(starts[:, k])
        for i in range(len(st)):
            g = (st[i] // S) * S
            if i >= 1 and st[i - 1] >= g:
                return False
    return True


def test_refold_group_is_largest_valid_power_of_two():
    rng = np.random.default_rng(0)
    for _ in range(40):
        T, K = int(rng.integers(5, 300)), int(rng.integers(1, 6))
        starts = np.zeros((T, K), dtype=bool)
        starts[0] = True
        for k in range(K):
            t = 0
            gap = int(rng.integers(1, 20))
            while True:
                t += int(rng.integers(gap, gap + 5))
                if t >= T:
                    break
                starts[t, k] = True
        S = UnevenBatchSolverETD2._refold_group(starts)
        assert S & (S - 1) == 0 and 1 <= S <= 64
        assert _brute_ok(starts, S)
        if S < 64:
            assert not _brute_ok(starts, 2 * S)


def test_refold_group_of_the_refresh_gate():
    """dz = 1e-3 with the 0.01 refresh threshold starts a window every 10 or 11
    steps (|z - z_cached| > 0.01 in floating point): groups of 8 steps hold at
    most one start per column."""
    from prince_amd.solvers.propagation import refresh_schedule, z_grid_for

    cols = []
    for z in (0.1, 0.537, 1.0, 2.9):
        win, _ = refresh_schedule(z_grid_for(z, 0.0, 1e-3), 0.01)
        cols.append(win)
    T = max(len(w) for w in cols)
    starts = np.zeros((T, len(cols)), dtype=bool)
    for c, w in enumerate(cols):
        starts[0, c] = True
        starts[1:len(w), c] = w[1:] != w[:-1]
    gaps = np.concatenate([np.diff(np.flatnonzero(starts[:, c])) for c in range(len(cols))])
    assert set(np.unique(gaps)) == {10, 11}
    assert UnevenBatchSolverETD2._refold_group(starts) == 8
