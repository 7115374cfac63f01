"""Finite-difference continuous-loss operator D.

Restates ``solvers/partial_diff.py:28-110``: per species a d x d matrix on
ln E with the 4th-order upwind-biased bulk stencil

    rows 3..d-4: offsets [-1, 0, 1, 2, 3], coefficients [-3, -10, 18, -6, 1]/(12 h)
    rows 1, 2  : offsets [-1, 0, 1, 2],    coefficients [-2, -3, 6, -1]/(6 h)
    row 0      : offsets [1, 2, 3],        coefficients [-3, 4, -1]/(2 h)

with the right edge mirrored (offsets and coefficients negated, reversed),
h = ln(bins[1]/bins[0]), every row scaled by 1/E_centre, and the blocks placed
on the diagonal for every species.  Rows 0 and d-1 have no diagonal entry.
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import block_diag, coo_matrix

# (row selector, offsets, integer coefficients, denominator)
_LEFT = (
    ([1, 2, 3], [-3, 4, -1], 2.0),
    ([-1, 0, 1, 2], [-2, -3, 6, -1], 6.0),
    ([-1, 0, 1, 2], [-2, -3, 6, -1], 6.0),
)
_BULK = ([-1, 0, 1, 2, 3], [-3, -10, 18, -6, 1], 12.0)


def _mirror(stencil):
    offs, coeffs, den = stencil
    return ([-o for o in offs[::-1]], [-c for c in coeffs[::-1]], den)


def single_block(grid):
    """Dense d x d FD operator for one species on ``grid`` (EnergyGrid)."""
    h = np.log(grid.bins[1] / grid.bins[0])
    d = grid.d
    last = d - 1
    op = np.zeros((d, d))
    for r, st in enumerate(_LEFT):
        offs, coeffs, den = st
        op[r, r + np.asarray(offs)] = np.asarray(coeffs) / (den * h)
    for r, st in enumerate(_LEFT):
        offs, coeffs, den = _mirror(st)
        op[last - r, last - r + np.asarray(offs)] = np.asarray(coeffs) / (den * h)
    offs, coeffs, den = _BULK
    for row in range(3, d - 3):
        op[row, row + np.asarray(offs)] = np.asarray(coeffs) / (den * h)
    return op * (1.0 / grid.grid)[:, None]


class DifferentialOperator:
    def __init__(self, cr_grid, nspec):
        self.ebins = cr_grid.bins
        self.egrid = cr_grid.grid
        self.dim_e = cr_grid.d
        self.log_width = np.log(self.ebins[1] / self.ebins[0])
        self.nspec = nspec
        self.block = single_block(cr_grid)
        self.operator = block_diag(nspec * [coo_matrix(self.block)]).tocsr()
