"""What the HIP solvers read from a ``PriNCeRun``, by the reference's names only.

The solvers take their inputs from the same objects the reference's ETD2
solver uses, through the attributes those objects carry in the reference:

* fold kernel: ``int_rates._batch_matrix`` -- a dense array
  (``batch_matrix_format = "dense"``) or a scipy CSR matrix with float32
  values (the reference default ``"csr"`` with ``xs_dtype = "float32"``,
  ``interaction_rates.py:134-215, 596-631``); rows in the CSR order of
  ``int_rates.coupling_mat`` (``interaction_rates.py:648-697``).  Densified
  to float64 here; a missing kernel is an error;
* photon vector on the photon grid: ``int_rates.photon_vector(z, pfield)``
  (``interaction_rates.py:105-116``);
* pair-production loss: ``pair_loss_rates_grid.photon_vector(z, pfield)``,
  ``.phi_xi2`` and ``.xi`` give the per-energy rate
  ``trapz(n(m_e xi / 2 gamma, z) phi(xi) / xi^2, xi)`` that every nucleus
  block shares, and ``.scale_vec`` the per-state factor
  ``alpha r_e^2 m_e^2 Z^2 / A`` (``interaction_rates.py:1277-1332``);
* FD loss operator: ``DifferentialOperator(cr_grid, nspec).operator``, the
  block-diagonal CSR matrix of one ``d x d`` stencil per species
  (``solvers/partial_diff.py:4-46``);
* species and grids: ``spec_man.species_refs`` (``princeidx``, ``is_nucleus``,
  ``charge``, ``A``, ``lidx()``), ``spec_man.nspec``, ``cr_grid``
  (``grid``, ``d``), ``dim_states``, ``photon_field``;
* EM cascade: ``enable_em_cascade`` and ``em_grid``
  (``propagation.py:256-316``).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.integrate import trapezoid


def fold_kernel_fp64(int_rates):
    """The rate kernel ``_batch_matrix`` as a C-contiguous float64 array [R, dph]."""
    B = getattr(int_rates, "_batch_matrix", None)
    if B is None:
        raise ValueError("int_rates has no _batch_matrix: the rate kernel was not built")
    if sp.issparse(B):
        B = B.toarray()
    B = np.ascontiguousarray(B, dtype=np.float64)
    nnz = int(int_rates.coupling_mat.nnz)
    if B.ndim != 2 or B.shape[0] != nnz:
        raise ValueError(f"_batch_matrix has shape {B.shape}, expected ({nnz}, dph)")
    return B


def pair_rate_per_energy(pair, z, pfield=None):
    """Per-nucleon-energy pair-production rate [d] at z (the trapezoid of
    ``loss_vector`` before the per-species scale)."""
    return trapezoid(pair.photon_vector(z, pfield=pfield) * pair.phi_xi2, pair.xi, axis=1)


def pair_species_scale(pair, spec_man):
    """``scale_vec`` per species ([nspec]; constant over a species block, zero
    for non-nuclei)."""
    sv = np.asarray(pair.scale_vec, dtype=np.float64)
    out = np.zeros(spec_man.nspec)
    for s in spec_man.species_refs:
        out[s.princeidx] = sv[s.lidx()]
    return out


def fd_block(operator, nspec, d):
    """The ``d x d`` stencil block of a block-diagonal FD operator whose blocks
    are all equal (the single-grid operator of the reference)."""
    Dcsr = sp.csr_matrix(operator)
    if Dcsr.shape != (nspec * d, nspec * d):
        raise ValueError(f"FD operator shape {Dcsr.shape} != ({nspec * d}, {nspec * d})")
    block = Dcsr[:d, :d].toarray()
    diff = Dcsr - sp.kron(sp.identity(nspec, format="csr"), sp.csr_matrix(block), format="csr")
    diff.eliminate_zeros()
    if diff.nnz:
        raise NotImplementedError("FD operators with different blocks per species")
    return block


def em_cascade_grid(prince_run):
    """The EM grid when the run co-evolves the EM cascade, else None."""
    if not bool(getattr(prince_run, "enable_em_cascade", False)):
        return None
    eg = getattr(prince_run, "em_grid", None)
    if eg is None:
        raise ValueError("enable_em_cascade is set but the run has no em_grid")
    return eg


#: EM species whose blocks follow the nuclear states: e+, e-, photon
EM_PDGS = (-11, 11, 22)


def em_slices(prince_run):
    """{pdg: slice} of the EM blocks of the solver state (after the nuclear states)."""
    eg = em_cascade_grid(prince_run)
    if eg is None:
        return {}
    n, d = prince_run.dim_states, eg.d
    return {p: slice(n + i * d, n + (i + 1) * d) for i, p in enumerate(EM_PDGS)}
