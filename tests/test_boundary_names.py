"""The solver inputs are read by the reference's attribute names only
(``prince_amd/solvers/run_inputs.py``), checked on a stand-in run that has
nothing else."""

import numpy as np
import pytest
import scipy.sparse as sp

from prince_amd import workloads
from prince_amd.model.partial_diff import DifferentialOperator
from prince_amd.solvers import run_inputs as ri

from .reference_names import reference_view


@pytest.mark.parametrize("kernel", ["dense", "csr32"])
def test_inputs_from_reference_names(kernel):
    run = workloads.nuclear_run()
    view = reference_view(run, kernel)
    with pytest.raises(AttributeError):
        view.int_rates.batch_matrix  # noqa: B018
    with pytest.raises(AttributeError):
        view.pair_loss_rates_grid.rate_single  # noqa: B018
    B = ri.fold_kernel_fp64(view.int_rates)
    assert B.dtype == np.float64 and B.flags.c_contiguous
    if kernel == "dense":
        assert np.array_equal(B, run.int_rates._batch_matrix)
    else:  # the reference's float32 storage, widened exactly
        assert np.array_equal(B, run.int_rates._batch_matrix.astype(np.float32).astype(np.float64))
    pair = view.pair_loss_rates_grid
    for z in (0.0, 1.3):
        r = ri.pair_rate_per_energy(pair, z)
        scale = ri.pair_species_scale(pair, run.spec_man)
        full = np.repeat(scale, run.cr_grid.d) * np.tile(r, run.spec_man.nspec)
        assert np.allclose(full, pair.loss_vector(z), rtol=1e-15, atol=0)
    D = DifferentialOperator(run.cr_grid, run.spec_man.nspec).operator
    block = ri.fd_block(D, run.spec_man.nspec, run.cr_grid.d)
    assert np.array_equal(sp.csr_matrix(D)[:64, :64].toarray(), block)


def test_fold_kernel_rejects_missing_or_misshaped():
    run = workloads.proton_run()

    class NoKernel:
        coupling_mat = run.int_rates.coupling_mat
        _batch_matrix = None

    with pytest.raises(ValueError):
        ri.fold_kernel_fp64(NoKernel())

    class Wrong:
        coupling_mat = run.int_rates.coupling_mat
        _batch_matrix = np.zeros((3, 72))

    with pytest.raises(ValueError):
        ri.fold_kernel_fp64(Wrong())


def test_fd_block_rejects_unequal_blocks():
    run = workloads.proton_run()
    D = DifferentialOperator(run.cr_grid, 3).operator.tolil()
    D[70, 71] = 123.0
    with pytest.raises(NotImplementedError):
        ri.fd_block(D.tocsr(), 3, run.cr_grid.d)


def test_em_slices_follow_nuclear_states():
    run = workloads.proton_em_run()
    sl = ri.em_slices(run)
    d = run.em_grid.d
    assert sl[-11] == slice(run.dim_states, run.dim_states + d)
    assert sl[22] == slice(run.dim_states + 2 * d, run.dim_states + 3 * d)
    assert ri.em_slices(workloads.proton_run()) == {}
