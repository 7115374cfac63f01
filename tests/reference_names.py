"""A stand-in PriNCeRun that exposes only the attribute names the reference's
objects carry (``core.py:19-257``, ``interaction_rates.py:60-1386``,
``config.py:371-407``): any other attribute read raises AttributeError, so a
solver built on it shows that it needs nothing beyond the reference's names.

The rate kernel can be handed over as the reference's default storage, a
scipy CSR matrix with float32 values (``batch_matrix_format = "csr"``,
``xs_dtype = "float32"``).
"""

import numpy as np
import scipy.sparse as sp


class Only:
    """Read-only view of ``obj`` restricted to ``names`` (writes stay local)."""

    def __init__(self, obj, names, **override):
        object.__setattr__(self, "_obj", obj)
        object.__setattr__(self, "_names", frozenset(names) | frozenset(override))
        object.__setattr__(self, "_override", dict(override))

    def __getattr__(self, name):
        if name in self._override:
            return self._override[name]
        if name not in self._names:
            raise AttributeError(f"{name!r} is not an attribute of the reference's object")
        return getattr(self._obj, name)


RUN_NAMES = ("backend", "cr_grid", "ph_grid", "spec_man", "dim_states", "dim_bins",
             "int_rates", "adia_loss_rates_grid", "pair_loss_rates_grid", "photon_field",
             "enable_em_cascade", "em_grid", "cross_sections")
INT_RATES_NAMES = ("_batch_matrix", "_batch_rows", "_batch_cols", "coupling_mat",
                   "photon_vector", "e_photon", "e_cosmicray", "get_hadr_jacobian")
PAIR_NAMES = ("xi", "phi_xi2", "photon_vector", "photon_grid", "scale_vec", "loss_vector",
              "e_cosmicray")
ADIA_NAMES = ("energy_vector", "loss_vector")


def reference_view(run, kernel="dense"):
    """The run seen through the reference's names; ``kernel="csr32"`` stores the
    rate kernel as the reference's default float32 CSR matrix."""
    ir = run.int_rates
    B = ir._batch_matrix
    if kernel == "csr32":
        B = sp.csr_matrix(B.astype(np.float32))
    elif kernel != "dense":
        raise ValueError(kernel)
    return Only(
        run, RUN_NAMES,
        int_rates=Only(ir, INT_RATES_NAMES, _batch_matrix=B),
        pair_loss_rates_grid=Only(run.pair_loss_rates_grid, PAIR_NAMES),
        adia_loss_rates_grid=Only(run.adia_loss_rates_grid, ADIA_NAMES),
        enable_em_cascade=bool(getattr(run, "enable_em_cascade", False)),
        em_grid=getattr(run, "em_grid", None),
    )
