"""On-disk cache of the photo-nuclear rate kernel.

The reference keeps whole pickled ``PriNCeRun`` objects as kernel caches
(``tests/conftest.py:163-222``, ``tests/_build_kernel_cache.py``) because the
kernel construction dominates start-up.  Here the cache is a plain ``.npz``
(no pickled objects, loaded with ``allow_pickle=False``) holding what the
construction produces -- the dense kernel ``_batch_matrix`` in CSR row order,
``_batch_rows`` / ``_batch_cols`` and the ``coupling_mat`` pattern -- plus the
inputs that determine it: the species list, the cosmic-ray and photon grid
edges, the x-cuts and the cross-section model's class and tables' shapes.  A
cache whose inputs differ from the run's is refused, never silently used.
"""

from __future__ import annotations

import os

import numpy as np

from . import config

FORMAT_VERSION = 1


def cache_key(run):
    """Arrays and scalars that determine the rate kernel of ``run``."""
    cs = run.cross_sections
    return {
        "format_version": np.int64(FORMAT_VERSION),
        "species": np.asarray(run.spec_man.known_species, dtype=np.int64),
        "cr_bins": np.asarray(run.cr_grid.bins, dtype=np.float64),
        "ph_bins": np.asarray(run.ph_grid.bins, dtype=np.float64),
        "x_cuts": np.array([config.x_cut, config.x_cut_proton], dtype=np.float64),
        "xsec_model": np.array(type(cs).__name__),
        "xsec_egrid": np.asarray(getattr(cs, "egrid", np.zeros(0)), dtype=np.float64),
        "xsec_xbins": np.asarray(getattr(cs, "xbins", np.zeros(0)), dtype=np.float64),
    }


def save(int_rates, path):
    """Write the kernel of ``int_rates`` (a ``PhotoNuclearInteractionRate``) to ``path``."""
    cm = int_rates.coupling_mat
    payload = dict(cache_key(int_rates.prince_run))
    payload.update(
        batch_matrix=int_rates._batch_matrix, batch_rows=int_rates._batch_rows,
        batch_cols=int_rates._batch_cols, indptr=cm.indptr, indices=cm.indices,
    )
    tmp = path + ".tmp.npz"
    np.savez(tmp, **payload)
    os.replace(tmp, path)


def load(run, path):
    """Kernel arrays of ``path`` after checking that they were built for ``run``."""
    with np.load(path, allow_pickle=False) as f:
        data = {k: f[k] for k in f.files}
    want = cache_key(run)
    for key, val in want.items():
        if key not in data or not np.array_equal(data[key], val):
            raise ValueError(f"kernel cache {path}: '{key}' differs from this run's; rebuild it")
    n = run.dim_states
    if data["indptr"].size != n + 1 or data["batch_matrix"].shape[0] != data["indices"].size:
        raise ValueError(f"kernel cache {path}: inconsistent pattern for {n} states")
    return data
