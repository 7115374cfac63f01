"""Rate-kernel cache on disk (.npz, no pickles): round trip, refusal of a
cache built for other inputs."""

import numpy as np
import pytest

from prince_amd import workloads
from prince_amd.model.core import PriNCeRun
from prince_amd.model.photonfields import CMBPhotonSpectrum


def test_kernel_cache_round_trip_and_refusal(tmp_path):
    path = str(tmp_path / "kernel_m14.npz")
    a = PriNCeRun(max_mass=14, photon_field=CMBPhotonSpectrum(), kernel_cache=path)
    assert not a.int_rates.loaded_from_cache
    b = PriNCeRun(max_mass=14, photon_field=CMBPhotonSpectrum(), kernel_cache=path)
    assert b.int_rates.loaded_from_cache
    for name in ("_batch_matrix", "_batch_rows", "_batch_cols"):
        assert np.array_equal(getattr(a.int_rates, name), getattr(b.int_rates, name))
    ca, cb = a.int_rates.coupling_mat, b.int_rates.coupling_mat
    assert np.array_equal(ca.indptr, cb.indptr) and np.array_equal(ca.indices, cb.indices)
    z = 0.7
    assert np.array_equal(a.int_rates.get_hadr_jacobian(z).data, b.int_rates.get_hadr_jacobian(z).data)
    with pytest.raises(ValueError, match="species"):
        PriNCeRun(max_mass=12, photon_field=CMBPhotonSpectrum(), kernel_cache=path)
    with pytest.raises(ValueError, match="cr_bins"):
        PriNCeRun(max_mass=14, photon_field=CMBPhotonSpectrum(), cosmic_ray_grid=(6, 13, 8),
                  kernel_cache=path)
    assert workloads.proton_run().int_rates.loaded_from_cache is False
