"""Physics constants and numerical knobs of the propagation model.

Values mirror the reference configuration module (``config.py:125-148`` for
cosmology / CMB, ``config.py:225-284`` for grids, x-cuts and the rate-cache
threshold) so a run built here has the same numerical semantics.
"""

from __future__ import annotations

from dataclasses import dataclass, field

#: Hubble constant in 1/s and density parameters (config.py:128-135).
H_0s = 2.28475e-18
Omega_m = 0.27
Omega_Lambda = 0.73

#: CMB temperature today as k_B*T0 in GeV (config.py:138).
E_CMB = 2.34823e-13

#: Cosmic-ray and photon grids (log10 E_lo, log10 E_hi, bins/decade).
cosmic_ray_grid = (6, 14, 8)
photon_grid = (-15, -6, 8)

#: Redistribution cuts on x = E_daughter/E_mother (config.py:257-258).
x_cut = 1e-4
x_cut_proton = 1e-1

#: Rate-cache refresh threshold in z (config.py:284).
update_rates_z_threshold = 0.01


@dataclass
class BackendConfig:
    """Per-run backend handle (reference ``config.py:370-407``).

    ``linear_algebra_backend`` selects the solver implementation that the
    ``UHECRPropagationSolverETD2`` facade dispatches to.  This package
    implements the ``"hip"`` backend only.
    """

    linear_algebra_backend: str = "hip"
    update_rates_z_threshold: float = update_rates_z_threshold
    use_hip_graphs: bool = True
    device: int = 0
    extra: dict = field(default_factory=dict)
