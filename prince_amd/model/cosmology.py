"""Cosmology: expansion rate and source-evolution shapes.

Restates ``cosmology.py:10-46`` of the reference: flat LambdaCDM
``H(z) = H0 sqrt(Om (1+z)^3 + OL)`` and the Hopkins-Beacom star formation
rate used by float-valued source evolutions.
"""

from __future__ import annotations

import math

import numpy as np

from . import config


def H(z, H0=config.H_0s):
    """Expansion rate in 1/s (cosmology.py:10-21)."""
    return H0 * np.sqrt(config.Omega_m * (1 + z) ** 3 + config.Omega_Lambda)


def hubble_steps(zs, H0=config.H_0s):
    """``H(z)`` for every z of ``zs``, point by point with the scalar
    arithmetic of ``H`` (``math.sqrt`` is correctly rounded like numpy's) on
    Python floats: the same bits without numpy's per-scalar dispatch."""
    om, ol = config.Omega_m, config.Omega_Lambda
    return [H0 * math.sqrt(om * (1 + z) ** 3 + ol)
            for z in np.asarray(zs, dtype=np.float64).tolist()]


def star_formation_rate(z, z_inhom=0.0):
    """Star formation rate normalised to 1 at z=0 (cosmology.py:24-46)."""
    if z < z_inhom:
        return 0.0
    elif z <= 0.97:
        return (1.0 + z) ** 3.44
    elif 0.97 < z <= 4.48:
        return 10.0**1.09 * (1.0 + z) ** -0.26
    else:
        return 10.0**6.66 * (1.0 + z) ** -7.8
