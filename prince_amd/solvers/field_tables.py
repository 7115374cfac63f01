"""Photon vectors and pair-production rates of many windows, on the device.

For a CMB [+ synthetic EBL] field the per-window inputs of the rate fold and
of the pair-loss term are evaluated by ``prh_field_tables``: the CMB Planck
spectrum at the window's redshift plus the z-separable EBL tabulated at
z = 0, at

* the photon-grid centres -> ``photon_vector(z)`` of the reference
  (``interaction_rates.py:105-116``), [nwin][dph];
* the (nucleon energy, xi) samples of the pair-loss integral, contracted with
  the trapezoid weights of xi times phi(xi)/xi^2 -> the per-energy pair rate
  ``trapz(photon_vector * phi_xi2, xi)`` (``interaction_rates.py:1277-1312``),
  [nwin][d].

Both stay in HBM; the host only supplies the window redshifts and scales.
"""

from __future__ import annotations

import numpy as np

from .. import _lib
from ..model import config as mconfig
from ..model.photonfields import PLANCK_PREF, CMBPhotonSpectrum, CombinedPhotonField, SyntheticEBL


def _torch():
    import torch

    return torch


def device_field(field):
    """(CMB, EBL or None) when the device can evaluate ``field``, else None."""
    if isinstance(field, CMBPhotonSpectrum):
        return field, None
    if isinstance(field, CombinedPhotonField):
        cmb = [m for m in field.model_list if isinstance(m, CMBPhotonSpectrum)]
        ebl = [m for m in field.model_list if isinstance(m, SyntheticEBL)]
        if len(cmb) == 1 and len(ebl) <= 1 and len(field.model_list) == len(cmb) + len(ebl):
            return cmb[0], (ebl[0] if ebl else None)
    return None


def trapz_weights(x):
    """w with sum(w * y) == trapezoid(y, x)."""
    w = np.empty_like(x)
    w[1:-1] = 0.5 * (x[2:] - x[:-2])
    w[0] = 0.5 * (x[1] - x[0])
    w[-1] = 0.5 * (x[-1] - x[-2])
    return w


class DeviceFieldTables:
    """Per-window photon vectors and pair rates of one run on one device."""

    def __init__(self, run, device):
        torch = _torch()
        parts = device_field(run.photon_field)
        if parts is None:
            raise NotImplementedError("device field tables need a CMB [+ SyntheticEBL] field")
        _, self.ebl = parts
        self.lib = _lib.load()
        self.device = device
        ir = run.int_rates
        pair = run.pair_loss_rates_grid
        dv = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device)  # noqa: E731
        eph = ir.e_photon.grid
        pg = pair.photon_grid  # [d][nxi]
        W = np.broadcast_to(trapz_weights(pair.xi) * pair.phi_xi2, pg.shape)
        ebl0 = (lambda E: np.zeros_like(E)) if self.ebl is None else self.ebl.density_z0
        self.ph = {"P": dv(eph), "W": dv(np.ones_like(eph)), "E0": dv(ebl0(eph)),
                   "ni": eph.size, "nm": 1}
        self.pr = {"P": dv(pg), "W": dv(W), "E0": dv(ebl0(pg)), "ni": pg.shape[0], "nm": pg.shape[1]}

    def _table(self, spec, zt, st, planck, stream):
        torch = _torch()
        out = torch.empty((zt.numel(), spec["ni"]), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.prh_field_tables(
            zt.data_ptr(), st.data_ptr(), zt.numel(), spec["P"].data_ptr(), spec["W"].data_ptr(),
            spec["E0"].data_ptr(), spec["ni"], spec["nm"], planck, mconfig.E_CMB,
            SyntheticEBL.Z_POWER, out.data_ptr(), stream), "field tables")
        return out

    def tables(self, zs, scales=None, cmb=True, ebl=True):
        """(photon vectors [nw][dph], pair rates [nw][d]) at redshifts ``zs`` with
        the EBL scaled by ``scales`` per window (its own scale times 1 by
        default); ``cmb`` / ``ebl`` switch the components."""
        torch = _torch()
        zs = np.ascontiguousarray(zs, dtype=np.float64).reshape(-1)
        s = np.ones(zs.size) if scales is None else np.asarray(scales, dtype=np.float64).reshape(-1)
        escale = 0.0 if (self.ebl is None or not ebl) else self.ebl.scale
        zt = torch.from_numpy(zs).to(self.device)
        st = torch.from_numpy(s * escale).to(self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        planck = PLANCK_PREF if cmb else 0.0
        return (self._table(self.ph, zt, st, planck, stream),
                self._table(self.pr, zt, st, planck, stream))
