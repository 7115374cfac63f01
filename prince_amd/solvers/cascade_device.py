"""EM-cascade window operators built on the device.

For a list of windows (redshift z_w, EBL scale s_w) of a CMB + synthetic-EBL
field this builds, on the GPU, what ``model/cascade.py`` builds on the host
(reference ``cascade/cascade.py:284-589``, ``cascade/bethe_heitler.py:219-351``):

* T_gamma(z_w), T_electron(z_w) [w][n][n]: the kinetic single-scatter cascade
  transfer of one step of |dz| (``prh_cascade_transfer``);
* R_BH(z_w) [w][n][d]: the Bethe-Heitler e+- shapes per nucleon bin
  (``prh_cascade_bh``).

The field-free tensors (IC emission and loss, gamma-gamma pair injection,
opacity, Bethe-Heitler kernel) come from ``CascadeKernels`` once per grid and
stay in HBM.  Per window the target-photon densities at the tensors' photon
grids are evaluated on the device (``prh_field_tables``: CMB Planck spectrum
plus the z-separable EBL), contracted with the tensors by the MFMA GEMM of
``prh_fold_gemm``, and the per-window builder turns the contractions into
the transfer.  Windows are processed in batches sized to a workspace budget.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from ..model import cascade as C
from ..model import config as mconfig
from ..model.cosmology import H
from ..model.photonfields import PLANCK_PREF, CMBPhotonSpectrum, CombinedPhotonField, SyntheticEBL
from ..model.units import PRINCE_UNITS

#: bytes of per-window builder workspace (9 padded 144 x 144 FP64 matrices)
WORK_PER_WINDOW = 9 * 144 * 144 * 8


def _torch():
    import torch

    return torch


def field_parts(field):
    """(CMB, EBL or None) of a CMB [+ SyntheticEBL] photon field."""
    if isinstance(field, CMBPhotonSpectrum):
        return field, None
    if isinstance(field, CombinedPhotonField):
        cmb = [m for m in field.model_list if isinstance(m, CMBPhotonSpectrum)]
        ebl = [m for m in field.model_list if isinstance(m, SyntheticEBL)]
        if len(cmb) == 1 and len(ebl) <= 1 and len(field.model_list) == len(cmb) + len(ebl):
            return cmb[0], (ebl[0] if ebl else None)
    raise NotImplementedError("device cascade builds need a CMB [+ SyntheticEBL] photon field")


class DeviceCascade:
    """Field-free cascade tensors resident on one device, and the window builds."""

    def __init__(self, kernels, device):
        torch = _torch()
        self.lib = _lib.load()
        self.device = device
        ck = kernels
        n = ck.n
        if n > 144:
            raise NotImplementedError("the device cascade builder takes at most 144 EM bins")
        self.n = n
        self.d = ck.E_cr.size
        eps = C.EPS_CASCADE
        we = C.trapz_weights(eps)
        # contraction rows: emission [n*n] | pair [n*n] | IC power [n] | IC rate [n] | sum we n | sum we eps n
        power = np.einsum("jm,jmk->jk", ck.w1 * ck.E1, ck.B_loss)
        rate = np.einsum("jm,jmk->jk", ck.w1, ck.B_loss)
        kern_e = np.concatenate([
            ck.A_emis.reshape(n * n, eps.size), ck.A_pair.reshape(n * n, eps.size),
            power, rate, we[None, :], (we * eps)[None, :],
        ])
        self.rows_e = kern_e.shape[0]
        dv = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dt)  # noqa: E731
        self.kern_e = dv(kern_e)
        self.kern_o = dv(ck.K_tau)
        self.kern_b = dv(ck.K_bh.reshape(n * self.d, C.EPS_BH.size))
        # degraded-electron cloud-in-cell map with -1 outside the grid
        b = np.where(ck._cic_valid, ck._cic_b, -1).astype(np.int32)
        self.cic_b = dv(b, torch.int32)
        self.cic_f = dv(ck._cic_f)
        self.E = dv(ck.E)
        self.dE = dv(ck.dE)
        self.pref_emis = dv(ck.pref_emis)
        self.pref_pair = dv(ck.pref_pair)
        self.Ecr = dv(ck.E_cr)
        self.eps_b = dv(C.EPS_BH)
        self.grids = {"e": C.EPS_CASCADE, "o": C.EPS_OPACITY, "b": C.EPS_BH}
        self.grids_dev = {k: dv(v) for k, v in self.grids.items()}
        self.ones = {k: dv(np.ones_like(v)) for k, v in self.grids.items()}
        self.cs = _lib.PrhCascade(
            n=n, d=self.d, nb=C.EPS_BH.size, max_scatter=C.MAX_SCATTER, tol=C.CASCADE_TOL,
            cool_floor=C.COOL_FLOOR_FRAC, bh_emin=C.BH_E_P_MIN,
            E=self.E.data_ptr(), dE=self.dE.data_ptr(), pref_emis=self.pref_emis.data_ptr(),
            pref_pair=self.pref_pair.data_ptr(), cic_b=self.cic_b.data_ptr(),
            cic_f=self.cic_f.data_ptr(), Ecr=self.Ecr.data_ptr(), eps_b=self.eps_b.data_ptr(),
            kern_b=self.kern_b.data_ptr(),
        )
        self._ebl_cache = {}

    def _ebl_z0(self, ebl):
        """EBL densities at z = 0 on the three photon grids (scale 1)."""
        key = None if ebl is None else (tuple(ebl.amplitudes),)
        if key not in self._ebl_cache:
            torch = _torch()
            out = {}
            for k, g in self.grids.items():
                v = np.zeros_like(g) if ebl is None else ebl.density_z0(g)
                out[k] = torch.from_numpy(v).to(self.device)
            self._ebl_cache[key] = out
        return self._ebl_cache[key]

    def _densities(self, key, zt, st, e0, stream):
        torch = _torch()
        g = self.grids_dev[key]
        out = torch.empty((zt.numel(), g.numel()), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.prh_field_tables(
            zt.data_ptr(), st.data_ptr(), zt.numel(), g.data_ptr(), self.ones[key].data_ptr(),
            e0[key].data_ptr(), g.numel(), 1, PLANCK_PREF, mconfig.E_CMB,
            SyntheticEBL.Z_POWER, out.data_ptr(), stream), "cascade photon densities")
        return out

    def build(self, zs, field, dz, scales=None, out=None, batch=2048, stream=None):
        """T_gamma, T_electron [nw][n][n], R_BH [nw][n][d] and the iteration
        counts [nw] for windows at redshifts ``zs`` (EBL scaled by ``scales``
        per window, default 1) and step |dz|.  ``out`` may hold preallocated
        (tg, te, rbh, iters) device tensors."""
        torch = _torch()
        _, ebl = field_parts(field)
        zs = np.asarray(zs, dtype=np.float64).reshape(-1)
        nw = zs.size
        s = np.ones(nw) if scales is None else np.asarray(scales, dtype=np.float64).reshape(-1)
        ebl_scale = 0.0 if ebl is None else ebl.scale
        dev = self.device
        stream = torch.cuda.current_stream(dev).cuda_stream if stream is None else stream
        n, d = self.n, self.d
        if out is None:
            tg = torch.empty((nw, n, n), dtype=torch.float64, device=dev)
            te = torch.empty_like(tg)
            rbh = torch.empty((nw, n, d), dtype=torch.float64, device=dev)
            iters = torch.empty(nw, dtype=torch.int32, device=dev)
        else:
            tg, te, rbh, iters = out
        e0 = self._ebl_z0(ebl)
        dl = PRINCE_UNITS.c * abs(dz) / ((1.0 + zs) * np.array([H(z) for z in zs]))
        bsz = max(1, min(batch, nw))
        work = torch.empty(bsz * WORK_PER_WINDOW // 8, dtype=torch.float64, device=dev)
        contr = torch.empty((bsz, self.rows_e), dtype=torch.float64, device=dev)
        dtau = torch.empty((bsz, n), dtype=torch.float64, device=dev)
        for b0 in range(0, nw, bsz):
            b1 = min(nw, b0 + bsz)
            m = b1 - b0
            zt = torch.from_numpy(zs[b0:b1].copy()).to(dev)
            st = torch.from_numpy(s[b0:b1] * ebl_scale).to(dev)
            dlt = torch.from_numpy(dl[b0:b1].copy()).to(dev)
            ne = self._densities("e", zt, st, e0, stream)
            no = self._densities("o", zt, st, e0, stream)
            nb = self._densities("b", zt, st, e0, stream)
            _lib.check(self.lib.prh_fold_gemm(
                self.kern_e.data_ptr(), self.rows_e, ne.shape[1], ne.data_ptr(), m,
                contr.data_ptr(), self.rows_e, stream), "cascade contraction")
            _lib.check(self.lib.prh_fold_gemm(
                self.kern_o.data_ptr(), n, no.shape[1], no.data_ptr(), m, dtau.data_ptr(), n,
                stream), "cascade opacity")
            _lib.check(self.lib.prh_cascade_transfer(
                ctypes.byref(self.cs), contr.data_ptr(), self.rows_e, dtau.data_ptr(), dlt.data_ptr(), m,
                work.data_ptr(), work.numel() * 8, tg[b0:b1].data_ptr(), te[b0:b1].data_ptr(),
                iters[b0:b1].data_ptr(), stream), "cascade transfer")
            _lib.check(self.lib.prh_cascade_bh(ctypes.byref(self.cs), nb.data_ptr(), m, rbh[b0:b1].data_ptr(),
                                               stream), "Bethe-Heitler shapes")
        return tg, te, rbh, iters
