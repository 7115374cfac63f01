"""GPU: EM-cascade window operators built on the device (prh_cascade_transfer,
prh_cascade_bh, the large-K contraction of prh_fold_gemm) against the host
builders of model/cascade.py, window by window."""

import numpy as np
import pytest

from prince_amd import workloads
from prince_amd.model import cascade as C
from prince_amd.model.photonfields import CMBPhotonSpectrum, CombinedPhotonField, SyntheticEBL
from prince_amd.solvers.cascade_device import DeviceCascade

pytestmark = pytest.mark.gpu


def _field_with_scale(field, s):
    if isinstance(field, CMBPhotonSpectrum):
        return field
    ebl = [m for m in field.model_list if isinstance(m, SyntheticEBL)][0]
    return CombinedPhotonField([CMBPhotonSpectrum(),
                                SyntheticEBL(amplitudes=ebl.amplitudes, scale=ebl.scale * s)])


def _close(dev, host, rtol):
    scale = np.abs(host).max()
    return float(np.abs(dev - host).max() / scale)


@pytest.mark.parametrize("which", ["proton_em", "nuclear_em"])
def test_device_cascade_matches_host_builders(which):
    import torch

    run = workloads.proton_em_run() if which == "proton_em" else workloads.nuclear_em_run()
    ck = C.cascade_kernels(run.em_grid.grid, run.cr_grid.grid)
    dc = DeviceCascade(ck, torch.device("cuda", 0))
    zs = np.array([0.0, 0.37, 1.2, 3.0, 4.9])
    scales = np.array([1.0, 0.6, 1.4, 1.0, 0.8])
    dz = 1e-3
    tg, te, rbh, iters = dc.build(zs, run.photon_field, dz, scales=scales, batch=3)
    torch.cuda.synchronize()
    tg, te, rbh, iters = (t.cpu().numpy() for t in (tg, te, rbh, iters))
    worst = {"tg": 0.0, "te": 0.0, "rbh": 0.0}
    for w, (z, s) in enumerate(zip(zs, scales)):
        f = _field_with_scale(run.photon_field, s)
        Tg, Te, it = ck.transfer_one(z, f, dz)
        R = ck.bh_shape_one(z, f)
        assert iters[w] == it, (w, iters[w], it)
        worst["tg"] = max(worst["tg"], _close(tg[w], Tg, 0))
        worst["te"] = max(worst["te"], _close(te[w], Te, 0))
        if np.abs(R).max() > 0:
            worst["rbh"] = max(worst["rbh"], _close(rbh[w], R, 0))
        else:
            assert not np.any(rbh[w])
    print(which, "max deviation relative to each matrix's largest entry:", worst)
    assert worst["tg"] < 1e-11 and worst["te"] < 1e-11 and worst["rbh"] < 1e-11


def test_large_k_contraction_matches_numpy():
    import torch

    from prince_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(5)
    for rows, kdim, nvec in ((37, 400, 5), (100, 129, 33), (16, 256, 16)):
        kern = rng.normal(size=(rows, kdim))
        vec = rng.normal(size=(nvec, kdim))
        kt = torch.from_numpy(kern).cuda()
        vt = torch.from_numpy(vec).cuda()
        out = torch.full((nvec, rows + 3), 7.0, dtype=torch.float64, device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.prh_fold_gemm(kt.data_ptr(), rows, kdim, vt.data_ptr(), nvec,
                                     out.data_ptr(), rows + 3, stream), "contract")
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.allclose(got[:, :rows], vec @ kern.T, rtol=1e-12, atol=1e-12)
        assert np.all(got[:, rows:] == 7.0)
