"""GPU unit checks of the C-ABI kernels (fold GEMM, row gather)."""

import numpy as np
import pytest

from prince_amd import _lib

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


@pytest.mark.parametrize("nrows,dph,nwin", [(37, 72, 5), (1024, 72, 288), (16, 4, 16), (33, 70, 17)])
def test_fold_gemm_matches_numpy(nrows, dph, nwin):
    torch = _torch()
    lib = _lib.load()
    rng = np.random.default_rng(nrows + dph + nwin)
    B = rng.standard_normal((nrows, dph))
    N = rng.standard_normal((nwin, dph))
    stride = nrows + 3
    dev = torch.device("cuda", 0)
    Bd = torch.from_numpy(B).to(dev)
    Nd = torch.from_numpy(N).to(dev)
    out = torch.full((nwin, stride), -7.0, dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.prh_fold_gemm(Bd.data_ptr(), nrows, dph, Nd.data_ptr(), nwin,
                                 out.data_ptr(), stride, s), "fold")
    got = out.cpu().numpy()
    ref = N @ B.T
    assert np.allclose(got[:, :nrows], ref, rtol=1e-13, atol=1e-12)
    assert np.all(got[:, nrows:] == -7.0)


def test_gather_rows():
    torch = _torch()
    lib = _lib.load()
    rng = np.random.default_rng(3)
    src = rng.standard_normal((50, 72))
    sel = np.array([3, -1, 49, 0, -1, 7], dtype=np.int64)
    dev = torch.device("cuda", 0)
    srcd = torch.from_numpy(src).to(dev)
    seld = torch.from_numpy(sel).to(dev)
    dst = torch.empty((sel.size, 72), dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.prh_gather_rows(srcd.data_ptr(), seld.data_ptr(), sel.size, 72,
                                   dst.data_ptr(), s), "gather")
    got = dst.cpu().numpy()
    ref = np.where(sel[:, None] >= 0, src[np.maximum(sel, 0)], 0.0)
    assert np.array_equal(got, ref)


def test_library_error_path():
    lib = _lib.load()
    st = lib.prh_fold_gemm(None, 10, 0, None, 1, None, 10, None)
    assert st != 0
    assert b"dph" in lib.prh_last_error()
