"""fe_esp_forward_noalloc (the forward that allocates nothing) against
fe_esp_forward on the same device buffers: two calls in a row, the second with
fewer points, FP64 and mixed, plus its size and state errors."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import have_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    if not have_gpu():
        pytest.skip("no GPU")
    import torch
    from fastesp_amd import _lib, synth
    from fastesp_amd.table import build_table
    ctx = _lib.Context.get(0)
    w = synth.config0(1000)
    tab = build_table(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    return torch, _lib, ctx, w, tab


def _run(lib, fn, tab, pts, n, prec, V, sh):
    if fn == "fe_esp_forward":
        return lib.fe_esp_forward(tab._h, pts.data_ptr(), n, prec, 3, 1, sh, V.data_ptr())
    return lib.fe_esp_forward_noalloc(tab._h, pts.data_ptr(), n, prec, 3, sh, V.data_ptr())


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_forward_two_calls(setup, prec):
    torch, _lib, ctx, w, tab = setup
    lib = ctx.lib
    dev = torch.device("cuda:0")
    sh = torch.cuda.current_stream(dev).cuda_stream
    pts = torch.from_numpy(np.ascontiguousarray(w.points)).to(dev)
    if prec == 1:
        # the FP32 coefficients are built by the first mixed forward
        V0 = torch.empty(len(w.points), dtype=torch.float64, device=dev)
        _lib.check(_run(lib, "fe_esp_forward", tab, pts, len(w.points), prec, V0, sh), "fwd")
    for n in (1000, 700):
        print(f"noalloc: config0, {n} points, precision {prec}", flush=True)
        p = pts[:n].contiguous()
        Va = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
        Vb = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
        _lib.check(_run(lib, "fe_esp_forward", tab, p, n, prec, Va, sh), "fe_esp_forward")
        _lib.check(_run(lib, "fe_esp_forward_noalloc", tab, p, n, prec, Vb, sh),
                   "fe_esp_forward_noalloc")
        torch.cuda.synchronize()
        a, b = Va.cpu().numpy(), Vb.cpu().numpy()
        assert np.isfinite(b).all()
        assert np.abs(a - b).max() <= 1e-12, np.abs(a - b).max()


def test_noalloc_rejects_too_many_points(setup):
    torch, _lib, ctx, w, tab = setup
    lib = ctx.lib
    n = _lib.FE_NOALLOC_MAX_POINTS + 1
    dev = torch.device("cuda:0")
    pts = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    V = torch.empty(n, dtype=torch.float64, device=dev)
    rc = lib.fe_esp_forward_noalloc(tab._h, pts.data_ptr(), n, 0, 3, None, V.data_ptr())
    assert rc == -1
    assert b"at most" in lib.fe_last_error()


def test_noalloc_mixed_needs_fp32_table(setup):
    torch, _lib, ctx, w, _ = setup
    from fastesp_amd.table import build_table
    fresh = build_table(w.wfn, w.basis, w.density, w.points[:600], tol=1e-12)
    lib = ctx.lib
    dev = torch.device("cuda:0")
    pts = torch.from_numpy(np.ascontiguousarray(w.points[:600])).to(dev)
    V = torch.empty(600, dtype=torch.float64, device=dev)
    rc = lib.fe_esp_forward_noalloc(fresh._h, pts.data_ptr(), 600, 1, 3, None, V.data_ptr())
    assert rc == -6
    assert b"FP32" in lib.fe_last_error()
