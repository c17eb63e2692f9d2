"""The device-resident table build and its in-place rebuild.

* A rebuild for a new density and new points gives the same table as a fresh
  fe_table_build with those inputs (counts, every exported field in reference
  order, sum rule) and V equal to the oracle's.
* A step of rebuild -> forward + field -> backward recorded as a graph replays
  bit for bit like the eager step, for new points loaded between replays
  (every replay is checked, with eager work in between).
* ESPModel reuses its table across calls, and its gradients equal those of a
  model that builds a fresh table per call.
"""
import ctypes as C

import numpy as np
import pytest

from fastesp_amd import _lib, synth
from fastesp_amd.builder import build_table_hip, export_flat
from oracle import engine, evaluate

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def w0():
    return synth.config0(1000)


def _density2(w):
    rng = np.random.default_rng(7)
    A = rng.normal(0.0, 0.02, w.density.shape)
    return np.ascontiguousarray(0.9 * w.density + (A + A.T))


def _flat_equal(a, b):
    for k in ("pp_nt", "pp_off", "pp_nmono", "pp_monooff", "mo_parent", "mo_axis", "tm_aidx",
              "tm_m", "pp_exp", "pp_cx", "pp_cy", "pp_cz", "pp_K", "tm_g"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


@pytest.mark.parametrize("ptr_kind", [_lib.FE_PTR_HOST, _lib.FE_PTR_DEVICE])
def test_rebuild_equals_fresh_build(w0, ptr_kind):
    import torch
    P2 = _density2(w0)
    pts2 = np.ascontiguousarray(w0.points[::3] + 0.05)
    say(f"config0: build, then rebuild (ptr_kind {ptr_kind}) with a new density and points")
    t = build_table_hip(w0.wfn, w0.basis, w0.density, w0.points, tol=1e-12)
    fresh = build_table_hip(w0.wfn, w0.basis, P2, pts2, tol=1e-12)
    lib = _lib.load()
    if ptr_kind == _lib.FE_PTR_HOST:
        rc = lib.fe_table_rebuild(t._h, P2.ctypes.data, len(pts2), pts2.ctypes.data, 1e-12, 0,
                                  ptr_kind, None)
    else:
        dP = torch.from_numpy(P2).cuda()
        dpts = torch.from_numpy(pts2).cuda()
        rc = lib.fe_table_rebuild(t._h, dP.data_ptr(), len(pts2), dpts.data_ptr(), 1e-12, 0,
                                  ptr_kind, torch.cuda.current_stream().cuda_stream)
    assert rc == 0, lib.fe_last_error()
    a, b = t.info(), fresh.info()
    assert a["pairs_per_L"] == b["pairs_per_L"] and a["n_terms"] == b["n_terms"]
    assert a["n_candidates"] == b["n_candidates"]
    assert abs(t.stats()["sum_rule"] - fresh.stats()["sum_rule"]) < 1e-12
    t._flat = None
    _flat_equal(export_flat(t), export_flat(fresh))
    ref = engine.build_table(w0.wfn, w0.basis, P2, pts2, 1e-12)
    assert t.n_pairs == ref.n_pairs
    V = t.esp(pts2)
    assert np.abs(V - evaluate.esp(ref, pts2)).max() < 1e-10
    assert np.abs(V - fresh.esp(pts2)).max() < 1e-13
    # a second rebuild back to the first inputs restores the first table
    say("rebuild back to the original inputs")
    rc = lib.fe_table_rebuild(t._h, w0.density.ctypes.data, len(w0.points),
                              np.ascontiguousarray(w0.points).ctypes.data, 1e-12, 0,
                              _lib.FE_PTR_HOST, None)
    assert rc == 0, lib.fe_last_error()
    first = build_table_hip(w0.wfn, w0.basis, w0.density, w0.points, tol=1e-12)
    t._flat = None
    _flat_equal(export_flat(t), export_flat(first))


def test_rebuild_keeps_mixed_coefficients_current(w0):
    """The FP32 copies of a table follow its rebuilds."""
    P2 = _density2(w0)
    pts = np.ascontiguousarray(w0.points[:600])
    t = build_table_hip(w0.wfn, w0.basis, w0.density, pts, tol=1e-12)
    t.esp(pts, precision=_lib.FE_PREC_MIXED)  # builds the FP32 copies
    say("mixed forward after rebuilds")
    lib = _lib.load()
    for P in (P2, w0.density, P2):
        assert lib.fe_table_rebuild(t._h, np.ascontiguousarray(P).ctypes.data, len(pts),
                                    pts.ctypes.data, 1e-12, 0, _lib.FE_PTR_HOST, None) == 0
        ref = engine.build_table(w0.wfn, w0.basis, P, pts, 1e-12)
        want = evaluate.esp(ref, pts)
        assert np.abs(t.esp(pts, precision=_lib.FE_PREC_MIXED) - want).max() < 1e-8
        assert np.abs(t.esp(pts, precision=_lib.FE_PREC_FP64) - want).max() < 1e-10


def test_rebuild_forward_backward_graph_replays(w0):
    import torch
    dev = torch.device("cuda:0")
    lib = _lib.load()
    n, nf = 900, w0.basis.n
    pts_all = np.ascontiguousarray(w0.points)
    t = build_table_hip(w0.wfn, w0.basis, w0.density, pts_all[:n], tol=1e-12)
    dens = torch.from_numpy(_density2(w0)).to(dev)
    pts = torch.from_numpy(pts_all[:n].copy()).to(dev)
    wts = torch.linspace(0.5, 1.0, n, dtype=torch.float64, device=dev)
    target = torch.zeros(n, dtype=torch.float64, device=dev)
    V = torch.empty(n, dtype=torch.float64, device=dev)
    F = torch.empty((n, 3), dtype=torch.float64, device=dev)
    dP = torch.empty((nf, nf), dtype=torch.float64, device=dev)

    def step():
        sh = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.fe_table_rebuild(t._h, dens.data_ptr(), n, pts.data_ptr(), 1e-12, 0,
                                        _lib.FE_PTR_DEVICE, sh), "rebuild")
        _lib.check(lib.fe_esp_forward_grad(t._h, pts.data_ptr(), n, _lib.FE_PREC_FP64,
                                           _lib.FE_PTR_DEVICE, sh, V.data_ptr(), F.data_ptr()),
                   "forward_grad")
        gV = 2.0 * wts * (V - target)
        _lib.check(lib.fe_esp_backward(t._h, pts.data_ptr(), n, gV.data_ptr(), _lib.FE_PREC_FP64,
                                       _lib.FE_PTR_DEVICE, sh, None, dP.data_ptr()), "backward")

    say("graph: eager warm-up (the first rebuild allocates)")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    say("graph: capture rebuild -> forward + field -> backward")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    rng = np.random.default_rng(11)
    for rep in range(4):
        sel = rng.choice(len(pts_all), n, replace=False)
        new = torch.from_numpy(pts_all[sel] + rng.normal(0, 0.05, (n, 3))).to(dev)
        pts.copy_(new)
        say(f"graph replay {rep}")
        g.replay()
        torch.cuda.synchronize()
        got = (V.clone(), F.clone(), dP.clone())
        step()  # eager, same inputs
        torch.cuda.synchronize()
        assert torch.equal(got[0], V) and torch.equal(got[1], F) and torch.equal(got[2], dP)
        if rep == 0:
            ref = engine.build_table(w0.wfn, w0.basis, dens.cpu().numpy(), new.cpu().numpy(),
                                     1e-12)
            assert np.abs(V.cpu().numpy() - evaluate.esp(ref, new.cpu().numpy())).max() < 1e-10


def test_esp_model_reuses_its_table(w0):
    import torch
    from fastesp_amd.autograd import ESPModel
    dev = torch.device("cuda:0")
    m1 = ESPModel(w0.wfn, w0.basis, tol=1e-12)
    m2 = ESPModel(w0.wfn, w0.basis, tol=1e-12, reuse_table=False)
    for it, sl in enumerate((slice(0, 700), slice(300, 1000), slice(100, 900))):
        say(f"ESPModel call {it}")
        grads = []
        for m in (m1, m2):
            pts = torch.from_numpy(np.ascontiguousarray(w0.points[sl])).to(dev).requires_grad_()
            P = torch.from_numpy(_density2(w0) if it % 2 else w0.density).to(dev).requires_grad_()
            V = m(pts, P)
            (V * V).sum().backward()
            grads.append((V.detach(), pts.grad, P.grad))
        for x, y in zip(*grads):
            assert torch.allclose(x, y, rtol=0, atol=1e-12 * max(1.0, float(y.abs().max())))
    assert m1.last_table is not None and m1.generation == 3
    # two forwards before a backward: the first graph sees a rebuilt table
    pts = torch.from_numpy(np.ascontiguousarray(w0.points[:500])).to(dev).requires_grad_()
    P = torch.from_numpy(w0.density).to(dev)
    Va = m1(pts, P)
    m1(pts, P)
    with pytest.raises(RuntimeError, match="rebuilt"):
        Va.sum().backward()
