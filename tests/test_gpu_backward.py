"""GPU parity of the adjoint: field dV/dC, dL/dG and dL/dP.

Pinned against the oracle's analytic adjoint (itself pinned by central finite
differences in tests/test_oracle_table.py), against central finite
differences of the GPU forward, and at full size through identities that do
not depend on the size: V is linear and homogeneous in P (and in G), so
<dL/dP, P> = <dL/dG, G> = sum_i gV_i (V_i - nuc_i) exactly.
"""
from pathlib import Path

import numpy as np
import pytest

from fastesp_amd import synth
from fastesp_amd.autograd import ESPModel, esp_backward, esp_with_grad
from fastesp_amd.builder import build_table_hip, export_flat
from oracle import engine, evaluate
from tests.helpers import cf_toy

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_field_and_dP_match_oracle_on_toy():
    g = np.load(GOLD / "grad_toy.npz")
    wfn, basis, P, _ = cf_toy()
    pts = g["pts"]
    say("toy forward_grad")
    t = build_table_hip(wfn, basis, P, pts, tol=0.0)
    V, G = esp_with_grad(t, pts)
    assert np.abs(V - g["V"]).max() < 1e-10
    assert np.abs(G - g["field"]).max() < 1e-9
    say("toy backward")
    dG, dP = esp_backward(t, pts, g["gV"], n_func=basis.n)
    assert np.abs(dP - g["dP"]).max() < 1e-9 * max(1.0, np.abs(g["dP"]).max())
    np.testing.assert_allclose(dP, dP.T, rtol=0, atol=1e-13)
    # dL/dG against the oracle on the same (exported) table
    tab = engine.build_table(wfn, basis, P, pts, tol=0.0)
    dG_ref = evaluate.grad_G(tab, pts, g["gV"])
    assert np.abs(dG - dG_ref).max() < 1e-9 * max(1.0, np.abs(dG_ref).max())


def test_field_matches_finite_differences_of_gpu_forward():
    w = synth.config0(200)
    say("cfg0 field vs FD")
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    V, G = esp_with_grad(t, w.points)
    np.testing.assert_allclose(V, t.esp(), rtol=0, atol=1e-12)
    h = 1e-4
    for ax in range(3):
        d = np.zeros(3)
        d[ax] = h
        fd = (t.esp(w.points + d) - t.esp(w.points - d)) / (2 * h)
        assert np.abs(G[:, ax] - fd).max() < 2e-7 * max(1.0, np.abs(fd).max())


def test_autograd_function_matches_oracle_gradients():
    import torch
    g = np.load(GOLD / "grad_toy.npz")
    wfn, basis, P, _ = cf_toy()
    dev = torch.device("cuda:0")
    pts = torch.tensor(g["pts"], dtype=torch.float64, device=dev, requires_grad=True)
    dens = torch.tensor(P, dtype=torch.float64, device=dev, requires_grad=True)
    model = ESPModel(wfn, basis, tol=0.0)
    say("autograd toy")
    V = model(pts, dens)
    w = torch.tensor(g["w"], device=dev)
    target = torch.tensor(g["target"], device=dev)
    loss = (w * (V - target) ** 2).sum()
    loss.backward()
    assert np.abs(V.detach().cpu().numpy() - g["V"]).max() < 1e-10
    gpts = pts.grad.cpu().numpy()
    np.testing.assert_allclose(gpts, g["gV"][:, None] * g["field"], rtol=1e-8, atol=1e-12)
    assert np.abs(dens.grad.cpu().numpy() - g["dP"]).max() < 1e-9 * max(1.0, np.abs(g["dP"]).max())


def test_homogeneity_identities_at_full_size():
    """<dL/dP, P> = <dL/dG, G> = sum gV (V - nuc) at 30000 points."""
    w = synth.config1(30000)
    rng = np.random.default_rng(5)
    say("cfg1 30000 build")
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    V = t.esp()
    nuc = t.nuclear()
    gV = rng.uniform(0.5, 1.0, len(V)) * 2.0 * rng.normal(0, 1e-3, len(V))
    say("cfg1 30000 backward")
    dG, dP = esp_backward(t, w.points, gV, n_func=w.basis.n)
    ref = float((gV * (V - nuc)).sum())
    flat = export_flat(t)
    lhs_G = float((dG * flat.tm_g).sum())
    lhs_P = float((dP * w.density).sum())
    assert lhs_G == pytest.approx(ref, rel=1e-9, abs=1e-12)
    assert lhs_P == pytest.approx(ref, rel=1e-9, abs=1e-12)
    np.testing.assert_allclose(dP, dP.T, rtol=0, atol=1e-12 * np.abs(dP).max())


@pytest.mark.skipif(not (GOLD / "cfg2_sub.npz").exists(), reason="cfg2_sub fixture not generated")
def test_config2_subset_matches_oracle():
    g = np.load(GOLD / "cfg2_sub.npz")
    w = synth.config1(30000)
    np.testing.assert_allclose(w.fingerprint(), g["fp"], rtol=1e-9)
    pts = w.points[g["idx"]]
    say("cfg2 subset build")
    t = build_table_hip(w.wfn, w.basis, w.density, pts, tol=1e-12)
    assert t.n_pairs == int(g["npairs"])
    V, G = esp_with_grad(t, pts)
    assert np.abs(V - g["V"]).max() < 1e-10
    assert np.abs(G[:64] - g["field64"]).max() < 1e-9
    say("cfg2 subset backward")
    _, dP = esp_backward(t, pts, g["gV"], want_dG=False, n_func=w.basis.n)
    scale = float(g["dP_absmax"])  # relative to the largest entry
    ij = g["dP_ij"]
    assert np.abs(dP[ij[:, 0], ij[:, 1]] - g["dP_sample"]).max() < 1e-9 * scale
    assert np.abs(np.diag(dP) - g["dP_diag"]).max() < 1e-9 * scale
    assert np.abs(dP.sum(axis=1) - g["dP_rowsum"]).max() < 1e-8 * scale
    assert abs(np.abs(dP).max() - float(g["dP_absmax"])) < 1e-9 * scale


def test_phase_profile_counts_every_class_launch():
    from fastesp_amd import _lib as L_
    from fastesp_amd.builder import build_table_hip as bth
    ctx = L_.Context.get(0)
    w = synth.config0(1500)
    t = bth(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    present = [L for L, c in enumerate(t.info()["pairs_per_L"]) if c]
    say(f"profile phases, classes {present}")
    ctx.profile_start()
    V, _ = esp_with_grad(t, w.points)
    esp_backward(t, w.points, np.ones_like(V), want_dG=False)
    ph = ctx.profile_stop()
    for L in present:
        assert ph[8 + L][1] == 1 and ph[8 + L][0] > 0
        assert ph[16 + L][1] == 1 and ph[24 + L][1] == 1
    # dL/dP runs once per class that has shell pairs (each class's shell pairs
    # start as soon as that class's dL/dG is reduced)
    ls = sorted({s.l for s in w.wfn.shells})
    sp_classes = {a + b for a in ls for b in ls}
    assert ph[1][1] == 1 and ph[2][1] == 1 and ph[3][1] == len(sp_classes)
    ctx.profile_start()
    assert ctx.profile_stop() == {}


def test_backward_without_points_is_zero():
    """No points: dL/dG = 0 and dL/dP = 0 (every output written)."""
    wfn, basis, P, _ = cf_toy()
    t = build_table_hip(wfn, basis, P, np.zeros((4, 3)) + 3.0, tol=0.0)
    dG, dP = esp_backward(t, np.zeros((0, 3)), np.zeros(0), n_func=basis.n)
    assert not dG.any() and not dP.any()
    V, F = esp_with_grad(t, np.zeros((0, 3)))
    assert V.shape == (0,) and F.shape == (0, 3)


def test_backward_without_points_accepts_null_arrays():
    """n = 0 with NULL points and weights, host outputs: nothing is read from
    them and dL/dG, dL/dP come back zero."""
    import ctypes as C
    from fastesp_amd import _lib
    wfn, basis, P, _ = cf_toy()
    t = build_table_hip(wfn, basis, P, np.zeros((4, 3)) + 3.0, tol=0.0)
    n_terms = t.n_terms
    dG = np.full(n_terms, np.nan)
    dP = np.full((basis.n, basis.n), np.nan)
    lib = _lib.load()
    say("backward n=0 with NULL arrays")
    rc = lib.fe_esp_backward(t._h, None, 0, None, _lib.FE_PREC_FP64, _lib.FE_PTR_HOST, None,
                             dG.ctypes.data, dP.ctypes.data)
    assert rc == 0, lib.fe_last_error()
    assert not dG.any() and not dP.any()


def test_trim_releases_workspace_and_next_call_works():
    from fastesp_amd import _lib
    wfn, basis, P, _ = cf_toy()
    pts = np.array([[2.0, 1.0, 3.0], [-2.5, 0.5, 1.0]])
    t = build_table_hip(wfn, basis, P, pts, tol=0.0)
    V0, F0 = esp_with_grad(t, pts)
    say("fe_ctx_trim")
    _lib.Context.get().trim()
    V1, F1 = esp_with_grad(t, pts)
    assert np.array_equal(V0, V1) and np.array_equal(F0, F1)


def test_stream_modes_agree():
    """The dL/dG classes on two streams (default), three, or one at a time
    give the same bits; the evaluation classes on two streams (tuning key 7,
    own partial rows) change only the summation order."""
    from fastesp_amd import _lib
    ctx = _lib.Context.get(0)
    w = synth.config0(1500)
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    gV = np.linspace(-1.0, 1.0, len(w.points))
    out = {}
    try:
        for dg2, ev2 in ((1, 0), (0, 0), (3, 0), (1, 1)):
            say(f"dL/dG tuning {dg2}, evaluation streams {1 + ev2}")
            ctx.set_tuning(6, dg2)
            ctx.set_tuning(7, ev2)
            V, F = esp_with_grad(t, w.points)
            dG, dP = esp_backward(t, w.points, gV, want_dG=True, n_func=w.basis.n)
            out[(dg2, ev2)] = (V, F, dG, dP)
    finally:
        ctx.set_tuning(6, 1)
        ctx.set_tuning(7, 2)
    for a, b, c in zip(out[(1, 0)], out[(0, 0)], out[(3, 0)]):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    V0, F0, dG0, dP0 = out[(1, 0)]
    V2, F2, dG2, dP2 = out[(1, 1)]
    assert np.abs(V2 - V0).max() < 1e-12 and np.abs(F2 - F0).max() < 1e-12
    assert np.array_equal(dG2, dG0) and np.array_equal(dP2, dP0)
