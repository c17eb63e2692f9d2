"""Batched conformers (fe_batch_*): parity with the single-table path, which is
itself held to the FP64 oracle, plus one conformer against the oracle directly.

Bars: FP64 batched vs single-table V and field within 1e-11 a.u. (the two
differ only in summation order), dL/dP within 1e-10 of the single-table
adjoint relative to its largest entry, mixed within 1e-8 a.u. of FP64; a
replayed hipGraph equals the eager step bit for bit, also after new
coordinates are loaded into the recorded buffers.
Edge cases: a conformer with no points, one with a single point, uneven
sizes, every class L present.
"""
import numpy as np
import pytest

from fastesp_amd import _lib, synth
from fastesp_amd.autograd import esp_backward, esp_with_grad
from fastesp_amd.batch import Batch
from fastesp_amd.builder import build_table_hip
from oracle import engine, evaluate

pytestmark = pytest.mark.gpu

# (atoms, points) of the conformers: uneven, with the empty and 1-point cases
SHAPES = [(10, 700), (14, 1), (22, 2500), (12, 0), (30, 1800), (16, 513)]


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def maxabs(a) -> float:
    return float(np.abs(a).max()) if np.size(a) else 0.0


@pytest.fixture(scope="module")
def confs():
    ws = [synth.conformer(i, n_atoms=a, n_points=n) for i, (a, n) in enumerate(SHAPES)]
    tabs = []
    for w in ws:
        say(f"build conformer atoms={w.meta['n_atoms']} fns={w.meta['n_func']} pts={len(w.points)}")
        tabs.append(build_table_hip(w.wfn, w.basis, w.density,
                                    w.points if len(w.points) else np.zeros((1, 3)), tol=1e-12))
    return ws, tabs


def loss_data(ws, seed=7):
    rng = np.random.default_rng(seed)
    w = [rng.uniform(0.5, 1.0, len(x.points)) for x in ws]
    T = [rng.normal(0.0, 0.05, len(x.points)) for x in ws]
    return w, T


def test_batch_forward_matches_single_tables(confs):
    ws, tabs = confs
    b = Batch(tabs, [w.points for w in ws])
    say("batch fp64 forward")
    b.forward("fp64")
    for w, t, V in zip(ws, tabs, b.values()):
        assert V.shape == (len(w.points),)
        if len(w.points):
            assert np.abs(V - t.esp(w.points)).max() < 1e-11
    info = b.info()
    assert info["n_conf"] == len(SHAPES) and info["n_points"] == sum(n for _, n in SHAPES)


def test_batch_conformer_against_oracle(confs):
    ws, tabs = confs
    w = ws[0]
    b = Batch(tabs, [x.points for x in ws])
    b.forward("fp64")
    tab = engine.build_table(w.wfn, w.basis, w.density, w.points, 1e-12)
    say(f"oracle conformer 0: {tab.n_pairs} pairs")
    assert tabs[0].n_pairs == tab.n_pairs
    assert np.abs(b.values()[0] - evaluate.esp(tab, w.points)).max() < 1e-10


def test_batch_step_matches_single_table_adjoint(confs):
    ws, tabs = confs
    b = Batch(tabs, [w.points for w in ws])
    wts, T = loss_data(ws)
    b.set_loss(wts, T)
    say("batch fp64 step (field + dL/dP)")
    b.step("fp64", field=True, dldp=True)
    Vs, Fs, gVs, dPs, losses = b.values(), b.field(), b.gV(), b.dldp(), b.losses()
    for i, (w, t) in enumerate(zip(ws, tabs)):
        n = len(w.points)
        if n == 0:
            assert losses[i] == 0.0
            assert np.abs(dPs[i]).max() == 0.0
            continue
        V, G = esp_with_grad(t, w.points)
        assert np.abs(Vs[i] - V).max() < 1e-11
        assert np.abs(Fs[i] - G).max() < 1e-11
        d = Vs[i] - T[i]
        np.testing.assert_allclose(gVs[i], 2 * wts[i] * d, rtol=1e-14, atol=0)
        assert losses[i] == pytest.approx(float(np.sum(wts[i] * d * d)), rel=1e-12)
        _, dP = esp_backward(t, w.points, gVs[i], want_dG=False)
        scale = np.abs(dP).max()
        assert np.abs(dPs[i] - dP).max() <= 1e-10 * scale
        assert np.abs(dPs[i] - dPs[i].T).max() <= 1e-12 * scale


@pytest.mark.parametrize("fp32_classes", ["default", "all"])
def test_batch_mixed_within_bar(confs, fp32_classes):
    ws, tabs = confs
    b = Batch(tabs, [w.points for w in ws])
    if fp32_classes == "all":
        for L in range(7):
            b.set_tuning(_lib.FE_BT_MIXED0 + L, 1)
    say(f"batch mixed forward + field, FP32 classes: {fp32_classes}")
    b.forward("mixed", field=True)
    for w, t, V, G in zip(ws, tabs, b.values(), b.field()):
        if len(w.points):
            V64, G64 = esp_with_grad(t, w.points)
            assert np.abs(V - V64).max() < 1e-8
            assert np.abs(G - G64).max() < 1e-7


def test_graph_replay_is_bitwise_and_takes_new_points(confs):
    ws, tabs = confs
    pts = [w.points for w in ws]
    b = Batch(tabs, pts)
    wts, T = loss_data(ws, seed=11)
    b.set_loss(wts, T)
    b.step("mixed", field=True, dldp=True)
    ref = (b.values(), b.field(), b.dldp(), b.losses())
    say("capture mixed step")
    b.capture("mixed", field=True, dldp=True)
    assert b.info()["captured"] == 1
    for r in range(3):
        say(f"replay {r}")
        b.launch()
        got = (b.values(), b.field(), b.dldp(), b.losses())
        for x, y in zip(got, ref):
            for u, v in zip(x if isinstance(x, list) else [x], y if isinstance(y, list) else [y]):
                np.testing.assert_array_equal(u, v)
        tabs[0].esp(ws[0].points[:50])  # unrelated eager work between replays
    # new coordinates through the recorded buffers
    rng = np.random.default_rng(5)
    moved = [p + rng.normal(0, 0.05, p.shape) for p in pts]
    b.set_points(moved)
    say("replay after set_points")
    b.launch()
    Vg = b.values()
    b.step("mixed", field=True, dldp=True)
    for u, v, t, p in zip(Vg, b.values(), tabs, moved):
        np.testing.assert_array_equal(u, v)
        if len(p):
            assert np.abs(u - t.esp(p)).max() < 1e-8


def test_tuning_changes_only_rounding(confs):
    ws, tabs = confs
    b = Batch(tabs, [w.points for w in ws])
    b.forward("fp64", field=True)
    V0, F0 = b.values(), b.field()
    for key, val in ((_lib.FE_BT_ITEMS_PER_WG, 2), (_lib.FE_BT_ITEMS_PER_WG, 32),
                     (_lib.FE_BT_PPL0 + 2, 4), (_lib.FE_BT_PPL0 + 4, 1)):
        say(f"batch tuning key={key} value={val}")
        b.set_tuning(key, val)
        b.forward("fp64", field=True)
        for u, v in zip(b.values(), V0):
            assert maxabs(u - v) < 1e-12
        for u, v in zip(b.field(), F0):
            assert maxabs(u - v) < 1e-12
    with pytest.raises(_lib.HipBackendError):
        b.set_tuning(_lib.FE_BT_PPL0, 3)
    with pytest.raises(_lib.HipBackendError):
        b.launch()  # nothing captured
