"""Batched steps that build their tables again on the device (FE_BATCH_REBUILD).

With ``set_density`` the step runs, per conformer, the table build from the
batch's density and current points, then forward + field + loss + dL/dP; the
whole chain is one recorded hipGraph.  Checked against tables built afresh
from the same densities and points, evaluated one by one on the single-table
path (itself held to the FP64 oracle): the kept pairs per class equal, V and
the field within 1e-11 a.u., dL/dP within 1e-10 of its largest entry; replays
equal the eager step bit for bit, also after new points and new densities.
Edge cases: a conformer with no points, one with a single point.
"""
import numpy as np
import pytest

from fastesp_amd import synth
from fastesp_amd.autograd import esp_backward, esp_with_grad
from fastesp_amd.batch import Batch
from fastesp_amd.builder import build_table_hip

pytestmark = pytest.mark.gpu

SHAPES = [(10, 700), (14, 1), (22, 2500), (12, 0), (30, 1800), (16, 513)]


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def perturbed(P, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(0.0, 0.02, P.shape)
    return np.ascontiguousarray(0.9 * P + (A + A.T))


def tables_for(ws, dens, pts):
    return [build_table_hip(w.wfn, w.basis, P, p if len(p) else np.zeros((0, 3)), tol=1e-12)
            for w, P, p in zip(ws, dens, pts)]


def check_against_fresh(b, ws, dens, pts):
    fresh = tables_for(ws, dens, pts)
    V, F, gV, dP = b.values(), b.field(), b.gV(), b.dldp()
    for c, (t, p) in enumerate(zip(fresh, pts)):
        assert b.tables[c].info()["pairs_per_L"] == t.info()["pairs_per_L"], c
        if not len(p):
            assert V[c].shape == (0,)
            continue
        Vs, Fs = esp_with_grad(t, p)
        assert np.abs(V[c] - Vs).max() < 1e-11, c
        assert np.abs(F[c] - Fs).max() < 1e-11, c
        _, dPs = esp_backward(t, p, gV[c], want_dG=False, n_func=w_nfunc(ws[c]))
        scale = max(np.abs(dPs).max(), 1e-300)
        assert np.abs(dP[c] - dPs).max() <= 1e-10 * scale, c


def w_nfunc(w):
    return w.basis.n


@pytest.fixture(scope="module")
def setup():
    ws = [synth.conformer(i, n_atoms=a, n_points=n) for i, (a, n) in enumerate(SHAPES)]
    pts = [w.points for w in ws]
    say("build the batch's own tables")
    tabs = tables_for(ws, [w.density for w in ws], pts)
    b = Batch(tabs, pts)
    rng = np.random.default_rng(3)
    b.set_loss([rng.uniform(0.5, 1.0, len(p)) for p in pts],
               [rng.normal(0.0, 0.05, len(p)) for p in pts])
    return ws, pts, b


def test_eager_rebuild_step_matches_fresh_tables(setup):
    ws, pts, b = setup
    dens = [perturbed(w.density, 10 + c) for c, w in enumerate(ws)]
    say("set_density (first call: capacity layouts) + eager rebuild step")
    b.set_density(dens, tol=1e-12)
    b.step("fp64", field=True, dldp=True, rebuild=True)
    check_against_fresh(b, ws, dens, pts)


def test_recorded_rebuild_step_replays_and_takes_new_inputs(setup):
    ws, pts, b = setup
    dens = [perturbed(w.density, 20 + c) for c, w in enumerate(ws)]
    b.set_density(dens, tol=1e-12)
    b.step("fp64", field=True, dldp=True, rebuild=True)
    ref = (b.values(), b.field(), b.dldp(), b.losses())
    say("capture build -> forward -> loss -> backward")
    b.capture("fp64", field=True, dldp=True, rebuild=True)
    other = build_table_hip(ws[0].wfn, ws[0].basis, ws[0].density, ws[0].points, tol=1e-12)
    for r in range(3):
        say(f"replay {r}")
        b.launch()
        got = (b.values(), b.field(), b.dldp(), b.losses())
        for x, y in zip(got, ref):
            for u, v in zip(x if isinstance(x, list) else [x], y if isinstance(y, list) else [y]):
                np.testing.assert_array_equal(u, v)
        other.esp(ws[0].points[:64])  # unrelated eager work between replays
    # new points and new densities through the recorded buffers
    rng = np.random.default_rng(5)
    moved = [p + rng.normal(0, 0.05, p.shape) for p in pts]
    dens2 = [perturbed(w.density, 30 + c) for c, w in enumerate(ws)]
    b.set_points(moved)
    b.set_density(dens2, tol=1e-12)
    say("replay after set_points + set_density")
    b.launch()
    Vg = b.values()
    check_against_fresh(b, ws, dens2, moved)
    b.step("fp64", field=True, dldp=True, rebuild=True)
    for u, v in zip(Vg, b.values()):
        np.testing.assert_array_equal(u, v)


def test_mixed_rebuild_step_within_bar(setup):
    ws, pts, b = setup
    dens = [perturbed(w.density, 40 + c) for c, w in enumerate(ws)]
    b.set_density(dens, tol=1e-12)
    say("mixed rebuild step against the FP64 rebuild step")
    b.step("fp64", field=True, dldp=False, rebuild=True)
    V64 = b.values()
    b.capture("mixed", field=True, dldp=False, rebuild=True)
    b.launch()
    for u, v in zip(b.values(), V64):
        if len(v):
            assert np.abs(u - v).max() < 1e-8


def test_rebuild_flag_needs_densities():
    w = synth.conformer(0, n_atoms=10, n_points=300)
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    b = Batch([t], [w.points])
    with pytest.raises(RuntimeError, match="fe_batch_set_density"):
        b.step("fp64", rebuild=True)
    with pytest.raises(ValueError):
        b.set_density([np.zeros((3, 3))])
