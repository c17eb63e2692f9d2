"""The configurations at their real sizes, against the oracle.

* Config 1 at 3000 points: device build + FP64 and mixed forward against the
  C++ restatement's V at every point (tests/golden/cfg1_3k.npz).
* Config 2 on the GPU path: the field against central finite differences of
  the GPU forward (h = 1e-4 Bohr) and dL/dP entries against central finite
  differences of the GPU loss in single density entries (h = 1e-6), on the
  config-1 molecule with its 30000 points.  V is linear in P, so the P
  difference quotient is exact up to rounding once the screen is held fixed
  (tol = 0 keeps every candidate).
* Config 4 in full: the 256-conformer batch (10-80 atoms, 517-19988 points),
  forward + field + loss + dL/dP, with eight conformers against the oracle's
  sample (tests/golden/cfg4.npz: FP64 <= 1e-10, mixed <= 1e-8 a.u.), every
  conformer against its single-table evaluation, and for every conformer the
  identity <dL/dP, P> = sum_i gV_i (V_i - nuc_i) (V is linear and homogeneous
  in P).
"""
from pathlib import Path

import numpy as np
import pytest

from fastesp_amd import _lib, synth
from fastesp_amd.autograd import esp_backward, esp_with_grad
from fastesp_amd.builder import build_table_hip

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_config1_3000_points_against_oracle():
    g = np.load(GOLD / "cfg1_3k.npz")
    w = synth.config1(3000)
    np.testing.assert_allclose(w.fingerprint(), g["fp"], rtol=1e-9)
    say("config1 3000 points: build + fp64/mixed forward")
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    assert t.n_pairs == int(g["npairs"])
    assert t.info()["n_candidates"] == int(g["n_candidates"])
    assert t.stats()["sum_rule"] == pytest.approx(float(g["sum_rule"]), abs=1e-8)
    V = t.esp()
    assert np.abs(V - g["V"]).max() < 1e-10
    Vm = t.esp(precision=_lib.FE_PREC_MIXED)
    assert np.abs(Vm - g["V"]).max() < 1e-8
    V2, F = esp_with_grad(t, w.points)
    assert np.abs(V2 - g["V"]).max() < 1e-10


@pytest.fixture(scope="module")
def cfg1_molecule():
    return synth.config1(30000)


def test_config2_field_matches_finite_differences(cfg1_molecule):
    w = cfg1_molecule
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    rng = np.random.default_rng(2202)
    sel = rng.choice(len(w.points), 16, replace=False)
    say("config2: field at 16 of 30000 points vs central differences, h = 1e-4")
    V, F = esp_with_grad(t, w.points)
    h = 1e-4
    C = w.points[sel]
    shifted = np.concatenate([C + s * h * np.eye(3)[k] for k in range(3) for s in (1, -1)])
    Vs = t.esp(shifted).reshape(3, 2, len(sel))
    fd = ((Vs[:, 0] - Vs[:, 1]) / (2 * h)).T
    err = np.abs(fd - F[sel]).max()
    say(f"field vs FD: max {err:.2e} (|field| max {np.abs(F[sel]).max():.2e})")
    assert err < 1e-7


def test_config2_dP_matches_finite_differences(cfg1_molecule):
    w = cfg1_molecule
    rng = np.random.default_rng(2203)
    wts = rng.uniform(0.5, 1.0, len(w.points))
    t0 = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=0.0)
    V0 = t0.esp()
    target = V0 + rng.normal(0, 1e-3, len(V0))

    def esp_at(P):
        return build_table_hip(w.wfn, w.basis, P, w.points, tol=0.0).esp()

    gV = 2.0 * wts * (V0 - target)
    _, dP = esp_backward(t0, w.points, gV, want_dG=False, n_func=w.basis.n)
    # the three largest entries of dL/dP and three random large density entries
    flat = np.argsort(np.abs(dP), axis=None)[::-1][:3]
    picks = [tuple(np.unravel_index(k, dP.shape)) for k in flat]
    big = np.argwhere(np.abs(w.density) > 1e-2)
    picks += [tuple(big[k]) for k in rng.choice(len(big), 3, replace=False)]
    h = 1e-6
    scale = np.abs(dP).max()
    say(f"config2: dL/dP at {len(picks)} entries vs central differences, h = 1e-6, tol 0")
    for i, j in picks:
        Pp = w.density.copy()
        Pm = w.density.copy()
        Pp[i, j] += h
        Pm[i, j] -= h
        Vp, Vm = esp_at(Pp), esp_at(Pm)
        # L(P+) - L(P-) = sum w (V+ - V-)(V+ + V- - 2T), formed from V+ - V- so
        # that the difference keeps its digits
        fd = float(np.sum(wts * (Vp - Vm) * (Vp + Vm - 2.0 * target))) / (2 * h)
        say(f"  dL/dP[{i},{j}] = {dP[i, j]:+.12e}  FD {fd:+.12e}")
        # rounding in V+ and V- (two builds) leaves ~2e-9 of the largest entry
        assert abs(fd - dP[i, j]) <= 1e-7 * abs(dP[i, j]) + 1e-8 * scale
    assert np.abs(dP - dP.T).max() <= 1e-12 * scale


@pytest.fixture(scope="module")
def cfg4():
    g = np.load(GOLD / "cfg4.npz")
    say("config4: 256 conformers (threads)")
    ws = synth.conformer_set(256, threads=True)
    tabs = []
    for k, w in enumerate(ws):
        if k % 64 == 0:
            say(f"config4: building tables {k}..{min(k + 63, 255)}")
        tabs.append(build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12))
    return g, ws, tabs


def test_config4_full_batch_against_oracle_and_single_tables(cfg4):
    from fastesp_amd.batch import Batch
    g, ws, tabs = cfg4
    sizes = [(w.meta["n_atoms"], len(w.points)) for w in ws]
    assert min(a for a, _ in sizes) == 10 and max(a for a, _ in sizes) == 80
    assert min(n for _, n in sizes) == 517 and max(n for _, n in sizes) == 19988
    b = Batch(tabs, [w.points for w in ws])
    say(f"config4: batch fp64 forward + field, {b.info()['n_points']} points")
    b.forward("fp64", field=True)
    V64, F64 = b.values(), b.field()
    for c in g["conformers"]:
        c = int(c)
        w, t = ws[c], tabs[c]
        np.testing.assert_allclose(w.fingerprint(), g[f"c{c}_fp"], rtol=1e-9)
        assert t.n_pairs == int(g[f"c{c}_npairs"])
        assert t.info()["n_candidates"] == int(g[f"c{c}_n_candidates"])
        assert t.stats()["sum_rule"] == pytest.approx(float(g[f"c{c}_sum_rule"]), abs=1e-8)
        idx = g[f"c{c}_idx"]
        err = np.abs(V64[c][idx] - g[f"c{c}_V"]).max()
        ferr = np.abs(F64[c][idx[:64]] - g[f"c{c}_field"]).max()
        say(f"  conformer {c}: {len(w.points)} points, |dV| {err:.2e}, |dfield| {ferr:.2e}")
        assert err < 1e-10 and ferr < 1e-9
    say("config4: every conformer against its single-table evaluation")
    for c, (w, t) in enumerate(zip(ws, tabs)):
        V, F = esp_with_grad(t, w.points)
        assert np.abs(V64[c] - V).max() < 1e-11, c
        assert np.abs(F64[c] - F).max() < 1e-11, c
    say("config4: batch mixed forward")
    b.forward("mixed", field=True)
    Vm = b.values()
    for c in g["conformers"]:
        c = int(c)
        assert np.abs(Vm[c][g[f"c{c}_idx"]] - g[f"c{c}_V"]).max() < 1e-8
    for c in range(len(ws)):
        assert np.abs(Vm[c] - V64[c]).max() < 1e-8, c
    # loss + backward: <dL/dP, P> = sum gV (V - nuc) for every conformer
    rng = np.random.default_rng(4404)
    wts = [rng.uniform(0.5, 1.0, len(v)) for v in V64]
    T = [v + rng.normal(0.0, 1e-3, len(v)) for v in V64]
    b.set_loss(wts, T)
    say("config4: fp64 step with dL/dP")
    b.step("fp64", field=True, dldp=True)
    Vs, gVs, dPs = b.values(), b.gV(), b.dldp()
    for c, (w, t) in enumerate(zip(ws, tabs)):
        nuc = t.nuclear(w.points)
        lhs = float(np.sum(dPs[c] * w.density))
        rhs = float(np.sum(gVs[c] * (Vs[c] - nuc)))
        assert abs(lhs - rhs) <= 1e-9 * max(1.0, abs(rhs)), (c, lhs, rhs)
        assert np.abs(dPs[c] - dPs[c].T).max() <= 1e-12 * max(1e-300, np.abs(dPs[c]).max())
    b.close()
