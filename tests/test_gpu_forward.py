"""GPU parity of the FP64 forward evaluation (fe_esp_forward) and the device Boys.

The evaluator is fed tables built by the oracle or by the reference's own C++
builder (``oracle/_ref/build_pairs``) and compared with the oracle's FP64
evaluation (golden fixtures under tests/golden/) to 1e-10 a.u.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from fastesp_amd import _lib, synth
from fastesp_amd.formats import basis_payload, write_basis_payload
from fastesp_amd.table import PairTable
from oracle import engine, evaluate
from tests.helpers import cf_toy, oh_toy

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
REF_BUILDER = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "build_pairs"
TOL_FP64 = 1e-10  # a.u., FP64 path vs the FP64 oracle


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def ctx():
    return _lib.Context.get(0)


def test_boys_matches_mpmath(ctx):
    g = np.load(GOLD / "boys.npz")
    T, Fref = g["T"], g["F"]
    for M in range(9):
        say(f"boys M={M} n={len(T)}")
        F = ctx.boys(T, M)
        rel = np.abs(F - Fref[:, :M + 1]) / Fref[:, :M + 1]
        assert rel.max() < 3e-15, (M, rel.max(), T[np.argmax(rel.max(1))])


def test_forward_toys_all_classes(ctx):
    for name, mk in (("oh", oh_toy), ("cf", cf_toy)):
        wfn, basis, P, _ = mk()
        rng = np.random.default_rng(4)
        pts = rng.normal(0, 3.0, (777, 3))
        tab = engine.build_table(wfn, basis, P, pts, tol=0.0)
        say(f"toy {name} pairs={tab.n_pairs} L={sorted(set(tab.pp_L.tolist()))}")
        ref = evaluate.esp(tab, pts)
        got = PairTable.from_flat(tab).esp()
        assert np.abs(got - ref).max() < TOL_FP64


def test_forward_cfg0_tol0_and_tol12(ctx):
    w = synth.config0()
    g = np.load(GOLD / "cfg0.npz")
    np.testing.assert_allclose(w.fingerprint(), g["fp"], rtol=1e-9)
    prims = engine.build_prim_pairs(w.basis, w.density)
    for key, tol in (("V_tol0", 0.0), ("V_tol12", 1e-12)):
        tab = engine.build_table(w.wfn, w.basis, w.density, w.points, tol, prims=prims)
        say(f"cfg0 tol={tol} pairs={tab.n_pairs}")
        t = PairTable.from_flat(tab)
        V = t.esp()
        assert np.abs(V - g[key]).max() < TOL_FP64
        # electron + nuclear decomposition is consistent
        np.testing.assert_allclose(t.nuclear() - t.electron(), V, rtol=0, atol=1e-12)


def test_tuning_knobs_keep_results(ctx):
    """Launch shape (points per lane, chunking, point order) only changes the
    summation order; a repeated call with the same shape is bitwise identical."""
    w = synth.config0(1500)
    tab = engine.build_table(w.wfn, w.basis, w.density, w.points, 1e-12)
    t = PairTable.from_flat(tab)
    ref = evaluate.esp(tab)
    say("ppl=1")
    ctx.set_tuning(0, 1)
    V1 = t.esp()
    assert np.array_equal(V1, t.esp())
    say("ppl=2")
    ctx.set_tuning(0, 2)
    V2 = t.esp()
    ctx.set_tuning(0, 1)
    say("target_wg=64")
    ctx.set_tuning(1, 64)
    V3 = t.esp()
    ctx.set_tuning(1, 8 * 256)
    say("unsorted points")
    ctx.set_tuning(3, 0)
    V4 = t.esp()
    ctx.set_tuning(3, 1)
    say("pairs in reference order")
    ctx.set_tuning(5, 0)
    V5 = PairTable.from_flat(tab).esp()
    ctx.set_tuning(5, 1)
    for V in (V1, V2, V3, V4, V5):
        assert np.abs(V - ref).max() < TOL_FP64
        assert np.abs(V - V1).max() < 1e-12


@pytest.mark.skipif(not REF_BUILDER.exists(), reason="reference builder not compiled")
def test_reference_built_table_drops_in(ctx, tmp_path):
    """pairs.bin written by the reference's own C++ builder evaluates unchanged."""
    w = synth.config1(300)
    g = np.load(GOLD / "cfg1_300.npz")
    np.testing.assert_allclose(w.fingerprint(), g["fp"], rtol=1e-9)
    b, p = tmp_path / "basis.bin", tmp_path / "pairs.bin"
    write_basis_payload(b, basis_payload(w.wfn, w.basis, w.density, w.points))
    say("reference build_pairs cfg1 300 tol=0")
    r = subprocess.run([str(REF_BUILDER), str(b), str(p), "0"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    say("evaluate reference table")
    V = PairTable.read(p).esp()
    assert np.abs(V - g["V_tol0"]).max() < TOL_FP64


def test_bad_tables_are_rejected(ctx):
    wfn, basis, P, _ = oh_toy()
    tab = engine.build_table(wfn, basis, P, np.zeros((2, 3)), tol=0.0)
    bad = engine.OracleTable(**{k: getattr(tab, k).copy() if hasattr(getattr(tab, k), "copy")
                                else getattr(tab, k) for k in (
        "pp_exp", "pp_cx", "pp_cy", "pp_cz", "pp_K", "pp_nt", "pp_off", "pp_nmono",
        "pp_monooff", "mo_parent", "mo_axis", "tm_aidx", "tm_m", "tm_g", "pts", "Z", "xyz")})
    bad.tm_m[0] = 3
    with pytest.raises(_lib.HipBackendError, match="canonical"):
        PairTable.from_flat(bad)
    bad2 = engine.OracleTable(**{k: getattr(tab, k) for k in (
        "pp_exp", "pp_cx", "pp_cy", "pp_cz", "pp_K", "pp_nt", "pp_off", "pp_nmono",
        "pp_monooff", "mo_parent", "mo_axis", "tm_aidx", "tm_m", "tm_g", "pts", "Z", "xyz")})
    bad2.pp_nt = tab.pp_nt.copy()
    bad2.pp_nt[0] = 7  # no class has 7 terms
    with pytest.raises(_lib.HipBackendError, match="class"):
        PairTable.from_flat(bad2)


def test_empty_and_single_point(ctx):
    wfn, basis, P, _ = oh_toy()
    pts = np.array([[2.0, 1.0, 0.5]])
    tab = engine.build_table(wfn, basis, P, pts, tol=0.0)
    t = PairTable.from_flat(tab)
    say("single point")
    assert abs(t.esp()[0] - evaluate.esp(tab)[0]) < TOL_FP64
    say("zero points")
    assert t.esp(np.zeros((0, 3))).shape == (0,)
