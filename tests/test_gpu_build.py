"""GPU parity of the device table build (fe_table_build) against the oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from fastesp_amd import _lib, synth
from fastesp_amd.builder import SCREEN_CPP, build_table_hip, export_flat
from fastesp_amd.formats import basis_payload, read_pairs, write_basis_payload
from fastesp_amd.table import PairTable, build_table
from oracle import engine
from tests.helpers import cf_toy, oh_toy

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
REF_BUILDER = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "build_pairs"
TOL_FP64 = 1e-10


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _compare_tables(flat, tab, rtol=1e-13):
    """Same pairs in the same (reference) order with the same p, P, K, G."""
    assert flat.n_pairs == tab.n_pairs and flat.n_terms == tab.n_terms
    np.testing.assert_allclose(flat.pp_exp, tab.pp_exp, rtol=1e-15)
    for a, b in ((flat.pp_cx, tab.pp_cx), (flat.pp_cy, tab.pp_cy), (flat.pp_cz, tab.pp_cz)):
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(flat.pp_K, tab.pp_K, rtol=1e-13)
    assert (flat.pp_nt == tab.pp_nt).all() and (flat.pp_off == tab.pp_off).all()
    assert (flat.tm_aidx == tab.tm_aidx).all() and (flat.tm_m == tab.tm_m).all()
    assert (flat.mo_parent == tab.mo_parent).all() and (flat.mo_axis == tab.mo_axis).all()
    # G relative to the largest |G| of its pair
    scale = np.repeat(np.array([np.abs(tab.tm_g[o:o + n]).max() if n else 1.0
                                for o, n in zip(tab.pp_off, tab.pp_nt)]), tab.pp_nt)
    assert (np.abs(flat.tm_g - tab.tm_g) / np.maximum(scale, 1e-300)).max() < rtol


@pytest.mark.parametrize("mk", [oh_toy, cf_toy])
def test_build_toys_match_oracle(mk):
    wfn, basis, P, _ = mk()
    rng = np.random.default_rng(8)
    pts = rng.normal(0, 3.0, (300, 3))
    say(f"toy build {mk.__name__}")
    t = build_table_hip(wfn, basis, P, pts, tol=0.0)
    tab = engine.build_table(wfn, basis, P, pts, tol=0.0)
    _compare_tables(export_flat(t), tab)
    assert t.stats()["sum_rule"] == pytest.approx(tab.sum_rule, abs=1e-11)


def test_build_cfg0_matches_oracle_and_golden():
    w = synth.config0()
    g = np.load(GOLD / "cfg0.npz")
    prims = engine.build_prim_pairs(w.basis, w.density)
    for key, tol in (("V_tol0", 0.0), ("V_tol12", 1e-12)):
        say(f"cfg0 build tol={tol}")
        t = build_table(w.wfn, w.basis, w.density, w.points, tol=tol)
        tab = engine.build_table(w.wfn, w.basis, w.density, w.points, tol, prims=prims)
        _compare_tables(export_flat(t), tab)
        assert t.n_pairs == int(g["npairs_" + key[2:]])
        V = t.esp()
        assert np.abs(V - g[key]).max() < TOL_FP64
    assert t.stats()["sum_rule"] == pytest.approx(w.wfn.nelec, abs=1e-9)


def test_build_cfg1_300_matches_golden():
    w = synth.config1(300)
    g = np.load(GOLD / "cfg1_300.npz")
    np.testing.assert_allclose(w.fingerprint(), g["fp"], rtol=1e-9)
    for key in ("V_tol0", "V_tol12"):
        tol = 0.0 if key == "V_tol0" else 1e-12
        say(f"cfg1 300 build tol={tol}")
        t = build_table(w.wfn, w.basis, w.density, w.points, tol=tol)
        assert t.n_pairs == int(g["npairs_" + key[2:]])
        assert np.abs(t.esp() - g[key]).max() < TOL_FP64
    assert t.stats()["sum_rule"] == pytest.approx(float(g["sum_rule"]), abs=1e-9)


@pytest.mark.parametrize("rule,tol", [("numpy", 1e-12), ("numpy", 1e-8), ("engine", 1e-7)])
def test_point_split_screen_matches_oracle(rule, tol):
    """More than 1024 points and few candidates: the screen splits the points
    over blocks and combines the per-slice minima (maxima for the engine rule)."""
    from fastesp_amd.builder import SCREEN_ENGINE
    w = synth.config0(n_points=5000)
    say(f"cfg0 5000 points build rule={rule} tol={tol}")
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=tol,
                        rule=SCREEN_ENGINE if rule == "engine" else 0)
    tab = engine.build_table(w.wfn, w.basis, w.density, w.points, tol, rule=rule)
    assert 0 < tab.n_pairs
    _compare_tables(export_flat(t), tab)


def test_point_split_screen_cpp_rule_matches_port():
    from oracle.cpu_port import CpuTable
    w = synth.config0(n_points=5000)
    say("cfg0 5000 points build rule=cpp tol=1e-9")
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-9, rule=SCREEN_CPP)
    ref = CpuTable(w.wfn, w.basis, w.density, w.points, tol=1e-9, rule=1)
    p6, _ = ref.export()
    flat = export_flat(t)
    assert flat.n_pairs == ref.n_pairs
    mine = np.stack([flat.pp_exp, flat.pp_cx, flat.pp_cy, flat.pp_cz], 1)
    order = lambda a: a[np.lexsort(a.T[::-1])]  # noqa: E731
    np.testing.assert_allclose(order(mine), order(p6[:, :4]), rtol=1e-13, atol=1e-13)


@pytest.mark.skipif(not REF_BUILDER.exists(), reason="reference builder not compiled")
def test_cpp_screen_rule_matches_reference_builder(tmp_path):
    """Rule 1 reproduces build_pairs' kept set at a tolerance that drops pairs."""
    w = synth.config0()
    b, p = tmp_path / "b.bin", tmp_path / "p.bin"
    write_basis_payload(b, basis_payload(w.wfn, w.basis, w.density, w.points))
    r = subprocess.run([str(REF_BUILDER), str(b), str(p), "1e-9"], capture_output=True, text=True)
    assert r.returncode == 0
    ref = read_pairs(p)
    say("cfg0 build rule=cpp tol=1e-9")
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-9, rule=SCREEN_CPP)
    flat = export_flat(t)
    assert flat.n_pairs == ref.n_pairs
    key = lambda f, i: (round(f.pp_exp[i], 10), round(f.pp_cx[i], 8), round(f.pp_cy[i], 8),  # noqa
                        round(f.pp_cz[i], 8))
    assert sorted(key(flat, i) for i in range(flat.n_pairs)) == sorted(
        key(ref, i) for i in range(ref.n_pairs))


def test_written_pairs_bin_roundtrip(tmp_path):
    wfn, basis, P, _ = cf_toy()
    pts = np.random.default_rng(1).normal(0, 3, (50, 3))
    say("write/read pairs.bin")
    t = build_table_hip(wfn, basis, P, pts, tol=0.0)
    t.write(tmp_path / "pairs.bin")
    t2 = PairTable.read(tmp_path / "pairs.bin")
    assert np.array_equal(t.esp(), t2.esp())


def test_build_rejects_bad_input():
    wfn, basis, P, _ = oh_toy()
    with pytest.raises(_lib.HipBackendError):
        build_table_hip(wfn, basis, P, np.zeros((1, 3)), tol=-1.0)


def _permuted_payload(pl, perm):
    """The same basis and density with the functions in the order `perm`."""
    q = dict(pl)
    for k in ("fn_atom", "fn_nprim", "fn_off", "fn_lx", "fn_ly", "fn_lz", "fn_shell"):
        q[k] = np.ascontiguousarray(pl[k][perm])
    q["fc"] = np.ascontiguousarray(pl["fc"].reshape(-1, 3)[perm].ravel())
    q["dens"] = np.ascontiguousarray(pl["dens"][np.ix_(perm, perm)])
    return q


def test_build_takes_cartesian_components_in_any_order():
    """A basis whose shells list their Cartesian components in another order
    (here reversed in every shell with l >= 1), with the density permuted to
    match, gives the same pairs, G and ESP: the build reads each shell's
    block through its components, not their positions."""
    from fastesp_amd.builder import build_table_payload
    wfn, basis, P, _ = cf_toy()
    pts = np.random.default_rng(21).normal(0, 3.0, (200, 3))
    pl = basis_payload(wfn, basis, P, pts)
    perm = []
    for s in np.unique(pl["fn_shell"]):
        idx = np.flatnonzero(pl["fn_shell"] == s)
        perm.extend(idx[::-1])
    perm = np.array(perm)
    assert not np.array_equal(perm, np.arange(len(perm)))
    say("cf toy build, components reversed in every shell")
    a = build_table_payload(pl, tol=0.0)
    b = build_table_payload(_permuted_payload(pl, perm), tol=0.0)
    fa, fb = export_flat(a), export_flat(b)
    assert fa.n_pairs == fb.n_pairs and fa.n_terms == fb.n_terms
    np.testing.assert_array_equal(fa.pp_exp, fb.pp_exp)
    np.testing.assert_array_equal(fa.pp_K, fb.pp_K)
    np.testing.assert_allclose(fb.tm_g, fa.tm_g, rtol=1e-14, atol=1e-14 * np.abs(fa.tm_g).max())
    np.testing.assert_allclose(b.esp(), a.esp(), rtol=0, atol=1e-13)


def test_build_rejects_a_shell_with_a_repeated_component():
    """Every Cartesian component of a shell exactly once: a d shell listing xx
    twice (and no yy) is a malformed basis."""
    from fastesp_amd.builder import build_table_payload
    wfn, basis, P, _ = oh_toy()
    pl = basis_payload(wfn, basis, P, np.zeros((4, 3)))
    d = np.flatnonzero((pl["fn_lx"] + pl["fn_ly"] + pl["fn_lz"]) == 2)
    assert len(d) == 6
    pl["fn_lx"] = pl["fn_lx"].copy()
    pl["fn_ly"] = pl["fn_ly"].copy()
    yy = d[(pl["fn_ly"][d] == 2)][0]
    pl["fn_lx"][yy], pl["fn_ly"][yy] = 2, 0
    say("oh toy build with a repeated d component")
    with pytest.raises(_lib.HipBackendError, match="components"):
        build_table_payload(pl, tol=0.0)
