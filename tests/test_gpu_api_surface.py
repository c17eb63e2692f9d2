"""Caller code written against the reference package runs unchanged on this one.

The reference exports ``boys`` and ``primitive_pair_integral``
(``fastesp/__init__.py:29,38``) and its ``PairTable`` exposes the flat-table
fields, ``max_mono`` and ``screen(tol)`` (``fastesp/table.py:41-70,107-116``).
Values are pinned against the oracle: the mpmath Boys fixture, the oracle's
primitive integral (itself pinned by the s-s closed form and the reference's
quadrature cases) and the oracle's table in reference order.
"""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_boys_like_the_reference():
    from fastesp_amd import boys
    g = np.load(GOLD / "boys.npz")
    T, F = g["T"], g["F"]
    say("boys(m, T) arrays")
    for m in range(9):
        got = boys(m, T)
        assert got.shape == T.shape
        assert (np.abs(got - F[:, m]) / F[:, m]).max() < 3e-15
    # broadcasting of m against T, and a scalar comes back as a float
    got = boys(np.arange(9)[:, None], T[None, :5])
    assert got.shape == (9, 5)
    np.testing.assert_allclose(got, F[:5, :9].T, rtol=3e-15, atol=0)
    v = boys(3, float(T[7]))
    assert isinstance(v, float) and abs(v - F[7, 3]) <= 3e-15 * F[7, 3]
    with pytest.raises(ValueError):
        boys(9, 1.0)
    with pytest.raises(ValueError):
        boys(1, -1.0)


def test_primitive_pair_integral_like_the_reference():
    from fastesp_amd import primitive_pair_integral
    from oracle import ints as oints
    # the geometry of the reference's quadrature cases (tests/test_integrals.py:72-85)
    A = np.array([0.3, -0.4, 0.2])
    B = A + np.array([1.1, 0.5, -0.6])
    C = np.array([0.1, 1.6, 1.1])
    say("primitive_pair_integral, 8 component pairs")
    for la, lb in [((0, 0, 0), (0, 0, 0)), ((1, 0, 0), (0, 0, 0)), ((0, 1, 0), (0, 0, 1)),
                   ((2, 0, 0), (0, 1, 0)), ((1, 1, 0), (0, 0, 1)), ((2, 1, 0), (1, 1, 1)),
                   ((0, 0, 3), (1, 2, 0)), ((3, 0, 0), (0, 3, 0))]:
        got = primitive_pair_integral(0.45, A, la, 0.72, B, lb, C)
        want = oints.primitive_pair_integral(0.45, A, la, 0.72, B, lb, C)
        assert isinstance(got, float)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (la, lb, got, want)
    # several points at once
    Cs = np.array([C, C + 0.5, C - 1.0])
    got = primitive_pair_integral(0.45, A, (1, 0, 0), 0.72, B, (0, 1, 0), Cs)
    want = [oints.primitive_pair_integral(0.45, A, (1, 0, 0), 0.72, B, (0, 1, 0), c) for c in Cs]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_pair_table_fields_like_the_reference(tmp_path):
    from fastesp_amd import PairTable, build_table, synth
    from oracle import engine
    w = synth.config0(300)
    say("config0 300 points: device table fields vs the oracle's")
    t = build_table(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    ref = engine.build_table(w.wfn, w.basis, w.density, w.points, 1e-12)
    assert t.n_pairs == ref.n_pairs and t.n_terms == ref.n_terms
    for k in ("pp_nt", "pp_off", "pp_nmono", "pp_monooff", "mo_parent", "mo_axis", "tm_aidx",
              "tm_m"):
        assert np.array_equal(getattr(t, k), getattr(ref, k)), k
    for k in ("pp_exp", "pp_cx", "pp_cy", "pp_cz", "pp_K"):
        np.testing.assert_allclose(getattr(t, k), getattr(ref, k), rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(t.tm_g, ref.tm_g, rtol=0, atol=1e-13 * np.abs(ref.tm_g).max())
    assert t.max_mono == ref.max_mono
    assert t.pts.shape == (300, 3) and len(t.Z) == len(t.xyz) == len(w.wfn.symbols)
    # read-only views, as the reference's frombuffer arrays of a read table
    with pytest.raises(ValueError):
        t.tm_g[0] = 1.0
    # screen(tol) returns the table itself, as the reference's does
    assert t.screen(1e-12) is t
    # pairs.bin round trip keeps every field
    t.write(tmp_path / "pairs.bin")
    r = PairTable.read(tmp_path / "pairs.bin")
    for k in ("pp_exp", "pp_K", "tm_g", "tm_aidx", "mo_parent"):
        assert np.array_equal(getattr(r, k), getattr(t, k)), k
    np.testing.assert_allclose(r.esp(), t.esp(), rtol=0, atol=1e-13)
