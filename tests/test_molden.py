"""MOLDEN input and the pure -> Cartesian density (CPU).

Pins:
* the reference's own outputs on ``tests/golden/toy.molden``
  (``make_molden_golden.py``): the transform matrices T_l (l <= 4, which fix the
  order and phase convention), the overlap blocks, and the Cartesian density
  of ``read_molden`` + ``Basis.density_matrix_cartesian``;
* the invariants of ``tests/test_pipeline.py:89-126``: T S T^T = 1, C^T S C = 1,
  Tr(P S) = N_elec, the table against a quadrature of rho / |C - r|;
* every row of T is a harmonic polynomial (the pure subspace).
"""
from pathlib import Path

import numpy as np
import pytest

from fastesp_amd import synth
from fastesp_amd.basis import (BOHR, Basis, cart_order, cartesian_overlap_block,
                               overlap_matrix, pure_to_cartesian, solid_harmonics)
from fastesp_amd.molden import read_molden, write_molden
from oracle import engine, evaluate
from tests.helpers import quad_electron

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def ref():
    return np.load(GOLD / "molden_ref.npz")


def test_transform_matches_reference(ref):
    for l in range(5):
        np.testing.assert_allclose(pure_to_cartesian(l), ref[f"T{l}"], rtol=0, atol=2e-15)
        np.testing.assert_allclose(cartesian_overlap_block(l), ref[f"S{l}"], rtol=1e-15, atol=0)


def test_spherical_transform_is_orthonormal():
    """``tests/test_pipeline.py:89-93``."""
    for l in range(6):
        T = pure_to_cartesian(l)
        assert np.abs(T @ cartesian_overlap_block(l) @ T.T - np.eye(2 * l + 1)).max() < 1e-14


def test_solid_harmonics_are_harmonic():
    """Laplacian of every row vanishes, and the rows span 2l+1 dimensions."""
    for l in range(6):
        rows = solid_harmonics(l)
        assert len(rows) == 2 * l + 1
        for poly in rows:
            lap = {}
            for (a, b, c), v in poly.items():
                for k, e in enumerate((a, b, c)):
                    if e >= 2:
                        key = [a, b, c]
                        key[k] -= 2
                        lap[tuple(key)] = lap.get(tuple(key), 0) + v * e * (e - 1)
            assert all(v == 0 for v in lap.values()), (l, poly)
        comps = {c: i for i, c in enumerate(cart_order(l))}
        M = np.zeros((len(rows), len(comps)))
        for r, poly in enumerate(rows):
            for k, v in poly.items():
                M[r, comps[k]] = v
        assert np.linalg.matrix_rank(M) == 2 * l + 1


def test_read_molden_and_density_match_reference(ref):
    w = read_molden(GOLD / "toy.molden")
    np.testing.assert_array_equal(w.coords, ref["coords"])
    np.testing.assert_array_equal(w.mo_occ, ref["mo_occ"])
    assert w.mo_coeff.shape == tuple(ref["mo_shape"])
    assert [w.spherical[l] for l in range(5)] == list(ref["spherical"])
    assert [s.l for s in w.shells] == list(ref["shell_l"])
    b = Basis(w)
    assert b.n == int(ref["n_func"])
    P = b.density_matrix_cartesian()
    assert np.abs(P - ref["P"]).max() < 1e-13 * np.abs(ref["P"]).max()


def test_mos_orthonormal_and_electron_count():
    """``tests/test_pipeline.py:96-111``."""
    w = read_molden(GOLD / "toy.molden")
    b = Basis(w)
    S = overlap_matrix(b)
    T = b.block_transform()
    C = w.mo_coeff[:, w.mo_occ > 1e-10]
    assert np.abs(C.T @ (T.T @ S @ T) @ C - np.eye(C.shape[1])).max() < 1e-10
    assert float((b.density_matrix_cartesian() * S).sum()) == pytest.approx(w.nelec, abs=1e-9)


def test_table_from_molden_matches_quadrature():
    """``tests/test_pipeline.py:114-126`` through the MOLDEN route."""
    w = read_molden(GOLD / "toy.molden")
    b = Basis(w)
    P = b.density_matrix_cartesian()
    pts = np.array([[3.4, 0.4, 0.3]])
    got = evaluate.electron(engine.build_table(w, b, P, pts, tol=0.0))
    want = quad_electron(b, P, pts, ngl=120, ncos=110, nphi=220)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-10)


def test_write_read_round_trip(tmp_path):
    w0 = synth.config0(10).wfn
    synth.pure_orbitals(w0, 12, np.random.default_rng(5))
    w = read_molden(write_molden(tmp_path / "c0.molden", w0))
    assert w.symbols == w0.symbols
    np.testing.assert_allclose(w.coords, w0.coords, rtol=0, atol=1e-13)
    assert len(w.shells) == len(w0.shells)
    for a, b in zip(w.shells, w0.shells):
        assert (a.atom, a.l) == (b.atom, b.l)
        np.testing.assert_allclose(a.exps, b.exps, rtol=1e-15)
        np.testing.assert_allclose(a.coeffs, b.coeffs, rtol=1e-15)
    np.testing.assert_allclose(w.mo_coeff, w0.mo_coeff, rtol=1e-15, atol=1e-300)
    assert w.nelec == 12.0
    assert all(w.spherical[l] for l in (2, 3, 4)) and not w.spherical[1]


def test_angstrom_atoms_fortran_exponents_and_combined_flags(tmp_path):
    text = """[Molden Format]
[Atoms] Angs
H 1 1 0.0 0.0 0.0
H 2 1 0.74 0.0 0.0
[GTO]
 1 0
s 1 1.00
  1.0D+00 0.7127054D+00

 2 0
s 1 1.00
  1.0D+00 0.7127054D+00

[5D7F]
[MO]
 Sym= 1a
 Ene= -0.5
 Spin= Alpha
 Occup= 2.0
   1 0.5
   2 0.5
"""
    p = tmp_path / "h2.molden"
    p.write_text(text)
    w = read_molden(p)
    assert w.coords[1, 0] == pytest.approx(0.74 / BOHR, rel=1e-15)
    assert w.shells[1].exps[0] == 1.0 and w.shells[0].coeffs[0] == pytest.approx(0.7127054)
    assert w.spherical[2] and w.spherical[3] and not w.spherical[4]
    assert w.mo_coeff.shape == (2, 1) and w.nelec == 2.0


def test_cartesian_d_input_is_rejected(tmp_path):
    w0 = synth.config0(10).wfn
    synth.pure_orbitals(w0, 12, np.random.default_rng(5))
    w0.mo_coeff = np.vstack([w0.mo_coeff, np.zeros((3, w0.mo_coeff.shape[1]))])
    w = read_molden(write_molden(tmp_path / "x.molden", w0))
    with pytest.raises(ValueError, match="Cartesian"):
        Basis(w).density_matrix_cartesian()
