"""GPU parity of the electron density (fe_density_*) and the surface points
(fe_vdw_candidates) against the oracle (``oracle/density.py``)."""
from pathlib import Path

import numpy as np
import pytest

from fastesp_amd import synth
from fastesp_amd.basis import Basis
from fastesp_amd.density import ElectronDensity, vdw_surface_points
from fastesp_amd.molden import read_molden
from oracle import density

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _check(got, want):
    # FP64 sums of n^2 terms of both signs: relative to the largest |rho|
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max() + 1e-300


def test_density_toy_molden():
    w = read_molden(GOLD / "toy.molden")
    b = Basis(w)
    P = b.density_matrix_cartesian()
    pts = np.vstack([np.random.default_rng(2).normal(0.0, 2.5, (500, 3)), w.coords])
    say("density toy.molden (44 functions, one block)")
    got = ElectronDensity(w, b, P).rho(pts)
    _check(got, density.rho(b, P, pts))


def test_density_config0_blocks_and_chunks():
    """~150 functions (3 blocks: the doubled upper blocks), 40000 points
    (two chunks of the device loop)."""
    wl = synth.config0(10)
    pts = np.random.default_rng(5).uniform(-2.0, 12.0, (40000, 3))
    say("density config0 40000 points")
    ed = ElectronDensity(wl.wfn, wl.basis, wl.density)
    got = ed.rho(pts)
    want = density.rho(wl.basis, wl.density, pts)
    _check(got, want)
    assert ed.info()["n_pad"] % 64 == 0
    ed.set_tile128(False)   # the 64 x 64 tile kernel
    _check(ed.rho(pts), want)
    # only the symmetric part of P enters
    A = np.random.default_rng(6).normal(size=wl.density.shape) * 1e-3
    asym = ElectronDensity(wl.wfn, wl.basis, wl.density + A - A.T).rho(pts[:2000])
    np.testing.assert_allclose(asym, got[:2000], rtol=0, atol=1e-13 * np.abs(got).max())


def test_density_config1_subset():
    wl = synth.config1(300)
    say("density config1 (1634 functions, 26 blocks) 300 points")
    got = ElectronDensity(wl.wfn, wl.basis, wl.density).rho(wl.points)
    _check(got, density.rho(wl.basis, wl.density, wl.points))


def test_surface_points_match_oracle():
    wl = synth.config1(10)
    say("vdw surface points config1 molecule")
    for n, seed in ((3000, 0), (30000, 7)):
        got = vdw_surface_points(wl.wfn, n, seed=seed)
        want = density.vdw_surface_points(wl.wfn.symbols, wl.wfn.coords, n, seed=seed)
        assert got.shape == want.shape
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_density_edge_cases():
    """No points, one point, points on the nuclei (r = 0: every s function at
    its maximum, every other component zero) and a ragged count (not a
    multiple of the 64- or 128-point tiles)."""
    w = read_molden(GOLD / "toy.molden")
    b = Basis(w)
    P = b.density_matrix_cartesian()
    ed = ElectronDensity(w, b, P)
    say("density edge cases")
    assert ed.rho(np.zeros((0, 3))).shape == (0,)
    for pts in (w.coords[:1], w.coords, np.random.default_rng(9).normal(0, 2, (131, 3))):
        _check(ed.rho(pts), density.rho(b, P, pts))
