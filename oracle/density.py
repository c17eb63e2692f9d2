"""Oracle electron density and vdW-surface points (FP64 NumPy) -- test
infrastructure only.

* ``rho(basis, P, pts)`` restates ``our_density`` of
  ``benchmarks/check_density_mwfn.py:60-75``: phi_j(C) = (x-A)^lx (y-A)^ly
  (z-A)^lz sum_k c_k exp(-a_k |C-A|^2) per Cartesian function, then
  rho = sum_ij phi_i P_ij phi_j.
* ``rho_from_orbitals(basis, wfn, pts)``: the same density from the MOs,
  sum_i occ_i |sum_mu T c_i phi_mu|^2 -- an independent route used to pin ``rho``.
* ``vdw_surface_points(symbols, coords, n, seed)`` restates
  ``examples/vdw_surface_esp.py:38-52``.

Pins (tests/test_density.py): the MO route; the integral of rho over space
equals Tr(P S) = N_elec (radial x Lebedev-free product quadrature on a
one-atom toy); the surface points lie on their own sphere and outside every
other one.
"""
from __future__ import annotations

import numpy as np

BONDI_EXAMPLE = {"H": 1.20, "C": 1.70, "N": 1.55, "O": 1.52, "S": 1.80, "P": 1.80}
BOHR = 0.529177210903


def basis_values(basis, pts) -> np.ndarray:
    pts = np.asarray(pts, float).reshape(-1, 3)
    phi = np.empty((len(pts), basis.n))
    for j in range(basis.n):
        d = pts - basis.centers[j]
        r2 = (d * d).sum(1)
        lx, ly, lz = map(int, basis.ang[j])
        acc = np.zeros(len(pts))
        for a, c in zip(basis.exps[j], basis.coefs[j]):
            acc += c * np.exp(-a * r2)
        phi[:, j] = d[:, 0] ** lx * d[:, 1] ** ly * d[:, 2] ** lz * acc
    return phi


def rho(basis, P, pts) -> np.ndarray:
    phi = basis_values(basis, pts)
    return np.einsum("ni,ij,nj->n", phi, np.asarray(P, float), phi)


def rho_from_orbitals(basis, wfn, pts, min_occ: float = 1e-10) -> np.ndarray:
    phi = basis_values(basis, pts)
    occ = np.asarray(wfn.mo_occ) > min_occ
    psi = phi @ (basis.block_transform() @ wfn.mo_coeff[:, occ])
    return (psi * psi) @ np.asarray(wfn.mo_occ)[occ]


def _fibonacci_sphere(n: int) -> np.ndarray:
    i = np.arange(n) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / n)
    theta = np.pi * (1.0 + 5.0 ** 0.5) * i
    return np.stack([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)], 1)


def vdw_surface_points(symbols, coords, n_points: int, seed: int = 0) -> np.ndarray:
    coords = np.asarray(coords, float)
    rad = np.array([BONDI_EXAMPLE[s] / BOHR for s in symbols])
    share = (4 * np.pi * rad ** 2) / (4 * np.pi * rad ** 2).sum()
    rng = np.random.default_rng(seed)
    out = []
    for a in range(len(symbols)):
        n = int(np.ceil(n_points * share[a] * 3.0)) + 50
        p = coords[a] + rad[a] * _fibonacci_sphere(n)
        d = np.linalg.norm(p[:, None, :] - coords[None, :, :], axis=2)
        out.append(p[(d >= rad[None, :] - 1e-9).all(axis=1)])
    cand = np.concatenate(out)
    if len(cand) > n_points:
        cand = cand[rng.choice(len(cand), n_points, replace=False)]
    return cand
