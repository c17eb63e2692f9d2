"""Small molecule builders shared by the tests."""
from __future__ import annotations

import numpy as np

from fastesp_amd.basis import Basis, Shell, Wavefunction, Z_OF
from fastesp_amd.synth import normalised_coeffs, random_density


def make_mol(atoms, shells, seed=0, nelec=None):
    """atoms: [(symbol, (x,y,z))]; shells: per atom [(l, exps, raw_coefs)]."""
    syms = [a[0] for a in atoms]
    coords = np.array([a[1] for a in atoms], float)
    sh = []
    for ia, per in enumerate(shells):
        for (l, exps, raw) in per:
            e = np.asarray(exps, float)
            sh.append(Shell(atom=ia, l=l, exps=e,
                            coeffs=normalised_coeffs(l, e, np.asarray(raw, float))))
    wfn = Wavefunction(symbols=syms, coords=coords, shells=sh)
    basis = Basis(wfn)
    if nelec is None:
        nelec = int(sum(Z_OF[s] for s in syms))
        nelec += nelec % 2
    rng = np.random.default_rng(seed)
    P, S = random_density(basis, nelec, rng)
    wfn.nelec_override = nelec
    return wfn, basis, P, S


def oh_toy(seed=0):
    """O + H with s/p/d shells (the toy of ``tests/test_pipeline.py:114-126`` plus a d)."""
    atoms = [("O", (0.0, 0.0, 0.0)), ("H", (1.8, 0.0, 0.0))]
    shells = [[(0, [3.5, 0.7], [0.5, 0.6]), (1, [1.2], [0.7]), (2, [0.9], [1.0])],
              [(0, [1.0], [0.6]), (1, [0.8], [1.0])]]
    return make_mol(atoms, shells, seed=seed, nelec=4)


def cf_toy(seed=1):
    """C + F-like pair with an f shell so every class L <= 6 appears."""
    atoms = [("C", (0.0, 0.0, 0.0)), ("O", (1.1, 0.9, -0.4))]
    shells = [[(0, [5.0, 1.0], [0.4, 0.5]), (1, [0.8], [0.6]), (2, [0.6], [0.4]),
               (3, [0.7, 0.3], [0.5, 0.5])],
              [(0, [2.0], [1.0]), (1, [1.1, 0.4], [0.5, 0.5]), (2, [0.9], [1.0]),
               (3, [0.5], [1.0])]]
    return make_mol(atoms, shells, seed=seed, nelec=8)


def quad_electron(basis, density, pts, R=14.0, ngl=160, ncos=140, nphi=280):
    """int rho(r)/|C-r| dr by spherical quadrature about C
    (the recipe of ``tests/test_pipeline.py:252-275``)."""
    from numpy.polynomial.legendre import leggauss
    xs, ws = leggauss(ngl)
    s = 0.5 * R * (xs + 1.0)
    wr = 0.5 * R * ws
    cz, wz = leggauss(ncos)
    phi_a = (np.arange(nphi) + 0.5) * (2 * np.pi / nphi)
    st = np.sqrt(np.clip(1 - cz ** 2, 0, None))
    nx = (st[None, :] * np.cos(phi_a)[:, None])[..., None]
    ny = (st[None, :] * np.sin(phi_a)[:, None])[..., None]
    nz = np.broadcast_to(cz[None, :, None], (nphi, ncos, 1))
    ss = s[None, None, :]
    wang = wz[None, :, None] * (2 * np.pi / nphi)
    out = np.zeros(len(pts))
    for ip, C in enumerate(pts):
        X, Y, Z = C[0] + ss * nx, C[1] + ss * ny, C[2] + ss * nz
        phis = []
        for j in range(basis.n):
            x, y, z = X - basis.centers[j][0], Y - basis.centers[j][1], Z - basis.centers[j][2]
            lx, ly, lz = map(int, basis.ang[j])
            acc = np.zeros(X.shape)
            for a, c in zip(basis.exps[j], basis.coefs[j]):
                acc += c * np.exp(-a * (x * x + y * y + z * z))
            phis.append((x ** lx) * (y ** ly) * (z ** lz) * acc)
        pf = np.stack(phis).reshape(basis.n, -1)
        rho = np.einsum("an,ab,bn->n", pf, density, pf).reshape(X.shape)
        out[ip] = float((rho * ss * wr[None, None, :] * wang).sum())
    return out
