"""Seeded synthetic molecules, densities and ESP point sets.

Every workload is generated from ``numpy.random.default_rng(seed)`` (PCG64,
platform stable), so the same seed gives bit-identical inputs on every host.

* atoms: uniform in a cubic box with a minimum pair separation (rejection);
* basis: per-element shell recipes, exponents log-uniform in per-shell
  ranges, contraction coefficients carrying N_s(alpha, l) and normalised so the
  (l,0,0) component has unit norm (the convention of ``fastesp/basis.py:53-66``);
* density: P = 2 C C^T with C (n_fn x n_occ) Gaussian random and
  orthonormalised in the overlap metric, so Tr(P S) = N_elec exactly;
* points: uniform on the exposed part of the union of k x Bondi spheres
  (the construction of ``examples/vdw_surface_esp.py:38-52`` generalised with
  a scale factor k and a guaranteed point count).
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

from .basis import BOHR, Basis, Shell, Wavefunction, Z_OF, overlap_matrix, primitive_norm

BONDI_ANG = {"H": 1.20, "C": 1.70, "N": 1.55, "O": 1.52, "S": 1.80, "P": 1.80}


@dataclass
class Workload:
    name: str
    seed: int
    wfn: Wavefunction
    basis: Basis
    density: np.ndarray
    points: np.ndarray
    meta: dict = field(default_factory=dict)

    def fingerprint(self) -> np.ndarray:
        """Checksums of the inputs.  The density and the point rotations come out
        of BLAS/LAPACK, whose last-bit rounding differs between hosts, so inputs
        regenerated elsewhere are compared with a tolerance, not bit for bit."""
        c, P, x = self.wfn.coords, self.density, self.points
        e = np.concatenate([np.asarray(s.exps, float) for s in self.wfn.shells])
        k = np.concatenate([np.asarray(s.coeffs, float) for s in self.wfn.shells])
        i = np.arange(P.shape[0], dtype=float)
        return np.array([c.sum(), (c * c).sum(), P.sum(), (P * P).sum(), np.abs(P).sum(),
                         (P @ i).sum() / max(len(i), 1), x.sum(), (x * x).sum(), np.log(e).sum(),
                         np.abs(k).sum(), float(len(x)), float(P.shape[0])])

    def input_hash(self) -> str:
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.wfn.coords).tobytes())
        for sh in self.wfn.shells:
            h.update(np.asarray(sh.exps, float).tobytes())
            h.update(np.asarray(sh.coeffs, float).tobytes())
        h.update(np.ascontiguousarray(self.density).tobytes())
        h.update(np.ascontiguousarray(self.points).tobytes())
        return h.hexdigest()[:16]


# ---------------------------------------------------------------------------
def place_atoms(rng, n: int, box: float, min_sep: float) -> np.ndarray:
    """n points uniform in [0, box)^3 with all pair distances >= min_sep."""
    out = np.zeros((0, 3))
    tries = 0
    while len(out) < n:
        c = rng.uniform(0.0, box, 3)
        tries += 1
        if tries > 200000:
            raise RuntimeError("cannot place atoms with the requested separation")
        if len(out) and np.min(np.linalg.norm(out - c, axis=1)) < min_sep:
            continue
        out = np.vstack([out, c])
    return out - out.mean(axis=0)


def normalised_coeffs(l: int, exps: np.ndarray, raw: np.ndarray) -> np.ndarray:
    """Coefficients c_k = raw_k N_s(a_k, l) scaled so <(l,0,0)|(l,0,0)> = 1."""
    c = raw * primitive_norm(exps, l)
    p = exps[:, None] + exps[None, :]
    dfact = float(np.prod(np.arange(2 * l - 1, 0, -2))) if l > 0 else 1.0
    s = (c[:, None] * c[None, :] * (np.pi / p) ** 1.5 * dfact / (2.0 * p) ** l).sum()
    return c / np.sqrt(s)


def _shell(rng, atom: int, l: int, nprim: int, lo: float, hi: float) -> Shell:
    e = np.sort(np.exp(rng.uniform(np.log(lo), np.log(hi), nprim)))[::-1].copy()
    raw = rng.uniform(0.2, 1.0, nprim)
    return Shell(atom=atom, l=l, exps=e, coeffs=normalised_coeffs(l, e, raw))


# Shell recipes: (l, nprim, exponent range).  Config-0 style: minimal-ish
# basis with exponents over [0.1, 5000]; config-1 style: triple-zeta-like
# s/p/d/f with diffuse functions down to 0.02.
def _recipe_small(rng, sym: str):
    if sym == "H":
        return [(0, 3, 0.1, 5000.0), (0, 1, 0.1, 5000.0), (1, 1, 0.1, 5000.0)]
    return ([(0, int(rng.integers(1, 7)), 0.1, 5000.0) for _ in range(3)]
            + [(1, int(rng.integers(1, 7)), 0.1, 5000.0) for _ in range(2)]
            + [(2, int(rng.integers(1, 7)), 0.1, 5000.0)])


def _recipe_tz(rng, sym: str, extra_d: bool = False):
    if sym == "H":
        sh = [(0, 3, 0.5, 40.0), (0, 1, 0.15, 0.5), (0, 1, 0.02, 0.06),
              (1, 1, 0.5, 1.5), (1, 1, 0.1, 0.4), (2, 1, 0.5, 1.5)]
    elif sym == "S":
        sh = [(0, 7, 20.0, 1.0e4), (0, 3, 1.0, 20.0), (0, 1, 0.3, 1.0),
              (0, 1, 0.1, 0.3), (0, 1, 0.05, 0.1), (0, 1, 0.02, 0.05),
              (1, 5, 1.0, 300.0), (1, 2, 0.3, 1.0), (1, 1, 0.1, 0.3),
              (1, 1, 0.05, 0.1), (1, 1, 0.02, 0.05),
              (2, 1, 1.0, 3.0), (2, 1, 0.3, 1.0), (2, 1, 0.1, 0.3),
              (3, 1, 0.3, 1.0)]
    else:  # C / N / O
        sh = [(0, 6, 10.0, 5000.0), (0, 2, 1.0, 20.0), (0, 1, 0.3, 1.5),
              (0, 1, 0.1, 0.4), (0, 1, 0.02, 0.06),
              (1, 4, 0.5, 40.0), (1, 1, 0.2, 0.6), (1, 1, 0.08, 0.25),
              (1, 1, 0.02, 0.06),
              (2, 1, 1.0, 3.0), (2, 1, 0.3, 1.0), (2, 1, 0.1, 0.3),
              (3, 1, 0.5, 1.5)]
    if extra_d:
        sh = sh + [(2, 1, 0.03, 0.08)]
    return sh


def build_wavefunction(rng, symbols, coords, recipe, extra_d_atoms=()) -> Wavefunction:
    shells = []
    for a, s in enumerate(symbols):
        rec = recipe(rng, s) if recipe is _recipe_small else recipe(rng, s, a in extra_d_atoms)
        for (l, npr, lo, hi) in rec:
            shells.append(_shell(rng, a, l, npr, lo, hi))
    return Wavefunction(symbols=list(symbols), coords=np.asarray(coords, float),
                        shells=shells)


def random_orbitals(basis: Basis, nelec: int, rng, S: np.ndarray | None = None):
    """C (n_fn x n_occ) Gaussian random with C^T S C = 1 (n_occ = nelec/2)."""
    if nelec % 2:
        raise ValueError("odd electron count")
    nocc = nelec // 2
    if S is None:
        S = overlap_matrix(basis)
    C = rng.standard_normal((basis.n, nocc))
    M = C.T @ S @ C
    w, v = np.linalg.eigh(M)
    return C @ (v @ np.diag(w ** -0.5) @ v.T), S


def density_from_orbitals(C: np.ndarray) -> np.ndarray:
    P = 2.0 * C @ C.T
    return 0.5 * (P + P.T)


def random_density(basis: Basis, nelec: int, rng, S: np.ndarray | None = None):
    """P = 2 C C^T, C Gaussian random with C^T S C = 1 (n_occ = nelec/2)."""
    C, S = random_orbitals(basis, nelec, rng, S)
    return density_from_orbitals(C), S


def pure_orbitals(wfn: Wavefunction, nelec: int, rng, S: np.ndarray | None = None):
    """MO data of a closed-shell wavefunction stored in pure functions, as a
    MOLDEN file holds it: C_pure ((sum 2l+1) x n_occ) Gaussian random with
    C^T (T^T S T) C = 1, occupations 2.  Sets ``wfn.mo_coeff``, ``mo_occ`` and
    ``mo_energy`` and returns the Cartesian density T C C^T T^T * 2 (the
    construction of ``tests/test_pipeline.py:57-78``)."""
    if nelec % 2:
        raise ValueError("odd electron count")
    basis = Basis(wfn)
    if S is None:
        S = overlap_matrix(basis)
    T = basis.block_transform()
    Sp = T.T @ S @ T
    nocc = nelec // 2
    C = rng.standard_normal((T.shape[1], nocc))
    w, v = np.linalg.eigh(C.T @ Sp @ C)
    C = C @ (v @ np.diag(w ** -0.5) @ v.T)
    wfn.mo_coeff = C
    wfn.mo_occ = np.full(nocc, 2.0)
    wfn.mo_energy = -np.linspace(1.0, 0.1, nocc)
    return basis.density_matrix_cartesian()


# ---------------------------------------------------------------------------
def _fibonacci_sphere(n: int) -> np.ndarray:
    i = np.arange(n) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / n)
    theta = np.pi * (1.0 + 5.0 ** 0.5) * i
    return np.stack([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta),
                     np.cos(phi)], axis=1)


def vdw_surface_points(symbols, coords, n: int, scale: float, rng) -> np.ndarray:
    """n points on the exposed surface of the union of scale x Bondi spheres (Bohr)."""
    rad = np.array([BONDI_ANG[s] * scale / BOHR for s in symbols])
    share = rad ** 2 / (rad ** 2).sum()
    over = 3.0
    while True:
        cand = []
        for a in range(len(symbols)):
            m = int(np.ceil(n * share[a] * over)) + 50
            # random rotation per atom so Fibonacci lattices of neighbours differ
            q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            p = coords[a] + rad[a] * (_fibonacci_sphere(m) @ q.T)
            d = np.linalg.norm(p[:, None, :] - coords[None, :, :], axis=2)
            keep = (d >= rad[None, :] - 1e-9).all(axis=1)
            cand.append(p[keep])
        cand = np.concatenate(cand)
        if len(cand) >= n:
            return cand[rng.choice(len(cand), n, replace=False)]
        over *= 2.0


def _even_electrons(rng, draw):
    while True:
        syms = draw()
        ne = int(round(sum(Z_OF[s] for s in syms)))
        if ne % 2 == 0:
            return syms, ne


# ---------------------------------------------------------------------------
def config0(n_points: int = 1000, seed: int = 1000) -> Workload:
    """12 atoms, Z in {1,6,7,8}, 10 Bohr box, small s/p/d basis, 1.4 x vdW points."""
    rng = np.random.default_rng(seed)
    coords = place_atoms(rng, 12, 10.0, 1.8)
    syms, ne = _even_electrons(
        rng, lambda: [("H", "C", "N", "O")[int(k)] for k in rng.integers(0, 4, 12)])
    wfn = build_wavefunction(rng, syms, coords, _recipe_small)
    basis = Basis(wfn)
    P, _ = random_density(basis, ne, rng)
    wfn.nelec_override = ne
    pts = vdw_surface_points(syms, wfn.coords, n_points, 1.4, rng)
    return Workload("cfg0", seed, wfn, basis, P, pts,
                    dict(n_atoms=12, n_func=basis.n, n_prim=basis.n_prim, nelec=ne))


def _tz_molecule(rng, n_atoms: int, n_s: int, n_heavy: int, box: float,
                 extra_d_atoms=(), orbitals: bool = False):
    coords = place_atoms(rng, n_atoms, box, 1.8)

    def draw():
        heavy = [("C", "N", "O")[int(k)] for k in rng.integers(0, 3, n_heavy)]
        syms = ["S"] * n_s + heavy + ["H"] * (n_atoms - n_s - n_heavy)
        perm = rng.permutation(n_atoms)
        return [syms[i] for i in perm]

    syms, ne = _even_electrons(rng, draw)
    s_atoms = [i for i, s in enumerate(syms) if s == "S"]
    extra = tuple(s_atoms[:len(extra_d_atoms)]) if extra_d_atoms else ()
    wfn = build_wavefunction(rng, syms, coords, _recipe_tz, extra)
    basis = Basis(wfn)
    C, _ = random_orbitals(basis, ne, rng)
    wfn.nelec_override = ne
    return wfn, basis, (C if orbitals else density_from_orbitals(C)), ne


def config1(n_points: int = 30000, seed: int = 1001) -> Workload:
    """54 atoms (C/N/O/H + 2 S), 20 Bohr box, TZ-like s/p/d/f basis: 1634 functions."""
    rng = np.random.default_rng(seed)
    wfn, basis, P, ne = _tz_molecule(rng, 54, 2, 25, 20.0, extra_d_atoms=(0,))
    prng = np.random.default_rng(seed * 7919 + n_points)
    pts = vdw_surface_points(wfn.symbols, wfn.coords, n_points, 1.4, prng)
    return Workload("cfg1", seed, wfn, basis, P, pts,
                    dict(n_atoms=54, n_func=basis.n, n_prim=basis.n_prim, nelec=ne))


def config3(n_points: int = 200000, seed: int = 1003) -> Workload:
    """150 atoms, ~4500 functions up to f, points on 1.4/1.6/1.8/2.0 x vdW shells."""
    rng = np.random.default_rng(seed)
    wfn, basis, P, ne = _tz_molecule(rng, 150, 2, 74, 28.0)
    prng = np.random.default_rng(seed * 7919 + n_points)
    per = [n_points // 4 + (1 if i < n_points % 4 else 0) for i in range(4)]
    pts = np.concatenate([vdw_surface_points(wfn.symbols, wfn.coords, k, s, prng)
                          for k, s in zip(per, (1.4, 1.6, 1.8, 2.0))])
    return Workload("cfg3", seed, wfn, basis, P, pts,
                    dict(n_atoms=150, n_func=basis.n, n_prim=basis.n_prim, nelec=ne))


def conformer(index: int, seed: int = 1004, n_atoms: int | None = None,
              n_points: int | None = None, orbitals: bool = False) -> Workload:
    """One member of the batched set: 10-80 atoms, 500-20000 points, TZ recipe.
    With ``orbitals`` the ``density`` field holds the occupied factor C
    (P = 2 C C^T, ``density_from_orbitals``) instead of P."""
    rng = np.random.default_rng([seed, index])
    na = int(rng.integers(10, 81)) if n_atoms is None else n_atoms
    npts = int(rng.integers(500, 20001)) if n_points is None else n_points
    n_s = 2 if na >= 20 else 0
    n_heavy = max(1, int(round(0.48 * na)) - n_s)
    box = 20.0 * (na / 54.0) ** (1.0 / 3.0)
    wfn, basis, P, ne = _tz_molecule(rng, na, n_s, n_heavy, box, orbitals=orbitals)
    pts = vdw_surface_points(wfn.symbols, wfn.coords, npts, 1.4, rng)
    return Workload(f"conf{index}", seed, wfn, basis, P, pts,
                    dict(n_atoms=na, n_func=basis.n, n_prim=basis.n_prim, nelec=ne))


def _conformer_orbitals(args):
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):
        return conformer(*args, orbitals=True)


def conformer_set(n: int, seed: int = 1004, workers: int = 0, threads: bool = False) -> list:
    """conformer(0..n-1) generated in parallel.  Workers return the occupied
    factor C, a few MB instead of the n_fn^2 density; P = 2 C C^T is formed
    here.  Processes (fork) by default: call this before the process touches
    the GPU.  ``threads=True`` uses a thread pool instead (single-threaded
    BLAS in each, so the result is the same), which is safe in a process that
    already uses the GPU."""
    import multiprocessing as mp
    import os
    if workers <= 0:
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
    if workers == 1 or n < 4:
        return [conformer(i, seed) for i in range(n)]
    if threads:
        from concurrent.futures import ThreadPoolExecutor
        from threadpoolctl import threadpool_limits
        with threadpool_limits(1), ThreadPoolExecutor(workers) as ex:
            out = list(ex.map(lambda i: conformer(i, seed, orbitals=True), range(n)))
    else:
        with mp.get_context("fork").Pool(workers) as pool:
            out = pool.map(_conformer_orbitals, [(i, seed) for i in range(n)], chunksize=1)
    for w in out:
        w.density = density_from_orbitals(w.density)
    return out
