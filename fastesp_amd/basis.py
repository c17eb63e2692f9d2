"""Cartesian Gaussian basis objects consumed by ``build_table``.

Field names and conventions follow the reference so that a caller holding a
fastesp ``Wavefunction``/``Basis`` pair can hand it over unchanged:

* ``Shell(atom, l, exps, coeffs)`` and ``Wavefunction(symbols, coords, shells,
  ...)`` mirror ``fastesp/molden.py:20-49``.
* ``Basis`` mirrors ``fastesp/basis.py:207-238``: one entry per Cartesian
  component, components of a shell in ``cart_order(l)`` order, every component
  sharing the shell's contraction coefficients (which already carry the radial
  prefactor ``N_s(alpha, l)``, ``fastesp/basis.py:53-66``).

* ``pure_to_cartesian(l)`` / ``Basis.block_transform()`` /
  ``Basis.density_matrix_cartesian()`` re-express a MOLDEN wavefunction stored
  in pure (spherical) functions in the Cartesian span of each shell
  (``fastesp/basis.py:1-13,105-180,240-270``): the density is
  P_cart = T P_pure T^T with T built shell by shell from S-orthonormal real
  solid harmonics.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

BOHR = 0.529177210903  # angstrom per bohr (CODATA 2018)

Z_OF = {"H": 1.0, "He": 2.0, "Li": 3.0, "Be": 4.0, "B": 5.0, "C": 6.0, "N": 7.0,
        "O": 8.0, "F": 9.0, "Ne": 10.0, "Na": 11.0, "Mg": 12.0, "Al": 13.0,
        "Si": 14.0, "P": 15.0, "S": 16.0, "Cl": 17.0, "Ar": 18.0, "Br": 35.0,
        "I": 53.0}


@lru_cache(maxsize=None)
def cart_order(l: int) -> tuple:
    """Cartesian components of an l-shell, x-major: (l,0,0), (l-1,1,0), (l-1,0,1), ...

    Same order as ``fastesp/basis.py:27-39``; the table builder and the binary
    basis payload depend on it.
    """
    return tuple((lx, ly, l - lx - ly)
                 for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1))


def primitive_norm(alpha, l: int):
    """Radial prefactor N_s(alpha, l) = (2 alpha/pi)^{3/4} (4 alpha)^{l/2}.

    The coefficient convention of ``fastesp/basis.py:53-66``: stored contraction
    coefficients include this factor but not the angular (2l-1)!! factor.
    """
    alpha = np.asarray(alpha, float)
    return (2.0 * alpha / np.pi) ** 0.75 * (4.0 * alpha) ** (0.5 * l)


@dataclass
class Shell:
    """One contracted Gaussian shell on one atom (``fastesp/molden.py:20-26``)."""
    atom: int
    l: int
    exps: np.ndarray
    coeffs: np.ndarray


@dataclass
class Wavefunction:
    """Atoms + shells (+ optional MO data), ``fastesp/molden.py:29-49``."""
    symbols: list
    coords: np.ndarray
    shells: list
    mo_energy: np.ndarray = field(default_factory=lambda: np.zeros(0))
    mo_occ: np.ndarray = field(default_factory=lambda: np.zeros(0))
    mo_coeff: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    spherical: dict = field(default_factory=dict)
    nelec_override: float | None = None

    @property
    def natom(self) -> int:
        return len(self.symbols)

    @property
    def nao(self) -> int:
        return self.mo_coeff.shape[0]

    @property
    def nelec(self) -> float:
        if self.nelec_override is not None:
            return float(self.nelec_override)
        return float(self.mo_occ.sum())

    @property
    def Z(self) -> np.ndarray:
        return np.array([Z_OF[s] for s in self.symbols], np.float64)


# ---------------------------------------------------------------------------
# Pure (spherical) -> Cartesian
# ---------------------------------------------------------------------------
def _odd_dfact(n: int) -> int:
    """n!! for odd n, (-1)!! = 1."""
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


@lru_cache(maxsize=None)
def cartesian_overlap_block(l: int) -> np.ndarray:
    """Overlap of the Cartesian components of one l-shell with itself, one
    primitive of exponent 1 and coefficient N_s(1, l) (``fastesp/basis.py:
    94-113``).  It does not depend on the exponent: N_s^2 carries (2 alpha)^{l+3/2}.

    1-D moments: int x^n exp(-2 x^2) dx = (n-1)!! / 4^{n/2} sqrt(pi/2), n even.
    """
    comps = cart_order(l)
    pref = float(primitive_norm(1.0, l)) ** 2 * (np.pi / 2.0) ** 1.5

    def mom(n: int) -> float:
        return 0.0 if n % 2 else _odd_dfact(n - 1) / 4.0 ** (n // 2)

    return np.array([[pref * mom(a[0] + b[0]) * mom(a[1] + b[1]) * mom(a[2] + b[2])
                      for b in comps] for a in comps])


def _poly_mul(p: dict, q: dict) -> dict:
    out: dict = {}
    for (a, b, c), u in p.items():
        for (d, e, f), v in q.items():
            k = (a + d, b + e, c + f)
            out[k] = out.get(k, 0) + u * v
    return {k: v for k, v in out.items() if v != 0}


@lru_cache(maxsize=None)
def solid_harmonics(l: int) -> tuple:
    """Real regular solid harmonics of degree l as integer polynomials
    {(ax, ay, az): coefficient}, in MOLDEN order m = 0, +1, -1, +2, -2, ...

    S_l^{+-m} = Pi_l^m(z, r^2) * {Re, Im} (x + i y)^m with
    Pi_l^m = sum_t (-1)^t C(l,t) C(2l-2t,l) (l-2t)!/(l-2t-m)! r^{2t} z^{l-2t-m}
    (Helgaker, Jorgensen & Olsen, Molecular Electronic-Structure Theory, sec.
    6.4.2; no Condon-Shortley phase; the common factor of each row is left out).

    Phase convention: the rows m = +3 and m = -3 of l = 3 are negated.  The
    reference's table (``fastesp/basis.py:119-148``) carries these two signs,
    which it read back from an ORCA MOLDEN file; a wavefunction's density
    depends on the phases of the functions its coefficients refer to.
    """
    from math import comb, factorial

    rows = []
    for m in [0] + [s * k for k in range(1, l + 1) for s in (1, -1)]:
        a = abs(m)
        pi: dict = {}
        for t in range((l - a) // 2 + 1):
            c = (-1) ** t * comb(l, t) * comb(2 * l - 2 * t, l) * factorial(l - 2 * t) \
                // factorial(l - 2 * t - a)
            r2t = {(0, 0, 0): 1}
            for _ in range(t):
                r2t = _poly_mul(r2t, {(2, 0, 0): 1, (0, 2, 0): 1, (0, 0, 2): 1})
            term = _poly_mul(r2t, {(0, 0, l - 2 * t - a): c})
            for k, v in term.items():
                pi[k] = pi.get(k, 0) + v
        if m >= 0:   # Re (x + i y)^a
            ang = {(a - 2 * k, 2 * k, 0): (-1) ** k * comb(a, 2 * k) for k in range(a // 2 + 1)}
        else:        # Im (x + i y)^a
            ang = {(a - 2 * k - 1, 2 * k + 1, 0): (-1) ** k * comb(a, 2 * k + 1)
                   for k in range((a - 1) // 2 + 1)}
        poly = _poly_mul({k: v for k, v in pi.items() if v != 0}, ang)
        if l == 3 and a == 3:
            poly = {k: -v for k, v in poly.items()}
        rows.append(poly)
    return tuple(rows)


@lru_cache(maxsize=None)
def pure_to_cartesian(l: int) -> np.ndarray:
    """T (2l+1, ncart): rows are the solid harmonics of ``solid_harmonics(l)`` in
    the N_s-normalised Cartesian components, scaled so that T S T^T = 1 with
    S = ``cartesian_overlap_block(l)`` (``fastesp/basis.py:151-169``).  The
    identity for l <= 1."""
    comps = cart_order(l)
    if l <= 1:
        return np.eye(len(comps))
    col = {c: i for i, c in enumerate(comps)}
    v = np.zeros((2 * l + 1, len(comps)))
    for r, poly in enumerate(solid_harmonics(l)):
        for k, c in poly.items():
            v[r, col[k]] = c
    S = cartesian_overlap_block(l)
    return v / np.sqrt(np.einsum("ri,ij,rj->r", v, S, v))[:, None]


class Basis:
    """Cartesian basis, one entry per component (``fastesp/basis.py:207-238``)."""

    def __init__(self, wfn: Wavefunction, use_primitive_norm: bool = True):
        self.wfn = wfn
        # the MOLDEN coefficients already carry N_s(alpha, l); the flag exists
        # for call compatibility (the reference stores it and does not use it)
        self.use_primitive_norm = use_primitive_norm
        centers, atoms, ang, exps, coefs, shell_of = [], [], [], [], [], []
        self.shells = []
        for sh in wfn.shells:
            c = np.asarray(wfn.coords[sh.atom], float)
            e = np.asarray(sh.exps, float)
            cf = np.asarray(sh.coeffs, float)
            for comp in cart_order(sh.l):
                centers.append(c)
                atoms.append(sh.atom)
                ang.append(comp)
                exps.append(e)
                coefs.append(cf)
                shell_of.append(len(self.shells))
            self.shells.append(sh)
        self.n = len(centers)
        self.centers = np.asarray(centers, float).reshape(-1, 3)
        self.atoms = np.asarray(atoms, np.int32)
        self.ang = ang
        self.ang_arr = np.asarray(ang, np.int32).reshape(-1, 3)
        self.exps = exps
        self.coefs = coefs
        self.shell_of = np.asarray(shell_of, np.int32)

    @property
    def n_prim(self) -> int:
        """Primitive count with one copy per Cartesian component (``fastesp/cli.py:177``)."""
        return int(sum(len(e) for e in self.exps))

    def shell_offsets(self) -> np.ndarray:
        """First basis-function index of every shell."""
        out = np.zeros(len(self.shells) + 1, np.int64)
        for i, sh in enumerate(self.shells):
            out[i + 1] = out[i] + len(cart_order(sh.l))
        return out

    def block_transform(self) -> np.ndarray:
        """(n_cart, n_pure) map from stored pure coefficients to Cartesian
        components, block-diagonal over the shells (``fastesp/basis.py:240-262``).
        Every stored shell is taken as 2l+1 pure functions, as the reference does."""
        npure = [2 * sh.l + 1 for sh in self.wfn.shells]
        T = np.zeros((self.n, int(sum(npure))))
        oc = op = 0
        for sh, n_p in zip(self.wfn.shells, npure):
            nc = len(cart_order(sh.l))
            T[oc:oc + nc, op:op + n_p] = pure_to_cartesian(sh.l).T
            oc += nc
            op += n_p
        return T

    def density_matrix_cartesian(self, min_occ: float = 1e-10) -> np.ndarray:
        """P = T (sum_i occ_i c_i c_i^T) T^T over the MOs with occ > min_occ
        (``fastesp/basis.py:264-270``)."""
        w = self.wfn
        T = self.block_transform()
        if w.mo_coeff.shape[0] != T.shape[1]:
            raise ValueError(
                f"MO coefficients have {w.mo_coeff.shape[0]} rows, the shells {T.shape[1]} pure "
                "functions: Cartesian d/f input (no [5D]/[7F]) is not supported")
        occ = np.asarray(w.mo_occ) > min_occ
        c = w.mo_coeff[:, occ] * np.sqrt(np.asarray(w.mo_occ)[occ])[None, :]
        Tc = T @ c
        return Tc @ Tc.T


# ---------------------------------------------------------------------------
# Overlap (needed to build synthetic densities with Tr(P S) = N_elec)
# ---------------------------------------------------------------------------
def _os_1d(la: int, lb: int, xpa, xpb, inv2p):
    """1-D Obara-Saika overlap factors s[:, i, j] for i <= la, j <= lb.

    s_{i+1,j} = X_PA s_{ij} + (i s_{i-1,j} + j s_{i,j-1}) / (2p), s_00 = 1
    (the sqrt(pi/p) factor is left out).  Vectorised over primitive pairs.
    """
    n = xpa.shape[0]
    s = np.zeros((n, la + 1, lb + 1))
    s[:, 0, 0] = 1.0
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:
                v = xpa * s[:, i - 1, j]
                if i > 1:
                    v = v + (i - 1) * inv2p * s[:, i - 2, j]
                if j > 0:
                    v = v + j * inv2p * s[:, i - 1, j - 1]
            else:
                v = xpb * s[:, i, j - 1]
                if j > 1:
                    v = v + (j - 1) * inv2p * s[:, i, j - 2]
            s[:, i, j] = v
    return s


def overlap_matrix(basis: Basis) -> np.ndarray:
    """Full Cartesian overlap matrix S (n x n), vectorised per (la, lb) class.

    Same quantity as ``fastesp/basis.py:273-297`` (contracted Cartesian
    Gaussians with the stored coefficients, no extra normalisation).
    """
    shells = basis.shells
    coords = basis.wfn.coords
    off = basis.shell_offsets()
    n = basis.n
    S = np.zeros((n, n))
    by_l: dict[int, list[int]] = {}
    for i, sh in enumerate(shells):
        by_l.setdefault(sh.l, []).append(i)
    for la, ia in by_l.items():
        for lb, ib in by_l.items():
            # primitive lists of both shell classes, then their outer product
            pa_sh = np.concatenate([np.full(len(shells[a].exps), a) for a in ia])
            pa_e = np.concatenate([np.asarray(shells[a].exps, float) for a in ia])
            pa_c = np.concatenate([np.asarray(shells[a].coeffs, float) for a in ia])
            pb_sh = np.concatenate([np.full(len(shells[b].exps), b) for b in ib])
            pb_e = np.concatenate([np.asarray(shells[b].exps, float) for b in ib])
            pb_c = np.concatenate([np.asarray(shells[b].coeffs, float) for b in ib])
            na, nb = len(pa_e), len(pb_e)
            sa_idx = np.repeat(pa_sh, nb)
            sb_idx = np.tile(pb_sh, na)
            ea = np.repeat(pa_e, nb)
            eb = np.tile(pb_e, na)
            cc = np.repeat(pa_c, nb) * np.tile(pb_c, na)
            atom_of = np.array([sh.atom for sh in shells])
            Ac = coords[atom_of[sa_idx]]
            Bc = coords[atom_of[sb_idx]]
            p = ea + eb
            P = (ea[:, None] * Ac + eb[:, None] * Bc) / p[:, None]
            ab2 = ((Ac - Bc) ** 2).sum(1)
            pref = cc * np.exp(-ea * eb / p * ab2) * (np.pi / p) ** 1.5
            inv2p = 0.5 / p
            sx = _os_1d(la, lb, P[:, 0] - Ac[:, 0], P[:, 0] - Bc[:, 0], inv2p)
            sy = _os_1d(la, lb, P[:, 1] - Ac[:, 1], P[:, 1] - Bc[:, 1], inv2p)
            sz = _os_1d(la, lb, P[:, 2] - Ac[:, 2], P[:, 2] - Bc[:, 2], inv2p)
            ca_ = cart_order(la)
            cb_ = cart_order(lb)
            # shell-pair id for the scatter
            key = sa_idx * len(shells) + sb_idx
            ukey, inv = np.unique(key, return_inverse=True)
            blocks = np.zeros((len(ukey), len(ca_), len(cb_)))
            for u, (ax, ay, az) in enumerate(ca_):
                for v, (bx, by, bz) in enumerate(cb_):
                    val = pref * sx[:, ax, bx] * sy[:, ay, by] * sz[:, az, bz]
                    blocks[:, u, v] = np.bincount(inv, weights=val, minlength=len(ukey))
            a_of, b_of = np.divmod(ukey, len(shells))
            rows = off[a_of][:, None, None] + np.arange(len(ca_))[None, :, None]
            cols = off[b_of][:, None, None] + np.arange(len(cb_))[None, None, :]
            S[rows, cols] = blocks
    return S


def build_basis(path, use_primitive_norm: bool = True):
    """read_molden + Basis (``fastesp/basis.py:300-303``)."""
    from .molden import read_molden
    wfn = read_molden(path)
    return wfn, Basis(wfn, use_primitive_norm=use_primitive_norm)
