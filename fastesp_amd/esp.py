"""``ESPEngine``: the reference's unflattened evaluator, on the GPU.

Surface of ``fastesp/esp.py:150-296``:

* ``ESPEngine(basis, density)`` holds the contracted primitive pairs of every
  significant shell pair (``_build``/``_make_pair``, ``esp.py:168-257``);
* ``esp_electron(points, tol=1e-15, verbose=False)`` skips a pair when
  ``K bound / max(dmax - reff, 0.05) < tol``, dmax being the largest distance
  from the pair centre to any of the points (``esp.py:260-283``), and sums
  ``K sum_k g_k R^a_k F_mk`` over the rest;
* ``esp(points, tol, verbose)`` is the nuclear term with the charges of
  ``esp.py:293`` (H, C, N, O, S, P only; other symbols raise KeyError) minus
  ``esp_electron``.

Here every call builds the device table of its points with that skip rule
(screen rule 2 of ``fe_table_build``) and evaluates it with the same kernels as
``PairTable``.  ``prims`` materialises the host ``PrimPair`` list, in the
reference's order, from an unscreened device build.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .builder import SCREEN_ENGINE, build_table_payload, export_flat
from .formats import basis_payload

# fastesp/esp.py:293
ENGINE_Z = {"H": 1.0, "C": 6.0, "N": 7.0, "O": 8.0, "S": 16.0, "P": 15.0}

class PrimPair:
    """One contracted primitive pair (``fastesp/esp.py:150-156``)."""
    __slots__ = ("p", "center", "K", "alpha", "m", "coef", "bound", "npow", "reff")

    def __init__(self, p, center, K, alpha, m, coef, bound, npow, reff):
        self.p, self.center, self.K = p, center, K
        self.alpha, self.m, self.coef = alpha, m, coef
        self.bound, self.npow, self.reff = bound, npow, reff


class ESPEngine:
    def __init__(self, basis, density: np.ndarray, ctx: _lib.Context | None = None):
        self.basis = basis
        self.density = np.ascontiguousarray(density, np.float64)
        self.pair_atoms: list = []
        self._ctx = ctx
        self._prims = None

    # -- the candidate list ------------------------------------------------
    @property
    def prims(self) -> list:
        """Every candidate primitive pair (|blk| >= 1e-15, K >= 1e-30) in the
        order of ``_build`` (``fastesp/esp.py:168-257``)."""
        if self._prims is None:
            t = build_table_payload(self._payload(np.zeros((0, 3))), 0.0, 0, self._ctx)
            f = export_flat(t)
            kb, reff = f.meta["kbound"], f.meta["reff"]
            out = []
            for i in range(f.n_pairs):
                o, n = int(f.pp_off[i]), int(f.pp_nt[i])
                mo, nm = int(f.pp_monooff[i]), int(f.pp_nmono[i])
                alpha = self._mono_exponents(f.mo_parent[mo:mo + nm], f.mo_axis[mo:mo + nm])
                alpha = alpha[f.tm_aidx[o:o + n]]
                m = f.tm_m[o:o + n].copy()
                K = float(f.pp_K[i])
                out.append(PrimPair(float(f.pp_exp[i]),
                                    np.array([f.pp_cx[i], f.pp_cy[i], f.pp_cz[i]]), K, alpha, m,
                                    f.tm_g[o:o + n].copy(), float(kb[i]) / K, m.copy(),
                                    float(reff[i])))
            self._prims = out
        return self._prims

    @staticmethod
    def _mono_exponents(parent, axis) -> np.ndarray:
        """Monomial exponents from the parent/axis schedule (``table.py:86-88``)."""
        e = np.zeros((len(parent), 3), np.int64)
        for j in range(1, len(parent)):
            e[j] = e[int(parent[j])]
            e[j, int(axis[j])] += 1
        return e

    # -- evaluation ----------------------------------------------------------
    def _payload(self, points) -> dict:
        return basis_payload(self.basis.wfn, self.basis, self.density, points)

    @staticmethod
    def _points(points) -> np.ndarray:
        pts = np.ascontiguousarray(np.asarray(points, float).reshape(-1, 3))
        if len(pts) == 0:
            # the reference's d.max() over an empty point set raises
            raise ValueError("zero-size array to reduction operation maximum which has no identity")
        return pts

    def _table(self, points, tol: float, Z=None):
        pts = self._points(points)
        pl = self._payload(pts)
        if Z is not None:
            pl["Z"] = Z
        return pts, build_table_payload(pl, float(tol), SCREEN_ENGINE, self._ctx)

    @staticmethod
    def _report(t, verbose: bool) -> None:
        if verbose:
            info = t.info()
            used = int(info["n_pairs"])
            print(f"    primitive pairs: used {used}, skipped {int(info['n_candidates']) - used}")

    def esp_electron(self, points, tol: float = 1e-15, verbose: bool = False) -> np.ndarray:
        """Positive electron integral at the points (``fastesp/esp.py:260-283``)."""
        pts, t = self._table(points, tol)
        self._report(t, verbose)
        return t.electron(pts)

    def esp(self, points, tol: float = 1e-15, verbose: bool = False) -> np.ndarray:
        """nuclear - electron in Hartree/e (``fastesp/esp.py:285-296``)."""
        self._points(points)   # the electron part fails first, as in the reference
        w = self.basis.wfn
        Z = np.array([ENGINE_Z[s] for s in w.symbols], np.float64)
        pts, t = self._table(points, tol, Z)
        self._report(t, verbose)
        return t.esp(pts)
