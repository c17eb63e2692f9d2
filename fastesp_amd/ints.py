"""The reference's integral entry points, evaluated on the GPU.

``fastesp`` exports ``boys`` and ``primitive_pair_integral`` from
``fastesp/ints.py:35-40,138-158`` (``fastesp/__init__.py:38``).  Here both run
through ``libfastesp_hip.so``:

* ``boys(m, T)`` evaluates F_m(T) with the device scheme (``fe_boys_eval``:
  tabulated Taylor expansion + downward recursion, asymptotic form at large T);
* ``primitive_pair_integral(alpha, A, la, beta, B, lb, C)`` is the electron
  term of a two-shell table whose density selects one Cartesian component pair
  (the device table build + evaluation), which for unit-coefficient primitives
  is exactly <g_A^la | 1/|r - C| | g_B^lb>.

There is no CPU fallback: without the library or a device both raise
``HipBackendError``.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def boys(m, T):
    """Boys function F_m(T) = int_0^1 t^{2m} exp(-T t^2) dt (``fastesp/ints.py:35-40``).

    ``m`` and ``T`` broadcast against each other; a scalar result comes back as
    a float.  The device scheme covers integer orders 0 <= m <= 8 and T >= 0;
    other arguments raise ``ValueError``."""
    m = np.asarray(m)
    T = np.asarray(T, dtype=np.float64)
    mb, Tb = np.broadcast_arrays(m, T)
    if mb.size == 0:
        return np.zeros(mb.shape)
    mi = mb.astype(np.int64)
    if not np.array_equal(mi, mb) or mi.min() < 0 or mi.max() > _lib.FE_BOYS_MMAX:
        raise ValueError(f"boys: orders must be integers in [0, {_lib.FE_BOYS_MMAX}]")
    if not np.isfinite(Tb).all() or Tb.min() < 0.0:
        raise ValueError("boys: T must be finite and >= 0")
    F = _lib.Context.get().boys(Tb.ravel(), int(mi.max()))
    out = F[np.arange(F.shape[0]), mi.ravel()].reshape(mb.shape)
    return out if out.ndim else float(out)


def pair_case(A, la, alpha, B, lb, beta):
    """(wfn, basis, P) of two unit-coefficient primitive shells on A and B and a
    density whose only non-zero entry couples component ``la`` (on A) with
    component ``lb`` (on B).  The table build folds P_ij + P_ji^T, so that pair
    carries weight 1 and every other block is zero."""
    from .basis import Basis, Shell, Wavefunction, cart_order
    la, lb = tuple(int(x) for x in la), tuple(int(x) for x in lb)
    shells = [Shell(0, sum(la), np.array([float(alpha)]), np.array([1.0])),
              Shell(1, sum(lb), np.array([float(beta)]), np.array([1.0]))]
    wfn = Wavefunction(symbols=["H", "H"], coords=np.array([A, B], float), shells=shells)
    basis = Basis(wfn)
    P = np.zeros((basis.n, basis.n))
    i = cart_order(sum(la)).index(la)
    j = len(cart_order(sum(la))) + cart_order(sum(lb)).index(lb)
    P[i, j] = 1.0
    return wfn, basis, P


def primitive_pair_integral(alpha, A, la, beta, B, lb, C, a_tab=None, E=None):
    """< g_A^la | 1/|r - C| | g_B^lb > for bare (unnormalised) primitives
    (``fastesp/ints.py:138-158``), on the GPU.

    ``la``/``lb`` are Cartesian exponent triples with l <= 3 each (the shells
    the device table covers).  ``C`` is one point (a float comes back) or an
    (n, 3) array (an (n,) array comes back).  ``a_tab`` and ``E`` are the
    reference's precomputed-table arguments; the device derives both itself,
    so they are accepted and ignored."""
    from .builder import build_table_hip
    if max(sum(la), sum(lb)) > 3 or min(min(la), min(lb)) < 0:
        raise ValueError("primitive_pair_integral: components up to f (l <= 3) only")
    wfn, basis, P = pair_case(A, la, alpha, B, lb, beta)
    Cs = np.asarray(C, float)
    pts = Cs.reshape(-1, 3)
    v = build_table_hip(wfn, basis, P, pts, tol=0.0).electron(pts)
    return float(v[0]) if Cs.ndim == 1 else v
