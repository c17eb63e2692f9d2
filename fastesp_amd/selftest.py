"""``python -m fastesp_amd selftest``: the reference's self-test on the GPU path.

``fastesp selftest`` (``fastesp/cli.py:230-282``) checks the analytic
integrals against independent values.  Here the integrals come from the
device: each case is a two-shell wavefunction whose density matrix selects one
Cartesian component pair, so the table's electron term at C is exactly the
primitive integral <a|1/|r-C||b> (the fold P_ij + P_ji^T of the table build
gives 1 for that pair and every other block is zero).

* Boys functions (``fe_boys_eval``): F_0(T) = sqrt(pi/T) erf(sqrt T)/2,
  F_m(0) = 1/(2m+1) and the recursion (2m+1) F_m = 2T F_{m+1} + exp(-T).
* s-s pairs: K (pi/p)^{3/2} erf(sqrt(p) R)/R at 8 random geometries.
* general pairs (shells up to f, the classes the device table covers): a
  spherical quadrature of the Gaussian product about C (Gauss-Legendre in r
  and cos(theta), uniform in phi), the reference's construction.
"""
from __future__ import annotations

import math

import numpy as np



def pair_case(A, la, al, B, lb, be):
    """(wfn, basis, P): two unit-coefficient primitive shells and a density
    whose only non-zero entry couples the components la (on A) and lb (on B)."""
    from .ints import pair_case as _case
    return _case(A, la, al, B, lb, be)


def _pair_integral(A, la, al, B, lb, be, C) -> float:
    """<a|1/|r-C||b> of two unit-coefficient Cartesian primitives, on the GPU."""
    from .ints import primitive_pair_integral
    return primitive_pair_integral(al, A, la, be, B, lb, np.asarray(C, float).reshape(3))


def quad(C, A, la, al, B, lb, be, R=14.0, ngl=240, ncos=220, nphi=440) -> float:
    from numpy.polynomial.legendre import leggauss
    xs, ws = leggauss(ngl)
    s = 0.5 * R * (xs + 1.0)
    wr = 0.5 * R * ws
    cz, wz = leggauss(ncos)
    phi = (np.arange(nphi) + 0.5) * (2.0 * np.pi / nphi)
    st = np.sqrt(np.clip(1.0 - cz ** 2, 0.0, None))
    total = 0.0
    for k in range(nphi):   # one azimuth at a time keeps the work arrays small
        n = np.stack([st * math.cos(phi[k]), st * math.sin(phi[k]), cz], 1)     # (ncos, 3)
        r = C[None, None, :] + s[None, :, None] * n[:, None, :]                   # (ncos, ngl, 3)
        da, db = r - A, r - B
        f = (np.prod(da ** np.asarray(la), -1) * np.exp(-al * (da * da).sum(-1))
             * np.prod(db ** np.asarray(lb), -1) * np.exp(-be * (db * db).sum(-1)))
        # the volume element s^2 ds dOmega over the 1/s of the potential
        total += float((f * (s * wr)[None, :] * wz[:, None]).sum())
    return total * (2.0 * np.pi / nphi)


def run(verbose: bool = True) -> int:
    from . import _lib
    ctx = _lib.Context.get()
    # Boys
    T = np.concatenate([[0.0], np.geomspace(1e-6, 200.0, 400)])
    F = ctx.boys(T, 8)
    Tp = T[1:]
    f0 = 0.5 * np.sqrt(np.pi / Tp) * np.array([math.erf(math.sqrt(t)) for t in Tp])
    worst_b = float(np.abs(F[1:, 0] - f0).max() / f0.max())
    worst_b = max(worst_b, float(np.abs(F[0] - 1.0 / (2 * np.arange(9) + 1)).max()))
    rec = (2 * Tp[:, None] * F[1:, 1:] + np.exp(-Tp)[:, None]) / (2 * np.arange(8) + 1)
    worst_b = max(worst_b, float((np.abs(F[1:, :8] - rec) / np.abs(F[1:, :8])).max()))
    # s-s closed form
    rng = np.random.default_rng(0)
    worst_ss = 0.0
    for _ in range(8):
        A = rng.normal(0, 1, 3)
        B = A + rng.normal(0, 1, 3)
        C = rng.normal(0, 3, 3)
        al = float(np.exp(rng.uniform(-2, 2.5)))
        be = float(np.exp(rng.uniform(-2, 2.5)))
        p = al + be
        P = (al * A + be * B) / p
        K = math.exp(-al * be / p * float(((A - B) ** 2).sum()))
        Rn = float(np.linalg.norm(C - P))
        ref = K * (np.pi / p) ** 1.5 * math.erf(math.sqrt(p) * Rn) / Rn
        got = _pair_integral(A, (0, 0, 0), al, B, (0, 0, 0), be, C)
        worst_ss = max(worst_ss, abs(got - ref) / abs(ref))
    # general pairs vs quadrature (the geometry of tests/test_integrals.py:72-85)
    A = np.array([0.3, -0.4, 0.2])
    B = A + np.array([1.1, 0.5, -0.6])
    C = np.array([0.1, 1.6, 1.1])
    al, be = 0.45, 0.72
    worst = 0.0
    for la, lb in [((1, 0, 0), (0, 0, 0)), ((2, 0, 0), (0, 1, 0)), ((1, 1, 0), (0, 0, 1)),
                   ((2, 1, 0), (1, 1, 1)), ((0, 0, 3), (1, 2, 0))]:
        ana = _pair_integral(A, la, al, B, lb, be, C)
        ref = quad(C, A, la, al, B, lb, be)
        rel = abs(ana - ref) / max(abs(ref), 1e-300)
        worst = max(worst, rel)
        if verbose:
            print(f"  <{la}|1/r|{lb}>  device {ana:+.10e}  quadrature {ref:+.10e}  rel {rel:.1e}")
    ok = worst_b < 1e-13 and worst_ss < 1e-13 and worst < 1e-11
    print(f"  Boys vs closed forms  worst rel {worst_b:.1e}")
    print(f"  s-s vs closed form    worst rel {worst_ss:.1e}")
    print(f"  general vs quadrature worst rel {worst:.1e}")
    print("  SELFTEST", "PASS" if ok else "FAIL")
    return 0 if ok else 1
