"""Pins of the oracle's integral primitives (CPU).

Known-answer tests restate the reference's own (``tests/test_integrals.py``)
and add closed forms, mpmath and identities so that a dropped term, a wrong
sign or a transposed index fails at least one of them.
"""
import math

import mpmath
import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

from oracle.ints import (M_ABS, a_table, boys, dF_dR, e1d, mono_schedule, pair_type,
                         primitive_pair_integral)


def boys_mp(m, T):
    if T == 0:
        return mpmath.mpf(1) / (2 * m + 1)
    return mpmath.gammainc(m + mpmath.mpf(0.5), 0, T) / (2 * mpmath.power(T, m + mpmath.mpf(0.5)))


def test_boys_known_values():
    # tests/test_integrals.py:13-27
    for m in range(6):
        assert boys(m, 1e-14) == pytest.approx(1.0 / (2 * m + 1), rel=1e-12)
    for T in (0.1, 1.0, 5.0, 30.0):
        assert boys(0, T) == pytest.approx(
            0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T)), rel=1e-13)
    assert boys(0, 1e4) == pytest.approx(0.5 * math.sqrt(math.pi / 1e4), rel=1e-6)
    T, h = 2.7, 1e-6
    assert (boys(1, T + h) - boys(1, T - h)) / (2 * h) == pytest.approx(-boys(2, T), rel=1e-6)


def test_boys_vs_mpmath():
    mpmath.mp.dps = 40
    Ts = [0.0, 1e-12, 1e-6, 0.01, 0.3, 1.0, 2.5, 4.9, 7.1, 12.0, 25.0, 36.0, 59.9, 60.1,
          120.0, 1e3, 1e5]
    worst = 0.0
    for m in range(9):
        for T in Ts:
            ref = float(boys_mp(m, mpmath.mpf(T)))
            worst = max(worst, abs(boys(m, T) - ref) / ref)
    assert worst < 2e-15


def test_a_table_closed_form():
    a = a_table(12)
    for t in range(13):
        for j in range(t // 2 + 1):
            ref = (-1) ** (t - j) * math.factorial(t) / (2 ** j * math.factorial(j)
                                                         * math.factorial(t - 2 * j))
            assert a[t, j] == ref


def test_a_table_matches_derivative_definition():
    # tests/test_integrals.py:88-111
    W = [-1 / 60, 3 / 20, -3 / 4, 0, 3 / 4, -3 / 20, 1 / 60]
    p, x, h = 0.83, 1.37, 3e-3
    A = a_table(8)

    def d(order, xx):
        if order == 0:
            return boys(0, p * xx * xx)
        return sum(W[k] * d(order - 1, xx + (k - 3) * h) for k in range(7)) / h

    for t in range(6):
        ana = sum(A[t, j] * (2 * p) ** (t - j) * x ** (t - 2 * j) * boys(t - j, p * x * x)
                  for j in range(t // 2 + 1))
        assert ana == pytest.approx(d(t, x), rel=2e-3, abs=1e-7)


def test_e1d_polynomial_identity():
    # sum_t E_t Lambda_t(u) = (u+xa)^i (u+xb)^j exp(-p u^2),
    # Lambda_t(u) = (-d/du)^t exp(-p u^2) = p^{t/2} H_t(sqrt(p) u) exp(-p u^2)
    from numpy.polynomial.hermite import hermval
    rng = np.random.default_rng(3)
    for i in range(5):
        for j in range(5):
            p, xa, xb = 0.7 + rng.random(), rng.normal(), rng.normal()
            E = e1d(i, j, p, xa, xb)[0, i, j]
            for u in (-1.3, -0.2, 0.4, 1.9):
                lhs = sum(E[t] * p ** (t / 2) * hermval(math.sqrt(p) * u, [0] * t + [1])
                          for t in range(i + j + 1))
                rhs = (u + xa) ** i * (u + xb) ** j
                assert lhs == pytest.approx(rhs, rel=1e-12, abs=1e-12)


def test_ss_closed_form():
    # tests/test_integrals.py:30-44
    rng = np.random.default_rng(0)
    for _ in range(8):
        A = rng.normal(0, 1, 3)
        B = A + rng.normal(0, 1, 3)
        C = rng.normal(0, 3, 3)
        al = float(np.exp(rng.uniform(-2, 2.5)))
        be = float(np.exp(rng.uniform(-2, 2.5)))
        p = al + be
        P = (al * A + be * B) / p
        K = math.exp(-al * be / p * float(((A - B) ** 2).sum()))
        R = float(np.linalg.norm(C - P))
        ref = K * (np.pi / p) ** 1.5 * math.erf(np.sqrt(p) * R) / R
        got = primitive_pair_integral(al, A, (0, 0, 0), be, B, (0, 0, 0), C)
        assert got == pytest.approx(ref, rel=1e-13)


def quad(C, A, la, al, B, lb, be, R=14.0, ngl=240, ncos=220, nphi=440):
    """Spherical quadrature about C (the recipe of tests/test_integrals.py:47-69)."""
    xs, ws = leggauss(ngl)
    s = 0.5 * R * (xs + 1.0)
    wr = 0.5 * R * ws
    cz, wz = leggauss(ncos)
    phi = (np.arange(nphi) + 0.5) * (2 * np.pi / nphi)
    st = np.sqrt(np.clip(1 - cz ** 2, 0, None))
    nx = (st[None, :] * np.cos(phi)[:, None])[..., None]
    ny = (st[None, :] * np.sin(phi)[:, None])[..., None]
    nz = np.broadcast_to(cz[None, :, None], (nphi, ncos, 1))
    ss = s[None, None, :]
    X, Y, Z = C[0] + ss * nx, C[1] + ss * ny, C[2] + ss * nz
    rho = ((X - A[0]) ** la[0] * (Y - A[1]) ** la[1] * (Z - A[2]) ** la[2]
           * (X - B[0]) ** lb[0] * (Y - B[1]) ** lb[1] * (Z - B[2]) ** lb[2]
           * np.exp(-al * ((X - A[0]) ** 2 + (Y - A[1]) ** 2 + (Z - A[2]) ** 2)
                    - be * ((X - B[0]) ** 2 + (Y - B[1]) ** 2 + (Z - B[2]) ** 2)))
    return float((rho * ss * wr[None, None, :] * (wz[None, :, None] * (2 * np.pi / nphi))).sum())


@pytest.mark.parametrize("la,lb", [
    ((0, 0, 0), (0, 0, 0)), ((1, 0, 0), (0, 0, 0)), ((0, 1, 0), (1, 0, 0)),
    ((1, 1, 0), (0, 0, 1)), ((2, 0, 0), (0, 1, 0)), ((1, 0, 0), (0, 0, 2)),
    ((2, 1, 0), (1, 0, 0)), ((2, 0, 0), (2, 0, 0)), ((3, 0, 0), (1, 1, 0)),
    ((2, 1, 0), (2, 1, 0)), ((3, 2, 0), (2, 1, 0)), ((3, 3, 3), (3, 3, 3)),
])
def test_general_pair_vs_quadrature(la, lb):
    # tests/test_integrals.py:72-85
    A = np.array([0.3, -0.4, 0.2])
    B = A + np.array([1.1, 0.5, -0.6])
    C = np.array([0.1, 1.6, 1.1])
    al, be = 0.45, 0.72
    ana = primitive_pair_integral(al, A, la, be, B, lb, C)
    ref = quad(C, A, la, al, B, lb, be)
    assert ana == pytest.approx(ref, rel=1e-11, abs=1e-14)


@pytest.mark.parametrize("L", range(7))
def test_pair_type_reproduces_hermite_potential(L):
    """sum_k M0[k,t] (2p)^m R^alpha F_m * 2pi/p == (-1)^|t| (d/dR)^t F for every t."""
    keys, M0, wmom, abs_t = pair_type(L)
    T = L + 1
    rng = np.random.default_rng(L)
    p = 0.9
    R = rng.normal(0, 0.8, 3)
    TT = p * float(R @ R)
    F = [boys(m, TT) for m in range(L + 1)]
    vals = np.array([(2 * p) ** k[3] * R[0] ** k[0] * R[1] ** k[1] * R[2] ** k[2] * F[k[3]]
                     for k in keys]) * 2 * np.pi / p
    for tx in range(T):
        for ty in range(T - tx):
            for tz in range(T - tx - ty):
                col = (tx * T + ty) * T + tz
                ref = (-1) ** (tx + ty + tz) * dF_dR((tx, ty, tz), R, p)
                assert vals @ M0[:, col] == pytest.approx(ref, rel=1e-12, abs=1e-13)
    assert abs_t.shape == (T ** 3,) and wmom.shape == (T ** 3,)


def test_term_and_mono_counts():
    # per-L sizes quoted for the reference's tables (nterm, nmono)
    assert [pair_type(L)[0].shape[0] for L in range(7)] == [1, 4, 11, 24, 46, 80, 130]
    assert [mono_schedule(L)[0].shape[0] for L in range(7)] == [1, 4, 10, 20, 35, 56, 84]
    for L in range(7):
        monos = mono_schedule(L)[0]
        assert len({tuple(m) for m in monos}) == len(monos)
        assert all(sum(m) <= L for m in monos)


def test_m_abs_moments():
    """The screening moments.  The reference's comment calls them
    int |He_n(x)| exp(-x^2/2) dx; that holds for n = 0, 1 only (for n >= 2 the
    stored values differ, e.g. 1.5958 vs 2.4261 at n = 2).  They enter only the
    heuristic screening bound, so the oracle keeps the reference's values; their
    effect is pinned by comparing screening decisions with the compiled
    reference builder (test_oracle_table.py)."""
    from numpy.polynomial.hermite_e import hermeval
    x = np.linspace(-40, 40, 2_000_001)
    w = np.exp(-x * x / 2)
    for n in (0, 1):
        v = np.trapezoid(np.abs(hermeval(x, [0] * n + [1])) * w, x)
        assert v == pytest.approx(M_ABS[n], rel=1e-9)
    v2 = np.trapezoid(np.abs(hermeval(x, [0, 0, 1])) * w, x)
    assert abs(v2 - M_ABS[2]) > 0.5  # documented discrepancy, kept on purpose
    assert M_ABS[2] == pytest.approx(2 * math.sqrt(2 / math.pi), rel=1e-15)


def test_generator_matches_oracle():
    """The kernels' generated per-L tables equal the oracle's keys and schedule."""
    import importlib.util
    from pathlib import Path
    spec = importlib.util.spec_from_file_location(
        "gen", Path(__file__).resolve().parents[1] / "fastesp_amd/csrc/gen_lclass.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    for L in range(7):
        keys, _ = gen.keys_of(L)
        order, parent, axis, aidx = gen.schedule_of(L)
        ok = pair_type(L)[0]
        om, op, oa, oi = mono_schedule(L)
        assert [tuple(k) for k in ok] == keys
        assert [tuple(m) for m in om] == order
        assert list(op) == parent and list(oa) == axis and list(oi) == aidx
