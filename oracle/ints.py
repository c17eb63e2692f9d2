"""Integral primitives of the oracle (FP64, NumPy/SciPy).

Restates ``fastesp/ints.py`` and the per-L bookkeeping of ``fastesp/esp.py``.
Pins (see tests/test_oracle_ints.py):

* ``boys``      -- closed forms F_m(0)=1/(2m+1), F_0 = sqrt(pi/T) erf(sqrt T)/2,
                   dF_m/dT = -F_{m+1}, mpmath at 40 digits; the reference KATs of
                   ``tests/test_integrals.py:13-27``.
* ``a_table``   -- the closed form a[t,j] = (-1)^(t-j) t! / (2^j j! (t-2j)!) and
                   the finite-difference check of ``tests/test_integrals.py:88-111``.
* ``e1d``       -- the polynomial identity sum_t E_t Lambda_t(u) = (u+xa)^i (u+xb)^j
                   e^{-p u^2} (``benchmarks/check_integrals.py:35-75``).
* ``primitive_pair_integral`` -- the s-s closed form and the 12 spherical
                   quadrature cases of ``tests/test_integrals.py:30-85``.
* ``pair_type`` -- (d/dR)^t F expanded through its keys equals ``dF_dR``.
* ``M_ABS``     -- the moments int |He_n(x)| e^{-x^2/2} dx recomputed by quadrature.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np
from scipy.special import hyp1f1

# int |He_n(x)| exp(-x^2/2) dx, n = 0..11: the screening-bound moments of
# fastesp/esp.py:25-28 (same values as fastesp/native/build_pairs.cpp:98-102).
M_ABS = np.array([2.5066282746310002, 2.0, 1.5957691216057308, 4.0,
                  10.212920779565312, 33.63284766793385, 136.5963625280984,
                  665.1416330443616, 3768.863619586618, 25506.330595911374,
                  202331.81842358326, 1872154.2572862155])


def boys(m, T):
    """F_m(T) = int_0^1 t^{2m} e^{-T t^2} dt via 1F1 (``fastesp/ints.py:35-40``)."""
    m = np.asarray(m)
    T = np.asarray(T, dtype=np.float64)
    out = hyp1f1(m + 0.5, m + 1.5, -T) / (2 * m + 1)
    return out if np.ndim(out) else float(out)


@lru_cache(maxsize=None)
def a_table(lmax: int) -> np.ndarray:
    """a[t, j] with d^t/dx^t F_0(p x^2) = sum_j a[t,j] (2p)^(t-j) x^(t-2j) F_{t-j}.

    Built by the recurrence a[t+1,j] = -a[t,j] + (t-2j+2) a[t,j-1]
    (``fastesp/ints.py:46-59``; ``fastesp/esp.py:31-41``).
    """
    a = np.zeros((lmax + 1, lmax // 2 + 2))
    a[0, 0] = 1.0
    for t in range(lmax):
        prev = a[t]
        nxt = -prev.copy()
        j = np.arange(1, a.shape[1])
        nxt[1:] += (t - 2 * j + 2) * prev[:-1]
        a[t + 1] = nxt
    a.setflags(write=False)
    return a


def e1d(i: int, j: int, p, xa, xb) -> np.ndarray:
    """Hermite coefficients E[..., i', j', t] of (u+xa)^i' (u+xb)^j' e^{-p u^2}.

    E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + xa E^{ij}_t + (t+1) E^{ij}_{t+1}
    (``fastesp/ints.py:62-81``, batched form ``fastesp/esp.py:108-147``).
    ``p, xa, xb`` are scalars or (B,) arrays; the result is (B, i+1, j+1, i+j+1).
    """
    p = np.atleast_1d(np.asarray(p, float))
    xa = np.atleast_1d(np.asarray(xa, float))
    xb = np.atleast_1d(np.asarray(xb, float))
    B = p.shape[0]
    tmax = i + j
    e = np.zeros((B, i + 1, j + 1, tmax + 1))
    e[:, 0, 0, 0] = 1.0
    h = 0.5 / p
    tt = np.arange(1, tmax + 1, dtype=float)
    for ii in range(i + 1):
        for jj in range(j + 1):
            if ii == 0 and jj == 0:
                continue
            if ii > 0:
                src, x = e[:, ii - 1, jj], xa
            else:
                src, x = e[:, ii, jj - 1], xb
            cur = x[:, None] * src
            cur[:, 1:] += src[:, :-1] * h[:, None]
            cur[:, :-1] += src[:, 1:] * tt[None, :]
            e[:, ii, jj] = cur
    return e


def E_coefficients(la, lb, p, A, B, P) -> dict:
    """3-D E_t for one primitive pair as {(tx,ty,tz): E} (``fastesp/ints.py:84-106``)."""
    out = {}
    ex = e1d(la[0], lb[0], p, P[0] - A[0], P[0] - B[0])[0, la[0], lb[0]]
    ey = e1d(la[1], lb[1], p, P[1] - A[1], P[1] - B[1])[0, la[1], lb[1]]
    ez = e1d(la[2], lb[2], p, P[2] - A[2], P[2] - B[2])[0, la[2], lb[2]]
    for tx, cx in enumerate(ex):
        for ty, cy in enumerate(ey):
            for tz, cz in enumerate(ez):
                v = cx * cy * cz
                if v != 0.0:
                    out[(tx, ty, tz)] = v
    return out


def dF_dR(t, R, p) -> float:
    """(d/dR)^t [(2 pi/p) F_0(p R^2)] (``fastesp/ints.py:109-135``)."""
    R = np.asarray(R, float)
    lt = int(sum(t))
    a = a_table(max(lt, 1))
    T = p * float(R @ R)
    tot = 0.0
    for jx in range(t[0] // 2 + 1):
        for jy in range(t[1] // 2 + 1):
            for jz in range(t[2] // 2 + 1):
                c = a[t[0], jx] * a[t[1], jy] * a[t[2], jz]
                if c == 0.0:
                    continue
                jj = jx + jy + jz
                tot += (c * (2.0 * p) ** (lt - jj) * R[0] ** (t[0] - 2 * jx)
                        * R[1] ** (t[1] - 2 * jy) * R[2] ** (t[2] - 2 * jz)
                        * boys(lt - jj, T))
    return 2.0 * np.pi / p * tot


def primitive_pair_integral(alpha, A, la, beta, B, lb, C) -> float:
    """<g_A^la | 1/|r-C| | g_B^lb> for bare primitives (``fastesp/ints.py:138-158``)."""
    A = np.asarray(A, float)
    B = np.asarray(B, float)
    C = np.asarray(C, float)
    p = alpha + beta
    P = (alpha * A + beta * B) / p
    K = math.exp(-alpha * beta / p * float((A - B) @ (A - B)))
    R = C - P
    tot = 0.0
    for t, et in E_coefficients(la, lb, p, A, B, P).items():
        tot += et * (-1.0) ** sum(t) * dF_dR(t, R, p)
    return K * tot


# ---------------------------------------------------------------------------
# per-L contraction map
# ---------------------------------------------------------------------------
@lru_cache(maxsize=None)
def pair_type(L: int):
    """Everything of a shell pair that depends on L = l_A + l_B only.

    Returns ``keys`` (nterm x 4 int: alpha_x, alpha_y, alpha_z, m, sorted
    lexicographically), ``M0`` (nterm x T^3: the map etab_t -> G_k including the
    (-1)^|t| sign), ``wmom`` (T^3: products of |He| moments), ``abs_t`` (T^3),
    with T = L+1 and t flattened as (tx*T + ty)*T + tz.  Restates
    ``fastesp/esp.py:54-105`` and ``fastesp/native/build_pairs.cpp:104-151``.
    """
    T = L + 1
    a = a_table(max(L, 1))
    acc: dict = {}
    for tx in range(T):
        for ty in range(T - tx):
            for tz in range(T - tx - ty):
                lt = tx + ty + tz
                col = (tx * T + ty) * T + tz
                sgn = -1.0 if lt % 2 else 1.0
                for jx in range(tx // 2 + 1):
                    for jy in range(ty // 2 + 1):
                        for jz in range(tz // 2 + 1):
                            c = a[tx, jx] * a[ty, jy] * a[tz, jz]
                            if c == 0.0:
                                continue
                            key = (tx - 2 * jx, ty - 2 * jy, tz - 2 * jz,
                                   lt - jx - jy - jz)
                            row = acc.setdefault(key, np.zeros(T ** 3))
                            row[col] += sgn * c
    keys = sorted(acc)
    M0 = np.array([acc[k] for k in keys])
    g = np.arange(T)
    abs_t = (g[:, None, None] + g[None, :, None] + g[None, None, :]).ravel()
    wmom = (M_ABS[:T, None, None] * M_ABS[None, :T, None] * M_ABS[None, None, :T]).ravel()
    return np.array(keys, np.int64).reshape(-1, 4), M0, wmom, abs_t


@lru_cache(maxsize=None)
def mono_schedule(L: int):
    """Monomial schedule of a class-L pair: closure of the key alphas under
    decrement, sorted; parent = drop one power along the first nonzero axis
    (``fastesp/table.py:198-221``; ``fastesp/native/build_pairs.cpp:330-369``).

    Returns (monos (nmono,3), parent (nmono,), axis (nmono,), aidx (nterm,)).
    """
    keys = pair_type(L)[0]
    grown = set(tuple(int(v) for v in k[:3]) for k in keys)
    stack = list(grown)
    while stack:
        u = stack.pop()
        for ax in range(3):
            if u[ax] > 0:
                w = list(u)
                w[ax] -= 1
                w = tuple(w)
                if w not in grown:
                    grown.add(w)
                    stack.append(w)
    order = sorted(grown)
    idx = {m: i for i, m in enumerate(order)}
    parent = np.zeros(len(order), np.int32)
    axis = np.zeros(len(order), np.int32)
    for i, m in enumerate(order):
        for ax in range(3):
            if m[ax] > 0:
                w = list(m)
                w[ax] -= 1
                parent[i] = idx[tuple(w)]
                axis[i] = ax
                break
    aidx = np.array([idx[tuple(int(v) for v in k[:3])] for k in keys], np.int32)
    return np.array(order, np.int32).reshape(-1, 3), parent, axis, aidx
