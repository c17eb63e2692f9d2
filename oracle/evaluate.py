"""Oracle per-point evaluation and its adjoint (FP64 NumPy).

* ``electron/nuclear/esp`` restate ``PairTable.electron/nuclear/esp``
  (``fastesp/table.py:73-104``): per pair R = C - P, T = p|R|^2, F_m(T) for
  m <= max m via scipy ``hyp1f1`` (``fastesp/table.py:85``), monomials by the
  parent/axis schedule (``fastesp/table.py:86-88``), acc = sum_k g_k mono F,
  electron += K acc; V = nuclear - electron.  Pairs are vectorised in groups
  of equal L instead of the reference's one-pair Python loop.
* ``field``   -- dV/dC.  d/dC [R^a F_m(p R^2)] = (d R^a/dR) F_m - 2p R R^a F_{m+1}
  and d/dC Z/|C-A| = -Z (C-A)/|C-A|^3.  No reference counterpart.
* ``grad_G``  -- dL/dG_k = -sum_i gV_i K R_i^a F_m(T_i) for every kept pair.
* ``grad_P``  -- dL/dP by the adjoint of the table contraction
  (``fastesp/esp.py:241-250``) including the density fold P_ij + P_ji^T
  (``fastesp/esp.py:185-187``): both triangles get the same value.

Pins: primitive integrals and quadrature (evaluator), central finite
differences of ``esp`` in C and in P (adjoint; V is linear in P), the s-s
closed-form field.
"""
from __future__ import annotations

import numpy as np

from fastesp_amd.basis import cart_order

from .engine import etab_of, shell_layout, shell_pair_hermite
from .ints import boys, mono_schedule, pair_type


def _groups(table):
    L = table.pp_L
    for l in np.unique(L):
        yield int(l), np.nonzero(L == l)[0]


def _chunks(idx, npts, budget=3_000_000):
    b = max(1, budget // max(npts, 1))
    for s in range(0, len(idx), b):
        yield idx[s:s + b]


def _pair_arrays(table, sel, L):
    nt = pair_type(L)[0].shape[0]
    g = np.stack([table.tm_g[table.pp_off[i]:table.pp_off[i] + nt] for i in sel])
    P = np.stack([table.pp_cx[sel], table.pp_cy[sel], table.pp_cz[sel]], 1)
    return P, table.pp_exp[sel], table.pp_K[sel], g


def _monos(R, L):
    """mono[..., j] built by the parent/axis schedule (``fastesp/table.py:86-88``)."""
    monos, parent, axis, _ = mono_schedule(L)
    out = np.empty(R.shape[:-1] + (len(monos),))
    out[..., 0] = 1.0
    for j in range(1, len(monos)):
        out[..., j] = out[..., parent[j]] * R[..., axis[j]]
    return out


def electron(table, points=None) -> np.ndarray:
    pts = table.pts if points is None else np.asarray(points, float).reshape(-1, 3)
    out = np.zeros(len(pts))
    for L, idx in _groups(table):
        keys = pair_type(L)[0]
        aidx = mono_schedule(L)[3]
        for sel in _chunks(idx, len(pts)):
            P, p, K, g = _pair_arrays(table, sel, L)
            R = pts[None, :, :] - P[:, None, :]
            T = p[:, None] * np.einsum("bnk,bnk->bn", R, R)
            F = np.stack([boys(m, T) for m in range(L + 1)], -1)
            mono = _monos(R, L)
            acc = np.zeros(T.shape)
            for k in range(len(keys)):
                acc += g[:, k, None] * mono[..., aidx[k]] * F[..., keys[k, 3]]
            out += (K[:, None] * acc).sum(0)
    return out


def nuclear(table, points=None) -> np.ndarray:
    pts = table.pts if points is None else np.asarray(points, float).reshape(-1, 3)
    d = np.linalg.norm(pts[:, None, :] - table.xyz[None, :, :], axis=2)
    return (table.Z[None, :] / d).sum(1)


def esp(table, points=None) -> np.ndarray:
    return nuclear(table, points) - electron(table, points)


# charges of ``ESPEngine.esp`` (``fastesp/esp.py:293``): only these six
# elements; any other symbol raises KeyError there
ENGINE_Z = {"H": 1.0, "C": 6.0, "N": 7.0, "O": 8.0, "S": 16.0, "P": 15.0}


def engine_nuclear(symbols, coords, points) -> np.ndarray:
    """sum_A Z_A/|C-R_A| with the charges of ``fastesp/esp.py:291-296``."""
    pts = np.asarray(points, float).reshape(-1, 3)
    nuc = np.zeros(len(pts))
    for sym, xyz in zip(symbols, np.asarray(coords, float)):
        nuc += ENGINE_Z[sym] / np.sqrt(((pts - xyz) ** 2).sum(axis=1))
    return nuc


# ---------------------------------------------------------------------------
def _dmono(R, keys, axis):
    """d R^a / dR_axis for every key alpha: a_axis R^(a - e_axis)."""
    a = keys[:, :3].copy()
    c = a[:, axis].astype(float)
    a[:, axis] = np.maximum(a[:, axis] - 1, 0)
    val = (R[..., None, 0] ** a[:, 0] * R[..., None, 1] ** a[:, 1]
           * R[..., None, 2] ** a[:, 2])
    return val * c


def field(table, points=None) -> np.ndarray:
    """dV/dC (N x 3)."""
    pts = table.pts if points is None else np.asarray(points, float).reshape(-1, 3)
    d = pts[:, None, :] - table.xyz[None, :, :]
    r = np.linalg.norm(d, axis=2)
    out = -(table.Z[None, :, None] * d / r[..., None] ** 3).sum(1)
    for L, idx in _groups(table):
        keys = pair_type(L)[0]
        for sel in _chunks(idx, len(pts), 1_000_000):
            P, p, K, g = _pair_arrays(table, sel, L)
            R = pts[None, :, :] - P[:, None, :]
            T = p[:, None] * np.einsum("bnk,bnk->bn", R, R)
            F = np.stack([boys(m, T) for m in range(L + 2)], -1)
            mono = (R[..., None, 0] ** keys[:, 0] * R[..., None, 1] ** keys[:, 1]
                    * R[..., None, 2] ** keys[:, 2])              # (B,N,nterm)
            Fm = F[..., keys[:, 3]]
            Fm1 = F[..., keys[:, 3] + 1]
            s1 = (g[:, None, :] * mono * Fm1).sum(-1)             # (B,N)
            for ax in range(3):
                dm = _dmono(R, keys, ax)
                gsum = (g[:, None, :] * dm * Fm).sum(-1) - 2.0 * p[:, None] * R[..., ax] * s1
                out[:, ax] -= (K[:, None] * gsum).sum(0)
    return out


def grad_G(table, points, gV) -> np.ndarray:
    """dL/dG as a flat (n_terms,) array aligned with ``tm_g``."""
    pts = np.asarray(points, float).reshape(-1, 3)
    gV = np.asarray(gV, float)
    out = np.zeros(table.n_terms)
    for L, idx in _groups(table):
        keys = pair_type(L)[0]
        nt = len(keys)
        for sel in _chunks(idx, len(pts), 1_000_000):
            P, p, K, _ = _pair_arrays(table, sel, L)
            R = pts[None, :, :] - P[:, None, :]
            T = p[:, None] * np.einsum("bnk,bnk->bn", R, R)
            F = np.stack([boys(m, T) for m in range(L + 1)], -1)
            mono = (R[..., None, 0] ** keys[:, 0] * R[..., None, 1] ** keys[:, 1]
                    * R[..., None, 2] ** keys[:, 2])
            dG = -K[:, None] * np.einsum("n,bnk->bk", gV, mono * F[..., keys[:, 3]])
            for b, i in enumerate(sel):
                out[table.pp_off[i]:table.pp_off[i] + nt] = dG[b]
    return out


def grad_P(table, basis, dG) -> np.ndarray:
    """dL/dP (n_fn x n_fn) from dL/dG through the table contraction."""
    coords = basis.wfn.coords
    layout = shell_layout(basis)
    n = basis.n
    out = np.zeros((n, n))
    by_pair: dict = {}
    for i in range(table.n_pairs):
        by_pair.setdefault(tuple(int(v) for v in table.pp_shells[i]), []).append(i)
    for (si, sj), rows in by_pair.items():
        oi, ni, shi = layout[si]
        oj, nj, shj = layout[sj]
        Ai = np.asarray(coords[shi.atom], float)
        Bj = np.asarray(coords[shj.atom], float)
        p, P, K, ii, jj, Ex, Ey, Ez = shell_pair_hermite(shi, shj, Ai, Bj)
        L = shi.l + shj.l
        keys, M0, _, _ = pair_type(L)
        nt = len(keys)
        # map kept rows to positions in the shell-pair primitive list
        pos = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(ii, jj))}
        sel = np.array([pos[tuple(int(v) for v in table.pp_prims[r])] for r in rows])
        dGk = np.stack([dG[table.pp_off[r]:table.pp_off[r] + nt] for r in rows])
        ps = p[sel]
        det = ((2.0 * np.pi / ps)[:, None] * (2.0 * ps)[:, None] ** keys[None, :, 3] * dGk) @ M0
        T = L + 1
        det = det.reshape(-1, T, T, T)
        cc = np.asarray(shi.coeffs, float)[ii[sel]] * np.asarray(shj.coeffs, float)[jj[sel]]
        blk = np.zeros((ni, nj))
        for mi, (ax, ay, az) in enumerate(cart_order(shi.l)):
            for mj, (bx, by, bz) in enumerate(cart_order(shj.l)):
                ex = Ex[sel, ax, bx, :]
                ey = Ey[sel, ay, by, :]
                ez = Ez[sel, az, bz, :]
                blk[mi, mj] = np.einsum("b,bxyz,bx,by,bz->", cc, det, ex, ey, ez)
        out[oi:oi + ni, oj:oj + nj] += blk
        if si != sj:
            out[oj:oj + nj, oi:oi + ni] += blk.T
    return out


def loss_and_gV(V, target, w):
    """L = sum_i w_i (V_i - T_i)^2 and gV = dL/dV (the config-2 loss)."""
    r = V - target
    return float((w * r * r).sum()), 2.0 * w * r
