"""Oracle pair-table build (FP64 NumPy).

Restates ``ESPEngine._build/_make_pair`` (``fastesp/esp.py:168-257``) and
``build_table`` (``fastesp/table.py:176-259``): shell pairs si <= sj, the folded
density block, the K < 1e-30 primitive skip, the Hermite contraction
G = (2 pi/p)(2p)^m M0 etab, the screening bound and reff, the NumPy screen
``keep = not (K bound / max(dmin, 0.05) < tol)``, the per-L monomial schedule
and the flat layout.  Besides the reference fields it records, per kept pair,
L, bound, reff and the (shell, primitive) indices the adjoint needs.

Pinned by: the compiled reference builder ``_ref/build_pairs`` (same G, K, p,
P per primitive pair), the sum rule Tr(P S) = N_elec, the per-pair check
against ``ints.primitive_pair_integral`` and the quadrature of rho/|C-r|
(``tests/test_pipeline.py:114-126``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from fastesp_amd.basis import cart_order

from .ints import e1d, mono_schedule, pair_type


@dataclass
class OracleTable:
    """Flat pair table in the layout of ``fastesp/table.py:40-62`` plus extras."""
    pp_exp: np.ndarray
    pp_cx: np.ndarray
    pp_cy: np.ndarray
    pp_cz: np.ndarray
    pp_K: np.ndarray
    pp_nt: np.ndarray
    pp_off: np.ndarray
    pp_nmono: np.ndarray
    pp_monooff: np.ndarray
    mo_parent: np.ndarray
    mo_axis: np.ndarray
    tm_aidx: np.ndarray
    tm_m: np.ndarray
    tm_g: np.ndarray
    pts: np.ndarray
    Z: np.ndarray
    xyz: np.ndarray
    max_mono: int = 0
    # extras (not in pairs.bin)
    pp_L: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    pp_bound: np.ndarray = field(default_factory=lambda: np.zeros(0))
    pp_reff: np.ndarray = field(default_factory=lambda: np.zeros(0))
    pp_shells: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    pp_prims: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    n_candidates: int = 0
    sum_rule: float = 0.0

    @property
    def n_pairs(self) -> int:
        return len(self.pp_exp)

    @property
    def n_terms(self) -> int:
        return len(self.tm_g)


@dataclass
class PrimPairs:
    """All candidate primitive pairs (before the point screen), in reference order."""
    p: np.ndarray
    P: np.ndarray
    K: np.ndarray
    L: np.ndarray
    G: list
    bound: np.ndarray
    reff: np.ndarray
    shells: np.ndarray
    prims: np.ndarray
    sum_rule: float


def shell_layout(basis):
    off = 0
    out = []
    for sh in basis.shells:
        nc = len(cart_order(sh.l))
        out.append((off, nc, sh))
        off += nc
    return out


def density_block(density, oi, ni, oj, nj, same: bool) -> np.ndarray:
    """P_ij (+ P_ji^T when si != sj): ``fastesp/esp.py:185-187``."""
    blk = np.array(density[oi:oi + ni, oj:oj + nj], dtype=float)
    if not same:
        blk = blk + density[oj:oj + nj, oi:oi + ni].T
    return blk


def shell_pair_hermite(shi, shj, Ai, Bj):
    """Primitive-pair quantities of one shell pair after the K >= 1e-30 cut.

    Returns (p, P, K, ii, jj, Ex, Ey, Ez) or None, following
    ``fastesp/esp.py:224-238``.
    """
    ea = np.asarray(shi.exps, float)
    eb = np.asarray(shj.exps, float)
    ii = np.repeat(np.arange(len(ea)), len(eb))
    jj = np.tile(np.arange(len(eb)), len(ea))
    al, be = ea[ii], eb[jj]
    p = al + be
    P = (al[:, None] * Ai[None, :] + be[:, None] * Bj[None, :]) / p[:, None]
    ab2 = float((Ai - Bj) @ (Ai - Bj))
    K = np.exp(-(al * be) / p * ab2)
    sel = K >= 1e-30
    if not sel.any():
        return None
    p, P, K, ii, jj = p[sel], P[sel], K[sel], ii[sel], jj[sel]
    XPA, XPB = P - Ai, P - Bj
    E = [e1d(shi.l, shj.l, p, XPA[:, k], XPB[:, k]) for k in range(3)]
    return p, P, K, ii, jj, E[0], E[1], E[2]


def etab_of(blk, shi, shj, ii, jj, Ex, Ey, Ez, L):
    """etab[b, tx, ty, tz] = sum_{mu nu} blk c_i c_j Ex Ey Ez (``fastesp/esp.py:241-247``)."""
    ci = cart_order(shi.l)
    cj = cart_order(shj.l)
    T = L + 1
    cc = np.asarray(shi.coeffs, float)[ii] * np.asarray(shj.coeffs, float)[jj]
    et = np.zeros((len(ii), T, T, T))
    for mi, (ax, ay, az) in enumerate(ci):
        for mj, (bx, by, bz) in enumerate(cj):
            w = blk[mi, mj]
            if w == 0.0:
                continue
            ex = Ex[:, ax, bx, :]
            ey = Ey[:, ay, by, :]
            ez = Ez[:, az, bz, :]
            et += (w * cc)[:, None, None, None] * (ex[:, :, None, None]
                                                   * ey[:, None, :, None]
                                                   * ez[:, None, None, :])
    return et


def build_prim_pairs(basis, density) -> PrimPairs:
    """Every primitive pair of every significant shell pair (reference order)."""
    coords = basis.wfn.coords
    layout = shell_layout(basis)
    p_l, P_l, K_l, L_l, G_l, b_l, r_l, sh_l, pr_l = [], [], [], [], [], [], [], [], []
    nelec = 0.0
    for si, (oi, ni, shi) in enumerate(layout):
        Ai = np.asarray(coords[shi.atom], float)
        for sj in range(si, len(layout)):
            oj, nj, shj = layout[sj]
            blk = density_block(density, oi, ni, oj, nj, si == sj)
            if np.abs(blk).max() < 1e-15:
                continue
            Bj = np.asarray(coords[shj.atom], float)
            hp = shell_pair_hermite(shi, shj, Ai, Bj)
            if hp is None:
                continue
            p, P, K, ii, jj, Ex, Ey, Ez = hp
            L = shi.l + shj.l
            keys, M0, wmom, abs_t = pair_type(L)
            et = etab_of(blk, shi, shj, ii, jj, Ex, Ey, Ez, L).reshape(len(p), -1)
            G = (2.0 * np.pi / p)[:, None] * (2.0 * p)[:, None] ** keys[None, :, 3] * (et @ M0.T)
            bound = (2.0 * p) ** -1.5 * (np.abs(et) * wmom[None, :]
                                         * (2.0 * p)[:, None] ** (abs_t[None, :] / 2.0)).sum(1)
            reff = 1.6 * np.sqrt((L + 2.0) / p)
            nelec += float((K * (np.pi / p) ** 1.5 * et[:, 0]).sum())
            p_l.append(p); P_l.append(P); K_l.append(K)
            L_l.append(np.full(len(p), L, np.int32))
            G_l.extend(list(G)); b_l.append(bound); r_l.append(reff)
            sh_l.append(np.tile([si, sj], (len(p), 1)))
            pr_l.append(np.stack([ii, jj], 1))
    cat = (lambda xs, shape: np.concatenate(xs) if xs else np.zeros(shape))
    return PrimPairs(p=cat(p_l, (0,)), P=cat(P_l, (0, 3)), K=cat(K_l, (0,)),
                     L=cat(L_l, (0,)).astype(np.int32), G=G_l, bound=cat(b_l, (0,)),
                     reff=cat(r_l, (0,)), shells=cat(sh_l, (0, 2)).astype(np.int32),
                     prims=cat(pr_l, (0, 2)).astype(np.int32), sum_rule=nelec)


def min_distance(P: np.ndarray, pts: np.ndarray, chunk: int = 4096) -> np.ndarray:
    """dmin[i] = min_k |pts_k - P_i| (``fastesp/table.py:190-192``)."""
    out = np.empty(len(P))
    pts = np.asarray(pts, float)
    for s in range(0, len(P), chunk):
        d2 = ((P[s:s + chunk, None, :] - pts[None, :, :]) ** 2).sum(-1)
        out[s:s + chunk] = np.sqrt(d2.min(axis=1)) if len(pts) else np.inf
    return out


def max_distance(P: np.ndarray, pts: np.ndarray, chunk: int = 4096) -> np.ndarray:
    """dmax[i] = max_k |pts_k - P_i| (the ``d.max()`` of ``fastesp/esp.py:266-267``)."""
    out = np.empty(len(P))
    pts = np.asarray(pts, float)
    for s in range(0, len(P), chunk):
        d2 = ((P[s:s + chunk, None, :] - pts[None, :, :]) ** 2).sum(-1)
        out[s:s + chunk] = np.sqrt(d2.max(axis=1))
    return out


def build_table(wfn, basis, density, points, tol: float = 1e-12,
                prims: PrimPairs | None = None, rule: str = "numpy") -> OracleTable:
    """``fastesp/table.py:176-259``: screen, monomial schedule and flatten.

    ``rule="engine"`` applies instead the per-pair skip of
    ``ESPEngine.esp_electron`` (``fastesp/esp.py:265-270``):
    skip if K bound / max(dmax - reff, 0.05) < tol, with dmax the largest
    distance from the pair centre to any point.  The evaluation of the kept
    pairs is the same sum, so ``evaluate.electron`` of that table is
    ``esp_electron``.
    """
    if prims is None:
        prims = build_prim_pairs(basis, density)
    pts = np.asarray(points, float).reshape(-1, 3)
    if rule == "numpy":
        dmin = min_distance(prims.P, pts)
        val = prims.K * prims.bound / np.maximum(dmin, 0.05)
    elif rule == "engine":
        dmax = max_distance(prims.P, pts)
        val = prims.K * prims.bound / np.maximum(dmax - prims.reff, 0.05)
    else:
        raise ValueError(f"rule {rule!r}")
    keep = ~(val < tol)
    idx = np.nonzero(keep)[0]
    L = prims.L[idx]
    nt = np.array([pair_type(int(l))[0].shape[0] for l in L], np.int32)
    nmono = np.array([mono_schedule(int(l))[0].shape[0] for l in L], np.int32)
    off = np.concatenate([[0], np.cumsum(nt)[:-1]]).astype(np.int32) if len(idx) else np.zeros(0, np.int32)
    moff = np.concatenate([[0], np.cumsum(nmono)[:-1]]).astype(np.int32) if len(idx) else np.zeros(0, np.int32)
    parent, axis, aidx, m, g = [], [], [], [], []
    for k, i in enumerate(idx):
        l = int(L[k])
        _, par, ax, ai = mono_schedule(l)
        parent.append(par); axis.append(ax); aidx.append(ai)
        m.append(pair_type(l)[0][:, 3].astype(np.int32))
        g.append(prims.G[i])
    cat = (lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt))
    return OracleTable(
        pp_exp=prims.p[idx], pp_cx=prims.P[idx, 0].copy(), pp_cy=prims.P[idx, 1].copy(),
        pp_cz=prims.P[idx, 2].copy(), pp_K=prims.K[idx], pp_nt=nt, pp_off=off,
        pp_nmono=nmono, pp_monooff=moff, mo_parent=cat(parent, np.int32),
        mo_axis=cat(axis, np.int32), tm_aidx=cat(aidx, np.int32), tm_m=cat(m, np.int32),
        tm_g=cat(g, np.float64), pts=pts, Z=wfn.Z, xyz=np.asarray(wfn.coords, float),
        max_mono=int(nmono.max()) if len(nmono) else 0,
        pp_L=L.astype(np.int32), pp_bound=prims.bound[idx], pp_reff=prims.reff[idx],
        pp_shells=prims.shells[idx], pp_prims=prims.prims[idx],
        n_candidates=len(prims.p), sum_rule=prims.sum_rule)
