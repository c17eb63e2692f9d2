"""Generate ``fe_lclass.inc``: straight-line per-L code for the HIP kernels.

For a primitive pair of class L = l_A + l_B the term list (alpha, m) of the
contracted table and its monomial schedule depend on L only
(``fastesp/esp.py:54-105``, ``fastesp/table.py:198-221``,
``fastesp/native/build_pairs.cpp:104-151,330-369``).  This script derives them
independently (own implementation of the a-table recurrence of
``fastesp/ints.py:46-59`` and the key enumeration) and emits, per L:

* host tables of the canonical keys and monomial schedule (used to validate
  uploaded tables and to build tables on the GPU);
* ``fe_contract<L>``: I(R) = sum_k g_k R^alpha_k F_mk with the monomials built
  once and the terms grouped by Boys order m (one FMA per term);
* ``fe_contract_grad<L>``: the same plus the gradient with respect to C;
* ``fe_dg_accum<L>``: the per-term products w_m R^alpha for dL/dG (and
  ``fe_dg_scale<L>``, the per-order factors applied once per pair);
* ``fe_contract_f32<L>``: the mixed-precision form (FP32 monomials/products);
* ``fe_fill_kelvin<L>``, ``fe_far_grad<L>``, ``fe_far_value<L>``: the far-field
  (Kelvin) block of a record and its evaluation;
* ``fe_det_from_dg<L>``: dL/dG -> d etab, the adjoint of the table contraction;
* ``fe_g_from_etab<L>``, ``fe_etab_bound<L>`` and, per pair of shell l's,
  ``fe_pair_etab<LI, LJ>``: the table build's etab, G and screening bound of
  one primitive pair.

Run: ``python gen_lclass.py [LMAX]`` (default 6).  Output is deterministic.
"""
from __future__ import annotations

import sys
from pathlib import Path


def a_table(lmax):
    a = [[0.0] * (lmax // 2 + 2) for _ in range(lmax + 1)]
    a[0][0] = 1.0
    for t in range(lmax):
        for j in range(lmax // 2 + 2):
            v = -a[t][j]
            if j >= 1:
                v += (t - 2 * j + 2) * a[t][j - 1]
            a[t + 1][j] = v
    return a


def keys_of(L):
    """Sorted (ax, ay, az, m) keys of class L and the t -> key coefficient map."""
    a = a_table(max(L, 1))
    T = L + 1
    acc = {}
    for tx in range(T):
        for ty in range(T - tx):
            for tz in range(T - tx - ty):
                lt = tx + ty + tz
                sgn = -1.0 if lt % 2 else 1.0
                for jx in range(tx // 2 + 1):
                    for jy in range(ty // 2 + 1):
                        for jz in range(tz // 2 + 1):
                            c = a[tx][jx] * a[ty][jy] * a[tz][jz]
                            if c == 0.0:
                                continue
                            k = (tx - 2 * jx, ty - 2 * jy, tz - 2 * jz, lt - jx - jy - jz)
                            acc.setdefault(k, {})
                            col = (tx * T + ty) * T + tz
                            acc[k][col] = acc[k].get(col, 0.0) + sgn * c
    keys = sorted(acc)
    return keys, acc


def schedule_of(L):
    keys, _ = keys_of(L)
    grown = set(k[:3] for k in keys)
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
    parent, axis = [], []
    for m in order:
        par, axv = 0, 0
        for ax in range(3):
            if m[ax] > 0:
                w = list(m)
                w[ax] -= 1
                par, axv = idx[tuple(w)], ax
                break
        parent.append(par)
        axis.append(axv)
    aidx = [idx[k[:3]] for k in keys]
    return order, parent, axis, aidx


def stride_of(nt):
    return 8 + ((nt + 7) // 8) * 8


def kelvin_of(L):
    """The far-field form of class L as a polynomial in u = R / |R|^2.

    Where T = p |R|^2 is past the Boys switch point, K F_m(T) equals
    S_m |R|^-(2m+1) with S_m = K sqrt(pi/p)/2 (2m-1)!! (2p)^-m, so a pair's
    electron integral is sum_k g_k S_m R^a |R|^-(2m+1).  With R^a = u^a |R|^(2|a|)
    and |u| = 1/|R| every term is |u| S_m g_k u^a (u.u)^(m-|a|); m >= |a| for
    every key, so I = |u| H(u) with H a polynomial of degree <= L (the Kelvin
    inversion of the pair's multipole potential).  Returns the monomials b of H
    with their contributions (key k, multinomial factor, order m)."""
    from math import factorial
    keys, _ = keys_of(L)
    h = {}
    for k, (ax, ay, az, m) in enumerate(keys):
        j = m - (ax + ay + az)
        assert j >= 0, (L, ax, ay, az, m)
        for i1 in range(j + 1):
            for i2 in range(j - i1 + 1):
                i3 = j - i1 - i2
                c = factorial(j) // (factorial(i1) * factorial(i2) * factorial(i3))
                b = (ax + 2 * i1, ay + 2 * i2, az + 2 * i3)
                h.setdefault(b, []).append((k, c, m))
    return h


def kelvin_monos(L):
    """Monomials of degree <= L - 1 in sorted order (each one's parent, the
    first nonzero axis lowered, comes earlier) with (parent, axis)."""
    order = sorted((a, b, c) for a in range(L) for b in range(L) for c in range(L)
                   if a + b + c <= L - 1)
    idx = {m: i for i, m in enumerate(order)}
    parent, axis = [], []
    for m in order:
        par, axv = 0, 0
        for ax in range(3):
            if m[ax] > 0:
                w = list(m)
                w[ax] -= 1
                par, axv = idx[tuple(w)], ax
                break
        parent.append(par)
        axis.append(axv)
    return order, parent, axis


def kelvin_top(L):
    """The degree-L monomials of H grouped by the axis of their parent (the
    first nonzero axis lowered, a monomial of degree L - 1): [(a, axis,
    parent index in kelvin_monos(L))], x group first."""
    kord = kelvin_monos(L)[0]
    idx = {m: i for i, m in enumerate(kord)}
    out = []
    for a in sorted((x, y, L - x - y) for x in range(L + 1) for y in range(L + 1 - x)):
        ax = next(d for d in range(3) if a[d] > 0)
        par = list(a)
        par[ax] -= 1
        out.append((a, ax, idx[tuple(par)]))
    return sorted(out, key=lambda t: t[1])


def nkel_of(L):
    """Length of the record's far-field block: the far-field thresholds on
    |R|^2 of the value and of the value + field evaluation, T_far(L)/p and
    T_far(L+1)/p; then h_0 for L = 0; else, over the monomials of degree
    <= L - 1, the coefficients of Q = H - (u . grad H)/L, of the three
    components of grad H and of H itself, then H's degree-L coefficients
    (kelvin_top order)."""
    if L == 0:
        return 3
    n1 = len(kelvin_monos(L)[0])
    return 2 + 5 * n1 + (L + 1) * (L + 2) // 2


def record_stride(L):
    """Doubles per device record of class L: header, g_k and c g_k (the
    candidate record the table build writes), then the far-field block."""
    keys, _ = keys_of(L)
    return stride_of(len(keys) + len(dcoef_of(L))) + ((nkel_of(L) + 7) // 8) * 8


def gen_kelvin(L, keys, koff):
    """fe_fill_kelvin<L>: the record's far-field block from its header and g_k
    (once per pair, at table build or upload); fe_far_grad<L>: H(u), grad H
    and D = u . grad H at one point (the far-field value + field).

    Field: with n = R/|R| and du/dR = (1 - 2 n n^T)/|R|^2, the gradient of
    |u| H(u) with respect to R is |u|^3 (grad H - (H + 2 D) R)."""
    o = []
    h = kelvin_of(L)
    ms = sorted({m for lst in h.values() for (_, _, m) in lst})
    o.append(f"template <> __host__ __device__ inline void fe_fill_kelvin<{L}>(double* r) {{")
    o.append("  const double hp = r[6];")
    o.append("  const double* g = r + 8;")
    # the evaluation tile's far test r^2 >= T_far(M)/p, formed as it was
    # formed per pair before (same operations, same bits)
    o.append(f"  r[{koff}] = fe_tfar({L}) * (hp + hp);")
    o.append(f"  r[{koff + 1}] = fe_tfar({L + 1}) * (hp + hp);")
    o.append(f"  double* kel = r + {koff + 2};")
    # s_m = K sqrt(pi/p)/2 (2m-1)!! (2p)^-m = s_{m-1} (2m-1) hp
    o.append("  const double s0 = r[7];")
    for m in range(1, max(ms) + 1):
        o.append(f"  const double s{m} = s{m - 1} * ({2 * m - 1}.0 * hp);")
    allb = sorted(h)
    bidx = {b: i for i, b in enumerate(allb)}

    def hname(b):
        return f"h{bidx[b]}" if b in bidx else "0.0"

    for b in allb:
        terms = [(f"g[{k}] * s{m}" if c == 1 else f"{c}.0 * g[{k}] * s{m}") for (k, c, m) in h[b]]
        o.append(f"  const double h{bidx[b]} = {' + '.join(terms)};")
    if L == 0:
        o.append("  kel[0] = h0;")
        o.append("}")
        return o
    kord, kpar, kax = kelvin_monos(L)
    n1 = len(kord)
    for j, b in enumerate(kord):
        d = sum(b)
        q = hname(b) if d == 0 else f"{(L - d) / L!r} * {hname(b)}"
        o.append(f"  kel[{j}] = {q};")
    for a in range(3):
        for j, b in enumerate(kord):
            up = list(b)
            up[a] += 1
            up = tuple(up)
            c = b[a] + 1
            o.append(f"  kel[{(a + 1) * n1 + j}] = " + (f"{hname(up)};" if c == 1 else f"{c}.0 * {hname(up)};"))
    top = kelvin_top(L)
    for j, b in enumerate(kord):
        o.append(f"  kel[{4 * n1 + j}] = {hname(b)};")
    for j, (a, _, _) in enumerate(top):
        o.append(f"  kel[{5 * n1 + j}] = {hname(a)};")
    o.append("}")

    o.append(f"template <> __device__ __forceinline__ void fe_far_grad<{L}>("
             "const double ux, const double uy, const double uz, fe_gdptr __restrict__ kel, "
             "double& H, double& Hx, double& Hy, double& Hz, double& D) {")
    U = ("ux", "uy", "uz")

    def mname(j):
        return U[kax[j]] if sum(kord[j]) == 1 else f"m{j}"

    # each monomial is formed and at once used by the four sums, so it stays
    # live only while later monomials still need it as a parent
    chains = (("Q", 0), ("Hx", n1), ("Hy", 2 * n1), ("Hz", 3 * n1))
    o.append("  double Q = kel[0];")
    for v, base in chains[1:]:
        o.append(f"  {v} = kel[{base}];")
    for j in range(1, n1):
        if sum(kord[j]) >= 2:
            o.append(f"  const double m{j} = {mname(kpar[j])} * {U[kax[j]]};")
        for v, base in chains:
            o.append(f"  {v} = fma(kel[{base + j}], {mname(j)}, {v});")
    o.append("  D = fma(uz, Hz, fma(uy, Hy, ux * Hx));")
    o.append(f"  H = " + ("Q + D;" if L == 1 else f"fma(D, {1.0 / L!r}, Q);"))
    o.append("}")

    # value only: H = sum over |b| <= L-1 of h_b u^b + ux Px + uy Py + uz Pz,
    # the degree-L terms through their parents (never forming degree L)
    o.append(f"template <> __device__ __forceinline__ double fe_far_value<{L}>("
             "const double ux, const double uy, const double uz, fe_gdptr __restrict__ kel) {")
    parents = {}
    for j, (a, ax, par) in enumerate(top):
        parents.setdefault(par, []).append((j, ax))
    o.append(f"  double H = kel[{4 * n1}];")
    init = set()

    def topacc(j, ax, mono):
        v = "P" + "xyz"[ax]
        k = f"kel[{5 * n1 + j}]"
        if v not in init:
            init.add(v)
            o.append(f"  double {v} = " + (f"{k};" if mono is None else f"{k} * {mono};"))
        else:
            o.append(f"  {v} = " + (f"{v} + {k};" if mono is None else f"fma({k}, {mono}, {v});"))

    for j, ax in parents.get(0, []):
        topacc(j, ax, None)
    for j in range(1, n1):
        if sum(kord[j]) >= 2:
            o.append(f"  const double m{j} = {mname(kpar[j])} * {U[kax[j]]};")
        o.append(f"  H = fma(kel[{4 * n1 + j}], {mname(j)}, H);")
        for jj, ax in parents.get(j, []):
            topacc(jj, ax, mname(j))
    for ax in range(3):
        v = "P" + "xyz"[ax]
        if v in init:
            o.append(f"  H = fma({U[ax]}, {v}, H);")
    o.append("  return H;")
    o.append("}")
    return o


DCOEF_LMAX = 3


def dcoef_of(L):
    """(key, multiplier) of every integer-scaled coefficient c g_k the field
    contraction uses (c = a_axis >= 2 of the key's monomial).  For L <= 3 they
    are stored in the record after the g_k, so the kernels do not rebuild them
    per pair; the longer records of L >= 4 would cost more in scalar loads and
    registers than the multiplies (measured: class 5 +52%, class 6 +110%), so
    those classes keep forming c g_k in the kernel."""
    if L > DCOEF_LMAX:
        return []
    keys, _ = keys_of(L)
    out = []
    for k, (ax, ay, az, m) in enumerate(keys):
        for al in (ax, ay, az):
            if al >= 2 and (k, al) not in out:
                out.append((k, al))
    return out


def emit_monos(out, L, order, parent, axis, T="double", pre="m", R=("Rx", "Ry", "Rz")):
    for j, m in enumerate(order):
        if j == 0:
            continue
        par = parent[j]
        src = f"{pre}{par}" if par != 0 else None
        r = R[axis[j]]
        out.append(f"  const {T} {pre}{j} = {src + ' * ' + r if src else r};")


def parent_chain(alpha):
    """alpha and its ancestors down to 0, each step lowering the first nonzero
    axis (the parent rule of schedule_of)."""
    out = [tuple(alpha)]
    a = list(alpha)
    while sum(a) > 0:
        for ax in range(3):
            if a[ax] > 0:
                a[ax] -= 1
                break
        out.append(tuple(a))
    return out


def dg_split(L, svec):
    """Split every key (a, m) of class L into gamma (ancestor of a at degree
    |a| - min(|a|, s_m)) and beta = a - gamma; return the product count, the
    splits, the monomials needed (closed under parents) and the weighted
    prefixes per order."""
    keys, _ = keys_of(L)
    ms = sorted({k[3] for k in keys})
    monos, weighted, terms = set(), {m: set() for m in ms}, []
    for k, (ax, ay, az, m) in enumerate(keys):
        a = (ax, ay, az)
        gamma = parent_chain(a)[min(sum(a), svec[ms.index(m)])]
        beta = tuple(a[i] - gamma[i] for i in range(3))
        terms.append((k, m, gamma, beta))
        weighted[m].update(parent_chain(gamma))
        monos.update(parent_chain(beta))
    cost = sum(1 for b in monos if sum(b) >= 2) + sum(len(v) for v in weighted.values())
    return cost, terms, monos, weighted


def dg_plan(L):
    """The cheapest split among s_m = min(m, c), c = 0..max degree (this family
    holds the optimum over all s vectors for L <= 6, checked by enumeration)."""
    keys, _ = keys_of(L)
    ms = sorted({k[3] for k in keys})
    maxd = max(sum(k[:3]) for k in keys)
    best = None
    for c in range(maxd + 1):
        r = dg_split(L, [min(m, c) for m in ms])
        if best is None or r[0] < best[0]:
            best = r
    return best


def value_polys(L, keys, parent, axis, aidx, T, fma):
    """Statements forming a_m = sum over the keys of order m of g_k R^a_k.

    The top order L is homogeneous of degree L and its monomials serve nothing
    else here, so a_L = Rx Px + Ry Py + Rz Pz with each key's term in the P of
    its parent axis, g_k R^(parent); the degree-L monomials are never formed."""
    out, init = [], set()

    def acc(var, k, mj):
        if var not in init:
            init.add(var)
            out.append(f"  {T} {var} = " + (f"g[{k}];" if mj == 0 else f"g[{k}] * m{mj};"))
        else:
            out.append(f"  {var} = " + (f"{var} + g[{k}];" if mj == 0 else f"{fma}(g[{k}], m{mj}, {var});"))

    split = L >= 2
    for k, (ax, ay, az, m) in enumerate(keys):
        mj = aidx[k]
        if split and m == L:
            acc(f"p{'xyz'[axis[mj]]}", k, parent[mj])
        else:
            acc(f"a{m}", k, mj)
    if split:
        e = None
        for d in "xyz":
            if f"p{d}" in init:
                e = f"R{d} * p{d}" if e is None else f"{fma}(R{d}, p{d}, {e})"
        out.append(f"  const {T} a{L} = {e};")
    return out


def gen_L(L):
    keys, acc = keys_of(L)
    order, parent, axis, aidx = schedule_of(L)
    nt, nm = len(keys), len(order)
    dc = dcoef_of(L)
    dci = {kc: nt + j for j, kc in enumerate(dc)}  # (key, multiplier) -> slot after the g_k
    o = []
    koff, nk = stride_of(nt + len(dc)), nkel_of(L)
    o.append(f"// ---- class L = {L}: nterm = {nt}, nmono = {nm}, scaled coefficients = {len(dc)}, "
             f"far-field coefficients = {nk}")
    o.append(f"template <> struct FeL<{L}> {{ static constexpr int NT = {nt}, NM = {nm}, "
             f"ND = {len(dc)}, KOFF = {koff}, KCOEF = {koff + 2}, NK = {nk}, "
             f"STRIDE = {record_stride(L)}; }};")
    o.append(f"template <> __host__ __device__ inline void fe_fill_dcoef<{L}>(double* g) {{")
    for j, (k, c) in enumerate(dc):
        o.append(f"  g[{nt + j}] = {c}.0 * g[{k}];")
    o.append("}")
    o.extend(gen_kelvin(L, keys, koff))

    def mono(j, pre="m", one="1.0"):
        return one if j == 0 else f"{pre}{j}"

    # ---------------- forward contraction, FP64
    o.append(f"template <> __device__ __forceinline__ double fe_contract<{L}>("
             "const double Rx, const double Ry, const double Rz, const double* F, "
             "fe_gdptr __restrict__ g) {")
    emit_monos(o, L, order, parent, axis)
    o.extend(value_polys(L, keys, parent, axis, aidx, "double", "fma"))
    seen = {k[3] for k in keys}
    expr = f"a{L} * F[{L}]" if L in seen else "0.0"
    for m in range(L - 1, -1, -1):
        if m in seen:
            expr = f"fma(a{m}, F[{m}], {expr})"
    o.append(f"  return {expr};")
    o.append("}")

    # ---------------- forward + gradient wrt C, FP64 (F has L+2 entries)
    # I = sum_k g R^a F_m ; dI/dC = sum_k g [ d(R^a) F_m - 2p R R^a F_{m+1} ]
    o.append(f"template <> __device__ __forceinline__ void fe_contract_grad<{L}>("
             "const double Rx, const double Ry, const double Rz, const double* F, "
             "fe_gdptr __restrict__ g, const double p, double& I, double& Gx, "
             "double& Gy, double& Gz) {")
    emit_monos(o, L, order, parent, axis)
    midx = {m: i for i, m in enumerate(order)}
    # accumulators: per m -> value poly a_m, and per (axis, m) derivative polys
    lines = []
    init = set()

    def coef(k, scale, lit=""):
        if scale == 1:
            return f"g[{k}]"
        return f"g[{dci[(k, scale)]}]" if (k, scale) in dci else f"({scale}.0{lit} * g[{k}])"

    def add(var, k, mono_j, scale=1):
        g = coef(k, scale)
        if var not in init:
            init.add(var)
            lines.append(f"  double {var} = " + (f"{g};" if mono_j == 0 else f"{g} * m{mono_j};"))
        else:
            lines.append(f"  {var} = " + (f"{var} + {g};" if mono_j == 0 else f"fma({g}, m{mono_j}, {var});"))

    # The top Boys order L holds exactly the keys of degree L (t = L, j = 0), so
    # a_L is homogeneous and Euler's identity gives it from its gradient, which
    # the field needs anyway: a_L = (R . grad a_L) / L.  Its degree-L
    # monomials are then never formed.
    euler = L >= 2
    for k, (ax, ay, az, m) in enumerate(keys):
        if not (euler and m == L):
            add(f"a{m}", k, aidx[k])
        for d, al in enumerate((ax, ay, az)):
            if al > 0:
                lo = [ax, ay, az]
                lo[d] -= 1
                add(f"d{'xyz'[d]}{m}", k, midx[tuple(lo)], al)
    if euler:
        lines.append(f"  const double a{L} = fma(Rz, dz{L}, fma(Ry, dy{L}, Rx * dx{L})) * {1.0 / L!r};")
    o.extend(lines)
    ms = sorted({k[3] for k in keys})
    # I = sum_m a_m F_m ; S1 = sum_m a_m F_{m+1}; each chain starts with a
    # product (not fma(., ., 0.0)) so a caller's F[0] = 1 folds away
    i_expr = f"a{ms[0]} * F[{ms[0]}]"
    s_expr = f"a{ms[0]} * F[{ms[0] + 1}]"
    for m in ms[1:]:
        i_expr = f"fma(a{m}, F[{m}], {i_expr})"
        s_expr = f"fma(a{m}, F[{m + 1}], {s_expr})"
    o.append(f"  I = {i_expr};")
    o.append(f"  const double s1 = -2.0 * p * ({s_expr});")
    for d in "xyz":
        e = f"s1 * R{d}"
        for m in ms:
            if f"d{d}{m}" in init:
                e = f"fma(d{d}{m}, F[{m}], {e})"
        o.append(f"  G{d} = {e};")
    o.append("}")
    # the same, accumulated: av += w I, a{x,y,z} += w G with w folded into
    # F first (wF_m = w F_m): L + 2 products instead of the 4 weighted
    # accumulations and the separate results, fewer operations for L <= 2
    o.append(f"template <> __device__ __forceinline__ void fe_contract_grad_acc<{L}>("
             "const double Rx, const double Ry, const double Rz, const double* F, "
             "fe_gdptr __restrict__ g, const double p, const double w, double& av, double& ax, "
             "double& ay, double& az) {")
    emit_monos(o, L, order, parent, axis)
    o.extend(lines)
    for m in range(L + 2):
        o.append(f"  const double wF{m} = w * F[{m}];")
    e = "av"
    for m in ms:
        e = f"fma(a{m}, wF{m}, {e})"
    o.append(f"  av = {e};")
    if L == 0:  # a_0 = g_0 is wave-uniform: -2p g_0 is one product per pair
        o.append("  const double s1 = (-2.0 * p * a0) * wF1;")
    else:
        s_expr = f"a{ms[0]} * wF{ms[0] + 1}"
        for m in ms[1:]:
            s_expr = f"fma(a{m}, wF{m + 1}, {s_expr})"
        o.append(f"  const double s1 = -2.0 * p * ({s_expr});")
    for d in "xyz":
        e = f"fma(s1, R{d}, a{d})"
        for m in ms:
            if f"d{d}{m}" in init:
                e = f"fma(d{d}{m}, wF{m}, {e})"
        o.append(f"  a{d} = {e};")
    o.append("}")

    # ---------------- dL/dG accumulation: acc[k] += w_mk * R^a_k
    # (w_m = wm[m], the caller's per-order weights.)
    # Each term is one fma(X, Y, acc) with X Y = w_m R^a, X = w_m R^gamma a
    # weighted prefix (gamma an ancestor of a on the parent chain) and Y = R^beta,
    # beta = a - gamma, a monomial of degree min(|a|, s_m).  s_m is chosen per
    # Boys order to minimise the products that build the X's and Y's (w_m R^a
    # for every term of a high order would rebuild long monomials once per
    # order; shared monomials alone would need every monomial of the class).
    o.append(f"template <> __device__ __forceinline__ void fe_dg_accum<{L}>("
             "const double Rx, const double Ry, const double Rz, const double* wm, "
             "double* __restrict__ acc) {")
    midx = {mm: i for i, mm in enumerate(order)}
    _, terms, mclos, weighted = dg_plan(L)
    def rname(j):  # a monomial by its schedule index: R_axis at degree 1
        return ("Rx", "Ry", "Rz")[axis[j]] if sum(order[j]) == 1 else f"m{j}"

    for j, mm in enumerate(order):  # monomials of degree >= 2 that a Y needs
        if sum(mm) >= 2 and mm in mclos:
            o.append(f"  const double m{j} = {rname(parent[j])} * {('Rx', 'Ry', 'Rz')[axis[j]]};")
    for m in ms:
        o.append(f"  const double w{m} = wm[{m}];")
        for j, mm in enumerate(order):  # weighted prefixes in parent-before-child order
            if sum(mm) >= 1 and mm in weighted[m]:
                src = f"w{m}" if parent[j] == 0 else f"v{m}_{parent[j]}"
                o.append(f"  const double v{m}_{j} = {src} * {('Rx', 'Ry', 'Rz')[axis[j]]};")

    def wname(m, gamma):
        return f"w{m}" if sum(gamma) == 0 else f"v{m}_{midx[gamma]}"

    def yname(beta):
        if sum(beta) == 1:
            return ("Rx", "Ry", "Rz")[beta.index(1)]
        return f"m{midx[beta]}"

    for k, m, gamma, beta in terms:
        x = wname(m, gamma)
        o.append(f"  acc[{k}] = " + (f"acc[{k}] + {x};" if sum(beta) == 0
                                     else f"fma({x}, {yname(beta)}, acc[{k}]);"))
    o.append("}")
    # acc[k] *= c[m_k]: the per-order factor the weights left out
    o.append(f"template <> __device__ __forceinline__ void fe_dg_scale<{L}>("
             "double* __restrict__ acc, const double* c) {")
    for k, (ax, ay, az, m) in enumerate(keys):
        o.append(f"  acc[{k}] *= c[{m}];")
    o.append("}")

    # ---------------- mixed: FP32 monomials and per-m polynomials, FP64 F and final sum
    o.append(f"template <> __device__ __forceinline__ double fe_contract_f32<{L}>("
             "const float Rx, const float Ry, const float Rz, const double* F, "
             "fe_gfptr __restrict__ g) {")
    emit_monos(o, L, order, parent, axis, T="float")
    o.extend(value_polys(L, keys, parent, axis, aidx, "float", "fmaf"))
    seen = {k[3] for k in keys}
    top = max(seen)
    expr = f"(double)a{top} * F[{top}]"
    for m in range(top - 1, -1, -1):
        if m in seen:
            expr = f"fma((double)a{m}, F[{m}], {expr})"
    o.append(f"  return {expr};")
    o.append("}")

    # ---------------- mixed gradient: FP32 monomials/polynomials, FP64 F, p and sums
    o.append(f"template <> __device__ __forceinline__ void fe_contract_grad_f32<{L}>("
             "const float Rx, const float Ry, const float Rz, const double* F, "
             "fe_gfptr __restrict__ g, const double p, const double Rxd, const double Ryd, "
             "const double Rzd, double& I, double& Gx, double& Gy, double& Gz) {")
    emit_monos(o, L, order, parent, axis, T="float")
    lines = []
    init = set()

    def add32(var, k, mono_j, scale=1):
        g = coef(k, scale, "f")
        if var not in init:
            init.add(var)
            lines.append(f"  float {var} = " + (f"{g};" if mono_j == 0 else f"{g} * m{mono_j};"))
        else:
            lines.append(f"  {var} = " + (f"{var} + {g};" if mono_j == 0 else f"fmaf({g}, m{mono_j}, {var});"))

    for k, (ax, ay, az, m) in enumerate(keys):  # a_L from Euler's identity, as above
        if not (euler and m == L):
            add32(f"a{m}", k, aidx[k])
        for d, al in enumerate((ax, ay, az)):
            if al > 0:
                lo = [ax, ay, az]
                lo[d] -= 1
                add32(f"d{'xyz'[d]}{m}", k, midx[tuple(lo)], al)
    if euler:
        lines.append(f"  const float a{L} = fmaf(Rz, dz{L}, fmaf(Ry, dy{L}, Rx * dx{L})) * {1.0 / L!r}f;")
    o.extend(lines)
    i_expr = f"(double)a{ms[0]} * F[{ms[0]}]"
    s_expr = f"(double)a{ms[0]} * F[{ms[0] + 1}]"
    for m in ms[1:]:
        i_expr = f"fma((double)a{m}, F[{m}], {i_expr})"
        s_expr = f"fma((double)a{m}, F[{m + 1}], {s_expr})"
    o.append(f"  I = {i_expr};")
    o.append(f"  const double s1 = -2.0 * p * ({s_expr});")
    for d in "xyz":
        e = f"s1 * R{d}d"
        for m in ms:
            if f"d{d}{m}" in init:
                e = f"fma((double)d{d}{m}, F[{m}], {e})"
        o.append(f"  G{d} = {e};")
    o.append("}")
    o.extend(gen_det(L, keys, acc, order))
    o.extend(gen_g_from_etab(L, keys, acc, order))
    return o, keys, order, parent, axis, aidx


def gen_det(L, keys, acc, order):
    """fe_det_from_dg<L>: the adjoint of G = (2 pi/p)(2p)^m M0 etab for one
    record, d etab_t = (2 pi/p) sum_k (2p)^{m_k} M0[k, t] dL/dG_k over the
    |t| <= L simplex in its sorted order (the order the etab columns use)."""
    T = L + 1
    o = [f"template <> __device__ __forceinline__ void fe_det_from_dg<{L}>("
         "const double* __restrict__ g, const double p, double* __restrict__ det) {"]
    ms = sorted({k[3] for k in keys})
    o.append("  const double tp = p + p;")
    o.append("  const double w0 = 6.283185307179586 / p;")
    for m in range(1, max(ms) + 1):
        o.append(f"  const double w{m} = w{m - 1} * tp;")
    for ti, (tx, ty, tz) in enumerate(order):
        col = (tx * T + ty) * T + tz
        by_m = {}
        for k, key in enumerate(keys):
            v = acc[key].get(col, 0.0)
            if v != 0.0:
                by_m.setdefault(key[3], []).append((k, v))
        parts = []
        for m, lst in sorted(by_m.items()):
            e = " + ".join((f"g[{k}]" if v == 1.0 else f"{v!r} * g[{k}]") for k, v in lst)
            parts.append(f"w{m} * ({e})")
        o.append(f"  det[{ti}] = " + (" + ".join(parts) if parts else "0.0") + ";")
    o.append("}")
    return o


def cart_order(l):
    """The Cartesian components of an l-shell, x-major ((l,0,0), (l-1,1,0), ...)."""
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def _fac(n):
    f = 1
    for k in range(2, n + 1):
        f *= k
    return f


def _binom(n, k):
    return _fac(n) // (_fac(k) * _fac(n - k))


def _lit(v):
    return repr(float(v))


def gen_pair_etab(LI, LJ):
    """fe_pair_etab<LI, LJ>: etab over the |t| <= L simplex of one primitive pair
    from its folded density block, by polynomial algebra instead of the 1-D
    Hermite recurrences.

    With u = r - P, (x - A)^a = sum_al C(a, al) PA^(a - al) u^al, so the block's
    polynomial sum_ij w_ij (r - A)^a_i (r - B)^b_j is a polynomial D(u) of
    degree L (D_n over the simplex), formed as sum_i Phi_i(u) Q_i(u) with
    Q_i = sum_j w_ij Psi_j.  A monomial u^n e^{-p u^2} expands in Hermite
    Gaussians with E^n_t = n! / (t! k!) (1/(4p))^k (1/(2p))^t for n = t + 2k
    (the E recurrence at PA = 0), so etab_t = cc sum_k D_{t+2k} prod_axes of
    those.  w is the block in the canonical component order of both shells
    (i-major), PA = P - A (shell i), PB = P - B (shell j), h = 1/(2p)."""
    L = LI + LJ
    ci, cj = cart_order(LI), cart_order(LJ)
    NJ = len(cj)
    simplex = sorted((a, b, c) for a in range(L + 1) for b in range(L + 1 - a)
                     for c in range(L + 1 - a - b))
    tpos = {t: i for i, t in enumerate(simplex)}
    qmon = [m for m in simplex if sum(m) <= LJ]
    o = [f"template <> __device__ __forceinline__ void fe_pair_etab<{LI}, {LJ}>("
         "const double* __restrict__ w, const double pax, const double pay, const double paz, "
         "const double pbx, const double pby, const double pbz, const double h, const double cc, "
         "double* __restrict__ et) {"]
    ax = "xyz"
    for d in range(3):
        for k in range(2, LI + 1):
            o.append(f"  const double pa{ax[d]}{k} = pa{ax[d]}{k - 1 if k > 2 else ''} * pa{ax[d]};")
        for k in range(2, LJ + 1):
            o.append(f"  const double pb{ax[d]}{k} = pb{ax[d]}{k - 1 if k > 2 else ''} * pb{ax[d]};")

    def pw(side, d, e):
        return None if e == 0 else f"p{side}{ax[d]}{e if e > 1 else ''}"

    def factor(side, comp, mon):
        """prod_d C(comp_d, mon_d) P{side}_d^(comp_d - mon_d) as (integer, [powers])."""
        c, ps = 1, []
        for d in range(3):
            c *= _binom(comp[d], mon[d])
            p = pw(side, d, comp[d] - mon[d])
            if p:
                ps.append(p)
        return c, ps

    def prod_expr(c, ps):
        if not ps:
            return _lit(c)
        e = " * ".join(ps)
        return e if c == 1 else f"{_lit(c)} * {e}"

    # Psi_j monomial coefficients (named when not a constant)
    psi = {}
    for j, bj in enumerate(cj):
        for m in qmon:
            if all(m[d] <= bj[d] for d in range(3)):
                c, ps = factor("b", bj, m)
                if ps and len(ps) + (c != 1) > 1:
                    nm = f"s{j}_{tpos[m]}"
                    o.append(f"  const double {nm} = {prod_expr(c, ps)};")
                    psi[(j, m)] = (1, [nm])
                else:
                    psi[(j, m)] = (c, ps)
    # per component i of shell A: Q_i over the monomials of degree <= LJ, then
    # D_n += Phi_i[al] Q_i[n - al] (D held as running sums, so only one Q_i
    # is live at a time)
    started = set()
    for i, a in enumerate(ci):
        for m in qmon:
            terms = []
            for j in range(NJ):
                if (j, m) in psi:
                    c, ps = psi[(j, m)]
                    wv = f"w[{i * NJ + j}]"
                    terms.append(wv if (c == 1 and not ps) else
                                 f"{wv} * {prod_expr(c, ps)}")
            o.append(f"  const double q{i}_{tpos[m]} = " + (" + ".join(terms) if terms else "0.0") + ";")
        dterms = {}
        for al in simplex:
            if not all(al[d] <= a[d] for d in range(3)):
                continue
            c, ps = factor("a", a, al)
            if ps and len(ps) + (c != 1) > 1:
                nm = f"f{i}_{tpos[al]}"
                o.append(f"  const double {nm} = {prod_expr(c, ps)};")
                fe = nm
            else:
                fe = prod_expr(c, ps)
            for m in qmon:
                n = tuple(al[d] + m[d] for d in range(3))
                qv = f"q{i}_{tpos[m]}"
                dterms.setdefault(n, []).append(qv if fe == "1.0" else f"{fe} * {qv}")
        for n in simplex:
            if n not in dterms:
                continue
            e = " + ".join(dterms[n])
            if n in started:
                o.append(f"  D{tpos[n]} += {e};")
            else:
                o.append(f"  double D{tpos[n]} = {e};")
                started.add(n)
    for n in simplex:
        if n not in started:
            o.append(f"  const double D{tpos[n]} = 0.0;")
    # powers of h = 1/(2p)
    for e in range(2, L + 1):
        o.append(f"  const double h{e} = h{e - 1 if e > 2 else ''} * h;")

    def hname(e):
        return None if e == 0 else ("h" if e == 1 else f"h{e}")

    for t in simplex:
        by_e = {}
        nt_ = sum(t)
        for kx in range((L - nt_) // 2 + 1):
            for ky in range((L - nt_) // 2 + 1 - kx):
                for kz in range((L - nt_) // 2 + 1 - kx - ky):
                    k = (kx, ky, kz)
                    n = tuple(t[d] + 2 * k[d] for d in range(3))
                    # E^n_t = n!/(t! k!) (1/(4p))^k (1/(2p))^t = n!/(t! k! 2^k) h^(t+k)
                    c = 1.0
                    for d in range(3):
                        c *= _fac(n[d]) / (_fac(t[d]) * _fac(k[d]) * 2 ** k[d])
                    by_e.setdefault(nt_ + sum(k), []).append((c, tpos[n]))
        parts = []
        for e, lst in sorted(by_e.items()):
            s = " + ".join(f"D{q}" if c == 1.0 else f"{_lit(c)} * D{q}" for c, q in lst)
            hn = hname(e)
            parts.append(s if hn is None and len(lst) == 1 else
                         f"({s})" if hn is None else f"{hn} * ({s})")
        o.append(f"  et[{tpos[t]}] = cc * (" + " + ".join(parts) + ");")
    o.append("}")
    return o


# The screening bound's per-axis weights for n = 0..11, the reference's table
# (fastesp/esp.py:25-28), kept as it is because it decides which pairs exist
# (it is labelled \int |He_n(x)| e^{-x^2/2} dx but equals that only for n <= 1).
MABS = [2.5066282746310002, 2.0, 1.5957691216057308, 4.0, 10.212920779565312,
        33.63284766793385, 136.5963625280984, 665.1416330443616, 3768.863619586618,
        25506.330595911374, 202331.81842358326, 1872154.2572862155]


def gen_g_from_etab(L, keys, acc, order):
    """fe_g_from_etab<L>: G_k = (2 pi/p)(2p)^m_k sum_t M0[k, t] etab_t for one
    primitive pair (the etab columns in the simplex order), and
    fe_etab_bound<L>: sum_t |etab_t| mabs(tx) mabs(ty) mabs(tz) (2p)^(|t|/2),
    the screening bound before its (2p)^-3/2 factor (fastesp/esp.py:248-252)."""
    T = L + 1
    o = [f"template <> __device__ __forceinline__ void fe_g_from_etab<{L}>("
         "const double* __restrict__ et, const double p, double* __restrict__ g) {"]
    ms = sorted({k[3] for k in keys})
    o.append("  const double tp = p + p;")
    o.append("  const double w0 = 6.283185307179586 / p;")
    for m in range(1, max(ms) + 1):
        o.append(f"  const double w{m} = w{m - 1} * tp;")
    pos = {((tx * T + ty) * T + tz): i for i, (tx, ty, tz) in enumerate(order)}
    for k, key in enumerate(keys):
        terms = [(pos[col], v) for col, v in sorted(acc[key].items()) if v != 0.0]
        e = " + ".join((f"et[{q}]" if v == 1.0 else f"{v!r} * et[{q}]") for q, v in terms)
        o.append(f"  g[{k}] = w{key[3]} * ({e if terms else '0.0'});")
    o.append("}")
    o.append(f"template <> __device__ __forceinline__ double fe_etab_bound<{L}>("
             "const double* __restrict__ et, const double sq) {")
    for e in range(2, L + 1):
        o.append(f"  const double sq{e} = sq{e - 1 if e > 2 else ''} * sq;")
    by_n = {}
    for q, t in enumerate(order):
        c = MABS[t[0]] * MABS[t[1]] * MABS[t[2]]
        by_n.setdefault(sum(t), []).append(f"{c!r} * fabs(et[{q}])")
    parts = []
    for n, lst in sorted(by_n.items()):
        s = " + ".join(lst)
        parts.append(f"({s})" if n == 0 else f"{'sq' if n == 1 else f'sq{n}'} * ({s})")
    o.append("  return " + " + ".join(parts) + ";")
    o.append("}")
    return o


def main():
    lmax = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    out = ["// GENERATED by gen_lclass.py -- do not edit by hand.",
           f"// Per-L straight-line code for L = 0..{lmax}.", "#pragma once",
           f"#define FE_LMAX {lmax}",
           "// Pair records are read-only while a kernel runs.  Reading them through",
           "// constant-address-space pointers keeps their wave-uniform loads scalar",
           "// (s_load) also when the record pointer itself is read from memory",
           "// (batched evaluation), where the compiler cannot prove them unclobbered.",
           "typedef const double __attribute__((address_space(4)))* fe_gdptr;",
           "typedef const float __attribute__((address_space(4)))* fe_gfptr;",
           "template <int L> struct FeL;",
           "template <int L> __device__ __forceinline__ double fe_contract(const double, const double, const double, const double*, fe_gdptr __restrict__);",
           "template <int L> __device__ __forceinline__ void fe_contract_grad(const double, const double, const double, const double*, fe_gdptr __restrict__, const double, double&, double&, double&, double&);",
           "template <int L> __device__ __forceinline__ void fe_dg_accum(const double, const double, const double, const double*, double* __restrict__);",
           "template <int L> __device__ __forceinline__ void fe_dg_scale(double* __restrict__, const double*);",
           "template <int L> __device__ __forceinline__ double fe_contract_f32(const float, const float, const float, const double*, fe_gfptr __restrict__);",
           "template <int L> __device__ __forceinline__ void fe_contract_grad_f32(const float, const float, const float, const double*, fe_gfptr __restrict__, const double, const double, const double, const double, double&, double&, double&, double&);",
           "template <int L> __host__ __device__ inline void fe_fill_dcoef(double*);",
           "template <int L> __device__ __forceinline__ void fe_contract_grad_acc(const double, const double, const double, const double*, fe_gdptr __restrict__, const double, const double, double&, double&, double&, double&);",
           "template <int L> __host__ __device__ inline void fe_fill_kelvin(double*);",
           "template <int L> __device__ __forceinline__ void fe_far_grad(const double, const double, const double, fe_gdptr __restrict__, double&, double&, double&, double&, double&);",
           "template <int L> __device__ __forceinline__ double fe_far_value(const double, const double, const double, fe_gdptr __restrict__);",
           "template <int L> __device__ __forceinline__ void fe_det_from_dg(const double* __restrict__, const double, double* __restrict__);",
           "template <int L> __device__ __forceinline__ void fe_g_from_etab(const double* __restrict__, const double, double* __restrict__);",
           "template <int L> __device__ __forceinline__ double fe_etab_bound(const double* __restrict__, const double);",
           "template <int LI, int LJ> __device__ __forceinline__ void fe_pair_etab(const double* __restrict__, const double, const double, const double, const double, const double, const double, const double, const double, double* __restrict__);",
           ""]
    host = ["// host-side canonical tables", "namespace fe_host {"]
    nts, nms, nds, koff, moff = [], [], [], [0], [0]
    allkeys, allmono = [], []
    for L in range(lmax + 1):
        o, keys, order, parent, axis, aidx = gen_L(L)
        out.extend(o)
        out.append("")
        nts.append(len(keys))
        nms.append(len(order))
        nds.append(len(dcoef_of(L)))
        for k, key in enumerate(keys):
            allkeys.append((*key, aidx[k]))
        for j in range(len(order)):
            allmono.append((*order[j], parent[j], axis[j]))
        koff.append(koff[-1] + len(keys))
        moff.append(moff[-1] + len(order))
    out.append("// etab of a primitive pair from its folded density block, per shell-l pair")
    out.append("// (l_i >= l_j; the build swaps the shells otherwise)")
    for LI in range(min(3, lmax) + 1):
        for LJ in range(LI + 1):
            if LI + LJ <= lmax:
                out.extend(gen_pair_etab(LI, LJ))
    out.append("")
    out.append("// record and candidate-record strides of a class known at run time")
    for fn, field in (("fe_stride", "STRIDE"), ("fe_cstride", "KOFF")):
        out.append(f"__host__ __device__ inline int {fn}(const int L) {{")
        out.append("  switch (L) {")
        for L in range(lmax + 1):
            out.append(f"    case {L}: return FeL<{L}>::{field};")
        out.append("    default: return 0;")
        out.append("  }")
        out.append("}")
    out.append("// the derived parts of a record of a class known at run time: the integer")
    out.append("// multiples c g_k (fe_fill_dcoef<L>) and the far-field block (fe_fill_kelvin<L>)")
    out.append("__host__ __device__ inline void fe_fill_record_any(const int L, double* r) {")
    out.append("  switch (L) {")
    for L in range(lmax + 1):
        out.append(f"    case {L}: fe_fill_dcoef<{L}>(r + 8); fe_fill_kelvin<{L}>(r); break;")
    out.append("    default: break;")
    out.append("  }")
    out.append("}")
    out.append("")
    host.append(f"static const int NTERM[{lmax + 1}] = {{{', '.join(map(str, nts))}}};")
    host.append(f"static const int NMONO[{lmax + 1}] = {{{', '.join(map(str, nms))}}};")
    host.append(f"static const int NDCOEF[{lmax + 1}] = {{{', '.join(map(str, nds))}}};")
    host.append("// record stride; the candidate records of the table build stop before the")
    host.append("// far-field block (CSTRIDE = its offset)")
    host.append(f"static const int STRIDE[{lmax + 1}] = "
                f"{{{', '.join(str(record_stride(L)) for L in range(lmax + 1))}}};")
    host.append(f"static const int CSTRIDE[{lmax + 1}] = "
                f"{{{', '.join(str(stride_of(n + d)) for n, d in zip(nts, nds))}}};")
    host.append(f"static const int NKEL[{lmax + 1}] = "
                f"{{{', '.join(str(nkel_of(L)) for L in range(lmax + 1))}}};")
    host.append(f"static const int KEY_OFF[{lmax + 2}] = {{{', '.join(map(str, koff))}}};")
    host.append(f"static const int MONO_OFF[{lmax + 2}] = {{{', '.join(map(str, moff))}}};")
    host.append("// ax, ay, az, m, aidx")
    host.append(f"static const int KEYS[{len(allkeys)}][5] = {{")
    host.extend("  {" + ", ".join(map(str, k)) + "}," for k in allkeys)
    host.append("};")
    host.append("// ax, ay, az, parent, axis")
    host.append(f"static const int MONOS[{len(allmono)}][5] = {{")
    host.extend("  {" + ", ".join(map(str, k)) + "}," for k in allmono)
    host.append("};")
    host.append("}  // namespace fe_host")
    out.extend(host)
    dst = Path(__file__).with_name("fe_lclass.inc")
    dst.write_text("\n".join(out) + "\n")
    print(f"wrote {dst} (L <= {lmax})")


if __name__ == "__main__":
    main()
