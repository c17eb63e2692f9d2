"""The generated per-class device code (fastesp_amd/csrc/fe_lclass.inc), run on
the CPU: each FP64 function body is a list of straight-line assignments, which
is evaluated here statement by statement and compared with the sums it must
form over the class's keys (a, m):

* fe_contract<L>:       sum_k g_k R^a F_m
* fe_contract_grad<L>:  the same, and sum_k g_k [d(R^a)/dR F_m - 2p R R^a F_{m+1}]
* fe_contract_grad_acc<L>: both, weighted and added to running sums
* fe_dg_accum<L>:       acc_k += w_m R^a (the product plan of the generator)
* fe_dg_scale<L>:       acc_k *= c_m
* the FP32 forms fe_contract_f32 / fe_contract_grad_f32, evaluated in FP64
"""
import re
from pathlib import Path

import numpy as np
import pytest

INC = Path(__file__).resolve().parents[1] / "fastesp_amd" / "csrc" / "fe_lclass.inc"


def body(src: str, fn: str, L: int) -> list:
    start = src.index(f"{fn}<{L}>(")
    start = src.index("{", start) + 1
    end = src.index("\n}\n", start)
    out = []
    for line in src[start:end].strip().splitlines():
        s = line.strip().rstrip(";")
        s = re.sub(r"^(const )?(double|float) ", "", s)
        s = s.replace("fmaf(", "fma(").replace("(double)", "")
        s = re.sub(r"(\d\.\d*)f\b", r"\1", s)  # float literals
        out.append(s)
    return out


def run(lines, env):
    env = dict(env, fma=lambda a, b, c: a * b + c)
    for s in lines:
        if s.startswith("return "):
            return eval(s[7:], {}, env)
        exec(s, {}, env)
    return env


def keys_of(src: str, L: int):
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen", INC.with_name("gen_lclass.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    return gen.keys_of(L)[0]


@pytest.fixture(scope="module")
def src():
    return INC.read_text()


@pytest.mark.parametrize("L", range(7))
def test_generated_contractions(src, L):
    keys = keys_of(src, L)
    rng = np.random.default_rng(40 + L)
    R = rng.uniform(-1.5, 1.5, 3)
    F = list(rng.uniform(0.1, 1.0, L + 2))
    nd = int(re.search(rf"struct FeL<{L}> {{[^}}]*ND = (\d+)", src).group(1))
    g = list(rng.normal(size=len(keys))) + [0.0] * nd
    run(body(src, "fe_fill_dcoef", L), dict(g=g))  # the record's scaled coefficients
    p, w = 0.7, 1.3
    env = dict(Rx=R[0], Ry=R[1], Rz=R[2], F=F, g=g, p=p)
    mono = [np.prod(R ** np.array(k[:3])) for k in keys]
    want = sum(g[i] * mono[i] * F[k[3]] for i, k in enumerate(keys))
    got = run(body(src, "fe_contract", L), env)
    assert got == pytest.approx(want, rel=1e-13, abs=1e-13)

    e = run(body(src, "fe_contract_grad", L), env)
    assert e["I"] == pytest.approx(want, rel=1e-13, abs=1e-13)
    for d, name in enumerate(("Gx", "Gy", "Gz")):
        gd = 0.0
        for i, k in enumerate(keys):
            a = np.array(k[:3])
            if a[d] > 0:
                lo = a.copy()
                lo[d] -= 1
                gd += g[i] * a[d] * np.prod(R ** lo) * F[k[3]]
            gd -= g[i] * 2 * p * R[d] * mono[i] * F[k[3] + 1]
        assert e[name] == pytest.approx(gd, rel=1e-12, abs=1e-12), name

    acc4 = dict(av=0.5, ax=-0.25, ay=0.125, az=1.0)
    e4 = run(body(src, "fe_contract_grad_acc", L), dict(env, w=w, **acc4))
    assert e4["av"] == pytest.approx(acc4["av"] + w * want, rel=1e-12, abs=1e-12)
    for name in ("x", "y", "z"):
        assert e4["a" + name] == pytest.approx(acc4["a" + name] + w * e["G" + name],
                                               rel=1e-12, abs=1e-12), name

    # the mixed-precision forms (FP32 in the device code; the algebra is the same)
    got32 = run(body(src, "fe_contract_f32", L), env)
    assert got32 == pytest.approx(want, rel=1e-12, abs=1e-12)
    e32 = run(body(src, "fe_contract_grad_f32", L), dict(env, Rxd=R[0], Ryd=R[1], Rzd=R[2]))
    assert e32["I"] == pytest.approx(want, rel=1e-12, abs=1e-12)
    for name in ("Gx", "Gy", "Gz"):
        assert e32[name] == pytest.approx(e[name], rel=1e-12, abs=1e-12), name

    acc = [0.25] * len(keys)
    wm = [w * f for f in F]
    run(body(src, "fe_dg_accum", L), dict(env, wm=wm, acc=acc))
    want_acc = [0.25 + w * mono[i] * F[k[3]] for i, k in enumerate(keys)]
    np.testing.assert_allclose(acc, want_acc, rtol=1e-13, atol=1e-14)
    c = list(rng.uniform(0.5, 2.0, L + 1))
    run(body(src, "fe_dg_scale", L), dict(acc=acc, c=c))
    np.testing.assert_allclose(acc, [a * c[k[3]] for a, k in zip(want_acc, keys)], rtol=1e-13)


def test_record_strides_match_the_flop_model(src):
    from fastesp_amd import flops
    m = re.search(r"static const int STRIDE\[\d+\] = \{([^}]*)\}", src)
    assert [int(x) for x in m.group(1).split(",")] == flops.STRIDE


def _gen():
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen", INC.with_name("gen_lclass.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    return gen


def _fill_record(src, L, g, p, K):
    """A record of class L as the device holds it: header, g_k, the c g_k
    (fe_fill_dcoef) and the far-field block (fe_fill_kelvin), the generated
    bodies run on the CPU over a NumPy record."""
    stride = int(re.search(rf"struct FeL<{L}> {{[^}}]*STRIDE = (\d+)", src).group(1))
    r = np.zeros(stride)
    r[3], r[4], r[6], r[7] = p, K, 0.5 / p, K * np.sqrt(np.pi / p) / 2
    r[8:8 + len(g)] = g
    run(body(src, "fe_fill_dcoef", L), dict(g=r[8:]))
    tfar = [35.5, 39.5, 43.0, 46.0, 49.0, 51.5, 54.0, 56.5, 59.0]  # fe_boys.h fe_tfar
    env = dict(r=r, fe_tfar=lambda M: tfar[M])
    for s in body(src, "fe_fill_kelvin", L):
        m = re.fullmatch(r"(?:double\* )?(\w+) = r \+ (\d+)", s.replace("const ", ""))
        if m:  # pointer into the record: a NumPy view, so writes land in r
            env[m.group(1)] = r[int(m.group(2)):]
        else:
            env = run([s], env)
    return r


def _far_keyed(keys, g, p, K, R):
    """The keyed far form: sum_k g_k K F_m(T) R^a with K F_m = S_m |R|^-(2m+1),
    S_m = K sqrt(pi/p)/2 (2m-1)!! (2p)^-m, and its gradient in R."""
    from math import prod
    r = np.sqrt(R @ R)
    I, G = 0.0, np.zeros(3)
    for gk, (ax, ay, az, m) in zip(g, keys):
        S = K * np.sqrt(np.pi / p) / 2 * prod(range(2 * m - 1, 0, -2)) * (2 * p) ** -m
        a = np.array([ax, ay, az])
        mono = np.prod(R ** a)
        I += gk * S * mono * r ** -(2 * m + 1)
        for d in range(3):
            if a[d]:
                lo = a.copy()
                lo[d] -= 1
                G[d] += gk * S * a[d] * np.prod(R ** lo) * r ** -(2 * m + 1)
            G[d] -= gk * S * (2 * m + 1) * mono * R[d] * r ** -(2 * m + 3)
    return I, G


@pytest.mark.parametrize("L", range(7))
def test_far_field_block_equals_the_keyed_far_form(src, L):
    """fe_fill_kelvin + fe_far_grad (as the evaluation tile combines them)
    against the keyed asymptotic form, term by term, at points of several
    distances and directions."""
    keys = keys_of(src, L)
    rng = np.random.default_rng(70 + L)
    g = rng.normal(size=len(keys))
    p, K = 0.83, 1.7
    r = _fill_record(src, L, g, p, K)
    koff = int(re.search(rf"struct FeL<{L}> {{[^}}]*KOFF = (\d+)", src).group(1))
    kcoef = int(re.search(rf"struct FeL<{L}> {{[^}}]*KCOEF = (\d+)", src).group(1))
    # the far tests of the value and the value + field tiles
    tfar = [35.5, 39.5, 43, 46, 49, 51.5, 54, 56.5, 59]
    assert r[koff] == tfar[L] * (r[6] + r[6]) and r[koff + 1] == tfar[L + 1] * (r[6] + r[6])
    kel = r[kcoef:]
    for _ in range(5):
        R = rng.normal(size=3) * rng.uniform(2.0, 9.0)
        want_I, want_G = _far_keyed(keys, g, p, K, R)
        ri = 1.0 / np.sqrt(R @ R)
        ri2, ri3 = ri * ri, ri ** 3
        if L == 0:
            H, Hg, D = kel[0], np.zeros(3), 0.0
        else:
            e = run(body(src, "fe_far_grad", L),
                    dict(ux=R[0] * ri2, uy=R[1] * ri2, uz=R[2] * ri2, kel=kel))
            H, Hg, D = e["H"], np.array([e["Hx"], e["Hy"], e["Hz"]]), e["D"]
            assert D == pytest.approx(Hg @ (R * ri2), rel=1e-13, abs=1e-300)
        got_I = ri * H
        got_G = ri3 * (Hg - (H + 2 * D) * R)
        if L > 0:  # the value-only form
            Hv = run(body(src, "fe_far_value", L),
                     dict(ux=R[0] * ri2, uy=R[1] * ri2, uz=R[2] * ri2, kel=kel))
            assert abs(ri * Hv - want_I) <= 1e-13 * (np.abs(want_I) + np.abs(want_G).max())
        scale = np.abs(want_I) + np.abs(want_G).max()
        assert abs(got_I - want_I) <= 1e-13 * scale
        np.testing.assert_allclose(got_G, want_G, rtol=0, atol=1e-13 * scale)


@pytest.mark.parametrize("L", range(1, 7))
def test_far_field_polynomial_is_harmonic_by_degree(L):
    """For coefficients g that come from Hermite expansion coefficients (the
    table's G = (2 pi/p) (2p)^m M0 etab), the far field is the potential of
    point multipoles at P, so each homogeneous part H_n of the Kelvin
    polynomial is harmonic.  A wrong multinomial factor, order factor S_m or a
    dropped key breaks this (checked on the coefficient map of kelvin_of)."""
    from math import prod
    gen = _gen()
    keys, acc = gen.keys_of(L)
    h = gen.kelvin_of(L)
    rng = np.random.default_rng(90 + L)
    T = L + 1
    p, K = 0.61, 1.0
    etab = {(tx, ty, tz): rng.normal() for tx in range(T) for ty in range(T) for tz in range(T)
            if tx + ty + tz <= L}
    g = []
    for (ax, ay, az, m) in keys:
        s = sum(v * etab[(col // (T * T), (col // T) % T, col % T)] for col, v in acc[(ax, ay, az, m)].items())
        g.append(2 * np.pi / p * (2 * p) ** m * s)
    S = [K * np.sqrt(np.pi / p) / 2 * prod(range(2 * m - 1, 0, -2)) * (2 * p) ** -m
         for m in range(L + 2)]
    coef = {b: sum(c * g[k] * S[m] for (k, c, m) in lst) for b, lst in h.items()}
    for n in range(L + 1):
        lap = {}
        for b, v in coef.items():
            if sum(b) != n:
                continue
            for d in range(3):
                if b[d] >= 2:
                    lo = list(b)
                    lo[d] -= 2
                    lap[tuple(lo)] = lap.get(tuple(lo), 0.0) + b[d] * (b[d] - 1) * v
        big = max(abs(v) for b, v in coef.items() if sum(b) == n)
        assert all(abs(v) <= 1e-12 * big for v in lap.values()), (n, lap)
    # and arbitrary g (not from a Hermite expansion) is not harmonic: the test has teeth
    g2 = rng.normal(size=len(keys))
    coef2 = {b: sum(c * g2[k] * S[m] for (k, c, m) in lst) for b, lst in h.items()}
    lap2 = {}
    for b, v in coef2.items():
        if sum(b) == 2:
            for d in range(3):
                if b[d] >= 2:
                    lap2[0] = lap2.get(0, 0.0) + b[d] * (b[d] - 1) * v
    if L >= 2:
        assert abs(lap2[0]) > 1e-6


@pytest.mark.parametrize("L", range(7))
def test_det_from_dg_is_the_adjoint_of_the_table_contraction(src, L):
    """fe_det_from_dg<L> (dL/dG -> d etab of one record) is the transpose of
    G_k = (2 pi/p)(2p)^m_k sum_t M0[k, t] etab_t: <det, etab> = <g, G(etab)>
    for random etab and g, with M0 from the key map of gen_lclass."""
    gen = _gen()
    keys, acc = gen.keys_of(L)
    order = gen.schedule_of(L)[0]
    T = L + 1
    rng = np.random.default_rng(120 + L)
    p = 0.77
    g = list(rng.normal(size=len(keys)))
    det = [0.0] * len(order)
    run(body(src, "fe_det_from_dg", L), dict(g=g, p=p, det=det))
    etab = rng.normal(size=len(order))
    pos = {((tx * T + ty) * T + tz): i for i, (tx, ty, tz) in enumerate(order)}
    G = [2 * np.pi / p * (2 * p) ** m * sum(v * etab[pos[col]] for col, v in acc[(ax, ay, az, m)].items())
         for (ax, ay, az, m) in keys]
    assert np.dot(det, etab) == pytest.approx(np.dot(g, G), rel=1e-12)


PAIR_LS = [(li, lj) for li in range(4) for lj in range(li + 1)]


@pytest.mark.parametrize("li,lj", PAIR_LS)
def test_pair_etab_matches_the_hermite_recurrence(src, li, lj):
    """fe_pair_etab<li, lj> (the polynomial-algebra etab of the device build)
    equals cc sum_ij w_ij Ex Ey Ez with the 1-D tables of the oracle's E
    recurrence (oracle/ints.py e1d), over the simplex in the class's order."""
    from fastesp_amd.basis import cart_order
    from oracle.ints import e1d
    gen = _gen()
    L = li + lj
    order = gen.schedule_of(L)[0]
    ci, cj = cart_order(li), cart_order(lj)
    rng = np.random.default_rng(300 + 10 * li + lj)
    for trial in range(3):
        p = float(rng.uniform(0.2, 9.0))
        pa = rng.uniform(-1.2, 1.2, 3)
        pb = rng.uniform(-1.2, 1.2, 3) if trial else -pa * 0.4
        cc = float(rng.normal())
        w = rng.normal(size=(len(ci), len(cj)))
        et = [0.0] * len(order)
        env = dict(w=list(w.ravel()), pax=pa[0], pay=pa[1], paz=pa[2], pbx=pb[0], pby=pb[1],
                   pbz=pb[2], h=0.5 / p, cc=cc, et=et)
        run(body(src, "fe_pair_etab", f"{li}, {lj}"), env)
        E = [e1d(li, lj, p, pa[d], pb[d])[0] for d in range(3)]
        want = np.zeros(len(order))
        for i, a in enumerate(ci):
            for j, b in enumerate(cj):
                for q, t in enumerate(order):
                    want[q] += w[i, j] * E[0][a[0], b[0], t[0]] * E[1][a[1], b[1], t[1]] \
                        * E[2][a[2], b[2], t[2]]
        want *= cc
        np.testing.assert_allclose(et, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("L", range(7))
def test_g_from_etab_and_bound_match_the_oracle_pair_type(src, L):
    """fe_g_from_etab<L> is G = (2 pi/p)(2p)^m M0 etab and fe_etab_bound<L> the
    screening bound's sum over |etab| (before (2p)^-3/2), both with the M0,
    moments and |t| of oracle/ints.py pair_type."""
    from oracle.ints import pair_type
    gen = _gen()
    order = gen.schedule_of(L)[0]
    T = L + 1
    keys, M0, wmom, abs_t = pair_type(L)
    rng = np.random.default_rng(400 + L)
    p = 1.37
    et = rng.normal(size=len(order))
    full = np.zeros(T ** 3)
    for q, (tx, ty, tz) in enumerate(order):
        full[(tx * T + ty) * T + tz] = et[q]
    g = [0.0] * len(keys)
    run(body(src, "fe_g_from_etab", L), dict(et=list(et), p=p, g=g))
    want = 2 * np.pi / p * (2 * p) ** keys[:, 3] * (M0 @ full)
    np.testing.assert_allclose(g, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    b = run(body(src, "fe_etab_bound", L), dict(et=list(et), sq=np.sqrt(2 * p), fabs=abs))
    wb = (np.abs(full) * wmom * (2 * p) ** (abs_t / 2.0)).sum()
    assert b == pytest.approx(wb, rel=1e-13)
