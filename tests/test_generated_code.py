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
