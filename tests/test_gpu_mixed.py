"""Mixed-precision path (FE_PREC_MIXED): error against the FP64 oracle.

Target: max |dV| <= 1e-8 a.u. against the FP64 oracle (the FP64 path itself is
held to 1e-10).  The routing threshold tau = 0 sends every pair to the FP64
contraction, which must then agree with the FP64 path to rounding.
"""
from pathlib import Path

import numpy as np
import pytest

from fastesp_amd import _lib, synth
from fastesp_amd.autograd import esp_with_grad
from fastesp_amd.builder import build_table_hip

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
TOL_MIXED = 1e-8


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_mixed_cfg0_and_cfg1_subset():
    ctx = _lib.Context.get(0)
    for L in range(7):  # every class through the FP32 contraction
        ctx.set_tuning(20 + L, 1)
    try:
        _mixed_cfg0_and_cfg1_subset()
    finally:
        for L in range(7):
            ctx.set_tuning(20 + L, int(L >= 3))


def _mixed_cfg0_and_cfg1_subset():
    for name, mk, gold in (("cfg0", synth.config0, "cfg0.npz"),
                           ("cfg1_300", lambda: synth.config1(300), "cfg1_300.npz")):
        w = mk()
        g = np.load(GOLD / gold)
        say(f"mixed {name}")
        t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12)
        Vm = t.esp(precision=_lib.FE_PREC_MIXED)
        V = t.esp()
        assert np.abs(V - g["V_tol12"]).max() < 1e-10
        err = np.abs(Vm - g["V_tol12"]).max()
        say(f"mixed {name} max err {err:.3e}")
        assert err < TOL_MIXED


@pytest.mark.parametrize("fp32_classes", ["default", "all"])
def test_mixed_full_size_against_oracle(fp32_classes):
    g = np.load(GOLD / "cfg1_30k.npz")
    w = synth.config1(30000)
    np.testing.assert_allclose(w.fingerprint(), g["fp"], rtol=1e-9)
    ctx = _lib.Context.get(0)
    if fp32_classes == "all":  # every class through the FP32 contraction
        for L in range(7):
            ctx.set_tuning(20 + L, 1)
    try:
        say(f"mixed cfg1 30000, FP32 classes: {fp32_classes}")
        t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12)
        Vm = t.esp(precision=_lib.FE_PREC_MIXED)
        err = np.abs(Vm[g["idx"]] - g["V"]).max()
        say(f"mixed cfg1 30000 max err {err:.3e}")
        assert err < TOL_MIXED
        V, G = esp_with_grad(t, w.points)
        Vm2, Gm = esp_with_grad(t, w.points, precision=_lib.FE_PREC_MIXED)
        assert np.abs(Vm2 - V).max() < TOL_MIXED
        assert np.abs(Gm - G).max() < 1e-7
    finally:
        for L in range(7):
            ctx.set_tuning(20 + L, int(L >= 3))


def test_mixed_with_zero_threshold_is_fp64():
    w = synth.config0(700)
    ctx = _lib.Context.get(0)
    t = build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    say("mixed tau=0")
    ctx.set_tuning(4, 0)
    try:
        Vm = t.esp(precision=_lib.FE_PREC_MIXED)
    finally:
        ctx.set_tuning(4, 10 ** 8)  # back to 1e-4
    assert np.abs(Vm - t.esp()).max() < 1e-12
