"""The C-ABI library loads and exports every entry point declared in
include/fastesp_hip.h (no compute calls, runs without a GPU)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from fastesp_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "fastesp_hip.h"


def declared():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fe_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_api():
    names = declared()
    for must in ("fe_init", "fe_finalize", "fe_last_error", "fe_table_build", "fe_table_from_pairs",
                 "fe_table_export", "fe_table_free", "fe_esp_forward", "fe_esp_forward_grad",
                 "fe_esp_backward", "fe_boys_eval", "fe_batch_create", "fe_batch_capture",
                 "fe_batch_launch", "fe_batch_step", "fe_batch_free"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert _lib.LIB_PATH.exists(), "build the library first (python -m fastesp_amd.build)"
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\bT (fe_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared()) <= exported


def test_python_binding_covers_the_header():
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_errors_are_reported_not_fatal():
    lib = _lib.load()
    assert lib.fe_version() >= 100
    # null arguments come back as FE_ERR_ARG with a message, never a crash
    assert lib.fe_init(0, None) == -1
    assert b"null" in lib.fe_last_error()
    assert lib.fe_table_free(None) == -1
    assert lib.fe_esp_forward(None, None, 0, 0, 3, 0, None, None) == -1
    assert lib.fe_batch_create(None, 1, None, None, None, 0, None, None) == -1
    assert lib.fe_batch_free(None) == 0
    assert lib.fe_batch_launch(None, None) == -1
    assert lib.fe_batch_step(None, 0, 3, None) == -1
    assert lib.fe_density_create(None, 1, 1, None, None, None, None, None, None, None, None,
                                 None, 0, None, None) == -1
    assert lib.fe_density_eval(None, None, 1, 0, None, None) == -1
    assert lib.fe_density_free(None) == 0
    assert lib.fe_vdw_candidates(None, 1, None, None, None, None, None) == -1


def test_no_device_is_a_clean_error():
    from tests.conftest import have_gpu
    if have_gpu():
        pytest.skip("a device is present")
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.fe_init(0, ctypes.byref(h))
    assert rc in (-1, -2)
    assert lib.fe_last_error()


def test_table_build_rejects_bad_arguments_without_a_device():
    """fe_table_build and fe_table_rebuild check their arguments before they
    touch a device: each bad argument comes back as FE_ERR_ARG with a message
    naming it."""
    import numpy as np
    lib = _lib.load()
    i = np.zeros(1, np.int32)
    d = np.zeros(3)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731

    def build(n_func=1, n_points=1, tol=1e-12, rule=0, ptr_kind=0, dens=True):
        out = ctypes.c_void_p()
        return lib.fe_table_build(None, n_func, 1, ip(i), ip(i), ip(i), ip(i), ip(i), ip(i), ip(i),
                                  dp(d), dp(d), dp(d), d.ctypes.data if dens else None, n_points,
                                  d.ctypes.data, 0, None, None, tol, rule, ptr_kind, None,
                                  ctypes.byref(out))

    for kw, msg in [({"rule": 5}, b"rule"), ({"tol": -1.0}, b"tol"), ({"n_points": -1}, b"point"),
                    ({"ptr_kind": 7}, b"ptr_kind"), ({"n_func": 0}, b"sizes"),
                    ({"dens": False}, b"null"), ({}, b"null argument")]:
        assert build(**kw) == -1, kw
        assert msg in lib.fe_last_error(), (kw, lib.fe_last_error())
    assert lib.fe_table_rebuild(None, None, 0, None, 1e-12, 0, 0, None) == -1
    assert b"null table" in lib.fe_last_error()
