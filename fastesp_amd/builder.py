"""Device table build: ``build_table(..., backend="hip")``.

Calls ``fe_table_build`` with the basis.bin fields of the basis
(``fastesp/table.py:268-293``); the Hermite contraction, the screen and the
compaction run on the GPU (``csrc/fe_build.hip``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .formats import FlatTable, basis_payload
from .table import PairTable

SCREEN_NUMPY = 0   # fastesp/table.py:190-194
SCREEN_CPP = 1     # fastesp/native/build_pairs.cpp:318-328
SCREEN_ENGINE = 2  # fastesp/esp.py:265-270 (ESPEngine.esp_electron)


def build_table_hip(wfn, basis, density, points, tol: float = 1e-12, progress: bool = False,
                    rule: int = SCREEN_NUMPY, ctx: _lib.Context | None = None) -> PairTable:
    pts = np.ascontiguousarray(np.asarray(points, float).reshape(-1, 3))
    t = build_table_payload(basis_payload(wfn, basis, density, pts), tol, rule, ctx)
    if progress:
        print(f"  pairs kept {t.n_pairs}")
    return t


def build_table_payload(pl: dict, tol: float = 1e-12, rule: int = SCREEN_NUMPY,
                        ctx: _lib.Context | None = None) -> PairTable:
    """Build from the basis.bin fields (``read_basis_payload`` / ``basis_payload``)."""
    ctx = ctx or _lib.Context.get()
    pts = np.ascontiguousarray(np.asarray(pl["pts"], float).reshape(-1, 3))
    i32 = {k: np.ascontiguousarray(pl[k], np.int32) for k in
           ("fn_atom", "fn_nprim", "fn_off", "fn_lx", "fn_ly", "fn_lz", "fn_shell")}
    f64 = {k: np.ascontiguousarray(pl[k], np.float64) for k in
           ("fc", "pexp", "pcoef", "dens", "pts", "Z", "xyz")}
    h = C.c_void_p()
    ip = lambda a: _lib.ptr(a, C.c_int32)  # noqa: E731
    _lib.check(ctx.lib.fe_table_build(
        ctx.handle, pl["n_func"], pl["n_prim"], ip(i32["fn_atom"]), ip(i32["fn_nprim"]),
        ip(i32["fn_off"]), ip(i32["fn_lx"]), ip(i32["fn_ly"]), ip(i32["fn_lz"]),
        ip(i32["fn_shell"]), _lib.ptr(f64["fc"]), _lib.ptr(f64["pexp"]), _lib.ptr(f64["pcoef"]),
        f64["dens"].ctypes.data, len(pts), f64["pts"].ctypes.data, len(f64["Z"]),
        _lib.ptr(f64["Z"]), _lib.ptr(f64["xyz"]), float(tol), int(rule), _lib.FE_PTR_HOST, None,
        C.byref(h)), "fe_table_build")
    t = PairTable(h, ctx, pts, f64["Z"], f64["xyz"])
    t.meta["sum_rule"] = t.stats()["sum_rule"]
    return t


def export_flat(t: PairTable) -> FlatTable:
    """Device table -> host flat table in reference order (``fe_table_export``)."""
    info = t.info()
    n, nt = info["n_pairs"], info["n_terms"]
    lib = t._ctx.lib
    f = {k: np.empty(n) for k in ("pp_exp", "pp_cx", "pp_cy", "pp_cz", "pp_K", "kb", "reff")}
    i = {k: np.empty(n, np.int32) for k in ("pp_nt", "pp_off", "pp_nmono", "pp_monooff")}
    nmono_cls = [1, 4, 10, 20, 35, 56, 84]
    n_mono = int(sum(c * nmono_cls[L] for L, c in enumerate(info["pairs_per_L"])))
    mo_p = np.empty(n_mono, np.int32)
    mo_a = np.empty(n_mono, np.int32)
    ta = np.empty(nt, np.int32)
    tm = np.empty(nt, np.int32)
    tg = np.empty(nt)
    ip = lambda a: _lib.ptr(a, C.c_int32)  # noqa: E731
    _lib.check(lib.fe_table_export(
        t._h, _lib.ptr(f["pp_exp"]), _lib.ptr(f["pp_cx"]), _lib.ptr(f["pp_cy"]),
        _lib.ptr(f["pp_cz"]), _lib.ptr(f["pp_K"]), ip(i["pp_nt"]), ip(i["pp_off"]),
        ip(i["pp_nmono"]), ip(i["pp_monooff"]), ip(mo_p), ip(mo_a), ip(ta), ip(tm),
        _lib.ptr(tg), _lib.ptr(f["kb"]), _lib.ptr(f["reff"])), "fe_table_export")
    flat = FlatTable(pp_exp=f["pp_exp"], pp_cx=f["pp_cx"], pp_cy=f["pp_cy"], pp_cz=f["pp_cz"],
                     pp_K=f["pp_K"], pp_nt=i["pp_nt"], pp_off=i["pp_off"],
                     pp_nmono=i["pp_nmono"], pp_monooff=i["pp_monooff"], mo_parent=mo_p,
                     mo_axis=mo_a, tm_aidx=ta, tm_m=tm, tm_g=tg, pts=t.pts, Z=t.Z, xyz=t.xyz,
                     max_mono=int(i["pp_nmono"].max()) if n else 0)
    flat.meta.update(kbound=f["kb"], reff=f["reff"])
    return flat
