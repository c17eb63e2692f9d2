"""ctypes wrapper of ``libcpu_port.so`` (the C++/OpenMP CPU port) -- test
infrastructure and CPU baseline only."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from fastesp_amd.formats import basis_payload

LIB = Path(__file__).resolve().parent / "libcpu_port.so"
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        lib = C.CDLL(str(LIB))
        lib.cp_build.restype = C.c_void_p
        lib.cp_build.argtypes = [C.c_int] + [C.c_void_p] * 7 + [C.c_void_p] * 4 + [
            C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int]
        lib.cp_info.argtypes = [C.c_void_p, C.c_void_p]
        lib.cp_sum_rule.argtypes = [C.c_void_p]
        lib.cp_sum_rule.restype = C.c_double
        lib.cp_export.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.cp_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        lib.cp_free.argtypes = [C.c_void_p]
        _lib = lib
    return _lib


def threads() -> int:
    return len(os.sched_getaffinity(0))


class CpuTable:
    def __init__(self, wfn, basis, density, points, tol=1e-12, rule=0, nthreads=0):
        lib = load()
        pl = basis_payload(wfn, basis, density, points)
        self._keep = {k: np.ascontiguousarray(pl[k], np.int32) for k in
                      ("fn_atom", "fn_nprim", "fn_off", "fn_lx", "fn_ly", "fn_lz", "fn_shell")}
        self._keep.update({k: np.ascontiguousarray(pl[k], np.float64) for k in
                           ("fc", "pexp", "pcoef", "dens", "pts", "Z", "xyz")})
        k = self._keep
        self.h = lib.cp_build(pl["n_func"], *(k[n].ctypes.data for n in (
            "fn_atom", "fn_nprim", "fn_off", "fn_lx", "fn_ly", "fn_lz", "fn_shell", "fc",
            "pexp", "pcoef", "dens")), pl["n_points"], k["pts"].ctypes.data, pl["n_atoms"],
            k["Z"].ctypes.data, k["xyz"].ctypes.data, float(tol), int(rule), int(nthreads))
        info = np.zeros(3, np.int64)
        lib.cp_info(self.h, info.ctypes.data)
        self.n_pairs, self.n_terms, self.n_candidates = map(int, info)
        self.sum_rule = lib.cp_sum_rule(self.h)

    def esp(self, points, nthreads=0) -> np.ndarray:
        pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        V = np.empty(len(pts))
        load().cp_eval(self.h, pts.ctypes.data, len(pts), V.ctypes.data, int(nthreads))
        return V

    def export(self):
        p6 = np.empty((self.n_pairs, 6))
        g = np.empty(self.n_terms)
        load().cp_export(self.h, p6.ctypes.data, g.ctypes.data)
        return p6, g

    def oracle_table(self, wfn, points):
        """The screened table as an oracle table (pair fields, terms in the
        class key order) for ``oracle.evaluate``."""
        from types import SimpleNamespace
        from .ints import pair_type
        p6, g = self.export()
        L = p6[:, 5].astype(np.int32)
        nt = np.array([pair_type(int(l))[0].shape[0] for l in range(7)])[L]
        off = np.concatenate([[0], np.cumsum(nt)[:-1]]).astype(np.int64)
        return SimpleNamespace(pp_exp=p6[:, 0], pp_cx=p6[:, 1], pp_cy=p6[:, 2], pp_cz=p6[:, 3],
                               pp_K=p6[:, 4], pp_L=L, pp_off=off, tm_g=g, n_terms=len(g),
                               n_pairs=len(L), pts=np.asarray(points, float),
                               xyz=np.asarray(wfn.coords, float), Z=wfn.Z)

    def __del__(self):
        try:
            if self.h:
                load().cp_free(self.h)
        except Exception:
            pass
