"""Electron density at points and vdW-surface points, on the GPU.

* ``ElectronDensity(wfn, basis, density).rho(points)`` -- rho(C) =
  sum_{mu nu} phi_mu(C) P_{mu nu} phi_nu(C), the quantity of ``our_density``
  (``benchmarks/check_density_mwfn.py:60-75``), with the basis values and the
  quadratic form on the device (``fe_density_*``: FP64 matrix cores, the
  row-dot fused into the matrix kernel).  ``electron_density(...)`` is the
  one-call form.
* ``vdw_surface_points(wfn, n_points, seed=0)`` -- ``examples/
  vdw_surface_esp.py:38-52``: Fibonacci spheres of Bondi radius on every
  atom, the points outside every sphere kept, then ``n_points`` of them drawn
  without replacement by ``numpy.random.default_rng(seed)``.  The candidates
  and the exposure test run on the device (``fe_vdw_candidates``); the draw
  stays on the host so the selection is the reference's, index for index.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .formats import basis_payload

# examples/vdw_surface_esp.py:24-25
BONDI = {"H": 1.20, "C": 1.70, "N": 1.55, "O": 1.52, "S": 1.80, "P": 1.80}
BOHR = 0.529177210903


class ElectronDensity:
    def __init__(self, wfn, basis, density, ctx: _lib.Context | None = None):
        self._ctx = ctx or _lib.Context.get()
        pl = basis_payload(wfn, basis, density, np.zeros((0, 3)))
        self._keep = {k: np.ascontiguousarray(pl[k], np.int32)
                      for k in ("fn_nprim", "fn_off", "fn_lx", "fn_ly", "fn_lz")}
        self._keep.update({k: np.ascontiguousarray(pl[k], np.float64)
                           for k in ("fc", "pexp", "pcoef", "dens")})
        k = self._keep
        ip = lambda a: _lib.ptr(a, C.c_int32)  # noqa: E731
        h = C.c_void_p()
        _lib.check(self._ctx.lib.fe_density_create(
            self._ctx.handle, pl["n_func"], pl["n_prim"], ip(k["fn_nprim"]), ip(k["fn_off"]),
            ip(k["fn_lx"]), ip(k["fn_ly"]), ip(k["fn_lz"]), _lib.ptr(k["fc"]),
            _lib.ptr(k["pexp"]), _lib.ptr(k["pcoef"]), k["dens"].ctypes.data,
            _lib.FE_PTR_HOST, None, C.byref(h)), "fe_density_create")
        self._h = h
        self.n_func = pl["n_func"]

    def rho(self, points) -> np.ndarray:
        pts = np.ascontiguousarray(np.asarray(points, float).reshape(-1, 3))
        out = np.empty(len(pts))
        _lib.check(self._ctx.lib.fe_density_eval(self._h, pts.ctypes.data, len(pts),
                                                 _lib.FE_PTR_HOST, None, out.ctypes.data),
                   "fe_density_eval")
        return out

    def rho_device(self, d_pts: int, n: int, d_rho: int, stream=None) -> None:
        """rho at n device-resident points (device pointers, enqueued on `stream`)."""
        _lib.check(self._ctx.lib.fe_density_eval(self._h, d_pts, n, _lib.FE_PTR_DEVICE, stream,
                                                 d_rho), "fe_density_eval")

    def set_tile128(self, on: bool) -> None:
        """128 x 64 or 64 x 64 (default) point x function tiles of the matrix kernel."""
        _lib.check(self._ctx.lib.fe_density_set_tuning(self._h, 0, int(bool(on))),
                   "fe_density_set_tuning")

    def info(self) -> dict:
        v = (C.c_int64 * 4)()
        _lib.check(self._ctx.lib.fe_density_info(self._h, v, 4), "fe_density_info")
        return dict(n_func=v[0], n_pad=v[1], mm_flops=v[2], chunk=v[3])

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._ctx.lib.fe_density_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def electron_density(wfn, basis, density, points) -> np.ndarray:
    return ElectronDensity(wfn, basis, density).rho(points)


def vdw_surface_points(wfn, n_points: int, seed: int = 0,
                       ctx: _lib.Context | None = None) -> np.ndarray:
    ctx = ctx or _lib.Context.get()
    xyz = np.ascontiguousarray(np.asarray(wfn.coords, float).reshape(-1, 3))
    rad = np.array([BONDI[s] / BOHR for s in wfn.symbols], np.float64)
    share = (4 * np.pi * rad ** 2) / (4 * np.pi * rad ** 2).sum()
    counts = np.array([int(np.ceil(n_points * share[a] * 3.0)) + 50 for a in range(len(rad))],
                      np.int64)
    n = int(counts.sum())
    pts = np.empty((n, 3))
    ok = np.empty(n, np.int32)
    _lib.check(ctx.lib.fe_vdw_candidates(ctx.handle, len(rad), _lib.ptr(xyz), _lib.ptr(rad),
                                         counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                         _lib.ptr(pts), _lib.ptr(ok, C.c_int32)),
               "fe_vdw_candidates")
    cand = pts[ok.astype(bool)]
    if len(cand) > n_points:
        cand = cand[np.random.default_rng(seed).choice(len(cand), n_points, replace=False)]
    return cand
