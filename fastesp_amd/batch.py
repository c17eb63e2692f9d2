"""Batched conformers: many pair tables evaluated together (``fe_batch_*``).

    tabs = [build_table(w.wfn, w.basis, w.density, w.points) for w in confs]
    b = Batch(tabs, [w.points for w in confs])
    b.set_loss(weights, targets)                    # per-conformer lists
    b.capture("mixed", field=True, dldp=True)       # one hipGraph per step
    b.launch()                                      # replay: V, dV/dC, loss, dL/dP
    V, dP = b.values(), b.dldp()

Per class L (l_A + l_B) one persistent launch covers every conformer, pulling
(conformer, pair range, point block) work items from a queue; the step --
forward (+ field), loss gradient gV = 2 w (V - T), dL/dG and dL/dP -- is
recorded once as a hipGraph and replayed.  New coordinates (same counts) and
new loss data are taken by the recorded graph without re-recording.  With
``set_density`` and ``rebuild=True`` the step first builds every table again
on the device from the batch's densities and current points, so one graph
runs build -> forward -> loss -> backward.  The reference has no batched
form: it runs one ``esp_gpu`` process per table (fastesp/cli.py:143-163).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .table import PairTable


def _prec(p) -> int:
    if p in ("fp64", _lib.FE_PREC_FP64):
        return _lib.FE_PREC_FP64
    if p in ("mixed", _lib.FE_PREC_MIXED):
        return _lib.FE_PREC_MIXED
    raise ValueError(f"unknown precision {p!r}")


def _flags(field: bool, dldp: bool, rebuild: bool = False) -> int:
    return ((_lib.FE_BATCH_FIELD if field else 0) | (_lib.FE_BATCH_DLDP if dldp else 0)
            | (_lib.FE_BATCH_REBUILD if rebuild else 0))


class Batch:
    """A set of tables (kept alive by the batch) and their points."""

    def __init__(self, tables: list[PairTable], points: list, stream=None):
        if not tables or len(tables) != len(points):
            raise ValueError("need one point array per table and at least one table")
        self.tables = list(tables)
        self._ctx = tables[0]._ctx
        if any(t._ctx is not self._ctx for t in tables):
            raise ValueError("all tables must belong to one context")
        arrs = [np.ascontiguousarray(np.asarray(p, np.float64).reshape(-1, 3)) for p in points]
        self.n_pts = np.array([len(a) for a in arrs], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.n_pts)]).astype(np.int64)
        pts = np.ascontiguousarray(np.concatenate(arrs)) if self.offsets[-1] else np.zeros((1, 3))
        n = len(tables)
        handles = (C.c_void_p * n)(*[t._h.value for t in tables])
        h = C.c_void_p()
        _lib.check(self._ctx.lib.fe_batch_create(
            self._ctx.handle, n, handles, pts.ctypes.data,
            self.n_pts.ctypes.data_as(C.POINTER(C.c_int64)), _lib.FE_PTR_HOST, stream,
            C.byref(h)), "fe_batch_create")
        self._h = h
        self._n_func = None

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._ctx.lib.fe_batch_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inputs -------------------------------------------------------------
    def _concat(self, parts, what: str) -> np.ndarray:
        if isinstance(parts, np.ndarray) and parts.ndim == 1 and len(parts) == self.offsets[-1]:
            return np.ascontiguousarray(parts, np.float64)
        if len(parts) != len(self.tables):
            raise ValueError(f"{what}: one array per conformer expected")
        out = np.concatenate([np.asarray(p, np.float64).ravel() for p in parts])
        if len(out) != self.offsets[-1]:
            raise ValueError(f"{what}: {len(out)} values for {self.offsets[-1]} points")
        return np.ascontiguousarray(out)

    def set_points(self, points: list, stream=None) -> None:
        arrs = [np.asarray(p, np.float64).reshape(-1, 3) for p in points]
        if [len(a) for a in arrs] != self.n_pts.tolist():
            raise ValueError("set_points: point counts must not change")
        pts = np.ascontiguousarray(np.concatenate(arrs)) if self.offsets[-1] else np.zeros((1, 3))
        _lib.check(self._ctx.lib.fe_batch_set_points(self._h, pts.ctypes.data, _lib.FE_PTR_HOST,
                                                     stream), "fe_batch_set_points")

    def set_loss(self, weights=None, targets=None, stream=None) -> None:
        w = self._concat(weights, "weights") if weights is not None else None
        t = self._concat(targets, "targets") if targets is not None else None
        _lib.check(self._ctx.lib.fe_batch_set_loss(
            self._h, w.ctypes.data if w is not None else None,
            t.ctypes.data if t is not None else None, _lib.FE_PTR_HOST, stream),
            "fe_batch_set_loss")

    def set_density(self, densities: list, tol: float = 1e-12, rule: int = 0,
                    stream=None) -> None:
        """Densities of every conformer (n_func x n_func each, in conformer
        order) and the screen for steps with ``rebuild=True``.  The first call
        rebuilds every table once (its capacity layout); afterwards the tables
        are rebuilt by the steps that ask for it."""
        nf = self.n_func
        if len(densities) != len(nf):
            raise ValueError("set_density: one density per conformer expected")
        parts = []
        for P, n in zip(densities, nf):
            P = np.asarray(P, np.float64)
            if P.shape != (n, n):
                raise ValueError(f"set_density: density of shape {P.shape}, expected {(n, n)}")
            parts.append(P.ravel())
        flat = np.ascontiguousarray(np.concatenate(parts))
        _lib.check(self._ctx.lib.fe_batch_set_density(self._h, flat.ctypes.data, float(tol),
                                                      int(rule), _lib.FE_PTR_HOST, stream),
                   "fe_batch_set_density")

    def set_tuning(self, key: int, value: int) -> None:
        _lib.check(self._ctx.lib.fe_batch_set_tuning(self._h, key, int(value)),
                   "fe_batch_set_tuning")

    # -- work ---------------------------------------------------------------
    def forward(self, precision="fp64", field: bool = False, stream=None) -> None:
        _lib.check(self._ctx.lib.fe_batch_forward(self._h, _prec(precision), _flags(field, False),
                                                  stream), "fe_batch_forward")

    def loss(self, stream=None) -> None:
        _lib.check(self._ctx.lib.fe_batch_loss(self._h, stream), "fe_batch_loss")

    def backward(self, gV=None, stream=None, device_ptr: int | None = None) -> None:
        if device_ptr is not None:
            _lib.check(self._ctx.lib.fe_batch_backward(self._h, device_ptr, _lib.FE_PTR_DEVICE,
                                                       stream), "fe_batch_backward")
            return
        g = self._concat(gV, "gV") if gV is not None else None
        _lib.check(self._ctx.lib.fe_batch_backward(
            self._h, g.ctypes.data if g is not None else None, _lib.FE_PTR_HOST, stream),
            "fe_batch_backward")

    def step(self, precision="fp64", field: bool = True, dldp: bool = True, stream=None,
             rebuild: bool = False) -> None:
        _lib.check(self._ctx.lib.fe_batch_step(self._h, _prec(precision),
                                               _flags(field, dldp, rebuild), stream),
                   "fe_batch_step")

    def capture(self, precision="fp64", field: bool = True, dldp: bool = True,
                rebuild: bool = False) -> None:
        _lib.check(self._ctx.lib.fe_batch_capture(self._h, _prec(precision),
                                                  _flags(field, dldp, rebuild)),
                   "fe_batch_capture")

    def launch(self, stream=None) -> None:
        _lib.check(self._ctx.lib.fe_batch_launch(self._h, stream), "fe_batch_launch")

    # -- outputs ------------------------------------------------------------
    def read(self, what: int, count: int, stream=None) -> np.ndarray:
        out = np.empty(max(count, 1))
        _lib.check(self._ctx.lib.fe_batch_read(self._h, what, _lib.FE_PTR_HOST, stream,
                                               out.ctypes.data), "fe_batch_read")
        return out[:count]

    def _split(self, flat: np.ndarray, width: int = 1) -> list:
        o = self.offsets
        return [flat[width * o[i]:width * o[i + 1]].reshape((-1, 3) if width == 3 else (-1,))
                for i in range(len(self.tables))]

    def values(self) -> list:
        return self._split(self.read(_lib.FE_BOUT_V, int(self.offsets[-1])))

    def field(self) -> list:
        return self._split(self.read(_lib.FE_BOUT_FIELD, 3 * int(self.offsets[-1])), 3)

    def gV(self) -> list:
        return self._split(self.read(_lib.FE_BOUT_GV, int(self.offsets[-1])))

    def losses(self) -> np.ndarray:
        return self.read(_lib.FE_BOUT_LOSS, len(self.tables))

    @property
    def n_func(self) -> list:
        if self._n_func is None:
            self._n_func = [int(t.stats()["n_func"]) for t in self.tables]
        return self._n_func

    def dldp(self) -> list:
        nf = self.n_func
        flat = self.read(_lib.FE_BOUT_DLDP, int(sum(n * n for n in nf)))
        out, o = [], 0
        for n in nf:
            out.append(flat[o:o + n * n].reshape(n, n))
            o += n * n
        return out

    def info(self) -> dict:
        v = (C.c_int64 * 12)()
        _lib.check(self._ctx.lib.fe_batch_info(self._h, v, 12), "fe_batch_info")
        keys = ("n_conf", "n_points", "n_pairs", "items", "dg_items", "partial_bytes",
                "dldp_doubles", "shell_pairs", "captured", "items_per_wg", "dg_items_per_wg",
                "class_mask")
        return dict(zip(keys, [int(x) for x in v]))
