"""Differentiable ESP: a ``torch.autograd.Function`` over the HIP kernels.

    model = ESPModel(wfn, basis, tol=1e-12)
    V = model(points, density)          # points (N,3), density (n,n), float64, cuda
    loss = (w * (V - target) ** 2).sum()
    loss.backward()                      # points.grad = gV * dV/dC, density.grad = dL/dP

Forward builds the pair table from ``density`` on the device: the first call
with ``fe_table_build`` (pointers handed over directly), later calls in place
with ``fe_table_rebuild``, which neither allocates nor synchronises with the
host, so a whole step (build, forward, backward) can be recorded as a graph.
It then evaluates V -- fused with the field dV/dC when ``points`` requires a gradient
(``fe_esp_forward_grad``).  Backward returns gV * dV/dC and dL/dP
(``fe_esp_backward``: dL/dG per pair with the table's G recomputed from the
stored Hermite data rather than stored per point, then the adjoint of the
table contraction).  The screen depends on the points and the density only
through discrete keep/drop decisions, which are held fixed by the backward
(their derivative is zero almost everywhere).

PyTorch is used here only for the autograd plumbing, device memory and the
current stream; every number is computed by ``libfastesp_hip.so``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .formats import basis_payload
from .table import PairTable


class _BasisArrays:
    """Host copies of the basis.bin arrays of a basis (built once per model)."""

    def __init__(self, wfn, basis):
        pl = basis_payload(wfn, basis, np.zeros((basis.n, basis.n)), np.zeros((0, 3)))
        self.n_func = pl["n_func"]
        self.n_prim = pl["n_prim"]
        self.i32 = {k: np.ascontiguousarray(pl[k], np.int32) for k in
                    ("fn_atom", "fn_nprim", "fn_off", "fn_lx", "fn_ly", "fn_lz", "fn_shell")}
        self.f64 = {k: np.ascontiguousarray(pl[k], np.float64) for k in
                    ("fc", "pexp", "pcoef", "Z", "xyz")}


def _prec(p) -> int:
    if p in ("fp64", _lib.FE_PREC_FP64):
        return _lib.FE_PREC_FP64
    if p in ("mixed", _lib.FE_PREC_MIXED):
        return _lib.FE_PREC_MIXED
    raise ValueError(f"unknown precision {p!r}")


class ESPModel:
    """Holds the basis; ``__call__(points, density)`` is differentiable in both.

    ``reuse_table=True`` (default) keeps one device table and rebuilds it in
    place on every call after the first (``fe_table_rebuild``); a backward must
    then run before the next forward (it checks, and raises otherwise).
    ``reuse_table=False`` builds a fresh table per call."""

    def __init__(self, wfn, basis, tol: float = 1e-12, precision="fp64", rule: int = 0,
                 device: int | None = None, reuse_table: bool = True):
        self.ctx = _lib.Context.get(device)
        self.arr = _BasisArrays(wfn, basis)
        self.tol = float(tol)
        self.precision = _prec(precision)
        self.rule = int(rule)
        self.reuse_table = bool(reuse_table)
        self.last_table: PairTable | None = None
        self.generation = 0  # number of builds of the held table

    def build(self, points: torch.Tensor, density: torch.Tensor) -> PairTable:
        a = self.arr
        stream = torch.cuda.current_stream(points.device).cuda_stream
        if self.reuse_table and self.last_table is not None:
            _lib.check(self.ctx.lib.fe_table_rebuild(
                self.last_table._h, density.data_ptr(), points.shape[0], points.data_ptr(),
                self.tol, self.rule, _lib.FE_PTR_DEVICE, stream), "fe_table_rebuild")
            self.generation += 1
            return self.last_table
        ip = lambda k: _lib.ptr(a.i32[k], C.c_int32)  # noqa: E731
        h = C.c_void_p()
        _lib.check(self.ctx.lib.fe_table_build(
            self.ctx.handle, a.n_func, a.n_prim, ip("fn_atom"), ip("fn_nprim"), ip("fn_off"),
            ip("fn_lx"), ip("fn_ly"), ip("fn_lz"), ip("fn_shell"), _lib.ptr(a.f64["fc"]),
            _lib.ptr(a.f64["pexp"]), _lib.ptr(a.f64["pcoef"]), density.data_ptr(),
            points.shape[0], points.data_ptr(), len(a.f64["Z"]), _lib.ptr(a.f64["Z"]),
            _lib.ptr(a.f64["xyz"]), self.tol, self.rule, _lib.FE_PTR_DEVICE, stream,
            C.byref(h)), "fe_table_build")
        t = PairTable(h, self.ctx, np.zeros((0, 3)), a.f64["Z"], a.f64["xyz"])
        t.precision = self.precision
        self.last_table = t
        self.generation += 1
        return t

    def __call__(self, points: torch.Tensor, density: torch.Tensor) -> torch.Tensor:
        return ESPFunction.apply(points, density, self)


def _check(x: torch.Tensor, name: str, shape=None) -> torch.Tensor:
    if x.dtype != torch.float64 or not x.is_cuda:
        raise TypeError(f"{name} must be a float64 CUDA tensor")
    if shape is not None and tuple(x.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(x.shape)}, expected {tuple(shape)}")
    return x.contiguous()


class ESPFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, density, model: ESPModel):
        n = points.shape[0]
        pts = _check(points.detach(), "points", (n, 3))
        dens = _check(density.detach(), "density", (model.arr.n_func, model.arr.n_func))
        table = model.build(pts, dens)
        ctx.generation = model.generation
        stream = torch.cuda.current_stream(pts.device).cuda_stream
        lib = model.ctx.lib
        V = torch.empty(n, dtype=torch.float64, device=pts.device)
        field = None
        if ctx.needs_input_grad[0]:
            field = torch.empty((n, 3), dtype=torch.float64, device=pts.device)
            _lib.check(lib.fe_esp_forward_grad(table._h, pts.data_ptr(), n, model.precision,
                                               _lib.FE_PTR_DEVICE, stream, V.data_ptr(),
                                               field.data_ptr()), "fe_esp_forward_grad")
        else:
            _lib.check(lib.fe_esp_forward(table._h, pts.data_ptr(), n, model.precision,
                                          _lib.FE_PART_ESP, _lib.FE_PTR_DEVICE, stream,
                                          V.data_ptr()), "fe_esp_forward")
        ctx.table = table
        ctx.model = model
        ctx.save_for_backward(pts, field if field is not None else torch.empty(0))
        return V

    @staticmethod
    def backward(ctx, gV):
        pts, field = ctx.saved_tensors
        model = ctx.model
        if model.reuse_table and ctx.generation != model.generation:
            raise RuntimeError("ESPModel: the table was rebuilt by a later forward; run each "
                               "backward before the next forward, or use reuse_table=False")
        g_pts = g_dens = None
        gV = gV.detach().to(torch.float64).contiguous()
        if ctx.needs_input_grad[0]:
            g_pts = gV[:, None] * field
        if ctx.needs_input_grad[1]:
            nf = model.arr.n_func
            g_dens = torch.empty((nf, nf), dtype=torch.float64, device=pts.device)
            stream = torch.cuda.current_stream(pts.device).cuda_stream
            _lib.check(model.ctx.lib.fe_esp_backward(
                ctx.table._h, pts.data_ptr(), pts.shape[0], gV.data_ptr(), model.precision,
                _lib.FE_PTR_DEVICE, stream, None, g_dens.data_ptr()), "fe_esp_backward")
        return g_pts, g_dens, None


def esp_with_grad(table: PairTable, points: np.ndarray, precision=_lib.FE_PREC_FP64):
    """Host convenience: (V, dV/dC) for a table (``fe_esp_forward_grad``)."""
    pts = np.ascontiguousarray(np.asarray(points, float).reshape(-1, 3))
    V = np.empty(len(pts))
    G = np.empty((len(pts), 3))
    _lib.check(table._ctx.lib.fe_esp_forward_grad(table._h, pts.ctypes.data, len(pts),
                                                  _prec(precision), _lib.FE_PTR_HOST, None,
                                                  V.ctypes.data, G.ctypes.data),
               "fe_esp_forward_grad")
    return V, G


def esp_backward(table: PairTable, points: np.ndarray, gV: np.ndarray, want_dG=True,
                 want_dP=True, n_func: int | None = None, precision=_lib.FE_PREC_FP64):
    """Host convenience: (dL/dG flat, dL/dP) for a table (``fe_esp_backward``)."""
    pts = np.ascontiguousarray(np.asarray(points, float).reshape(-1, 3))
    gV = np.ascontiguousarray(gV, np.float64)
    dG = np.empty(table.n_terms) if want_dG else None
    dP = None
    if want_dP:
        nf = n_func if n_func is not None else table.stats()["n_func"]
        dP = np.empty((nf, nf))
    _lib.check(table._ctx.lib.fe_esp_backward(
        table._h, pts.ctypes.data, len(pts), gV.ctypes.data, _prec(precision), _lib.FE_PTR_HOST,
        None, dG.ctypes.data if dG is not None else None,
        dP.ctypes.data if dP is not None else None), "fe_esp_backward")
    return dG, dP
