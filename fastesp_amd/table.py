"""``build_table(...).esp()`` with the HIP backend.

The public surface mirrors the reference (``fastesp/table.py:40-104,176-259``):

* ``build_table(wfn, basis, density, points, tol=1e-12, progress=False,
  backend="hip")`` returns a table whose ``esp()``, ``electron()`` and
  ``nuclear()`` evaluate on the GPU (``points=None`` means the build points);
* ``PairTable.read(path)`` / ``.write(path)`` keep pairs.bin compatibility, and
  a table read from a pairs.bin written by the reference's ``build_pairs``
  evaluates unchanged (``from_flat``).

All device work goes through ``libfastesp_hip.so``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .formats import FlatTable, read_pairs, write_pairs


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


class PairTable:
    """A device-resident pair table (``fastesp/table.py:40-104`` surface)."""

    def __init__(self, handle, ctx: _lib.Context, pts, Z, xyz, flat: FlatTable | None = None,
                 meta: dict | None = None):
        self._h = handle
        self._ctx = ctx
        self.pts = np.asarray(pts, float).reshape(-1, 3)
        self.Z = np.asarray(Z, float)
        self.xyz = np.asarray(xyz, float).reshape(-1, 3)
        self._flat = flat
        self.meta = dict(meta or {})
        self.precision = _lib.FE_PREC_FP64

    # -- construction ------------------------------------------------------
    @classmethod
    def from_flat(cls, t, ctx: _lib.Context | None = None) -> "PairTable":
        """Upload any object with the pairs.bin fields (``fastesp/table.py:40-62``)."""
        ctx = ctx or _lib.Context.get()
        f64 = {k: _c(getattr(t, k), np.float64) for k in
               ("pp_exp", "pp_cx", "pp_cy", "pp_cz", "pp_K", "tm_g")}
        i32 = {k: _c(getattr(t, k), np.int32) for k in
               ("pp_nt", "pp_off", "pp_nmono", "pp_monooff", "mo_parent", "mo_axis",
                "tm_aidx", "tm_m")}
        Z = _c(t.Z, np.float64)
        xyz = _c(t.xyz, np.float64).reshape(-1)
        h = C.c_void_p()
        lib = ctx.lib
        _lib.check(lib.fe_table_from_pairs(
            ctx.handle, len(f64["pp_exp"]), len(f64["tm_g"]),
            _lib.ptr(f64["pp_exp"]), _lib.ptr(f64["pp_cx"]), _lib.ptr(f64["pp_cy"]),
            _lib.ptr(f64["pp_cz"]), _lib.ptr(f64["pp_K"]),
            _lib.ptr(i32["pp_nt"], C.c_int32), _lib.ptr(i32["pp_off"], C.c_int32),
            _lib.ptr(i32["pp_nmono"], C.c_int32), _lib.ptr(i32["pp_monooff"], C.c_int32),
            _lib.ptr(i32["mo_parent"], C.c_int32), _lib.ptr(i32["mo_axis"], C.c_int32),
            _lib.ptr(i32["tm_aidx"], C.c_int32), _lib.ptr(i32["tm_m"], C.c_int32),
            _lib.ptr(f64["tm_g"]), len(Z), _lib.ptr(Z), _lib.ptr(xyz), C.byref(h)),
            "fe_table_from_pairs")
        flat = t if isinstance(t, FlatTable) else None
        return cls(h, ctx, t.pts, Z, xyz, flat=flat)

    @classmethod
    def read(cls, path) -> "PairTable":
        """Load a pairs.bin (``fastesp/table.py:143-170``) onto the device."""
        return cls.from_flat(read_pairs(path))

    def write(self, path) -> None:
        """pairs.bin in the reference layout (``fastesp/table.py:119-141``)."""
        write_pairs(path, self.to_flat())

    def to_flat(self) -> FlatTable:
        if self._flat is None:
            from .builder import export_flat
            self._flat = export_flat(self)
        return self._flat

    def stats(self) -> dict:
        v = (C.c_double * 4)()
        _lib.check(self._ctx.lib.fe_table_stats(self._h, v, 4), "fe_table_stats")
        return dict(sum_rule=v[0], n_pairs=int(v[1]), n_func=int(v[2]), n_shell_pairs=int(v[3]))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._ctx.lib.fe_table_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- the reference's flat-table fields (fastesp/table.py:41-62) ---------
    # Exported from the device on first use (fe_table_export, reference pair
    # order) and returned as read-only arrays, like the frombuffer views of a
    # table read from pairs.bin (fastesp/table.py:150-154).
    def _field(self, name: str) -> np.ndarray:
        v = getattr(self.to_flat(), name).view()
        v.flags.writeable = False
        return v

    pp_exp = property(lambda self: self._field("pp_exp"), doc="product exponents p")
    pp_cx = property(lambda self: self._field("pp_cx"), doc="product centre x")
    pp_cy = property(lambda self: self._field("pp_cy"), doc="product centre y")
    pp_cz = property(lambda self: self._field("pp_cz"), doc="product centre z")
    pp_K = property(lambda self: self._field("pp_K"), doc="exp(-ab/p |A-B|^2)")
    pp_nt = property(lambda self: self._field("pp_nt"), doc="terms per pair")
    pp_off = property(lambda self: self._field("pp_off"), doc="first term of each pair")
    pp_nmono = property(lambda self: self._field("pp_nmono"), doc="monomials per pair")
    pp_monooff = property(lambda self: self._field("pp_monooff"), doc="first monomial of each pair")
    mo_parent = property(lambda self: self._field("mo_parent"), doc="monomial schedule: parent")
    mo_axis = property(lambda self: self._field("mo_axis"), doc="monomial schedule: axis")
    tm_aidx = property(lambda self: self._field("tm_aidx"), doc="monomial index of each term")
    tm_m = property(lambda self: self._field("tm_m"), doc="Boys order of each term")
    tm_g = property(lambda self: self._field("tm_g"), doc="coefficient of each term")

    @property
    def max_mono(self) -> int:
        """Largest monomial count of a pair (header field of pairs.bin)."""
        return int(self.to_flat().max_mono)

    def screen(self, tol: float) -> "PairTable":
        """``PairTable.screen`` (``fastesp/table.py:107-116``): the reference's
        method keeps every pair and returns the table itself; the screen that
        drops pairs runs inside the device build (``tol`` of ``build_table``)."""
        return self

    # -- info --------------------------------------------------------------
    def info(self) -> dict:
        v = (C.c_int64 * 16)()
        _lib.check(self._ctx.lib.fe_table_info(self._h, v, 16), "fe_table_info")
        return dict(n_pairs=v[0], n_terms=v[1], n_atoms=v[2],
                    pairs_per_L=[v[3 + l] for l in range(_lib.FE_LMAX + 1)],
                    device_bytes=v[10], n_candidates=v[11])

    @property
    def n_pairs(self) -> int:
        return self.info()["n_pairs"]

    @property
    def n_terms(self) -> int:
        return self.info()["n_terms"]

    # -- evaluation --------------------------------------------------------
    def _eval(self, points, parts: int, precision: int | None = None) -> np.ndarray:
        pts = self.pts if points is None else points
        pts = _c(np.asarray(pts, float).reshape(-1, 3), np.float64)
        out = np.empty(len(pts))
        prec = self.precision if precision is None else precision
        _lib.check(self._ctx.lib.fe_esp_forward(self._h, pts.ctypes.data, len(pts), prec, parts,
                                                _lib.FE_PTR_HOST, None, out.ctypes.data),
                   "fe_esp_forward")
        return out

    def electron(self, points=None) -> np.ndarray:
        """Positive electron integral sum_pairs K sum_k g R^a F_m (``fastesp/table.py:73-93``)."""
        return -self._eval(points, _lib.FE_PART_ELECTRON)

    def nuclear(self, points=None) -> np.ndarray:
        """sum_A Z_A/|C-R_A| (``fastesp/table.py:95-100``)."""
        return self._eval(points, _lib.FE_PART_NUCLEAR)

    def esp(self, points=None, precision: int | None = None) -> np.ndarray:
        """nuclear - electron in Hartree/e (``fastesp/table.py:102-104``)."""
        return self._eval(points, _lib.FE_PART_ESP, precision)


def build_table(wfn, basis, density, points, tol: float = 1e-12, progress: bool = False,
                backend: str = "hip"):
    """Drop-in for ``fastesp.build_table`` (``fastesp/table.py:176-178``)."""
    if backend != "hip":
        raise ValueError(f"backend {backend!r}: this package provides the 'hip' backend only")
    from .builder import build_table_hip
    return build_table_hip(wfn, basis, density, points, tol=tol, progress=progress)
