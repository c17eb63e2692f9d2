"""fastesp wire formats: ``pairs.bin`` (the flat pair table) and ``basis.bin``.

Both are little-endian and byte-compatible with the reference:

pairs.bin (``fastesp/table.py:11-24,119-170``; writer
``fastesp/native/build_pairs.cpp:386-421``; reader ``fastesp/native/esp_gpu.cu:96-134``)::

    int32   n_pairs, n_terms, n_points, n_atoms, max_mono, 0, 0, 0
    f64     pp_exp, pp_cx, pp_cy, pp_cz, pp_K          [n_pairs] each
    i32     pp_nt, pp_off, pp_nmono, pp_monooff        [n_pairs] each
    i32     mo_parent, mo_axis                         [n_mono]  (n_mono = monooff[-1]+nmono[-1])
    i32     tm_aidx, tm_m                              [n_terms]
    f64     tm_g                                       [n_terms]
    f64     pts[3 n_points], atom_Z[n_atoms], atom_xyz[3 n_atoms]

basis.bin (``fastesp/table.py:268-316``; reader ``fastesp/native/build_pairs.cpp:48-70``)::

    int32   n_func, n_prim, n_points, n_atoms, max_prim, 0, 0, 0
    i32     fn_atom, fn_nprim, fn_off, fn_lx, fn_ly, fn_lz, fn_shell   [n_func] each
    f64     fc[3 n_func], pexp[n_prim], pcoef[n_prim], dens[n_func^2],
            pts[3 n_points], Z[n_atoms], xyz[3 n_atoms]

Errors are raised as ``ValueError`` (truncated or inconsistent files) instead
of the reference's ``exit(1)``.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .basis import Z_OF

HEADER = struct.Struct("<8i")

PAIR_F64 = ("pp_exp", "pp_cx", "pp_cy", "pp_cz", "pp_K")
PAIR_I32 = ("pp_nt", "pp_off", "pp_nmono", "pp_monooff")


@dataclass
class FlatTable:
    """The flat pair table with the field names of ``fastesp/table.py:40-62``."""
    pp_exp: np.ndarray
    pp_cx: np.ndarray
    pp_cy: np.ndarray
    pp_cz: np.ndarray
    pp_K: np.ndarray
    pp_nt: np.ndarray
    pp_off: np.ndarray
    pp_nmono: np.ndarray
    pp_monooff: np.ndarray
    mo_parent: np.ndarray
    mo_axis: np.ndarray
    tm_aidx: np.ndarray
    tm_m: np.ndarray
    tm_g: np.ndarray
    pts: np.ndarray
    Z: np.ndarray
    xyz: np.ndarray
    max_mono: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n_pairs(self) -> int:
        return len(self.pp_exp)

    @property
    def n_terms(self) -> int:
        return len(self.tm_g)


def write_pairs(path, t) -> None:
    """Write any object with the flat-table fields as pairs.bin."""
    with Path(path).open("wb") as f:
        f.write(HEADER.pack(len(t.pp_exp), len(t.tm_g), len(t.pts), len(t.Z),
                            int(t.max_mono), 0, 0, 0))
        for k in PAIR_F64:
            f.write(np.ascontiguousarray(getattr(t, k), "<f8").tobytes())
        for k in PAIR_I32 + ("mo_parent", "mo_axis", "tm_aidx", "tm_m"):
            f.write(np.ascontiguousarray(getattr(t, k), "<i4").tobytes())
        f.write(np.ascontiguousarray(t.tm_g, "<f8").tobytes())
        f.write(np.ascontiguousarray(t.pts, "<f8").tobytes())
        f.write(np.ascontiguousarray(t.Z, "<f8").tobytes())
        f.write(np.ascontiguousarray(t.xyz, "<f8").tobytes())


def read_pairs(path) -> FlatTable:
    """Read pairs.bin; every section is length-checked (``fastesp/table.py:143-170``)."""
    d = Path(path).read_bytes()
    if len(d) < HEADER.size:
        raise ValueError(f"{path}: truncated header")
    n, nt, npt, nat, maxmono = HEADER.unpack_from(d, 0)[:5]
    if min(n, nt, npt, nat) < 0:
        raise ValueError(f"{path}: negative count in header")
    o = HEADER.size

    def rd(dt, count):
        nonlocal o
        nb = np.dtype(dt).itemsize * count
        if o + nb > len(d):
            raise ValueError(f"{path}: truncated (need {o + nb} bytes, have {len(d)})")
        a = np.frombuffer(d, dtype=dt, count=count, offset=o).copy()
        o += nb
        return a

    out = {}
    for k in PAIR_F64:
        out[k] = rd("<f8", n)
    for k in PAIR_I32:
        out[k] = rd("<i4", n)
    n_mono = int(out["pp_monooff"][-1]) + int(out["pp_nmono"][-1]) if n else 0
    out["mo_parent"] = rd("<i4", n_mono)
    out["mo_axis"] = rd("<i4", n_mono)
    out["tm_aidx"] = rd("<i4", nt)
    out["tm_m"] = rd("<i4", nt)
    out["tm_g"] = rd("<f8", nt)
    out["pts"] = rd("<f8", 3 * npt).reshape(npt, 3)
    out["Z"] = rd("<f8", nat)
    out["xyz"] = rd("<f8", 3 * nat).reshape(nat, 3)
    return FlatTable(max_mono=maxmono, **out)


# ---------------------------------------------------------------------------
def basis_payload(wfn, basis, density, points) -> dict:
    """Flat basis + density + points arrays (``fastesp/table.py:268-293``)."""
    fn_nprim = np.array([len(e) for e in basis.exps], np.int32)
    shell_ids = np.repeat(np.arange(len(wfn.shells), dtype=np.int32),
                          [(sh.l + 1) * (sh.l + 2) // 2 for sh in wfn.shells])
    points = np.asarray(points, np.float64).reshape(-1, 3)
    return dict(
        n_func=basis.n, n_prim=int(fn_nprim.sum()), n_points=len(points),
        n_atoms=wfn.natom, max_prim=int(fn_nprim.max()),
        fn_atom=np.asarray(basis.atoms, np.int32), fn_nprim=fn_nprim,
        fn_off=np.concatenate([[0], np.cumsum(fn_nprim)[:-1]]).astype(np.int32),
        fn_lx=basis.ang_arr[:, 0].astype(np.int32),
        fn_ly=basis.ang_arr[:, 1].astype(np.int32),
        fn_lz=basis.ang_arr[:, 2].astype(np.int32),
        fn_shell=shell_ids,
        fc=np.asarray(basis.centers, np.float64).ravel(),
        pexp=np.concatenate(basis.exps).astype(np.float64),
        pcoef=np.concatenate(basis.coefs).astype(np.float64),
        dens=np.ascontiguousarray(density, np.float64),
        pts=points.ravel(),
        Z=np.array([Z_OF[s] for s in wfn.symbols], np.float64),
        xyz=np.ascontiguousarray(wfn.coords, np.float64),
    )


def write_basis_payload(path, payload: dict) -> None:
    """``fastesp/table.py:307-316``."""
    with Path(path).open("wb") as f:
        f.write(HEADER.pack(payload["n_func"], payload["n_prim"], payload["n_points"],
                            payload["n_atoms"], payload["max_prim"], 0, 0, 0))
        for k in ("fn_atom", "fn_nprim", "fn_off", "fn_lx", "fn_ly", "fn_lz", "fn_shell"):
            f.write(np.ascontiguousarray(payload[k], "<i4").tobytes())
        for k in ("fc", "pexp", "pcoef", "dens", "pts", "Z", "xyz"):
            f.write(np.ascontiguousarray(payload[k], "<f8").tobytes())


def read_basis_payload(path) -> dict:
    d = Path(path).read_bytes()
    if len(d) < HEADER.size:
        raise ValueError(f"{path}: truncated header")
    nf, npr, npt, nat, maxp = HEADER.unpack_from(d, 0)[:5]
    o = HEADER.size
    out = dict(n_func=nf, n_prim=npr, n_points=npt, n_atoms=nat, max_prim=maxp)

    def rd(dt, count):
        nonlocal o
        nb = np.dtype(dt).itemsize * count
        if o + nb > len(d):
            raise ValueError(f"{path}: truncated")
        a = np.frombuffer(d, dtype=dt, count=count, offset=o).copy()
        o += nb
        return a

    for k in ("fn_atom", "fn_nprim", "fn_off", "fn_lx", "fn_ly", "fn_lz", "fn_shell"):
        out[k] = rd("<i4", nf)
    out["fc"] = rd("<f8", 3 * nf)
    out["pexp"] = rd("<f8", npr)
    out["pcoef"] = rd("<f8", npr)
    out["dens"] = rd("<f8", nf * nf).reshape(nf, nf)
    out["pts"] = rd("<f8", 3 * npt)
    out["Z"] = rd("<f8", nat)
    out["xyz"] = rd("<f8", 3 * nat)
    return out
