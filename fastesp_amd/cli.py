"""Command-line drop-ins for the reference's native tools and its ESP driver.

The reference runs the path as two processes that exchange files
(``fastesp/cli.py:143-163``): ``build_pairs <basis.bin> <pairs.bin> [tol]``
(``fastesp/native/build_pairs.cpp:162-168``) and ``esp_gpu <pairs.bin>
<out.txt> [--repeats N] [--nchunk N] [--quiet]`` (``fastesp/native/
esp_gpu.cu:232-242``).  The same arguments, files and console protocol here:

    python -m fastesp_amd build-pairs basis.bin pairs.bin [tol]
    python -m fastesp_amd esp-gpu pairs.bin out.txt [--repeats N] [--quiet] [--precision mixed]
    python -m fastesp_amd esp basis.bin out.txt [--tol T] [--screen numpy|cpp] [--repeats N]
    python -m fastesp_amd esp WF.molden --points pts.txt [-o esp.txt] [--tol T] [--angstrom] [-q]
    python -m fastesp_amd build-native [--arch gfx950] [--force]
    python -m fastesp_amd selftest [-q]

* build-pairs builds the table on the GPU with the C++ tool's screen
  (``build_pairs.cpp:318-328``) and writes pairs.bin in reference order;
  stderr carries the tool's summary lines (shells, pairs, sum rule).
* esp-gpu evaluates a pairs.bin (from either builder), prints
  ``pairs=.. terms=.. points=.. atoms=.. max_mono=..``, one ``run i: X ms``
  line per repeat after a warm-up (the line ``fastesp/cli.py:166-168``
  parses), and writes ``N`` then `` x y z V`` lines in the reference's format
  (``esp_gpu.cu:318-323``).
* esp does both in one process (the table never leaves the device) from a
  basis.bin payload (``fastesp/table.py:268-316``), with the NumPy screen by
  default as ``build_table`` uses.
* esp with ``--points`` is the reference's own driver, ``fastesp esp
  WF.molden --points pts.txt -o esp.txt`` (``run_esp``, ``fastesp/cli.py:
  118-200,286-322``): MOLDEN reading, the pure -> Cartesian density, the
  point file (Bohr, or Angstrom with ``--angstrom``), the table build and the
  evaluation, the same report lines and the same output file format.

Errors print a message on stderr and return 1, as the native tools do.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib


def _write_out(path, pts: np.ndarray, V: np.ndarray) -> None:
    with Path(path).open("w") as f:
        f.write(f"{len(pts)}\n")
        for (x, y, z), v in zip(pts, V):
            f.write(f" {x:20.12f} {y:20.12f} {z:20.12f} {v:22.14e}\n")


def _timed_runs(table, pts: np.ndarray, precision: int, repeats: int, quiet: bool) -> np.ndarray:
    """Warm-up, then `repeats` timed evaluations with the points resident on
    the device (as esp_gpu times them: launch to synchronisation)."""
    import torch
    lib = table._ctx.lib
    n = len(pts)
    dev = torch.device("cuda", table._ctx.device)
    d_pts = torch.from_numpy(np.ascontiguousarray(pts)).to(dev)
    d_V = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def run():
        _lib.check(lib.fe_esp_forward(table._h, d_pts.data_ptr(), n, precision, _lib.FE_PART_ESP,
                                      _lib.FE_PTR_DEVICE, stream.cuda_stream, d_V.data_ptr()),
                   "fe_esp_forward")

    run()
    torch.cuda.synchronize(dev)
    for r in range(repeats):
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize(dev)
        if not quiet:
            print(f"run {r}: {(time.perf_counter() - t0) * 1e3:.3f} ms", flush=True)
    return d_V[:n].cpu().numpy()


def _precision(name: str) -> int:
    return {"fp64": _lib.FE_PREC_FP64, "mixed": _lib.FE_PREC_MIXED}[name]


def cmd_build_pairs(a) -> int:
    from .builder import SCREEN_CPP, build_table_payload
    from .formats import read_basis_payload
    pl = read_basis_payload(a.basis)
    n_shells = len(np.unique(pl["fn_shell"]))
    print(f"shells {n_shells}  functions {pl['n_func']}", file=sys.stderr)
    t = build_table_payload(pl, tol=a.tol, rule=SCREEN_CPP)
    flat = t.to_flat()
    t.write(a.pairs)
    print(f"pairs {flat.n_pairs} terms {flat.n_terms} mono {len(flat.mo_parent)}", file=sys.stderr)
    print(f"sum rule Tr(P S) = {t.stats()['sum_rule']:.9f}  (must equal the MOLDEN electron count)",
          file=sys.stderr)
    return 0


def cmd_esp_gpu(a) -> int:
    from .formats import read_pairs
    from .table import PairTable
    flat = read_pairs(a.pairs)
    if not a.quiet:
        print(f"pairs={flat.n_pairs} terms={flat.n_terms} points={len(flat.pts)} "
              f"atoms={len(flat.Z)} max_mono={flat.max_mono}", flush=True)
    t = PairTable.from_flat(flat)
    V = _timed_runs(t, flat.pts, _precision(a.precision), a.repeats, a.quiet)
    _write_out(a.out, flat.pts, V)
    if not a.quiet:
        print(f"wrote {a.out}")
    return 0


def cmd_esp(a) -> int:
    from .builder import SCREEN_CPP, SCREEN_NUMPY, build_table_payload
    from .formats import read_basis_payload
    pl = read_basis_payload(a.basis)
    pts = np.asarray(pl["pts"], float).reshape(-1, 3)
    t0 = time.perf_counter()
    t = build_table_payload(pl, tol=a.tol, rule=SCREEN_CPP if a.screen == "cpp" else SCREEN_NUMPY)
    t_build = time.perf_counter() - t0
    if not a.quiet:
        print(f"  backend      : hip", flush=True)
        print(f"  basis        : {pl['n_func']} Cartesian functions, {pl['n_prim']} primitives")
        print(f"  pair table   : {t_build:8.4f} s  ({t.n_pairs} pairs, point-independent)")
    V = _timed_runs(t, pts, _precision(a.precision), a.repeats, a.quiet)
    if not a.quiet:
        print(f"  ESP range    : {V.min() * 27.211386245988:+.4f} .. {V.max() * 27.211386245988:+.4f} V")
    _write_out(a.out, pts, V)
    if not a.quiet:
        print(f"  wrote        : {a.out}")
    return 0


def read_points(path, geom: str = "bohr") -> np.ndarray:
    """Rows of the first three numbers of every parsable line, Angstrom
    converted to Bohr (``fastesp/cli.py:211-227``)."""
    rows = []
    for line in Path(path).read_text().splitlines():
        f = line.split()
        if len(f) >= 3:
            try:
                rows.append([float(x) for x in f[:3]])
            except ValueError:
                continue
    if not rows:
        raise ValueError(f"cannot parse points from {path}")
    pts = np.asarray(rows, float)
    if geom == "angstrom":
        from .basis import BOHR
        pts = pts / BOHR
    return pts


def run_esp(molden, points_file, out_file=None, backend: str = "auto", tol: float = 1e-12,
            geom: str = "bohr", verbose: bool = True, precision: str = "fp64") -> np.ndarray:
    """``fastesp.cli.run_esp`` (``fastesp/cli.py:118-200``) on the HIP backend."""
    if backend not in ("auto", "hip"):
        raise ValueError(f"backend {backend!r}: this package provides the 'hip' backend only")
    from .basis import Basis
    from .builder import SCREEN_NUMPY, build_table_payload
    from .formats import basis_payload
    from .molden import read_molden
    t0 = time.perf_counter()
    wfn = read_molden(molden)
    basis = Basis(wfn)
    density = basis.density_matrix_cartesian()
    t_load = time.perf_counter() - t0
    pts = read_points(points_file, geom)
    t0 = time.perf_counter()
    table = build_table_payload(basis_payload(wfn, basis, density, pts), tol, SCREEN_NUMPY)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    V = table.esp(pts, precision=_precision(precision))
    t_solve = time.perf_counter() - t0
    if verbose:
        print("  backend      : hip")
        print(f"  basis        : {basis.n} Cartesian functions, {basis.n_prim} primitives")
        print(f"  load         : {t_load:8.3f} s")
        print(f"  pair table   : {t_build:8.4f} s  (point-independent)")
        print(f"  ESP solve    : {t_solve:8.4f} s")
        tot = t_build + t_solve
        print(f"  total (excl. load): {tot:.4f} s for {len(pts)} points "
              f"({1e3 * tot / len(pts):.4f} ms/point)")
        print(f"  ESP range    : {V.min() * 27.211386245988:+.4f} .. "
              f"{V.max() * 27.211386245988:+.4f} V")
    if out_file:
        _write_out(out_file, pts, V)
        if verbose:
            print(f"  wrote        : {out_file}")
    return V


def cmd_esp_any(a) -> int:
    if a.points is not None:
        out = a.out_opt or a.out
        run_esp(a.input, a.points, out, backend=a.backend, tol=a.tol,
                geom="angstrom" if a.angstrom else "bohr", verbose=not a.quiet,
                precision=a.precision)
        return 0
    if a.out is None:
        raise ValueError("esp: basis.bin input needs an output file (or give --points for MOLDEN)")
    a.basis = a.input
    return cmd_esp(a)


def has_gpu() -> bool:
    """A HIP device this library can open (the reference asks ``nvidia-smi -L``,
    ``fastesp/cli.py:203-208``): ``/dev/kfd`` present and ``fe_init`` succeeds."""
    if not Path("/dev/kfd").exists():
        return False
    try:
        _lib.Context.get()
        return True
    except _lib.HipBackendError:
        return False


def cmd_build_native(a) -> int:
    from .build import build_native
    path = build_native(force=a.force, arch=a.arch)
    print(f"  libfastesp_hip: {path}")
    print(f"  HIP device     : {'yes' if has_gpu() else 'no'}")
    return 0


def cmd_selftest(a) -> int:
    from .selftest import run
    return run(verbose=not a.quiet)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m fastesp_amd", description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build-pairs", help="basis.bin -> pairs.bin on the GPU (build_pairs)")
    b.add_argument("basis")
    b.add_argument("pairs")
    b.add_argument("tol", nargs="?", type=float, default=1e-12)
    b.set_defaults(fn=cmd_build_pairs)
    e = sub.add_parser("esp-gpu", help="pairs.bin -> out.txt (esp_gpu)")
    e.add_argument("pairs")
    e.add_argument("out")
    e.add_argument("--repeats", type=int, default=1)
    e.add_argument("--nchunk", type=int, default=0, help="accepted for compatibility; the "
                   "launch shape comes from the autotuned context")
    e.add_argument("--quiet", action="store_true")
    e.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    e.set_defaults(fn=cmd_esp_gpu)
    s = sub.add_parser("esp", help="basis.bin -> out.txt, or WF.molden --points pts.txt")
    s.add_argument("input", help="basis.bin payload, or a MOLDEN file when --points is given")
    s.add_argument("out", nargs="?", default=None)
    s.add_argument("--points", default=None,
                   help="text file, one 'x y z' per line (Bohr unless --angstrom): MOLDEN mode")
    s.add_argument("-o", "--out", dest="out_opt", default=None)
    s.add_argument("--backend", choices=["auto", "hip"], default="auto")
    s.add_argument("--angstrom", action="store_true", help="input points in Angstrom")
    s.add_argument("--tol", type=float, default=1e-12,
                   help="primitive pair screening threshold in Hartree/e")
    s.add_argument("--screen", choices=["numpy", "cpp"], default="numpy")
    s.add_argument("--repeats", type=int, default=1)
    s.add_argument("-q", "--quiet", action="store_true")
    s.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    s.set_defaults(fn=cmd_esp_any)
    n = sub.add_parser("build-native", help="compile libfastesp_hip.so if its sources changed")
    n.add_argument("--arch", default=None, help="offload target, default gfx950")
    n.add_argument("--force", action="store_true")
    n.set_defaults(fn=cmd_build_native)
    t = sub.add_parser("selftest", help="check the device integrals against closed forms "
                                        "and quadrature")
    t.add_argument("-q", "--quiet", action="store_true")
    t.set_defaults(fn=cmd_selftest)
    a = ap.parse_args(argv)
    try:
        return a.fn(a)
    except (OSError, ValueError, IndexError, KeyError, _lib.HipBackendError) as ex:
        print(f"fastesp_amd {a.cmd}: {ex}", file=sys.stderr)
        return 1
