"""Build ``libfastesp_hip.so`` in-tree for gfx950.

    python -m fastesp_amd.build [--arch gfx950] [--lmax 6]

Steps: regenerate the per-L code (``csrc/gen_lclass.py``), compile every HIP
source to an object in parallel (``_build/``), then link the shared library.
Packed-FP32 instructions are disabled for the device code (``-Xclang
-target-feature -Xclang -packed-fp32-ops``).

A second library, ``libfastesp_probe.so`` (``csrc/probe/``), holds the
benchmark's VALU-rate calibration kernel.  It is not part of the product ABI:
only ``bench.py`` loads it.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
SOURCES = ["fe_core.hip", "fe_eval.hip", "fe_mixed.hip", "fe_backward.hip", "fe_build.hip",
           "fe_sort.hip", "fe_batch.hip", "fe_density.hip", "fe_noalloc.hip", "fe_prim.hip"]
LIB = PKG / "libfastesp_hip.so"
PROBE_SOURCES = ["probe/fe_probe.hip"]
PROBE_LIB = PKG / "libfastesp_probe.so"
OBJ_DIR = PKG / "_build"


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.sep not in c or Path(c).exists()):
            return c
    raise RuntimeError("hipcc not found")


STAMP = LIB.with_name(LIB.name + ".stamp")
FLAGS = ["-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-Wno-unused-command-line-argument"]


def source_hash(arch: str = "gfx950", lmax: int = 6) -> str:
    """Key of a build: every source and header under csrc/, the flags, the
    target and LMAX (the cache key of ``fastesp/cli.py:51-67``: sources + arch)."""
    import hashlib
    h = hashlib.sha256()
    for p in sorted(CSRC.rglob("*")):
        if p.suffix in (".hip", ".h", ".py") and p.is_file():
            h.update(str(p.relative_to(CSRC)).encode())
            h.update(p.read_bytes())
    h.update(" ".join(FLAGS + [arch, str(lmax)]).encode())
    return h.hexdigest()[:16]


def _compile(cmd: list, verbose: bool) -> str:
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed: {' '.join(cmd)}\n{err[-6000:]}")
    return err


def _link(objs: list, dst: Path, arch: str, verbose: bool) -> Path:
    tmp = dst.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={arch}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    _compile(cmd, verbose)
    os.replace(tmp, dst)
    return dst


def build(arch: str = "gfx950", lmax: int = 6, verbose: bool = True, out: Path | None = None,
          defines: tuple = (), probe: bool = True) -> Path:
    """Compile the library (to `out`, default the in-tree LIB; `defines` are
    extra -D options, e.g. register budgets for a variant build).  Every
    source is compiled to its own object in parallel, then linked."""
    subprocess.run([sys.executable, str(CSRC / "gen_lclass.py"), str(lmax)], check=True,
                   capture_output=not verbose)
    dst = Path(out) if out is not None else LIB
    tag = dst.stem + ("_" + "_".join(d.replace("=", "-").replace(",", ".") for d in defines)
                      if defines else "")
    odir = OBJ_DIR / tag
    odir.mkdir(parents=True, exist_ok=True)
    jobs = []
    for s in SOURCES + (PROBE_SOURCES if probe and out is None else []):
        src = CSRC / s
        obj = odir / (Path(s).stem + ".o")
        cmd = [hipcc(), f"--offload-arch={arch}", *FLAGS, *(f"-D{d}" for d in defines), "-c",
               "-o", str(obj), str(src)]
        jobs.append((s, obj, cmd))
    if verbose:
        for _, _, cmd in jobs:
            print(" ".join(cmd), flush=True)
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0") or 0) or os.cpu_count() or 4, 16))
    with ThreadPoolExecutor(workers) as ex:
        errs = list(ex.map(lambda j: _compile(j[2], verbose), jobs))
    if verbose:
        for e in errs:
            if e.strip():
                print(e[-3000:])
    lib_objs = [o for s, o, _ in jobs if s in SOURCES]
    _link(lib_objs, dst, arch, verbose)
    probe_objs = [o for s, o, _ in jobs if s in PROBE_SOURCES]
    if probe_objs:
        _link(probe_objs, PROBE_LIB, arch, verbose)
    if out is None and not defines:
        STAMP.write_text(source_hash(arch, lmax) + "\n")
    return dst


def build_native(force: bool = False, arch: str | None = None, quiet: bool = False) -> Path:
    """``fastesp build-native`` (``fastesp/cli.py:81-115``): compile only when the
    library is missing or was built from other sources, flags or target."""
    arch = arch or os.environ.get("FASTESP_HIP_ARCH", "gfx950")
    key = source_hash(arch)
    if not force and LIB.exists() and STAMP.exists() and STAMP.read_text().strip() == key:
        if not quiet:
            print(f"  up to date ({key}, {arch})")
        return LIB
    return build(arch, verbose=not quiet)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default=os.environ.get("FASTESP_HIP_ARCH", "gfx950"))
    ap.add_argument("--lmax", type=int, default=6)
    a = ap.parse_args()
    print(build(a.arch, a.lmax))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
