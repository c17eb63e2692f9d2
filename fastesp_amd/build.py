"""Build ``libfastesp_hip.so`` in-tree for gfx950.

    python -m fastesp_amd.build [--arch gfx950] [--lmax 6]

Steps: regenerate the per-L code (``csrc/gen_lclass.py``), then one ``hipcc``
invocation over the HIP sources.  Packed-FP32 instructions are disabled for the
device code (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
SOURCES = ["fe_core.hip", "fe_eval.hip", "fe_mixed.hip", "fe_backward.hip", "fe_build.hip",
           "fe_sort.hip", "fe_batch.hip", "fe_density.hip", "fe_probe.hip"]
LIB = PKG / "libfastesp_hip.so"


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.sep not in c or Path(c).exists()):
            return c
    raise RuntimeError("hipcc not found")


STAMP = LIB.with_name(LIB.name + ".stamp")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-Wno-unused-command-line-argument"]


def source_hash(arch: str = "gfx950", lmax: int = 6) -> str:
    """Key of a build: every source and header under csrc/, the flags, the
    target and LMAX (the cache key of ``fastesp/cli.py:51-67``: sources + arch)."""
    import hashlib
    h = hashlib.sha256()
    for p in sorted(CSRC.iterdir()):
        if p.suffix in (".hip", ".h", ".py") and p.is_file():
            h.update(p.name.encode())
            h.update(p.read_bytes())
    h.update(" ".join(FLAGS + [arch, str(lmax)]).encode())
    return h.hexdigest()[:16]


def build(arch: str = "gfx950", lmax: int = 6, verbose: bool = True, out: Path | None = None,
          defines: tuple = ()) -> Path:
    """Compile the library (to `out`, default the in-tree LIB; `defines` are
    extra -D options, e.g. register budgets for a variant build)."""
    subprocess.run([sys.executable, str(CSRC / "gen_lclass.py"), str(lmax)], check=True,
                   capture_output=not verbose)
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    dst = Path(out) if out is not None else LIB
    tmp = dst.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={arch}", *FLAGS, *(f"-D{d}" for d in defines),
           "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{err[-6000:]}")
    if verbose and err.strip():
        print(err[-3000:])
    os.replace(tmp, dst)
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
