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
           "fe_sort.hip", "fe_batch.hip"]
LIB = PKG / "libfastesp_hip.so"


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.sep not in c or Path(c).exists()):
            return c
    raise RuntimeError("hipcc not found")


def build(arch: str = "gfx950", lmax: int = 6, verbose: bool = True) -> Path:
    subprocess.run([sys.executable, str(CSRC / "gen_lclass.py"), str(lmax)], check=True,
                   capture_output=not verbose)
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-munsafe-fp-atomics", "-Wno-unused-command-line-argument",
           "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{err[-6000:]}")
    if verbose and err.strip():
        print(err[-3000:])
    os.replace(tmp, LIB)
    return LIB


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default=os.environ.get("FASTESP_HIP_ARCH", "gfx950"))
    ap.add_argument("--lmax", type=int, default=6)
    a = ap.parse_args()
    print(build(a.arch, a.lmax))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
