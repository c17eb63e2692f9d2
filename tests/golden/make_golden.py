"""Write the golden fixtures of the GPU parity tests (run on the CPU).

    python tests/golden/make_golden.py [--only NAME]

Every fixture holds oracle outputs (``oracle/``: the NumPy restatement of the
reference's FP64 evaluation, pinned by tests/test_oracle_*.py) for seeded
synthetic inputs, plus the input hash so a test can prove it regenerated the
same inputs.  Inputs are regenerated from their seeds at test time, not stored.

Fixtures:
  boys.npz        F_m(T), m <= 8, on a T grid, from mpmath at 40 digits
  cfg0.npz        config 0 (12 atoms, 1000 points): V at tol 0 and 1e-12
  cfg1_300.npz    config 1 (1634 functions, 300 points): V at tol 0 and 1e-12
  cfg1_30k.npz    config 1 table screened against its 30000 points (tol 1e-12),
                  V at 256 sampled points
  cfg1_3k.npz     config 1 at 3000 points (tol 1e-12): V at every point (C++
                  restatement), pair counts and the sum rule
  cfg4.npz        8 conformers of the 256-conformer set (10-80 atoms, 517-19988
                  points): V at up to 256 sampled points (C++ restatement), the
                  field at 64 of them (NumPy oracle), pair counts, sum rules
  grad_toy.npz    field and dL/dP of the C+O toy (all classes L <= 6)
  cfg2_sub.npz    config-2 loss on a 400-point subset: V, field, dL/dG checksum,
                  dL/dP (1634 x 1634, float64, compressed)
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from fastesp_amd import synth  # noqa: E402
from oracle import engine, evaluate  # noqa: E402

OUT = Path(__file__).resolve().parent


def _esp_chunk(args):
    tab, pts = args
    return evaluate.esp(tab, pts)


def par_esp(tab, pts, workers=8):
    chunks = np.array_split(pts, max(1, min(workers * 4, len(pts))))
    with ProcessPoolExecutor(workers) as ex:
        return np.concatenate(list(ex.map(_esp_chunk, [(tab, c) for c in chunks])))


def make_boys():
    import mpmath
    mpmath.mp.dps = 40
    T = np.concatenate([[0.0, 1e-14, 1e-9, 1e-5], np.linspace(0.001, 70, 701),
                        np.geomspace(70, 1e7, 60),
                        np.arange(0, 60, 0.125) + 0.0625,  # grid midpoints (worst Taylor)
                        [59.999999, 60.0, 60.000001]])
    F = np.zeros((len(T), 9))
    for i, t in enumerate(T):
        for m in range(9):
            if t == 0:
                F[i, m] = 1.0 / (2 * m + 1)
            else:
                tt = mpmath.mpf(float(t))
                F[i, m] = float(mpmath.gammainc(m + mpmath.mpf(0.5), 0, tt)
                                / (2 * mpmath.power(tt, m + mpmath.mpf(0.5))))
    np.savez_compressed(OUT / "boys.npz", T=T, F=F)


def make_cfg0():
    w = synth.config0()
    prims = engine.build_prim_pairs(w.basis, w.density)
    out = dict(hash=w.input_hash())
    for name, tol in (("V_tol0", 0.0), ("V_tol12", 1e-12)):
        tab = engine.build_table(w.wfn, w.basis, w.density, w.points, tol, prims=prims)
        out[name] = par_esp(tab, w.points)
        out["npairs_" + name[2:]] = tab.n_pairs
    out["n_candidates"] = len(prims.p)
    out["sum_rule"] = prims.sum_rule
    out["fp"] = w.fingerprint()
    np.savez_compressed(OUT / "cfg0.npz", **out)


def make_cfg1_300():
    w = synth.config1(300)
    prims = engine.build_prim_pairs(w.basis, w.density)
    out = dict(hash=w.input_hash(), fp=w.fingerprint(), n_candidates=len(prims.p),
               sum_rule=prims.sum_rule)
    for name, tol in (("V_tol0", 0.0), ("V_tol12", 1e-12)):
        tab = engine.build_table(w.wfn, w.basis, w.density, w.points, tol, prims=prims)
        out[name] = par_esp(tab, w.points)
        out["npairs_" + name[2:]] = tab.n_pairs
    np.savez_compressed(OUT / "cfg1_300.npz", **out)


def make_cfg1_30k():
    w = synth.config1(30000)
    prims = engine.build_prim_pairs(w.basis, w.density)
    tab = engine.build_table(w.wfn, w.basis, w.density, w.points, 1e-12, prims=prims)
    rng = np.random.default_rng(77)
    idx = np.sort(rng.choice(len(w.points), 256, replace=False))
    V = par_esp(tab, w.points[idx])
    np.savez_compressed(OUT / "cfg1_30k.npz", hash=w.input_hash(), fp=w.fingerprint(), idx=idx, V=V,
                        npairs=tab.n_pairs, n_candidates=len(prims.p),
                        pairs_per_L=np.bincount(tab.pp_L, minlength=7))


def make_grad_toy():
    sys.path.insert(0, str(ROOT))
    from tests.helpers import cf_toy
    wfn, basis, P, _ = cf_toy()
    rng = np.random.default_rng(11)
    pts = rng.normal(0, 2.5, (64, 3)) + np.array([0.5, 0.4, -0.2])
    pts = pts[np.linalg.norm(pts - wfn.coords[0], axis=1) > 0.8]
    tab = engine.build_table(wfn, basis, P, pts, tol=0.0)
    V = evaluate.esp(tab, pts)
    w = rng.uniform(0.5, 1.0, len(pts))
    target = V + rng.normal(0, 1e-3, len(pts))
    _, gV = evaluate.loss_and_gV(V, target, w)
    dG = evaluate.grad_G(tab, pts, gV)
    dP = evaluate.grad_P(tab, basis, dG)
    np.savez_compressed(OUT / "grad_toy.npz", pts=pts, V=V, field=evaluate.field(tab, pts),
                        w=w, target=target, gV=gV, dP=dP)


def make_cfg2_sub():
    w = synth.config1(30000)
    rng = np.random.default_rng(2002)
    idx = np.sort(rng.choice(len(w.points), 400, replace=False))
    pts = w.points[idx]
    prims = engine.build_prim_pairs(w.basis, w.density)
    tab = engine.build_table(w.wfn, w.basis, w.density, pts, 1e-12, prims=prims)
    V = par_esp(tab, pts)
    wts = rng.uniform(0.5, 1.0, len(pts))
    target = V + rng.normal(0, 1e-3, len(pts))
    _, gV = evaluate.loss_and_gV(V, target, wts)
    fld = evaluate.field(tab, pts[:64])
    dG = evaluate.grad_G(tab, pts, gV)
    dP = evaluate.grad_P(tab, w.basis, dG)
    # a sample of dL/dP (the full 1634 x 1634 matrix would be ~20 MB): 4096
    # random entries, the diagonal and every row sum
    n = dP.shape[0]
    ij = np.stack([rng.integers(0, n, 4096), rng.integers(0, n, 4096)], axis=1)
    np.savez_compressed(OUT / "cfg2_sub.npz", hash=w.input_hash(), fp=w.fingerprint(), idx=idx,
                        V=V, w=wts, target=target, gV=gV, field64=fld, dP_ij=ij,
                        dP_sample=dP[ij[:, 0], ij[:, 1]], dP_diag=np.diag(dP).copy(),
                        dP_rowsum=dP.sum(axis=1), dP_absmax=np.abs(dP).max(),
                        npairs=tab.n_pairs)


def _port_table(ct, w):
    """The C++ port's screened table as an oracle table for ``evaluate``."""
    return ct.oracle_table(w.wfn, w.points)


def _field_chunk(args):
    tab, pts = args
    return evaluate.field(tab, pts), evaluate.esp(tab, pts)


def make_cfg3():
    """Config 3 (150 atoms, ~4500 functions, 200000 points): the NumPy oracle is
    too slow to screen ~1e6 pairs against 2e5 points, so the screen and V use the
    C++/OpenMP restatement (oracle/cpu_port.cpp, NumPy screen rule), which
    matches the NumPy oracle to 5e-14 a.u. on config 0 (test_cpu_port.py).  The
    field at the same 256 points is the NumPy oracle's analytic field over the
    port's table, whose NumPy V must agree with the port's V."""
    from oracle.cpu_port import CpuTable
    w = synth.config3()
    ct = CpuTable(w.wfn, w.basis, w.density, w.points, 1e-12, rule=0)
    rng = np.random.default_rng(303)
    idx = np.sort(rng.choice(len(w.points), 256, replace=False))
    V = ct.esp(w.points[idx])
    tab = _port_table(ct, w)
    chunks = np.array_split(w.points[idx], 8)
    with ProcessPoolExecutor(8) as ex:
        res = list(ex.map(_field_chunk, [(tab, c) for c in chunks]))
    field = np.concatenate([r[0] for r in res])
    Vnp = np.concatenate([r[1] for r in res])
    print(f"cfg3: NumPy V over the port's table vs the port: {np.abs(Vnp - V).max():.2e}")
    assert np.abs(Vnp - V).max() < 1e-11
    np.savez_compressed(OUT / "cfg3.npz", hash=w.input_hash(), fp=w.fingerprint(), idx=idx, V=V,
                        field=field, npairs=ct.n_pairs, n_candidates=ct.n_candidates,
                        sum_rule=ct.sum_rule, n_func=w.basis.n)


# conformers of the 256-set in the cfg4 fixture: the fewest and most atoms,
# the fewest and most points, and four between
CFG4_CONFORMERS = (0, 11, 52, 77, 117, 120, 132, 160)


def make_cfg1_3k():
    """Config 1 at 3000 points: V at every point from the C++ restatement
    (NumPy screen rule), checked against the NumPy oracle at 64 points."""
    from oracle.cpu_port import CpuTable
    w = synth.config1(3000)
    ct = CpuTable(w.wfn, w.basis, w.density, w.points, 1e-12, rule=0)
    V = ct.esp(w.points)
    tab = _port_table(ct, w)
    Vnp = par_esp(tab, w.points[:64])
    print(f"cfg1_3k: NumPy V over the port's table vs the port: {np.abs(Vnp - V[:64]).max():.2e}")
    assert np.abs(Vnp - V[:64]).max() < 1e-11
    np.savez_compressed(OUT / "cfg1_3k.npz", hash=w.input_hash(), fp=w.fingerprint(), V=V,
                        npairs=ct.n_pairs, n_candidates=ct.n_candidates, sum_rule=ct.sum_rule)


def make_cfg4():
    """Eight conformers of the batched set (synth.conformer_set(256), seed 1004),
    regenerated exactly as the set makes them (occupied factor, then P = 2 C C^T).
    V at up to 256 sampled points from the C++ restatement; the field at the
    first 64 of them from the NumPy oracle over the same table (whose V must
    agree with the restatement's)."""
    from oracle.cpu_port import CpuTable
    out = dict(conformers=np.array(CFG4_CONFORMERS))
    for c in CFG4_CONFORMERS:
        w = synth.conformer(c, 1004, orbitals=True)
        w.density = synth.density_from_orbitals(w.density)
        n = len(w.points)
        rng = np.random.default_rng([404, c])
        idx = np.sort(rng.choice(n, min(n, 256), replace=False))
        ct = CpuTable(w.wfn, w.basis, w.density, w.points, 1e-12, rule=0)
        V = ct.esp(w.points[idx])
        tab = _port_table(ct, w)
        chunks = np.array_split(w.points[idx[:64]], 8)
        with ProcessPoolExecutor(8) as ex:
            res = list(ex.map(_field_chunk, [(tab, ch) for ch in chunks]))
        field = np.concatenate([r[0] for r in res])
        Vnp = np.concatenate([r[1] for r in res])
        err = np.abs(Vnp - V[:64]).max()
        print(f"cfg4 conformer {c}: {w.meta['n_atoms']} atoms, {w.basis.n} functions, {n} points, "
              f"{ct.n_pairs} pairs; NumPy vs port V {err:.2e}", flush=True)
        assert err < 1e-11
        for k, v in dict(hash=np.frombuffer(w.input_hash().encode(), np.uint8), fp=w.fingerprint(),
                         idx=idx, V=V, field=field, npairs=ct.n_pairs,
                         n_candidates=ct.n_candidates, sum_rule=ct.sum_rule,
                         n_points=n).items():
            out[f"c{c}_{k}"] = v
    np.savez_compressed(OUT / "cfg4.npz", **out)


JOBS = dict(boys=make_boys, cfg0=make_cfg0, cfg1_300=make_cfg1_300, cfg1_30k=make_cfg1_30k,
            grad_toy=make_grad_toy, cfg2_sub=make_cfg2_sub, cfg3=make_cfg3, cfg1_3k=make_cfg1_3k,
            cfg4=make_cfg4)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=None)
    a = ap.parse_args()
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    for name, fn in JOBS.items():
        if a.only and name not in a.only:
            continue
        t = time.time()
        fn()
        print(f"{name}: {time.time() - t:.1f} s", flush=True)


if __name__ == "__main__":
    main()
