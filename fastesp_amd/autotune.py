"""Autotuner for the evaluation launch shape.

Knobs (``fe_set_tuning``):
  keys 10..16  points per lane of class L = key-10 (1, 2 or 4): more points per
               lane reuse each scalar-loaded pair record more often and hide
               its load latency, at the cost of VGPRs (occupancy);
  keys 20..26  under mixed precision, whether class L = key-20 takes the FP32
               contraction or the FP64 kernel (FP32 only pays where the
               contraction outweighs the conversion and routing work);
  key 1        target workgroups per class launch (pair-chunk count = target /
               point blocks): parallelism vs partial-row traffic.

``tune(table, points, ...)`` times candidates with HIP events on the current
stream (device-resident inputs): first the workgroup target, then the points
per lane of every class by coordinate descent.  The best shape is applied to
the context and recorded in a JSON cache keyed by the workload shape (pairs
per class and point count, rounded to powers of two), so later runs of the same
shape start tuned (``apply_cached``).
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from . import _lib

PPL_CHOICES = (1, 2, 4)
WG_CHOICES = (2048, 4096, 8192)
NCLS = _lib.FE_LMAX + 1
CACHE = Path(os.environ.get("FASTESP_TUNE_CACHE",
                            Path.home() / ".cache" / "fastesp_amd" / "tuning.json"))


def shape_key(pairs_per_L, n_points: int, op: str, precision: int) -> str:
    b = lambda x: int(2 ** round(np.log2(max(int(x), 1))))  # noqa: E731
    return f"{op}:p{precision}:n{b(n_points)}:" + ",".join(str(b(x) if x else 0)
                                                          for x in pairs_per_L)


def _load() -> dict:
    try:
        return json.loads(CACHE.read_text())
    except Exception:
        return {}


def _store(d: dict) -> None:
    try:
        CACHE.parent.mkdir(parents=True, exist_ok=True)
        CACHE.write_text(json.dumps(d, indent=1, sort_keys=True))
    except OSError:
        pass


DEFAULT_MIXED_L = (0, 0, 0, 1, 1, 1, 1)


def apply(ctx: _lib.Context, ppl_L, target_wg: int, mixed_L=None) -> None:
    for L, v in enumerate(ppl_L):
        ctx.set_tuning(10 + L, int(v))
    ctx.set_tuning(1, int(target_wg))
    if mixed_L is not None:
        for L, m in enumerate(mixed_L):
            ctx.set_tuning(20 + L, int(m))


def apply_cached(ctx, table, n_points: int, op: str = "fwdbwd",
                 precision: int = _lib.FE_PREC_FP64) -> bool:
    v = _load().get(shape_key(table.info()["pairs_per_L"], n_points, op, precision))
    if v:
        apply(ctx, v["ppl_L"], v["target_wg"], v.get("mixed_L"))
        return True
    return False


def tune(table, points_dev, op: str = "fwdbwd", precision: int = _lib.FE_PREC_FP64,
         reps: int = 3, verbose: bool = False, fix_mixed: bool = False) -> dict:
    """Time launch shapes on device-resident points (a torch float64 tensor).
    fix_mixed keeps the context's per-class FP32 choice instead of tuning it."""
    import torch
    ctx = table._ctx
    lib = ctx.lib
    n = int(points_dev.shape[0])
    dev = points_dev.device
    V = torch.empty(n, dtype=torch.float64, device=dev)
    G = torch.empty((n, 3), dtype=torch.float64, device=dev)
    gV = torch.ones(n, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    present = [L for L, c in enumerate(table.info()["pairs_per_L"]) if c]

    def run():
        if op == "fwd":
            _lib.check(lib.fe_esp_forward(table._h, points_dev.data_ptr(), n, precision,
                                          _lib.FE_PART_ESP, _lib.FE_PTR_DEVICE, sh, V.data_ptr()),
                       "fe_esp_forward")
        else:
            _lib.check(lib.fe_esp_forward_grad(table._h, points_dev.data_ptr(), n, precision,
                                               _lib.FE_PTR_DEVICE, sh, V.data_ptr(),
                                               G.data_ptr()), "fe_esp_forward_grad")
            _lib.check(lib.fe_esp_backward(table._h, points_dev.data_ptr(), n, gV.data_ptr(),
                                           precision, _lib.FE_PTR_DEVICE, sh, None, None),
                       "fe_esp_backward")

    def measure(ppl_L, wg, mixed_L=None):
        apply(ctx, ppl_L, wg, mixed_L)
        run()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        if verbose:
            print(f"  tune {op} ppl={list(ppl_L)} wg={wg}: {ms:.3f} ms", flush=True)
        return ms

    log = []
    ppl = [2] * NCLS
    mixed = precision == _lib.FE_PREC_MIXED and not fix_mixed
    mix = [1] * NCLS if fix_mixed else list(DEFAULT_MIXED_L)
    best_wg, best_ms = None, float("inf")
    for wg in WG_CHOICES:
        ms = measure(ppl, wg, mix)
        log.append(dict(ppl_L=list(ppl), target_wg=wg, ms=ms))
        if ms < best_ms:
            best_wg, best_ms = wg, ms
    for L in present:
        for m in ((0, 1) if mixed else (mix[L],)):
            for v in PPL_CHOICES:
                if v == ppl[L] and m == mix[L]:
                    continue
                trial, tmix = list(ppl), list(mix)
                trial[L], tmix[L] = v, m
                ms = measure(trial, best_wg, tmix)
                log.append(dict(ppl_L=trial, mixed_L=tmix, target_wg=best_wg, ms=ms))
                if ms < best_ms:
                    ppl, mix, best_ms = trial, tmix, ms
    apply(ctx, ppl, best_wg, mix)
    best = dict(ppl_L=ppl, target_wg=best_wg, ms=best_ms,
                mixed_L=mix if precision == _lib.FE_PREC_MIXED else None)
    k = shape_key(table.info()["pairs_per_L"], n, op, precision)
    d = _load()
    d[k] = dict(best, trials=len(log))
    _store(d)
    return dict(best=best, all=log, key=k)


# ---------------------------------------------------------------------------
# batched conformers
# ---------------------------------------------------------------------------
ITEMS_CHOICES = (8, 16, 32, 64, 128)
DG_ITEMS_CHOICES = (8, 16, 32, 64, 128)


def tune_batch(batch, precision="mixed", field: bool = True, dldp: bool = True, reps: int = 2,
               verbose: bool = False) -> dict:
    """Tile shape of a batch (``fe_batch_set_tuning``), timed with HIP events.

    1. points per lane of each class, timing that class's persistent launch
       alone (class mask), since the per-class launches are independent;
    2. forward work items per resident workgroup (item size: load balance
       against partial-row traffic), on the whole forward;
    3. dL/dG work items per resident workgroup, on the backward.
    The batch keeps the best shape; the graph must be captured afterwards.
    """
    import torch
    from .batch import _prec
    dev = torch.device("cuda", batch._ctx.device)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    prec = _prec(precision)
    present = [L for L in range(NCLS) if any(t.info()["pairs_per_L"][L] for t in batch.tables)]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    fwd = lambda: batch.forward(prec, field=field, stream=sh)  # noqa: E731
    log = []
    ppl, mix = {}, {}
    for L in present:
        batch.set_tuning(_lib.FE_BT_CLASS_MASK, 1 << L)
        best = (float("inf"), 2, DEFAULT_MIXED_L[L])
        for m in ((0, 1) if prec == _lib.FE_PREC_MIXED else (DEFAULT_MIXED_L[L],)):
            batch.set_tuning(_lib.FE_BT_MIXED0 + L, m)
            for v in PPL_CHOICES:
                batch.set_tuning(_lib.FE_BT_PPL0 + L, v)
                ms = timed(fwd)
                log.append(dict(cls=L, ppl=v, mixed=m, ms=ms))
                if verbose:
                    print(f"  tune batch class {L} ppl={v} fp32={m}: {ms:.3f} ms", flush=True)
                best = min(best, (ms, v, m))
        ppl[L], mix[L] = best[1], best[2]
        batch.set_tuning(_lib.FE_BT_PPL0 + L, best[1])
        batch.set_tuning(_lib.FE_BT_MIXED0 + L, best[2])
    batch.set_tuning(_lib.FE_BT_CLASS_MASK, (1 << NCLS) - 1)
    best_k = (float("inf"), 8)
    for k in ITEMS_CHOICES:
        batch.set_tuning(_lib.FE_BT_ITEMS_PER_WG, k)
        ms = timed(fwd)
        log.append(dict(items_per_wg=k, ms=ms))
        if verbose:
            print(f"  tune batch items/WG={k}: {ms:.3f} ms", flush=True)
        best_k = min(best_k, (ms, k))
    batch.set_tuning(_lib.FE_BT_ITEMS_PER_WG, best_k[1])
    best_kd = (None, None)
    if dldp:
        batch.forward(prec, field=field, stream=sh)
        batch.loss(stream=sh)
        best_kd = (float("inf"), 8)
        for k in DG_ITEMS_CHOICES:
            batch.set_tuning(_lib.FE_BT_DG_ITEMS_PER_WG, k)
            ms = timed(lambda: batch.backward(stream=sh))
            log.append(dict(dg_items_per_wg=k, ms=ms))
            if verbose:
                print(f"  tune batch dG items/WG={k}: {ms:.3f} ms", flush=True)
            best_kd = min(best_kd, (ms, k))
        batch.set_tuning(_lib.FE_BT_DG_ITEMS_PER_WG, best_kd[1])
    return dict(ppl_L=[ppl.get(L, 2) for L in range(NCLS)],
                mixed_L=[mix.get(L, DEFAULT_MIXED_L[L]) for L in range(NCLS)]
                if prec == _lib.FE_PREC_MIXED else None, items_per_wg=best_k[1],
                fwd_ms=best_k[0], dg_items_per_wg=best_kd[1], bwd_ms=best_kd[0],
                trials=len(log), all=log)
