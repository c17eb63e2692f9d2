#!/usr/bin/env python3
"""ESP hot-path benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--points 30000]
                    [--mode fwd|fwdbwd] [--precision fp64|mixed] [--no-cpu]

Workload: the 1634-function protein-fragment-sized synthetic molecule (54
atoms, C/N/O/H + 2 S, TZ-like s/p/d/f basis, seed 1001) with 30000 points on
its 1.4 x vdW surface.  The pair table is built on the GPU (timed and reported
separately); a step is one pass of forward + field + loss + backward to dL/dP
over all 30000 points (the config-2 loss) with the table and the points
already resident in HBM.  `value` is ms per 1e4 points of the FP64 step
(lower is better); the mixed-precision step is reported beside it.

Printed JSON (one line): the driver's contract fields plus
  roofline      -- the dominant launch of the headline step: its performed
                   FLOPs (each (point, pair) priced by the path it took, counted
                   by fe_eval_work; fastesp_amd.flops) over its HIP-event time,
                   against the FP64 vector peak; the model and executed
                   (hardware counter) figures beside it;
  roofline_step -- the same for the whole step;
  cpu_baseline  -- the NumPy restatement of the reference evaluation plus its
                   adjoint (forward + field + dL/dG), one core, on a fixed
                   point sample scaled to 1e4 points; cpu_baselines holds the
                   NumPy forward, the C++/OpenMP port on every core and the
                   reference's own build_pairs;
  build_plus_fwd -- table build + FP64 forward at 300 / 3000 / 30000 points,
                   the reference's published measurement;
  autograd      -- the differentiable step through ESPModel (rebuild in place,
                   forward + field, autograd backward) and the same step as a
                   recorded graph;
  batched, config3, density -- the 256-conformer graph-captured step, the
                   200000-point configuration, the electron density.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# the library's default points per lane of the value kernels and of the value
# + field kernels (fe_ctx in fastesp_amd/csrc/fe_internal.h), and its default
# target workgroups per launch (32 per compute unit)
LIB_PPL_FWD = [2, 2, 2, 2, 2, 2, 4]
LIB_PPL_GRAD = [2, 2, 2, 2, 2, 1, 1]

# published build + FP64 forward of the reference (RTX 5090 + 16 CPU cores),
# ms per 1e4 points (benchmarks/README.md:50-52)
REFERENCE_BUILD_PLUS_FWD = {300: 8959.0, 3000: 1598.0, 30000: 819.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=30000)
    ap.add_argument("--mode", choices=["fwd", "fwdbwd"], default="fwdbwd")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the C++/OpenMP port's evaluation sample")
    ap.add_argument("--numpy-points", type=int, default=300,
                    help="fixed point sample of the NumPy forward baseline")
    ap.add_argument("--numpy-adjoint-points", type=int, default=20,
                    help="fixed point sample of the NumPy forward + adjoint baseline")
    ap.add_argument("--ppl", type=int, default=0, help="points per lane (0 = library default)")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="skip the launch-shape autotuner (runs before the timed region)")
    ap.add_argument("--ppl-classes", default="",
                    help="explicit points per lane per class, e.g. 4,2,2,2,2,2,1 (implies --no-tune)")
    ap.add_argument("--wg", type=int, default=0, help="explicit target workgroups per launch")
    ap.add_argument("--fp32-classes", default="",
                    help="explicit 0/1 per class: FP32 contraction under mixed, e.g. 0,0,0,1,1,1,1")
    ap.add_argument("--legs", default="fwd_fp64,fwdbwd_fp64,fwdbwd_mixed,fwdbwd_mixed32",
                    help="comma list of mode_precision legs to time (mixed32: every class "
                         "through the FP32 contraction)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--batch", type=int, default=256,
                    help="conformers of the batched leg (0 = skip it)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-warmup", type=int, default=1)
    ap.add_argument("--batch-precisions", default="mixed,fp64")
    ap.add_argument("--density-steps", type=int, default=10,
                    help="timed evaluations of the electron-density leg (0 = skip it)")
    ap.add_argument("--config3-steps", type=int, default=3,
                    help="timed fwd+bwd steps of the config-3 leg, FP64 and mixed (0 = skip it)")
    ap.add_argument("--build-legs", default="300,3000,30000",
                    help="point counts of the build + FP64 forward legs ('' = skip)")
    ap.add_argument("--autograd-steps", type=int, default=10,
                    help="timed steps of the autograd / graph-captured leg (0 = skip it)")
    ap.add_argument("--eval-streams", type=int, choices=[0, 1, 2], default=0,
                    help="streams of the evaluation class launches (0 = library default: the "
                         "value kernels in two halves, with the field classes 5 and 6 beside "
                         "the others; 1 one stream; 2 two halves)")
    ap.add_argument("--dg-streams", type=int, choices=[1, 2, 3], default=2,
                    help="streams of the dL/dG class launches (2 = library default; 1 runs "
                         "them one at a time, so per-kernel counters and times are their own)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# phase ids of fe_profile
PHASE_NAMES = {0: "sort", 1: "reduce", 2: "pack_gV", 3: "dG_to_dP", 4: "density_phi",
               5: "density_mm", 6: "density_reduce"}
PHASE_NAMES.update({8 + L: f"eval_L{L}" for L in range(7)})
PHASE_NAMES.update({16 + L: f"dG_L{L}" for L in range(7)})
PHASE_NAMES.update({24 + L: f"dG_reduce_L{L}" for L in range(7)})


def work_counts(lib, tab, pts_ptr, n, grad, sh) -> dict:
    """fe_eval_work: far-field / tabulated (point, pair) evaluations per class of
    the evaluation and dL/dG launches, with the context's current tuning."""
    from fastesp_amd import _lib
    c = (ctypes.c_int64 * 28)()
    _lib.check(lib.fe_eval_work(tab._h, pts_ptr, n, int(grad), _lib.FE_PTR_DEVICE, sh, c),
               "fe_eval_work")
    v = list(c)
    return dict(eval_far=v[0:7], eval_near=v[7:14], dg_far=v[14:21], dg_near=v[21:28])


def launch_flops(counts: dict, grad: bool) -> dict:
    """Performed FLOPs of every class launch: {phase id: flops}."""
    from fastesp_amd import flops
    out = {}
    kind = "fwd_grad" if grad else "fwd"
    for L in range(7):
        out[8 + L] = flops.performed_flops(kind, L, counts["eval_far"][L], counts["eval_near"][L])
        if grad:
            out[16 + L] = flops.performed_flops("dg", L, counts["dg_far"][L], counts["dg_near"][L])
    return out


def profile_values(path: Path, kernel: str) -> dict:
    """The key = value lines of a committed rocprofv3 summary that follow its
    `kernel = <kernel>` line (a summary holds one such section per launch it
    covers), or {} if it does not cover `kernel`."""
    if not path.exists():
        return {}
    sections, cur = {}, None
    for line in path.read_text().splitlines():
        if "=" not in line or line.startswith(("#", "`", " ", "|")):
            continue
        k, v = (s.strip() for s in line.split("=", 1))
        if k == "kernel":
            cur = sections.setdefault(v, {"kernel": v})
        elif cur is not None:
            cur[k] = v
    return sections.get(kernel, {})


def roofline(phases, counts, grad, ppl_L, fp32_L, prec, pairs_per_L, n, traffic_file,
             counters_file, overlapped=frozenset()) -> dict:
    """The launch with the most time in the timed region: its performed FLOPs
    per launch over its own average event-timed duration, against the FP64
    vector peak (FP32 classes issue non-packed FP32, at the FP64 rate).
    Launches in `overlapped` run beside another launch on a second stream, so
    their event span is not their own duration; they are not candidates."""
    from fastesp_amd import flops
    fl = launch_flops(counts, grad)
    best = None
    for pid, (ms, cnt) in phases.items():
        if pid in fl and pid not in overlapped and (best is None or ms > best[1]):
            best = (pid, ms, cnt)
    if best is None:
        return None
    pid, ms, cnt = best
    avg_ms = ms / cnt
    if pid < 16:
        L = pid - 8
        fp32 = prec == "mixed" and bool(fp32_L[L])
        name = f"k_eval<{L}, {ppl_L[L]}, {str(grad).lower()}, {str(fp32).lower()}>"
        far, near = counts["eval_far"][L], counts["eval_near"][L]
        model = flops.kernel_flops("fwd_grad" if grad else "fwd", L, pairs_per_L[L], n)
    else:
        L = pid - 16
        name = f"k_dG<{L}>"
        far, near = counts["dg_far"][L], counts["dg_near"][L]
        model = flops.kernel_flops("dg", L, pairs_per_L[L], n)
    achieved = fl[pid] / (avg_ms * 1e-3)
    peak = flops.PEAK_FP64_VALU
    tv = profile_values(traffic_file, name)
    traffic = int(float(tv["traffic_bytes_per_launch"])) if "traffic_bytes_per_launch" in tv else None
    out = dict(bound="valu", achieved=round(achieved / 1e12, 3), peak=round(peak / 1e12, 2),
               unit="TFLOP/s", frac=round(achieved / peak, 4), traffic=traffic, kernel=name,
               launches=cnt, avg_launch_ms=round(avg_ms, 4), flops_per_launch=fl[pid],
               far_evaluations=int(far), tabulated_evaluations=int(near),
               flop_model="performed work: far and tabulated (point, pair) evaluations "
                          "(fe_eval_work) x fastesp_amd.flops.f_eval_path / f_dg_path",
               model_flops_per_launch=model,
               model_frac=round(model / (avg_ms * 1e-3) / peak, 4))
    cv = profile_values(counters_file, name)
    if "executed_f64_flops_per_launch" in cv:
        ex = float(cv["executed_f64_flops_per_launch"]) + float(cv.get("executed_f32_flops_per_launch", 0))
        out.update(executed_flops_per_launch=ex,
                   executed_tflops=round(ex / (avg_ms * 1e-3) / 1e12, 3),
                   executed_frac=round(ex / (avg_ms * 1e-3) / peak, 4),
                   valu_issue=float(cv.get("valu_issue", "nan")))
    return out


def probe_valu_rate(device: int, precision: int) -> float:
    """Sustained FMA rate (TFLOP/s) of FP64 (0) or non-packed FP32 (1) chains,
    from the benchmark's calibration library (not the product library)."""
    lib = ctypes.CDLL(str(ROOT / "fastesp_amd" / "libfastesp_probe.so"))
    lib.fe_probe_valu_rate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.fe_probe_last_error.restype = ctypes.c_char_p
    v = ctypes.c_double()
    if lib.fe_probe_valu_rate(device, precision, ctypes.byref(v)) != 0:
        raise RuntimeError(lib.fe_probe_last_error().decode())
    return v.value


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baselines(w, budget_s: float, n_np: int, n_adj: int, V_gpu) -> dict:
    """The CPU paths on this host, timed on fixed point samples and scaled
    linearly to 1e4 points (the evaluation is linear in the point count for a
    fixed table):
      numpy_fwd     -- the NumPy restatement of the reference evaluation
                       (fastesp/table.py:73-104), one process, one thread;
      numpy_fwdbwd  -- the same plus the NumPy adjoint (field and dL/dG; the
                       point-independent dL/dG -> dL/dP is left out);
      cpp_fwd       -- the C++/OpenMP port (builder of build_pairs.cpp, the
                       reference evaluation) on every core of the host;
      reference_build_pairs_s -- the reference's own builder, compiled here.
    The NumPy legs evaluate the C++ port's table (the NumPy screen rule; the
    two builders keep the same pairs)."""
    from threadpoolctl import threadpool_limits
    from oracle import evaluate
    from oracle.cpu_port import CpuTable, threads
    nt = threads()
    out = dict(cores_available=nt, cpu_model=cpu_model(),
               omp_num_threads=os.environ.get("OMP_NUM_THREADS"))
    t0 = time.perf_counter()
    ct = CpuTable(w.wfn, w.basis, w.density, w.points, 1e-12, nthreads=nt)
    t_build = time.perf_counter() - t0
    # C++ port: calibrate the sample so the evaluation takes about budget_s
    n = 32
    t0 = time.perf_counter()
    ct.esp(w.points[:n], nthreads=nt)
    dt = time.perf_counter() - t0
    n2 = int(min(len(w.points), max(n, n * budget_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    ct.esp(w.points[:n2], nthreads=nt)
    dt2 = time.perf_counter() - t0
    out["cpp_fwd"] = dict(value=round(dt2 * 1e3 / n2 * 1e4, 3), unit="ms/1e4 pts", cores=nt,
                          kind="port", build_s=round(t_build, 3),
                          sample=f"C++/OpenMP port (oracle/cpu_port.cpp): FP64 forward on the "
                                 f"first {n2} of {len(w.points)} points with its own table "
                                 f"({ct.n_pairs} pairs), scaled linearly to 1e4")
    tab = ct.oracle_table(w.wfn, w.points)
    rng = np.random.default_rng(300)
    sel = np.sort(rng.choice(len(w.points), n_np, replace=False))
    with threadpool_limits(1):
        t0 = time.perf_counter()
        Vn = evaluate.esp(tab, w.points[sel])
        dt = time.perf_counter() - t0
        out["numpy_fwd"] = dict(value=round(dt * 1e3 / n_np * 1e4, 1), unit="ms/1e4 pts",
                                cores=1, kind="port", seconds=round(dt, 2),
                                sample=f"NumPy restatement (oracle/evaluate.py) of "
                                       f"PairTable.esp: FP64 forward on a fixed {n_np}-point "
                                       f"subset (seed 300) of the {len(w.points)} points, one "
                                       f"thread, scaled linearly to 1e4")
        if V_gpu is not None:
            out["numpy_fwd"]["max_abs_dV_vs_gpu_au"] = float(np.abs(Vn - V_gpu[sel]).max())
        sel2 = sel[:n_adj]
        gV = np.ones(n_adj)
        t0 = time.perf_counter()
        evaluate.esp(tab, w.points[sel2])
        evaluate.field(tab, w.points[sel2])
        evaluate.grad_G(tab, w.points[sel2], gV)
        dt = time.perf_counter() - t0
        out["numpy_fwdbwd"] = dict(value=round(dt * 1e3 / n_adj * 1e4, 1), unit="ms/1e4 pts",
                                   cores=1, kind="port", seconds=round(dt, 2),
                                   sample=f"NumPy restatement: FP64 forward + field + dL/dG "
                                          f"(oracle/evaluate.py esp, field, grad_G) on the first "
                                          f"{n_adj} points of the same subset, one thread, "
                                          f"scaled linearly to 1e4 (dL/dG -> dL/dP not included)")
    ref = ROOT / "oracle" / "_ref" / "build_pairs"
    if ref.exists():
        from fastesp_amd.formats import basis_payload, write_basis_payload
        tmp = Path(os.environ.get("TMPDIR", "/tmp"))
        b, p = tmp / f"fe_bench_basis_{os.getpid()}.bin", tmp / f"fe_bench_pairs_{os.getpid()}.bin"
        write_basis_payload(b, basis_payload(w.wfn, w.basis, w.density, w.points))
        t0 = time.perf_counter()
        r = subprocess.run([str(ref), str(b), str(p), "1e-12"], capture_output=True, text=True)
        out["reference_build_pairs_s"] = round(time.perf_counter() - t0, 3)
        out["reference_build_pairs_rc"] = r.returncode
        for f in (b, p):
            f.unlink(missing_ok=True)
    return out


def build_plus_fwd_legs(counts, torch, ctx, stream, reps: int = 5) -> dict:
    """The reference's published measurement (benchmarks/benchmark.py:107-141):
    table build + FP64 forward at each point count on the config-1 molecule.
    Build: the public API from host arrays (build_table), fastest of `reps`
    (the first build of the basis, which also sets it up on the device, is
    reported too); forward: device-resident points, fastest of `reps` HIP-event
    timed runs, as the reference takes the minimum over its runs."""
    from fastesp_amd import _lib, synth
    from fastesp_amd.table import build_table
    out = {}
    sh = stream.cuda_stream
    for n in counts:
        w = synth.config1(n)
        ctx.trim()  # no basis cache from earlier legs: the first build sets it up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tab = build_table(w.wfn, w.basis, w.density, w.points, tol=1e-12)
        first = time.perf_counter() - t0
        bt = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tab = build_table(w.wfn, w.basis, w.density, w.points, tol=1e-12)
            bt.append(time.perf_counter() - t0)
        pts = torch.from_numpy(np.ascontiguousarray(w.points)).to("cuda:0")
        V = torch.empty(n, dtype=torch.float64, device="cuda:0")
        ft = []
        for r in range(reps + 1):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            _lib.check(ctx.lib.fe_esp_forward(tab._h, pts.data_ptr(), n, _lib.FE_PREC_FP64,
                                              _lib.FE_PART_ESP, _lib.FE_PTR_DEVICE, sh,
                                              V.data_ptr()), "fe_esp_forward")
            e1.record(stream)
            torch.cuda.synchronize()
            if r:
                ft.append(e0.elapsed_time(e1))
        ms = min(bt) * 1e3 + min(ft)
        leg = dict(n_points=n, n_pairs=int(tab.n_pairs), build_ms=round(min(bt) * 1e3, 3),
                   first_build_ms=round(first * 1e3, 3), fwd_ms=round(min(ft), 4),
                   build_plus_fwd_ms=round(ms, 3), ms_per_1e4_pts=round(ms / (n / 1e4), 3),
                   reference_ms_per_1e4_pts=REFERENCE_BUILD_PLUS_FWD.get(n))
        if leg["reference_ms_per_1e4_pts"]:
            leg["speedup_vs_reference"] = round(leg["reference_ms_per_1e4_pts"] / leg["ms_per_1e4_pts"], 1)
        out[str(n)] = leg
        log(f"build + fwd at {n} points: build {leg['build_ms']} ms (first {leg['first_build_ms']}),"
            f" fwd {leg['fwd_ms']} ms, {leg['ms_per_1e4_pts']} ms/1e4 pts")
        tab.close()
    return out


def autograd_leg(w, torch, ctx, stream, steps: int) -> dict:
    """Config 2 through the differentiable interface: ESPModel (table rebuilt
    in place from the density and points on every call, forward + field) and
    the autograd backward to points.grad and density.grad, eager; then the same
    step (rebuild -> forward + field -> loss -> dL/dP) recorded as one graph
    through the C ABI and replayed."""
    from fastesp_amd import _lib
    from fastesp_amd.autograd import ESPModel
    dev = torch.device("cuda:0")
    n, nf = len(w.points), w.basis.n
    model = ESPModel(w.wfn, w.basis, tol=1e-12)
    pts = torch.from_numpy(np.ascontiguousarray(w.points)).to(dev).requires_grad_()
    P = torch.from_numpy(np.ascontiguousarray(w.density)).to(dev).requires_grad_()
    rng = np.random.default_rng(2002)
    wts = torch.from_numpy(rng.uniform(0.5, 1.0, n)).to(dev)
    with torch.no_grad():
        V0 = model(pts.detach(), P.detach())
    target = V0 + torch.from_numpy(rng.normal(0, 1e-3, n)).to(dev)

    def step():
        pts.grad = None
        P.grad = None
        V = model(pts, P)
        loss = (wts * (V - target) ** 2).sum()
        loss.backward()

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    out = dict(n_points=n, eager_ms_per_step=round(ms, 3),
               eager_ms_per_1e4_pts=round(ms / (n / 1e4), 4),
               eager_step="ESPModel: fe_table_rebuild + fe_esp_forward_grad, loss, "
                          "autograd backward (points.grad = gV dV/dC, density.grad = dL/dP)")
    # the raw step as a graph
    lib = ctx.lib
    t = model.last_table
    Pd, pd = P.detach(), pts.detach()
    V = torch.empty(n, dtype=torch.float64, device=dev)
    F = torch.empty((n, 3), dtype=torch.float64, device=dev)
    dP = torch.empty((nf, nf), dtype=torch.float64, device=dev)

    def raw():
        sh = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.fe_table_rebuild(t._h, Pd.data_ptr(), n, pd.data_ptr(), 1e-12, 0,
                                        _lib.FE_PTR_DEVICE, sh), "fe_table_rebuild")
        _lib.check(lib.fe_esp_forward_grad(t._h, pd.data_ptr(), n, _lib.FE_PREC_FP64,
                                           _lib.FE_PTR_DEVICE, sh, V.data_ptr(), F.data_ptr()),
                   "fe_esp_forward_grad")
        gV = 2.0 * wts * (V - target)
        _lib.check(lib.fe_esp_backward(t._h, pd.data_ptr(), n, gV.data_ptr(), _lib.FE_PREC_FP64,
                                       _lib.FE_PTR_DEVICE, sh, None, dP.data_ptr()),
                   "fe_esp_backward")

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        raw()
        raw()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        raw()
    g.replay()
    torch.cuda.synchronize()
    Vg, dPg = V.clone(), dP.clone()
    raw()
    torch.cuda.synchronize()
    bitwise = bool(torch.equal(Vg, V) and torch.equal(dPg, dP))
    e0.record(stream)
    for _ in range(steps):
        g.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    gms = e0.elapsed_time(e1) / steps
    out.update(graph_ms_per_step=round(gms, 3), graph_ms_per_1e4_pts=round(gms / (n / 1e4), 4),
               graph_replay_equals_eager=bitwise,
               graph_step="one hipGraph: fe_table_rebuild -> fe_esp_forward_grad -> loss -> "
                          "fe_esp_backward (dL/dP)")
    log(f"autograd step {ms:.2f} ms, graph-captured rebuild+fwd+bwd {gms:.2f} ms "
        f"(replay == eager: {bitwise})")
    return out


def batch_leg(a, ws, torch, ctx, stream) -> dict:
    """Batched conformers: every table evaluated forward (+ field) + loss +
    dL/dP per step, recorded as one hipGraph and replayed; tile shape from the
    batch autotuner.  Errors are per conformer against the batched FP64 forward."""
    from fastesp_amd import autotune, flops
    from fastesp_amd.batch import Batch
    from fastesp_amd.builder import build_table_hip
    sh = stream.cuda_stream
    t0 = time.perf_counter()
    tabs = [build_table_hip(w.wfn, w.basis, w.density, w.points, tol=1e-12) for w in ws]
    t_build = time.perf_counter() - t0
    infos = [t.info() for t in tabs]
    b = Batch(tabs, [w.points for w in ws])
    n_tot = int(b.offsets[-1])
    fl_model = sum(flops.forward_flops(i["pairs_per_L"], len(w.points), i["n_atoms"])
                   + flops.backward_flops(i["pairs_per_L"], len(w.points), i["n_atoms"])
                   for i, w in zip(infos, ws))
    b.forward("fp64")
    V64 = b.values()
    rng = np.random.default_rng(4004)
    wts = [rng.uniform(0.5, 1.0, len(v)) for v in V64]
    T = [v + rng.normal(0.0, 1e-3, len(v)) for v in V64]
    b.set_loss(wts, T)
    log(f"batch: {len(ws)} conformers, {n_tot} points, {sum(i['n_pairs'] for i in infos)} pairs, "
        f"tables built in {t_build:.1f} s")
    out = dict(n_conformers=len(ws), n_points=n_tot,
               n_pairs=int(sum(i["n_pairs"] for i in infos)),
               atoms_range=[int(min(w.meta["n_atoms"] for w in ws)),
                            int(max(w.meta["n_atoms"] for w in ws))],
               functions_range=[int(min(w.basis.n for w in ws)), int(max(w.basis.n for w in ws))],
               points_range=[int(min(len(w.points) for w in ws)),
                             int(max(len(w.points) for w in ws))],
               table_build_s=round(t_build, 2), model_flops_per_step=fl_model, legs={})
    tuned = None
    for prec in a.batch_precisions.split(","):
        if tuned is None:
            t1 = time.perf_counter()
            tuned = autotune.tune_batch(b, prec, field=True, dldp=True)
            tuned["seconds"] = round(time.perf_counter() - t1, 1)
            log(f"batch autotune ({prec}): ppl_L={tuned['ppl_L']} fp32={tuned['mixed_L']} items/WG="
                f"{tuned['items_per_wg']} dG items/WG={tuned['dg_items_per_wg']} "
                f"({tuned['trials']} trials, {tuned['seconds']} s)")
            # performed work: every conformer's far / tabulated evaluation
            # counts with the single-table launch decomposition at the
            # batch's points per lane (the batch's items tile the same
            # Morton-sorted points and pair order)
            lib = ctx.lib
            autotune.apply(ctx, tuned["ppl_L"], 8192)
            fl = 0.0
            for t, w, i in zip(tabs, ws, infos):
                d = torch.from_numpy(np.ascontiguousarray(w.points)).to("cuda")
                c = work_counts(lib, t, d.data_ptr(), len(w.points), True, sh)
                fl += sum(launch_flops(c, True).values()) + len(w.points) * i["n_atoms"] * 28
            torch.cuda.synchronize()
            out["flops_per_step"] = fl
            out["flop_model"] = ("performed work (fe_eval_work counts x flops.f_eval_path / "
                                 "f_dg_path, nuclear 28 per (point, atom)); "
                                 "model_flops_per_step is the per-unit model")
        b.capture(prec, field=True, dldp=True)
        for _ in range(a.batch_warmup):
            b.launch(sh)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(a.batch_steps):
            b.launch(sh)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / a.batch_steps
        V = b.values()
        errs = np.array([np.abs(v - r).max() if len(v) else 0.0 for v, r in zip(V, V64)])
        # per-launch breakdown from one eager step (launches inside a graph are not timed)
        ctx.profile_start()
        b.step(prec, field=True, dldp=True, stream=sh)
        torch.cuda.synchronize()
        ph = ctx.profile_stop()
        leg = dict(ms_per_step=round(ms, 3), points_per_s=round(n_tot / (ms * 1e-3), 1),
                   ms_per_1e4_pts=round(ms / (n_tot / 1e4), 5),
                   achieved_tflops=round(out["flops_per_step"] / (ms * 1e-3) / 1e12, 2),
                   frac_fp64_peak=round(out["flops_per_step"] / (ms * 1e-3) / flops.PEAK_FP64_VALU, 4),
                   model_tflops=round(fl_model / (ms * 1e-3) / 1e12, 2),
                   max_abs_dV_vs_fp64_au=float(errs.max()),
                   median_conformer_max_dV_au=float(np.median(errs)),
                   eager_step_phases_ms={PHASE_NAMES.get(k, str(k)): round(v[0], 3)
                                         for k, v in sorted(ph.items())})
        out["legs"][prec] = leg
        log(f"batch {prec}: {ms:.1f} ms/step = {leg['points_per_s']:.3e} points/s, "
            f"max |dV| vs FP64 {errs.max():.2e}")
    out["tuning"] = {k: v for k, v in tuned.items() if k != "all"} if tuned else None
    # The same step with every table built again on the device inside the
    # graph (build -> forward -> loss -> backward): the densities go to the
    # batch once; the first call gives each table its capacity layout, after
    # which the plans are cut from capacities.  Timed with and without the
    # builds at those plans, so their difference is the builds' share.
    prec = a.batch_precisions.split(",")[0]
    t1 = time.perf_counter()
    b.set_density([w.density for w in ws], tol=1e-12)
    t_first = time.perf_counter() - t1
    rb = {}
    for flag in (False, True):
        b.capture(prec, field=True, dldp=True, rebuild=flag)
        for _ in range(a.batch_warmup):
            b.launch(sh)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(a.batch_steps):
            b.launch(sh)
        e1.record(stream)
        torch.cuda.synchronize()
        rb[flag] = e0.elapsed_time(e1) / a.batch_steps
    errs = np.array([np.abs(v - r).max() if len(v) else 0.0 for v, r in zip(b.values(), V64)])
    out["build_in_graph"] = dict(
        precision=prec, step_ms=round(rb[True], 3), step_without_builds_ms=round(rb[False], 3),
        builds_ms=round(rb[True] - rb[False], 3),
        builds_host_api_s=round(t_build, 3), first_set_density_s=round(t_first, 3),
        max_abs_dV_vs_fp64_au=float(errs.max()),
        graph="one hipGraph: 256 x fe_table_rebuild -> forward + field -> loss -> dL/dG -> dL/dP",
        note="the same densities as the tables' own builds, so the kept sets are the same")
    log(f"batch step with the 256 builds in the graph: {rb[True]:.1f} ms "
        f"(builds {rb[True] - rb[False]:.1f} ms; {t_build:.2f} s through the host API)")
    b.close()
    for t in tabs:
        t.close()
    return out


def density_leg(a, w, torch, ctx, stream, with_cpu: bool) -> dict:
    """Electron density rho(C) at the config-1 points (fe_density_*): basis
    values, then the quadratic form on the FP64 matrix cores.  The roofline of
    the matrix kernel is the FP64 matrix peak; the CPU baseline is the NumPy
    restatement of the reference's density (one process) on a point sample,
    which also checks the device values at those points."""
    from fastesp_amd.density import ElectronDensity
    n = len(w.points)
    ed = ElectronDensity(w.wfn, w.basis, w.density, ctx)
    pts = torch.from_numpy(np.ascontiguousarray(w.points)).to("cuda:0")
    rho = torch.empty(n, dtype=torch.float64, device="cuda:0")
    sh = stream.cuda_stream
    for _ in range(2):
        ed.rho_device(pts.data_ptr(), n, rho.data_ptr(), sh)
    torch.cuda.synchronize()
    ctx.profile_start()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(a.density_steps):
        ed.rho_device(pts.data_ptr(), n, rho.data_ptr(), sh)
    e1.record(stream)
    torch.cuda.synchronize()
    ph = ctx.profile_stop()
    ms = e0.elapsed_time(e1) / a.density_steps
    mm_ms = ph[5][0] / ph[5][1]
    info = ed.info()
    achieved = info["mm_flops"] / (mm_ms * 1e-3)
    out = dict(ms_per_eval=round(ms, 4), ms_per_1e4_pts=round(ms / (n / 1e4), 4),
               n_points=n, n_functions=int(w.basis.n),
               phases_ms={PHASE_NAMES[k]: round(v[0] / v[1], 4) for k, v in sorted(ph.items())},
               roofline=dict(kernel="k_density_mm", bound="mfma", unit="TFLOP/s",
                             achieved=round(achieved / 1e12, 3), peak=78.6,
                             frac=round(achieved / 78.6e12, 4),
                             flops_per_launch=int(info["mm_flops"]), avg_launch_ms=round(mm_ms, 4),
                             flop_model="N_pad x 2 x 64^2 x nJ (nJ + 1) / 2 (block-upper P)"))
    if with_cpu:
        sys.path.insert(0, str(ROOT))
        from oracle import density as odens
        sel = np.random.default_rng(77).choice(n, 1000, replace=False)
        t0 = time.perf_counter()
        want = odens.rho(w.basis, w.density, w.points[sel])
        dt = time.perf_counter() - t0
        got = rho.cpu().numpy()[sel]
        out["cpu_baseline"] = dict(
            value=round(dt * 1e3 / len(sel) * 1e4, 2), unit="ms/1e4 pts", cores=1, kind="port",
            sample=f"NumPy restatement of our_density (benchmarks/check_density_mwfn.py:60-75) "
                   f"on {len(sel)} sampled points, scaled linearly to 1e4")
        out["max_rel_err_vs_oracle"] = float(np.abs(got - want).max() / np.abs(want).max())
    ed.close()
    return out


C3_FP32_CLASSES = [0, 0, 0, 1, 1, 1, 1]  # the library default under mixed precision


def config3_leg(a, torch, ctx, stream, shape=None) -> dict:
    """Config 3: 150 atoms, 4538 functions up to f, 200000 points on four vdW
    shells.  Forward + field + backward (dL/dP) per step in FP64 and in mixed
    precision, at the points-per-lane and workgroup target tuned on config 1
    (`shape`) and, for mixed, the FP32 contraction in classes 3..6 (the library
    default, so the leg measures the FP32 error at this size whatever the
    config-1 tuner kept); errors of V and of the field against the FP64 oracle
    at the 256 points of tests/golden/cfg3.npz, and of mixed against FP64 at
    every point."""
    from fastesp_amd import _lib, synth
    if shape:
        for L, v in enumerate(shape["ppl_L"]):
            ctx.set_tuning(10 + L, int(v))
        ctx.set_tuning(1, int(shape["target_wg"]))
    for L, v in enumerate(C3_FP32_CLASSES):
        ctx.set_tuning(20 + L, v)
    from fastesp_amd.table import build_table
    g = np.load(ROOT / "tests" / "golden" / "cfg3.npz")
    w = synth.config3()
    n = len(w.points)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tab = build_table(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    t_build = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    sh = stream.cuda_stream
    lib = ctx.lib
    pts = torch.from_numpy(np.ascontiguousarray(w.points)).to(dev)
    V = torch.empty(n, dtype=torch.float64, device=dev)
    dVdC = torch.empty((n, 3), dtype=torch.float64, device=dev)
    dLdP = torch.empty((w.basis.n, w.basis.n), dtype=torch.float64, device=dev)
    rng = np.random.default_rng(3003)
    wts = torch.from_numpy(rng.uniform(0.5, 1.0, n)).to(dev)
    _lib.check(lib.fe_esp_forward(tab._h, pts.data_ptr(), n, _lib.FE_PREC_FP64, _lib.FE_PART_ESP,
                                  _lib.FE_PTR_DEVICE, sh, V.data_ptr()), "fe_esp_forward")
    target = V.clone() + torch.from_numpy(rng.normal(0, 1e-3, n)).to(dev)
    idx = torch.from_numpy(g["idx"]).to(dev)
    out = dict(molecule="config3 150 atoms", n_functions=int(w.basis.n), n_points=n,
               n_pairs=int(tab.n_pairs), table_build_ms=round(t_build * 1e3, 1),
               ppl_L=list(shape["ppl_L"]) if shape else None,
               target_wg=shape["target_wg"] if shape else None,
               mixed_fp32_classes=C3_FP32_CLASSES, legs={})
    res = {}
    for prec, p in (("fp64", _lib.FE_PREC_FP64), ("mixed", _lib.FE_PREC_MIXED)):
        def step():
            _lib.check(lib.fe_esp_forward_grad(tab._h, pts.data_ptr(), n, p, _lib.FE_PTR_DEVICE,
                                               sh, V.data_ptr(), dVdC.data_ptr()),
                       "fe_esp_forward_grad")
            gV = 2.0 * wts * (V - target)
            _lib.check(lib.fe_esp_backward(tab._h, pts.data_ptr(), n, gV.data_ptr(), p,
                                           _lib.FE_PTR_DEVICE, sh, None, dLdP.data_ptr()),
                       "fe_esp_backward")
        step()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(a.config3_steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / a.config3_steps
        Vs, Fs = V[idx].cpu().numpy(), dVdC[idx].cpu().numpy()
        res[prec] = (V.cpu().numpy(), dVdC.cpu().numpy(), dLdP.cpu().numpy())
        dv = Vs - g["V"]
        leg = dict(ms_per_step=round(ms, 2), ms_per_1e4_pts=round(ms / (n / 1e4), 4),
                   max_abs_dV_vs_oracle_au=float(np.abs(dv).max()),
                   rms_dV_vs_oracle_au=float(np.sqrt((dv ** 2).mean())))
        if "field" in g.files:
            leg["max_abs_dfield_vs_oracle_au"] = float(np.abs(Fs - g["field"]).max())
        out["legs"][prec] = leg
        log(f"config3 {prec}: {ms:.1f} ms/step, max |dV| vs oracle {leg['max_abs_dV_vs_oracle_au']:.2e}")
    Vd, Fd, Pd = (res["mixed"][k] - res["fp64"][k] for k in range(3))
    out["mixed_vs_fp64_all_points"] = dict(
        max_abs_dV_au=float(np.abs(Vd).max()), rms_dV_au=float(np.sqrt((Vd ** 2).mean())),
        max_abs_dfield_au=float(np.abs(Fd).max()),
        max_rel_ddLdP=float(np.abs(Pd).max() / np.abs(res["fp64"][2]).max()))
    out["speedup_mixed_vs_fp64_fwdbwd"] = round(
        out["legs"]["fp64"]["ms_per_step"] / out["legs"]["mixed"]["ms_per_step"], 3)
    tab.close()
    return out


def main() -> int:
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a "
                                   "design note, not an implemented path", "n_gpus": a.gpus}))
        return 2
    import torch
    from fastesp_amd import _lib, flops, synth
    from fastesp_amd.table import build_table

    batch_ws = None
    if a.batch:
        # before the GPU is touched: the generator forks worker processes
        t0 = time.perf_counter()
        batch_ws = synth.conformer_set(a.batch)
        log(f"generated {a.batch} conformers in {time.perf_counter() - t0:.1f} s")
    log(f"generating workload (config 1, {a.points} points)")
    w = synth.config1(a.points)
    ctx = _lib.Context.get(0)
    ctx.set_tuning(6, 0 if a.dg_streams == 1 else a.dg_streams)
    if a.eval_streams:
        ctx.set_tuning(7, int(a.eval_streams == 2))
    if a.ppl:
        ctx.set_tuning(0, a.ppl)
    if a.ppl_classes:
        a.tune = False
        for L, v in enumerate(int(x) for x in a.ppl_classes.split(",")):
            ctx.set_tuning(10 + L, v)
    if a.wg:
        ctx.set_tuning(1, a.wg)
    if a.fp32_classes:
        a.tune = False
        for L, v in enumerate(int(x) for x in a.fp32_classes.split(",")):
            ctx.set_tuning(20 + L, v)
    dev = torch.device("cuda:0")
    n = a.points
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    lib = ctx.lib
    # --- table build on the GPU (reported, not part of the step) ----------
    # public API from host arrays (first build of the basis, then with the
    # basis set-up cached), and from device-resident density and points
    t0 = time.perf_counter()
    build_table(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    t_first = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tab = build_table(w.wfn, w.basis, w.density, w.points, tol=1e-12)
    t_build = time.perf_counter() - t0
    info = tab.info()
    log(f"table: {info['n_pairs']} pairs, per L {info['pairs_per_L']}, build {t_build*1e3:.2f} ms "
        f"(first {t_first*1e3:.2f} ms)")
    pts = torch.from_numpy(np.ascontiguousarray(w.points)).to(dev)
    Pd = torch.from_numpy(np.ascontiguousarray(w.density)).to(dev)
    from fastesp_amd.autograd import ESPModel
    mdl = ESPModel(w.wfn, w.basis, tol=1e-12, reuse_table=False)
    mdl.build(pts, Pd)
    torch.cuda.synchronize()
    dt = []
    for _ in range(5):
        t0 = time.perf_counter()
        tb = mdl.build(pts, Pd)
        torch.cuda.synchronize()
        dt.append(time.perf_counter() - t0)
        tb.close()
    # in-place rebuild (device inputs, no host synchronisation): HIP events
    _lib.check(lib.fe_table_rebuild(tab._h, Pd.data_ptr(), n, pts.data_ptr(), 1e-12, 0,
                                    _lib.FE_PTR_DEVICE, sh), "fe_table_rebuild")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(5):
        _lib.check(lib.fe_table_rebuild(tab._h, Pd.data_ptr(), n, pts.data_ptr(), 1e-12, 0,
                                        _lib.FE_PTR_DEVICE, sh), "fe_table_rebuild")
    e1.record(stream)
    torch.cuda.synchronize()
    build = dict(host_api_ms=round(t_build * 1e3, 3), host_api_first_ms=round(t_first * 1e3, 3),
                 device_inputs_ms=round(min(dt) * 1e3, 3),
                 rebuild_in_place_ms=round(e0.elapsed_time(e1) / 5, 3),
                 note="host_api: build_table with NumPy inputs (21 MB density over the host "
                      "link), basis set-up cached after the first; device_inputs: fe_table_build "
                      "on HBM-resident density and points; rebuild_in_place: fe_table_rebuild, "
                      "event-timed, no host synchronisation")
    log(f"build: {build}")
    tab = build_table(w.wfn, w.basis, w.density, w.points, tol=1e-12)  # exact-size table again

    V = torch.empty(n, dtype=torch.float64, device=dev)
    dVdC = torch.empty((n, 3), dtype=torch.float64, device=dev)
    dLdP = torch.empty((w.basis.n, w.basis.n), dtype=torch.float64, device=dev)
    PREC = {"fp64": _lib.FE_PREC_FP64, "mixed": _lib.FE_PREC_MIXED}
    # config-2 loss: L = sum w (V - T)^2, w ~ U[0.5, 1], T = V_fp64 + N(0, 1e-3)
    rng = np.random.default_rng(2002)
    wts = torch.from_numpy(rng.uniform(0.5, 1.0, n)).to(dev)
    _lib.check(lib.fe_esp_forward(tab._h, pts.data_ptr(), n, PREC["fp64"], _lib.FE_PART_ESP,
                                  _lib.FE_PTR_DEVICE, sh, V.data_ptr()), "fe_esp_forward")
    target = V.clone() + torch.from_numpy(rng.normal(0, 1e-3, n)).to(dev)

    def make_step(mode, prec):
        p = PREC[prec]

        def fwd():
            _lib.check(lib.fe_esp_forward(tab._h, pts.data_ptr(), n, p, _lib.FE_PART_ESP,
                                          _lib.FE_PTR_DEVICE, sh, V.data_ptr()), "fe_esp_forward")

        def fwdbwd():
            _lib.check(lib.fe_esp_forward_grad(tab._h, pts.data_ptr(), n, p, _lib.FE_PTR_DEVICE,
                                               sh, V.data_ptr(), dVdC.data_ptr()),
                       "fe_esp_forward_grad")
            gV = 2.0 * wts * (V - target)
            _lib.check(lib.fe_esp_backward(tab._h, pts.data_ptr(), n, gV.data_ptr(), p,
                                           _lib.FE_PTR_DEVICE, sh, None, dLdP.data_ptr()),
                       "fe_esp_backward")
        return fwd if mode == "fwd" else fwdbwd

    gold = ROOT / "tests" / "golden" / "cfg1_30k.npz"
    g = np.load(gold) if (gold.exists() and n == 30000) else None
    from fastesp_amd import autotune
    tuned = {}

    def current_shape(mode):
        # without explicit values, the library's defaults for the mode
        ppl = ctx_ppl or (LIB_PPL_FWD if mode == "fwd" else LIB_PPL_GRAD)
        return dict(ppl_L=[ppl[L] for L in range(7)], target_wg=ctx_wg[0],
                    fp32_classes=[ctx_fp32[L] for L in range(7)])

    ctx_ppl = [a.ppl] * 7 if a.ppl else None
    if a.ppl_classes:
        ctx_ppl = [int(x) for x in a.ppl_classes.split(",")]
    ctx_wg = [a.wg or 32 * 256]
    ctx_fp32 = [0, 0, 0, 1, 1, 1, 1]
    if a.fp32_classes:
        ctx_fp32 = [int(x) for x in a.fp32_classes.split(",")]

    def timed(mode, prec):
        force32 = prec == "mixed32"
        if force32:
            prec = "mixed"
            for L in range(7):
                ctx.set_tuning(20 + L, 1)
        if a.tune:
            r = autotune.tune(tab, pts, op=mode, precision=PREC[prec], fix_mixed=force32)
            best = r["best"]
            shape = dict(ppl_L=best["ppl_L"], target_wg=best["target_wg"],
                         fp32_classes=best["mixed_L"] or ([1] * 7 if force32 else ctx_fp32),
                         trials=len(r["all"]))
            log(f"autotune {mode} {prec}: best ppl_L={best['ppl_L']} wg={best['target_wg']} "
                f"{best['ms']:.2f} ms (first {r['all'][0]['ms']:.2f} ms, {len(r['all'])} trials)")
        else:
            shape = current_shape(mode)
            if force32:
                shape["fp32_classes"] = [1] * 7
        autotune.apply(ctx, shape["ppl_L"], shape["target_wg"],
                       shape["fp32_classes"] if prec == "mixed" else None)
        step = make_step(mode, prec)
        for _ in range(a.warmup):
            step()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        # the step as a user runs it: no instrumentation
        e0.record(stream)
        for _ in range(a.steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / a.steps
        # the same K steps again with HIP events recorded inside the library
        # around every launch, each on the stream the launch goes to
        ctx.profile_start()
        e0.record(stream)
        for _ in range(a.steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        phases = ctx.profile_stop()
        ms_prof = e0.elapsed_time(e1) / a.steps
        res = dict(ms_per_step=round(ms, 4), ms_per_1e4_pts=round(ms / (n / 1e4), 4),
                   ms_per_step_profiled=round(ms_prof, 4),
                   tuning=shape,
                   phases_ms_per_step={PHASE_NAMES.get(k, str(k)): round(v[0] / a.steps, 4)
                                       for k, v in sorted(phases.items())})
        if prec == "mixed":
            res["fp32_classes_used"] = [L for L in range(7) if shape["fp32_classes"][L]]
        res["_phases"] = phases
        res["_shape"] = shape
        if g is not None:
            d = V.cpu().numpy()[g["idx"]] - g["V"]
            res["max_abs_dV_au"] = float(np.abs(d).max())
            res["rms_dV_au"] = float(np.sqrt((d * d).mean()))
        log(f"{mode} {prec}{' (FP32 in every class)' if force32 else ''}: {ms:.3f} ms/step = "
            f"{res['ms_per_1e4_pts']:.3f} ms/1e4 pts, "
            f"max |dV| {res.get('max_abs_dV_au', float('nan')):.2e}")
        if force32:
            for L in range(7):  # back to the library default
                ctx.set_tuning(20 + L, int(L >= 3))
        tuned[f"{mode}_{'mixed32' if force32 else prec}"] = shape
        return res

    legs = {}
    for leg in a.legs.split(","):
        mode, prec = leg.split("_")
        legs[leg] = timed(mode, prec)
    key = f"{a.mode}_{a.precision}"
    if key not in legs:
        legs[key] = timed(a.mode, a.precision)
    if a.copy_probe:
        # a plain copy of known size, so HBM byte counters can be calibrated
        src = torch.empty(a.copy_probe << 17, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {a.copy_probe} MiB read + {a.copy_probe} MiB written")
    head = legs[key]
    ms_step = head["ms_per_step"]
    grad = a.mode == "fwdbwd"

    # --- roofline of the headline step: performed work of its launches -------
    shape = head["_shape"]
    autotune.apply(ctx, shape["ppl_L"], shape["target_wg"],
                   shape["fp32_classes"] if a.precision == "mixed" else None)
    counts = work_counts(lib, tab, pts.data_ptr(), n, grad, sh)
    ppl = info["pairs_per_L"]
    traffic_file = ROOT / "profiles" / f"esp_cfg1_30k_{key}_hbm_traffic.md"
    counters_file = ROOT / "profiles" / f"esp_cfg1_30k_{key}_valu_counters.md"
    # with two streams the dL/dG classes run two at a time
    overlapped = frozenset(16 + L for L in range(7)) if grad and a.dg_streams >= 2 else frozenset()
    # the evaluation classes run two at a time when asked (2), or by the
    # library default (0): the value kernels in two halves, the value + field
    # kernels with classes 5 and 6 beside classes 0 and 1
    if a.eval_streams == 2 or (a.eval_streams == 0 and not grad):
        overlapped |= frozenset(8 + L for L in range(7))
    elif a.eval_streams == 0:
        overlapped |= frozenset({8, 9, 13, 14})
    roof = roofline(head["_phases"], counts, grad, shape["ppl_L"], shape["fp32_classes"],
                    a.precision, ppl, n, traffic_file, counters_file, overlapped)
    fl_launch = launch_flops(counts, grad)
    f_nuc = n * info["n_atoms"] * (11 + (17 if grad else 0))
    f_step = sum(fl_launch.values()) + f_nuc
    f_model = flops.forward_flops(ppl, n, info["n_atoms"]) + (
        flops.backward_flops(ppl, n, info["n_atoms"]) if grad else 0.0)
    roof_step = dict(achieved=round(f_step / (ms_step * 1e-3) / 1e12, 3),
                     peak=round(flops.PEAK_FP64_VALU / 1e12, 2), unit="TFLOP/s",
                     frac=round(f_step / (ms_step * 1e-3) / flops.PEAK_FP64_VALU, 4),
                     flops_per_step=f_step, model_flops_per_step=f_model,
                     model_frac=round(f_model / (ms_step * 1e-3) / flops.PEAK_FP64_VALU, 4),
                     far_share_eval=round(sum(counts["eval_far"]) /
                                          max(1, sum(counts["eval_far"]) + sum(counts["eval_near"])), 4),
                     far_share_dg=round(sum(counts["dg_far"]) /
                                        max(1, sum(counts["dg_far"]) + sum(counts["dg_near"])), 4),
                     per_launch={PHASE_NAMES[k]: (dict(
                         ms=round(head["_phases"][k][0] / head["_phases"][k][1], 4),
                         tflops=round(v / (head["_phases"][k][0] / head["_phases"][k][1] * 1e-3) / 1e12, 2))
                         if k not in overlapped else dict(
                         span_ms=round(head["_phases"][k][0] / head["_phases"][k][1], 4),
                         gflop=round(v / 1e9, 2), note="runs beside a launch on the other stream"))
                         for k, v in sorted(fl_launch.items()) if k in head["_phases"]},
                     work_counts=counts)
    for leg in legs.values():
        leg.pop("_phases", None)
        leg.pop("_shape", None)

    # sustained FMA rates of this device (independent chains): the FP32 the
    # mixed contraction can use is non-packed, so its per-instruction rate is
    # what bounds the mixed-precision gain
    valu_rate = {"fp64_tflops": round(probe_valu_rate(0, 0), 2),
                 "fp32_nonpacked_tflops": round(probe_valu_rate(0, 1), 2)}
    log(f"measured VALU FMA rate: {valu_rate}")

    # the public API from host memory beside the device-resident forward, same
    # tuning (the FP64 forward leg's), same timer, alternating in one loop
    host = None
    if "fwd_fp64" in legs:
        sf = legs["fwd_fp64"]["tuning"]
        autotune.apply(ctx, sf["ppl_L"], sf["target_wg"])
        reps = max(3, min(a.steps, 10))
        td, th = [], []
        for r in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _lib.check(lib.fe_esp_forward(tab._h, pts.data_ptr(), n, _lib.FE_PREC_FP64,
                                          _lib.FE_PART_ESP, _lib.FE_PTR_DEVICE, sh, V.data_ptr()),
                       "fe_esp_forward")
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            tab.esp(w.points, precision=_lib.FE_PREC_FP64)
            t2 = time.perf_counter()
            if r:
                td.append(t1 - t0)
                th.append(t2 - t1)
        host = dict(api="PairTable.esp(points): host NumPy points in, host NumPy V out, FP64",
                    calls=reps, ms_per_call=round(float(np.median(th)) * 1e3, 4),
                    ms_per_1e4_pts=round(float(np.median(th)) * 1e3 / (n / 1e4), 4),
                    device_resident_ms_per_call=round(float(np.median(td)) * 1e3, 4),
                    timer="perf_counter around each call (device call followed by a "
                          "synchronisation), medians, same launch shape")
        log(f"host-boundary forward: {host['ms_per_call']} ms per call, device-resident "
            f"{host['device_resident_ms_per_call']} ms")
    V_gpu = None
    if not a.no_cpu:
        _lib.check(lib.fe_esp_forward(tab._h, pts.data_ptr(), n, _lib.FE_PREC_FP64,
                                      _lib.FE_PART_ESP, _lib.FE_PTR_DEVICE, sh, V.data_ptr()),
                   "fe_esp_forward")
        V_gpu = V.cpu().numpy()
    autotune.apply(ctx, shape["ppl_L"], shape["target_wg"])

    ag = autograd_leg(w, torch, ctx, stream, a.autograd_steps) if a.autograd_steps else None
    batched = batch_leg(a, batch_ws, torch, ctx, stream) if batch_ws else None
    cfg3 = None
    if a.config3_steps:
        log("config 3 leg (200000 points)")
        cfg3 = config3_leg(a, torch, ctx, stream, tuned.get(key))
    dens = None
    if a.density_steps:
        log("electron density leg")
        dens = density_leg(a, w, torch, ctx, stream, with_cpu=not a.no_cpu)
    bpf = None
    if a.build_legs:
        autotune.apply(ctx, legs.get("fwd_fp64", head)["tuning"]["ppl_L"],
                       legs.get("fwd_fp64", head)["tuning"]["target_wg"])
        bpf = build_plus_fwd_legs([int(x) for x in a.build_legs.split(",")], torch, ctx, stream)

    cpu = cpus = None
    if not a.no_cpu:
        log("CPU baselines (C++/OpenMP port on every core, NumPy restatement on one)")
        cpus = cpu_baselines(w, a.cpu_seconds, a.numpy_points, a.numpy_adjoint_points, V_gpu)
        if a.mode == "fwdbwd":
            cpu = dict(cpus["numpy_fwdbwd"])
        else:
            cpu = dict(cpus["numpy_fwd"])
        cpu["cpu_model"] = cpus["cpu_model"]

    value = head["ms_per_1e4_pts"]
    mixed = legs.get(f"{a.mode}_mixed")
    out = {
        "metric": "ESP ms/1e4 pts (1634 fns) FP64 vs mixed, fwd+bwd; max err vs FP64 oracle",
        "value": value,
        "unit": "ms/1e4 pts",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded config-1 molecule, density and vdW-surface points)",
        "config": {"workload": f"esp_{a.mode}_{a.precision}", "molecule": "config1 54 atoms",
                   "n_functions": int(w.basis.n), "n_points": n,
                   "n_pairs_per_L": ppl, "tol": 1e-12, "parallelism": "single GPU",
                   "dg_streams": a.dg_streams},
        "max_abs_dV_vs_fp64_oracle_au": head.get("max_abs_dV_au"),
        "mixed": None if mixed is None else dict(
            ms_per_1e4_pts=mixed["ms_per_1e4_pts"], max_abs_dV_au=mixed.get("max_abs_dV_au"),
            fp32_classes_used=mixed.get("fp32_classes_used"),
            speedup_vs_fp64=round(head["ms_per_step"] / mixed["ms_per_step"], 3),
            note="the tuner keeps FP32 only where it times faster; non-packed FP32 issues at "
                 "the FP64 rate on this chip (valu_rate_measured), so mixed equals FP64 "
                 "within noise"),
        "legs": legs,
        "tuning": tuned,
        "table_build_ms": round(t_build * 1e3, 3),
        "table_build": build,
        "build_plus_fwd": bpf,
        "autograd": ag,
        "host_boundary_fwd": host,
        "roofline": roof,
        "roofline_step": roof_step,
        "valu_rate_measured": valu_rate,
        "batched": batched,
        "config3": cfg3,
        "density": dens,
        "cpu_baseline": cpu,
        "cpu_baselines": cpus,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
