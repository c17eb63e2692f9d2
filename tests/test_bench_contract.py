"""The bench.py contract: one JSON line with the metric, the roofline of the
dominant launch and the CPU baseline; one GPU only.

CPU: the multi-GPU refusal and the CPU-baseline leg on a small molecule.
GPU: a short run of the headline legs, checked key by key.
"""
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]

REQUIRED = {"metric": str, "value": float, "unit": str, "n_gpus": int, "steps": int,
            "warmup": int, "ms_per_step": float, "higher_is_better": bool, "scaling": str,
            "dtype": str, "data": str, "config": dict}


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_cpu_baseline_legs_on_config0():
    sys.path.insert(0, str(ROOT))
    import bench
    from fastesp_amd import synth
    w = synth.config0()
    out = bench.cpu_baselines(w, 0.2, 8, 2, None)
    assert out["cores_available"] >= 1 and out["cpu_model"]
    for leg in ("cpp_fwd", "numpy_fwd", "numpy_fwdbwd"):
        v = out[leg]
        assert v["kind"] == "port" and v["unit"] == "ms/1e4 pts", leg
        assert v["cores"] >= 1 and v["value"] > 0, leg
        assert "scaled linearly to 1e4" in v["sample"], leg
    assert out["numpy_fwd"]["cores"] == 1 and out["numpy_fwdbwd"]["cores"] == 1
    if (ROOT / "oracle" / "_ref" / "build_pairs").exists():
        assert out["reference_build_pairs_rc"] == 0


def test_performed_work_model_counts_the_generated_code():
    """The contraction counts come from the generated code (an FMA 2, any
    other operation 1), the far path is the cheaper one, and from L = 3 on the
    shared subexpressions put every path below the per-unit model's charge."""
    from fastesp_amd import flops
    assert flops.op_flops("a = fma(b, c, d);\n  e = a * b + c;") == 4
    assert flops.generated_flops("fe_contract_grad_acc", 2) == 68
    for L in range(7):
        for grad in (False, True):
            far, near = flops.f_eval_path(L, grad, True), flops.f_eval_path(L, grad, False)
            assert 0 < far <= near
            if L >= 3:
                assert near <= (flops.f_eval_grad(L) if grad else flops.f_fwd(L))
        assert 0 < flops.f_dg_path(L, True) <= flops.f_dg_path(L, False)


@pytest.mark.gpu
def test_bench_json_line():
    print("[gpu-case] bench.py short headline run", flush=True)
    cmd = [sys.executable, "-u", str(ROOT / "bench.py"), "--steps", "2", "--warmup", "1",
           "--no-tune", "--no-cpu", "--batch", "0", "--config3-steps", "0",
           "--density-steps", "0", "--legs", "fwd_fp64,fwdbwd_fp64,fwdbwd_mixed",
           "--build-legs", "300", "--autograd-steps", "2"]
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=600,
                       env=dict(os.environ))
    assert r.returncode == 0, r.stderr[-3000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    for k, t in REQUIRED.items():
        assert isinstance(d[k], t), k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["higher_is_better"] is False and d["vs_baseline"] is None
    assert d["config"]["workload"] == "esp_fwdbwd_fp64" and d["dtype"] == "f64"
    assert d["value"] > 0 and np.isclose(d["value"], d["ms_per_step"] / 3.0, rtol=1e-3)
    roof = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in roof, k
    assert roof["peak"] == 78.6 and roof["unit"] == "TFLOP/s"
    assert np.isclose(roof["frac"], roof["achieved"] / roof["peak"], rtol=1e-2)
    # a performed-work rate can never exceed the peak
    assert 0 < roof["frac"] <= 1 and 0 < d["roofline_step"]["frac"] <= 1
    assert roof["kernel"].startswith(("k_eval<", "k_dG<"))
    assert roof["far_evaluations"] + roof["tabulated_evaluations"] > 0
    assert d["cpu_baseline"] is None  # --no-cpu
    for leg in ("fwd_fp64", "fwdbwd_fp64", "fwdbwd_mixed"):
        assert d["legs"][leg]["max_abs_dV_au"] < 1e-10, leg
    assert d["mixed"]["max_abs_dV_au"] < 1e-8
    b = d["build_plus_fwd"]["300"]
    assert b["reference_ms_per_1e4_pts"] == 8959.0 and b["ms_per_1e4_pts"] > 0
    assert d["autograd"]["graph_replay_equals_eager"] is True
    assert d["table_build"]["rebuild_in_place_ms"] > 0
