"""The command-line drop-ins (python -m fastesp_amd): the reference's
build_pairs / esp_gpu file boundary (fastesp/cli.py:143-168) and the
in-process driver, checked against the golden config-0 values.

Files and console protocol as the native tools: pairs.bin in reference
layout, ``run i: X ms`` lines, output ``N`` then `` x y z V`` rows.
"""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from fastesp_amd import synth
from fastesp_amd.formats import basis_payload, read_pairs, write_basis_payload

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
GOLD = ROOT / "tests" / "golden"
REF_BUILDER = ROOT / "oracle" / "_ref" / "build_pairs"


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def run(*args, timeout=240):
    say("python -m fastesp_amd " + " ".join(str(a) for a in args))
    return subprocess.run([sys.executable, "-m", "fastesp_amd", *map(str, args)], cwd=ROOT,
                          capture_output=True, text=True, timeout=timeout)


def read_out(path):
    lines = Path(path).read_text().splitlines()
    n = int(lines[0])
    rows = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    assert rows.shape == (n, 4)
    return rows


@pytest.fixture(scope="module")
def cfg0_basis(tmp_path_factory):
    w = synth.config0()
    d = tmp_path_factory.mktemp("cli")
    b = d / "basis.bin"
    write_basis_payload(b, basis_payload(w.wfn, w.basis, w.density, w.points))
    return w, b


def test_esp_driver_matches_golden(cfg0_basis, tmp_path):
    w, b = cfg0_basis
    g = np.load(GOLD / "cfg0.npz")
    out = tmp_path / "out.txt"
    r = run("esp", b, out, "--repeats", 2)
    assert r.returncode == 0, r.stderr
    assert sum(l.startswith("run ") for l in r.stdout.splitlines()) == 2
    rows = read_out(out)
    np.testing.assert_allclose(rows[:, :3], w.points, rtol=0, atol=1e-11)
    # the text format keeps 15 significant digits of V
    assert np.abs(rows[:, 3] - g["V_tol12"]).max() < 1e-10


def test_build_pairs_then_esp_gpu(cfg0_basis, tmp_path):
    w, b = cfg0_basis
    g = np.load(GOLD / "cfg0.npz")
    pairs, out = tmp_path / "pairs.bin", tmp_path / "out.txt"
    r = run("build-pairs", b, pairs, "0")
    assert r.returncode == 0, r.stderr
    assert "sum rule Tr(P S) =" in r.stderr
    flat = read_pairs(pairs)
    r = run("esp-gpu", pairs, out, "--repeats", 3)
    assert r.returncode == 0, r.stderr
    head = r.stdout.splitlines()[0]
    assert head.startswith(f"pairs={flat.n_pairs} terms={flat.n_terms} points={len(w.points)}")
    times = [float(l.split()[-2]) for l in r.stdout.splitlines() if l.startswith("run ")]
    assert len(times) == 3 and min(times) > 0
    assert np.abs(read_out(out)[:, 3] - g["V_tol0"]).max() < 1e-10


@pytest.mark.skipif(not REF_BUILDER.exists(), reason="reference builder not compiled")
def test_esp_gpu_reads_reference_built_pairs(cfg0_basis, tmp_path):
    w, b = cfg0_basis
    g = np.load(GOLD / "cfg0.npz")
    pairs, out = tmp_path / "pairs.bin", tmp_path / "out.txt"
    say("reference build_pairs tol=0")
    r = subprocess.run([str(REF_BUILDER), str(b), str(pairs), "0"], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    r = run("esp-gpu", pairs, out, "--quiet")
    assert r.returncode == 0, r.stderr
    assert r.stdout == ""
    assert np.abs(read_out(out)[:, 3] - g["V_tol0"]).max() < 1e-10



def test_molden_driver_matches_oracle(tmp_path):
    """``fastesp esp WF.molden --points pts.txt -o esp.txt`` (``fastesp/cli.py:
    118-200``): the reference's Cartesian density of the MOLDEN fixture and
    the oracle's FP64 evaluation, points in Bohr and in Angstrom."""
    from fastesp_amd.basis import BOHR
    from fastesp_amd.molden import read_molden
    from oracle import engine, evaluate
    ref = np.load(GOLD / "molden_ref.npz")
    w = read_molden(GOLD / "toy.molden")
    from fastesp_amd.basis import Basis
    b = Basis(w)
    pts = np.random.default_rng(11).normal(0.0, 3.0, (200, 3)) + [0.4, 0.6, 0.0]
    want = evaluate.esp(engine.build_table(w, b, ref["P"], pts, tol=1e-12))
    for geom, scale in (("bohr", 1.0), ("angstrom", BOHR)):
        pf = tmp_path / f"pts_{geom}.txt"
        np.savetxt(pf, pts * scale, fmt="%.17g")
        out = tmp_path / f"esp_{geom}.txt"
        args = ["esp", GOLD / "toy.molden", "--points", pf, "-o", out]
        if geom == "angstrom":
            args.append("--angstrom")
        r = run(*args)
        assert r.returncode == 0, r.stderr
        for key in ("backend      : hip", "basis        : 44 Cartesian functions",
                    "pair table   :", "ESP solve    :", "ESP range    :", "wrote        :"):
            assert key in r.stdout
        rows = read_out(out)
        np.testing.assert_allclose(rows[:, :3], pts, rtol=0, atol=1e-11)
        assert np.abs(rows[:, 3] - want).max() < 1e-10


def test_selftest_passes():
    """``fastesp selftest`` (``fastesp/cli.py:230-282``) with the device integrals."""
    r = run("selftest")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SELFTEST PASS" in r.stdout


def test_build_native_reports_device():
    r = run("build-native")
    assert r.returncode == 0, r.stderr
    assert "HIP device     : yes" in r.stdout
