"""Command-line front end without a GPU: argument handling and error exits
(the native tools' convention: message on stderr, exit status 1)."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def run(*args):
    return subprocess.run([sys.executable, "-m", "fastesp_amd", *map(str, args)], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_help_lists_the_tools():
    r = run("--help")
    assert r.returncode == 0
    for name in ("build-pairs", "esp-gpu", "esp"):
        assert name in r.stdout


def test_truncated_files_exit_1(tmp_path):
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\x01\x02")
    for args in (("esp-gpu", bad, tmp_path / "o.txt"), ("build-pairs", bad, tmp_path / "p.bin"),
                 ("esp", bad, tmp_path / "o.txt")):
        r = run(*args)
        assert r.returncode == 1, args
        assert "truncated" in r.stderr


def test_usage_error_exit_nonzero():
    r = run("esp-gpu")
    assert r.returncode != 0
    assert "usage" in r.stderr


def test_selftest_pieces_against_the_oracle():
    """The self-test's two-shell construction selects exactly one primitive
    integral, and its quadrature reproduces the oracle's primitive integral."""
    import numpy as np
    from fastesp_amd.selftest import pair_case, quad
    from oracle import engine, evaluate
    from oracle.ints import primitive_pair_integral
    A = np.array([0.3, -0.4, 0.2])
    B = A + np.array([1.1, 0.5, -0.6])
    C = np.array([0.1, 1.6, 1.1])
    for la, lb in (((2, 1, 0), (1, 1, 1)), ((0, 0, 3), (1, 2, 0))):
        want = primitive_pair_integral(0.45, A, la, 0.72, B, lb, C)
        wfn, basis, P = pair_case(A, la, 0.45, B, lb, 0.72)
        got = evaluate.electron(engine.build_table(wfn, basis, P, C[None], tol=0.0))[0]
        assert abs(got - want) <= 1e-13 * abs(want)
    q = quad(C, A, (1, 0, 0), 0.45, B, (0, 0, 0), 0.72, ngl=120, ncos=110, nphi=220)
    want = primitive_pair_integral(0.45, A, (1, 0, 0), 0.72, B, (0, 0, 0), C)
    assert abs(q - want) <= 1e-8 * abs(want)


def test_build_native_is_keyed_by_sources(tmp_path, monkeypatch):
    """build-native recompiles only when the source hash / target differ."""
    from fastesp_amd import build as fb
    calls = []
    monkeypatch.setattr(fb, "LIB", tmp_path / "libfastesp_hip.so")
    monkeypatch.setattr(fb, "STAMP", tmp_path / "libfastesp_hip.so.stamp")

    def fake_build(arch="gfx950", lmax=6, verbose=True):
        calls.append(arch)
        fb.LIB.write_bytes(b"x")
        fb.STAMP.write_text(fb.source_hash(arch, lmax) + "\n")
        return fb.LIB

    monkeypatch.setattr(fb, "build", fake_build)
    fb.build_native(quiet=True)
    fb.build_native(quiet=True)
    assert calls == ["gfx950"]
    fb.build_native(quiet=True, arch="gfx942")
    fb.build_native(quiet=True, force=True, arch="gfx942")
    assert calls == ["gfx950", "gfx942", "gfx942"]
    assert fb.source_hash("gfx950") != fb.source_hash("gfx942")
