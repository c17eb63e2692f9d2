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
