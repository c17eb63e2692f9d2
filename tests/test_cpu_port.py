"""The C++/OpenMP restatement (oracle/cpu_port.cpp) against the NumPy oracle
and the compiled reference builder (CPU)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from fastesp_amd import synth
from fastesp_amd.formats import basis_payload, read_pairs, write_basis_payload

ROOT = Path(__file__).resolve().parents[1]
GOLD = ROOT / "tests" / "golden"
REF = ROOT / "oracle" / "_ref" / "build_pairs"

cpu_port = pytest.importorskip("oracle.cpu_port")
if not cpu_port.LIB.exists():
    pytest.skip("libcpu_port.so not built (make -C oracle)", allow_module_level=True)


def test_port_matches_numpy_oracle_on_config0():
    w = synth.config0()
    g = np.load(GOLD / "cfg0.npz")
    for key, tol in (("V_tol0", 0.0), ("V_tol12", 1e-12)):
        ct = cpu_port.CpuTable(w.wfn, w.basis, w.density, w.points, tol)
        assert ct.n_pairs == int(g["npairs_" + key[2:]])
        assert np.abs(ct.esp(w.points) - g[key]).max() < 1e-12
    assert ct.sum_rule == pytest.approx(w.wfn.nelec, abs=1e-9)


@pytest.mark.skipif(not REF.exists(), reason="reference builder not compiled")
def test_port_cpp_rule_matches_reference_builder(tmp_path):
    w = synth.config0()
    b, p = tmp_path / "b.bin", tmp_path / "p.bin"
    write_basis_payload(b, basis_payload(w.wfn, w.basis, w.density, w.points))
    subprocess.run([str(REF), str(b), str(p), "1e-9"], check=True, capture_output=True)
    ref = read_pairs(p)
    ct = cpu_port.CpuTable(w.wfn, w.basis, w.density, w.points, 1e-9, rule=1)
    assert ct.n_pairs == ref.n_pairs
    p6, g = ct.export()
    key = lambda e, x, y, z: (round(e, 10), round(x, 8), round(y, 8), round(z, 8))  # noqa: E731
    mine = sorted(key(*r[:4]) for r in p6)
    theirs = sorted(key(ref.pp_exp[i], ref.pp_cx[i], ref.pp_cy[i], ref.pp_cz[i])
                    for i in range(ref.n_pairs))
    assert mine == theirs
