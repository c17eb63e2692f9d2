"""Write the MOLDEN fixtures (run on the CPU, in a container holding the reference).

    python tests/golden/make_molden_golden.py

The pure -> Cartesian transform carries a phase convention (order m = 0, +1,
-1, ... and the signs of the f(+3) / f(-3) functions) that no mathematical
invariant fixes: T S T^T = 1, Tr(P S) = N_elec and the quadrature of the
density hold under any phases.  So this script runs the reference's own MOLDEN
path (``fastesp/molden.py``, ``fastesp/basis.py``) on a MOLDEN file written by
this package and stores its outputs:

  toy.molden          input: C, O, H with s/p/d/f shells and pure MOs
                      (written by ``fastesp_amd.molden.write_molden``)
  molden_ref.npz      T_l = pure_to_cartesian(l), l = 0..4; the Cartesian
                      overlap blocks; from read_molden: coords, mo_occ, the MO
                      matrix shape and the spherical flags; from
                      Basis.density_matrix_cartesian(): P (Cartesian)
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent
REF = Path("/root/reference")


def make_input() -> Path:
    sys.path.insert(0, str(ROOT))
    from fastesp_amd import synth
    from fastesp_amd.basis import Shell, Wavefunction
    from fastesp_amd.molden import write_molden

    def sh(atom, l, exps, raw):
        e = np.asarray(exps, float)
        return Shell(atom=atom, l=l, exps=e, coeffs=synth.normalised_coeffs(l, e, np.asarray(raw, float)))

    wfn = Wavefunction(
        symbols=["C", "O", "H"],
        coords=np.array([[0.0, 0.0, 0.0], [2.2, 0.4, -0.3], [-1.1, 1.6, 0.5]]),
        shells=[sh(0, 0, [9.0, 1.5], [0.4, 0.6]), sh(0, 1, [2.1, 0.5], [0.5, 0.5]),
                sh(0, 2, [0.8], [1.0]), sh(0, 3, [0.7], [1.0]),
                sh(1, 0, [12.0, 2.0], [0.5, 0.5]), sh(1, 1, [1.4], [1.0]),
                sh(1, 2, [1.1], [1.0]), sh(1, 3, [0.9], [1.0]),
                sh(2, 0, [1.3, 0.3], [0.6, 0.5]), sh(2, 1, [0.9], [1.0])])
    synth.pure_orbitals(wfn, 16, np.random.default_rng(2024))
    return write_molden(OUT / "toy.molden", wfn, title="molden fixture")


def main() -> int:
    path = make_input()
    sys.path.insert(0, str(REF))
    from fastesp.basis import Basis, cartesian_overlap_block, pure_to_cartesian
    from fastesp.molden import read_molden

    w = read_molden(path)
    b = Basis(w)
    P = b.density_matrix_cartesian()
    out = {f"T{l}": pure_to_cartesian(l) for l in range(5)}
    out.update({f"S{l}": cartesian_overlap_block(l) for l in range(5)})
    out.update(coords=w.coords, mo_occ=w.mo_occ, mo_shape=np.array(w.mo_coeff.shape),
               spherical=np.array([bool(w.spherical.get(l, False)) for l in range(5)]),
               shell_l=np.array([s.l for s in w.shells]), P=P, n_func=np.array(b.n))
    np.savez_compressed(OUT / "molden_ref.npz", **out)
    print(f"wrote {path.name} and molden_ref.npz (n_func {b.n}, Tr-weight {w.mo_occ.sum()})")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
