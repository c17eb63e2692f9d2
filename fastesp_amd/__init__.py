"""fastesp_amd: MI355X-native backend for fastesp's ESP hot path.

The package exports the names of the reference package (``fastesp/__init__.py:
26-42``) that this path provides, so ``from fastesp_amd import read_molden,
Basis, build_table`` replaces ``from fastesp import ...``::

    wfn = read_molden("mol.molden.input")
    basis = Basis(wfn)
    density = basis.density_matrix_cartesian()
    table = build_table(wfn, basis, density, points_bohr)   # device table
    esp = table.esp()                                        # Hartree/e

Every evaluation runs in ``libfastesp_hip.so`` on the GPU; the library is
loaded on first use and a missing library or device raises HipBackendError.
"""
from .basis import Basis, Wavefunction, cart_order, overlap_matrix
from .esp import ESPEngine
from .formats import basis_payload, write_basis_payload
from .ints import boys, primitive_pair_integral
from .molden import read_molden
from .table import PairTable, build_table

__version__ = "0.1.0"

__all__ = [
    "Wavefunction", "read_molden",
    "Basis", "cart_order", "overlap_matrix",
    "boys", "primitive_pair_integral",
    "ESPEngine", "PairTable", "build_table",
    "basis_payload", "write_basis_payload",
    "__version__",
]
