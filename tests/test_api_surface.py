"""The package's public names match the reference's (CPU: no device calls).

The reference exports ``boys`` and ``primitive_pair_integral``
(``fastesp/__init__.py:29,38``); its ``PairTable`` has the flat-table fields,
``max_mono``, ``screen(tol)`` and the evaluation methods
(``fastesp/table.py:41-116``).
"""
import inspect

import fastesp_amd
from fastesp_amd.formats import FlatTable
from fastesp_amd.table import PairTable

REFERENCE_ALL = ["Wavefunction", "read_molden", "Basis", "cart_order", "overlap_matrix", "boys",
                 "primitive_pair_integral", "PairTable", "build_table", "basis_payload",
                 "write_basis_payload", "__version__"]


def test_reference_names_are_exported():
    assert set(REFERENCE_ALL) <= set(fastesp_amd.__all__)
    for name in REFERENCE_ALL:
        assert hasattr(fastesp_amd, name), name


def test_signatures_follow_the_reference():
    p = list(inspect.signature(fastesp_amd.primitive_pair_integral).parameters)
    assert p == ["alpha", "A", "la", "beta", "B", "lb", "C", "a_tab", "E"]
    assert list(inspect.signature(fastesp_amd.boys).parameters) == ["m", "T"]
    b = inspect.signature(fastesp_amd.build_table).parameters
    assert list(b)[:6] == ["wfn", "basis", "density", "points", "tol", "progress"]
    assert b["tol"].default == 1e-12


def test_pair_table_has_the_reference_fields():
    fields = [f for f in FlatTable.__dataclass_fields__ if f != "meta"]
    for f in fields:
        assert hasattr(PairTable, f) or f in ("pts", "Z", "xyz"), f
    for m in ("electron", "nuclear", "esp", "screen", "write", "read", "n_pairs", "n_terms",
              "max_mono"):
        assert hasattr(PairTable, m), m
    assert list(inspect.signature(PairTable.screen).parameters) == ["self", "tol"]
