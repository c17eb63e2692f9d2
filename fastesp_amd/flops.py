"""Algorithmic work model of the ESP path (the single definition used for the
roofline figures).

Forward, per evaluated (point, pair) of class L:
    F_fwd(L) = 9 (R, T) + B(L) + (nmono(L) - 1) + 3 nterm(L) + 2,
    B(M) = 20 + 3M  (one exp/Taylor seed counted as 20, 3 per recursion step)
    -> 34, 49, 79, 131, 215, 341, 522 for L = 0..6;  nuclear: 11 per (point, atom).
Backward (field + dL/dG), per (point, pair):
    F_bwd(L) = 19 + B(L+1) + (nmono(L) - 1) + 14 nterm(L);  nuclear: 17 per (point, atom).
Mixed precision: (nmono - 1) + 3 nterm of the forward (and of the backward
contraction) are FP32, the rest FP64.

Bytes per pair record of the device table: 8 * STRIDE(L) (header, g and the
integer multiples c g the field contraction reads).
"""
from __future__ import annotations

NTERM = [1, 4, 11, 24, 46, 80, 130]
NMONO = [1, 4, 10, 20, 35, 56, 84]
STRIDE = [16, 16, 24, 48, 56, 88, 144]

PEAK_FP64_VALU = 78.6e12   # MI355X FP64 vector, FLOP/s
PEAK_FP32_PACKED = 157.3e12  # MI355X FP32 vector with packed (2-wide) FMAs, FLOP/s
# FP32 as this library is compiled (packed FP32 disabled): one FMA per lane per
# issue, the FP64 rate (fe_valu_rate measures both on the device)
PEAK_FP32_VALU = PEAK_FP64_VALU
PEAK_HBM = 8.0e12          # B/s


def boys_cost(M: int) -> int:
    return 20 + 3 * M


def f_fwd(L: int) -> int:
    return 9 + boys_cost(L) + (NMONO[L] - 1) + 3 * NTERM[L] + 2


def f_bwd(L: int) -> int:
    return 19 + boys_cost(L + 1) + (NMONO[L] - 1) + 14 * NTERM[L]


def f_eval_grad(L: int) -> int:
    """Evaluation launch that also forms the field (value + dV/dC in one pass):
    the forward plus the field's 9/term and the Boys-shift 3/term of F_bwd."""
    return f_fwd(L) + 12 * NTERM[L]


def f_dg(L: int) -> int:
    """dL/dG launch: the rest of F_bwd (its own R, T, Boys and monomials, 2/term)."""
    return f_bwd(L) - 12 * NTERM[L]


def f32_part(L: int, grad: bool) -> int:
    """FP32 share of a mixed evaluation launch per (point, pair): monomials and
    the per-term products of the value (3/term) and, with the field, 9/term more."""
    return (NMONO[L] - 1) + (12 if grad else 3) * NTERM[L]


def kernel_flops(kind: str, L: int, n_pairs_L: int, n_points: int) -> float:
    """Algorithmic FLOPs of one class launch: kind 'fwd', 'fwd_grad' or 'dg'."""
    f = {"fwd": f_fwd, "fwd_grad": f_eval_grad, "dg": f_dg}[kind](L)
    return float(n_points) * float(n_pairs_L) * f


def forward_flops(pairs_per_L, n_points: int, n_atoms: int) -> float:
    return float(n_points) * (sum(n * f_fwd(L) for L, n in enumerate(pairs_per_L))
                              + 11 * n_atoms)


def backward_flops(pairs_per_L, n_points: int, n_atoms: int) -> float:
    return float(n_points) * (sum(n * f_bwd(L) for L, n in enumerate(pairs_per_L))
                              + 17 * n_atoms)


def fp32_share_fwd(pairs_per_L) -> float:
    """Fraction of the forward pair work that the mixed path does in FP32."""
    tot = sum(n * f_fwd(L) for L, n in enumerate(pairs_per_L))
    f32 = sum(n * ((NMONO[L] - 1) + 3 * NTERM[L]) for L, n in enumerate(pairs_per_L))
    return f32 / tot if tot else 0.0


def table_bytes(pairs_per_L) -> int:
    return sum(8 * STRIDE[L] * n for L, n in enumerate(pairs_per_L))


assert [f_fwd(L) for L in range(7)] == [34, 49, 79, 131, 215, 341, 522]
assert all(f_eval_grad(L) + f_dg(L) == f_fwd(L) + f_bwd(L) for L in range(7))
