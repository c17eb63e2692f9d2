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

Bytes per pair record of the device table: 8 * STRIDE(L) (header, g, the
integer multiples c g the field contraction reads and the far-field block).

Performed work (``performed_flops``): the model above prices every (point,
pair) as a tabulated-Boys evaluation with a term-by-term contraction.  The
kernels take the exact far-field form whenever a whole wavefront is far
(no table, no exponential) and contract with straight-line code that shares
subexpressions between terms.  The performed-work count prices each
evaluation by the path it took, counted by ``fe_eval_work`` with the same
launch decompositions, at the operations the algorithm of that path needs:
an FMA is 2 FLOPs, any other arithmetic operation 1; the contraction's
operations are counted in the generated code (``csrc/fe_lclass.inc``).  It is
the figure the roofline divides by time.
"""
from __future__ import annotations

import re
from functools import lru_cache
from pathlib import Path

NTERM = [1, 4, 11, 24, 46, 80, 130]
NMONO = [1, 4, 10, 20, 35, 56, 84]
STRIDE = [24, 32, 56, 112, 176, 288, 456]

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


# ---------------------------------------------------------------------------
# performed work
# ---------------------------------------------------------------------------
_INC = Path(__file__).resolve().parent / "csrc" / "fe_lclass.inc"
GRAD_ACC_LMAX = 2   # classes whose value + field contraction folds the weight in (fe_internal.h)
KELVIN_LMIN = 0     # classes whose far value + field wavefronts use the far-field block (fe_internal.h)
DG_SPLIT = [True, True, True, True, True, False, True]  # dL/dG with per-order scaling (fe_adjoint.h)


def op_flops(code: str) -> int:
    """FLOPs of straight-line code: 2 per fma(...), 1 per other arithmetic operator."""
    n_fma = len(re.findall(r"\bfmaf?\(", code))
    n_bin = len(re.findall(r"\s[-+*/]=?\s", code))
    return 2 * n_fma + n_bin


@lru_cache(maxsize=None)
def generated_flops(name: str, L: int) -> int:
    """FLOPs of one call of the generated device function ``name<L>``."""
    txt = _INC.read_text()
    m = re.search(r"\b" + re.escape(name) + r"<" + str(L) + r">\([^)]*\)\s*\{(.*?)\n\}", txt, re.S)
    if not m:
        raise KeyError(f"{name}<{L}> not in {_INC.name}")
    return op_flops(m.group(1))


R_T = 8  # R = C - P (3), r^2 = fma(x, x, fma(y, y, z z)) (5)
RSQRT = 9  # hardware estimate + third-order Newton step (fe_rsqrt_pos)


def boys_near(M: int) -> int:
    """Tabulated Boys: T = p r^2, grid index and offset, 9-term Taylor of F_M,
    and for M > 0 exp(-T) from the row's value and a degree-9 polynomial,
    2T, and one fma + one multiply per downward step."""
    b = 1 + 2 + 2 + 16
    return b if M == 0 else b + 19 + 1 + 3 * M


def boys_far(M: int) -> int:
    """Far field: 1/|R| (rsqrt), x = 1/(2p) |R|^-2 (2), two multiplies per
    upward step, and the weight K F_0 = K sqrt(pi/p)/2 |R|^-1 (1)."""
    return RSQRT + 2 + 2 * M + 1


def f_eval_path(L: int, grad: bool, far: bool) -> int:
    """Performed FLOPs of one (point, pair) evaluation of class L on one path."""
    M = L + 1 if grad else L
    if not grad and far and L >= KELVIN_LMIN:
        # |u| H(u): 1/|R| and its square (1), u (3), the product and the
        # accumulation (2); class 0: H = h_0
        return R_T + RSQRT + (2 if L == 0 else generated_flops("fe_far_value", L) + 6)
    if grad and far and L >= KELVIN_LMIN:
        # |u| H(u) and |u|^3 (grad H - (H + 2D) R): 1/|R| and its square and
        # cube (2), u (3), t = H + 2D (2), the four accumulations (2 + 3 x 4);
        # class 0: H = h_0 wave-uniform, -H |R|^-3 (1) and three fma
        c = 11 if L == 0 else generated_flops("fe_far_grad", L) + 21
        return R_T + RSQRT + c
    b = boys_far(M) if far else boys_near(M)
    if grad and L >= KELVIN_LMIN:
        # tabulated wavefronts of the classes with a far-field block: the
        # keyed contraction and the shared weighted accumulation (4 fma)
        return R_T + b + generated_flops("fe_contract_grad", L) + 8
    if grad:
        if L <= GRAD_ACC_LMAX:
            c = generated_flops("fe_contract_grad_acc", L)
        else:
            c = generated_flops("fe_contract_grad", L) + 8  # w I and w G accumulated (4 fma)
    else:
        c = generated_flops("fe_contract", L) + 2  # w s accumulated
    return R_T + b + c


def f_dg_path(L: int, far: bool) -> int:
    """Performed FLOPs of one (pair, point) step of the dL/dG launch of class L."""
    acc = generated_flops("fe_dg_accum", L)
    if DG_SPLIT[L]:
        w = (RSQRT + 2 + L) if far else (boys_near(L) + 2 * (L + 1))
    else:
        w = (RSQRT + 2 + 2 * L + 1 + (L + 1)) if far else (boys_near(L) + 1 + (L + 1))
    return R_T + w + acc


def performed_flops(kind: str, L: int, n_far: int, n_near: int) -> float:
    """Performed FLOPs of a class launch from its far/tabulated evaluation
    counts (``fe_eval_work``): kind 'fwd', 'fwd_grad' or 'dg'."""
    if kind == "dg":
        return float(n_far) * f_dg_path(L, True) + float(n_near) * f_dg_path(L, False)
    grad = kind == "fwd_grad"
    return float(n_far) * f_eval_path(L, grad, True) + float(n_near) * f_eval_path(L, grad, False)


assert [f_fwd(L) for L in range(7)] == [34, 49, 79, 131, 215, 341, 522]
assert all(f_eval_grad(L) + f_dg(L) == f_fwd(L) + f_bwd(L) for L in range(7))
