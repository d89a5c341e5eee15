"""Stated FP64 tolerances, per algorithm level count.

Error metric: e = max|C_hat - C_exact| / (u * k * max|A| * max|B|), u = 2^-53,
with C_exact computed in extended precision (numpy longdouble).  Measured on
the CPU oracle (same product structure as the GPU kernel) at 576^3 and
1152^2 x 240: plain GEMM e <= 3.5, one level e <= 10.5, two levels e <= 18
for U[0,1) inputs (smaller for U[-1,1)).  The stated bounds leave ~3x margin;
a dropped term, wrong sign or wrong block gives e of order 1/u.
"""
import numpy as np

U64 = 2.0 ** -53
TOL_GEMM = 16.0
TOL_BY_LEVELS = {1: 32.0, 2: 128.0, 3: 512.0}


def tol_for(nlevels: int) -> float:
    return TOL_BY_LEVELS.get(nlevels, 4.0 ** (nlevels + 2))


def exact(A, B, C):
    ld = np.longdouble
    return C.astype(ld) + A.astype(ld) @ B.astype(ld)


def norm_err(got, A, B, C0):
    k = A.shape[1]
    ex = exact(A, B, C0)
    scale = U64 * max(k, 1) * max(np.abs(A).max(initial=0), 1e-300) * max(np.abs(B).max(initial=0), 1e-300)
    return float(np.abs(got.astype(np.longdouble) - ex).max(initial=0) / scale)


def rel_frob(got, A, B, C0):
    ex = exact(A, B, C0)
    den = float(np.sqrt((ex * ex).sum()))
    return float(np.sqrt(((got.astype(np.longdouble) - ex) ** 2).sum())) / max(den, 1e-300)


# Low-precision FMM (FP32 C).  Error metric: relative Frobenius error of the
# result against the FP64 product of the inputs as stored in the input type.
# Levels 0 (plain GEMM), 1, 2.  Measured with the numpy restatement of the
# kernels' rounding points (oracle.fmm_lowp) at 768^3 on U[-1,1) inputs:
#   f32     2.5e-8 / 1.3e-7 / 3.2e-7
#   bf16    2.5e-8 / 5.8e-3 / 1.2e-2     (GEMM products are exact in FP32)
#   f16     2.5e-8 / 7.7e-4 / 1.5e-3
#   bf16x3  3.8e-6 / 1.1e-5 / 2.2e-5
# The kernels accumulate in FP32 (the restatement in FP64), which adds up to
# ~1e-6 at the tested depths; the bounds below leave about 4x margin.
TOL_LOWP = {
    "f32": {0: 2e-6, 1: 4e-6, 2: 8e-6},
    "bf16": {0: 2e-6, 1: 2.5e-2, 2: 5e-2},
    "f16": {0: 2e-6, 1: 3e-3, 2: 6e-3},
    "bf16x3": {0: 2e-5, 1: 5e-5, 2: 1e-4},
}
# GPU result vs the restatement: FP32 accumulation differences only.
TOL_LOWP_VS_EMU = 4e-6


def tol_lowp(dtype: str, nlevels: int) -> float:
    return TOL_LOWP[dtype][min(nlevels, 2)]


# At depth k the kernels' FP32 accumulation adds a relative error that grows
# like sqrt(k) (random-sign terms); measured at 8192^3 (U[-1,1)): FP32 GEMM
# 1.6e-6 against the FP64 product.  The full-size tests add 4 * 2^-24 * sqrt(k)
# (2.2e-5 at k = 8192) to the per-level bound, and the same term bounds the
# GPU's distance from the restatement, which accumulates in FP64.
def acc_term(k: int) -> float:
    return 4.0 * 2.0 ** -24 * float(np.sqrt(max(k, 1)))


def tol_lowp_at(dtype: str, nlevels: int, k: int) -> float:
    return tol_lowp(dtype, nlevels) + acc_term(k)


def tol_lowp_vs_emu_at(k: int) -> float:
    return max(TOL_LOWP_VS_EMU, acc_term(k))
