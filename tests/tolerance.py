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
