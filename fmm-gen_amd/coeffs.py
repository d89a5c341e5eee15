"""[U,V,W] coefficient tables: parsing, multi-level composition, validation.

A fast-matrix-multiplication (FMM) algorithm <mx,kx,nx> of rank R is a triple
of coefficient matrices.  Row a of U names an A block, row b of V a B block,
row c of W a C block; column r is the r-th product

    M_r = (sum_a U[a,r] A_a) (sum_b V[b,r] B_b),   C_c += W[c,r] M_r.

File format (meta/common_util.py:84-104): whitespace-separated numbers, one
row per submatrix, one column per product; the three groups are separated by
lines starting with '#'.  A leading '#' line before any rows is a comment.

Row numbering (meta/common_mix.py:284-310, 94-106): U row i*kx+p is block
A(i,p), V row p*nx+j is B(p,j), W row i*nx+j is C(i,j).

Multi-level algorithms are Kronecker products of the per-level tables
(meta/common_mix.py:65-92, 324-341): composed row index = (earlier levels)
* rows(level) + row(level), composed product index likewise, so level 0 is
the most significant digit of both.

All coefficients in the shipped tables are dyadic rationals (0, +-1, +-1/2,
+-1/8); they are held as exact Fractions here, and every composed value is
exactly representable in FP64.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from fractions import Fraction
from typing import List, Sequence, Tuple

import numpy as np

ALG_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "algorithms")


class CoeffError(ValueError):
    """A coefficient table that cannot be used (bad syntax, shape, or value)."""


def parse_coeff(tok: str) -> Fraction:
    """Parse one table entry into an exact Fraction.

    The reference grammar (meta/common_util.py:37-82) also admits symbolic
    entries (a one-letter parameter ``x``, ``-x``, ``xi`` for 1/x, ``x2`` for
    x^2).  Those occur only in the approximate / parameterised tables the
    reference keeps under ``algorithms/ignore/`` and its generated C never binds
    the parameter (meta/common_util.py:56-60), so they are rejected here.
    """
    s = tok.strip()
    try:
        val = float(s)
    except ValueError:
        raise CoeffError(f"symbolic coefficient {tok!r} is not supported "
                         "(parameterised/approximate tables are out of scope)")
    if not np.isfinite(val):
        raise CoeffError(f"non-finite coefficient {tok!r}")
    return Fraction(s) if _is_decimal(s) else Fraction(val)


def _is_decimal(s: str) -> bool:
    try:
        Fraction(s)
        return True
    except (ValueError, ZeroDivisionError):
        return False


def read_coeffs(path: str) -> List[List[List[Fraction]]]:
    """Read a table file into [U_rows, V_rows, W_rows] (meta/common_util.py:84-104).

    Each group is a list of rows; each row a list of Fractions (one per
    product).  Blank lines are ignored.
    """
    groups: List[List[List[Fraction]]] = []
    cur: List[List[Fraction]] = []
    with open(path, "r") as fh:
        for line in fh:
            if line.startswith("#"):
                if cur:
                    groups.append(cur)
                    cur = []
                continue
            toks = line.split()
            if not toks:
                continue
            cur.append([parse_coeff(t) for t in toks])
    if cur:
        groups.append(cur)
    if len(groups) < 3:
        raise CoeffError(f"{path}: expected three coefficient groups, found {len(groups)}")
    return groups[:3]


@dataclass
class Algorithm:
    """One (possibly composed) FMM algorithm.

    ``U``, ``V``, ``W`` are integer-free exact tables in the reference's row
    order (level 0 = most significant digit); ``levels`` is the list of
    per-level (mx, kx, nx) partitions the rows are decoded with.
    """

    name: str
    levels: List[Tuple[int, int, int]]
    U: List[List[Fraction]]
    V: List[List[Fraction]]
    W: List[List[Fraction]]
    sources: List[str] = field(default_factory=list)

    @property
    def dims(self) -> Tuple[int, int, int]:
        mx = kx = nx = 1
        for (a, b, c) in self.levels:
            mx, kx, nx = mx * a, kx * b, nx * c
        return (mx, kx, nx)

    @property
    def rank(self) -> int:
        return len(self.U[0])

    def nnz(self) -> Tuple[int, int, int]:
        return tuple(sum(1 for row in T for x in row if x != 0) for T in (self.U, self.V, self.W))

    def as_arrays(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Dense float64 arrays (rows x R), reference row order.  Exact."""
        return tuple(np.array([[float(x) for x in row] for row in T], dtype=np.float64)
                     for T in (self.U, self.V, self.W))

    # --- row decoding --------------------------------------------------------
    def block_of_row(self, which: int, row: int) -> Tuple[int, int]:
        """Grid block (I, J) that composed row ``row`` of table ``which`` names.

        which = 0: U -> A(I=row block, J=k block); 1: V -> B(k block, n block);
        2: W -> C(m block, n block).  Digits are peeled off from the last level
        (least significant) as in meta/common_mix.py:296-310; inside each level
        digit d the block is (d // ncols, d % ncols).
        """
        a, b = {0: (0, 1), 1: (1, 2), 2: (0, 2)}[which]
        I = J = 0
        rbase = cbase = 1
        rest = row
        for lv in reversed(self.levels):
            nr, nc = lv[a], lv[b]
            d = rest % (nr * nc)
            rest //= nr * nc
            I += (d // nc) * rbase
            J += (d % nc) * cbase
            rbase *= nr
            cbase *= nc
        return I, J


def _check_shape(name: str, dims, U, V, W) -> None:
    mx, kx, nx = dims
    R = len(U[0]) if U else 0
    want = ((mx * kx, "U"), (kx * nx, "V"), (mx * nx, "W"))
    for (rows, lab), T in zip(want, (U, V, W)):
        if len(T) != rows:
            raise CoeffError(f"{name}: {lab} has {len(T)} rows, expected {rows} for <{mx},{kx},{nx}>")
        for row in T:
            if len(row) != R:
                raise CoeffError(f"{name}: ragged {lab} table (row of {len(row)} entries, rank {R})")
    if R == 0:
        raise CoeffError(f"{name}: empty table")


def load_algorithm(path: str, dims: Tuple[int, int, int], name: str | None = None) -> Algorithm:
    U, V, W = read_coeffs(path)
    nm = name or os.path.basename(path)
    _check_shape(nm, dims, U, V, W)
    return Algorithm(nm, [tuple(dims)], U, V, W, [path])


def phantom_kron(X: List[List[Fraction]], Y: List[List[Fraction]]) -> List[List[Fraction]]:
    """Kronecker product of two tables in row-list form (rows x products).

    Row (x*rows(Y) + y), column (r*cols(Y) + s) = X[x][r] * Y[y][s]; the same
    index convention as meta/common_mix.py:65-92 (which works on transposes).
    """
    ry, cy = len(Y), len(Y[0])
    out = [[Fraction(0)] * (len(X[0]) * cy) for _ in range(len(X) * ry)]
    for x, xrow in enumerate(X):
        for r, xv in enumerate(xrow):
            if xv == 0:
                continue
            for y, yrow in enumerate(Y):
                orow = out[x * ry + y]
                base = r * cy
                for s, yv in enumerate(yrow):
                    if yv != 0:
                        orow[base + s] = xv * yv
    return out


def compose(algs: Sequence[Algorithm], name: str | None = None) -> Algorithm:
    """Compose per-level algorithms (level 0 first) into one, as generateCoeffs
    does (meta/common_mix.py:324-341)."""
    if not algs:
        raise CoeffError("compose() needs at least one level")
    U, V, W = algs[0].U, algs[0].V, algs[0].W
    levels = list(algs[0].levels)
    for a in algs[1:]:
        U, V, W = phantom_kron(U, a.U), phantom_kron(V, a.V), phantom_kron(W, a.W)
        levels += list(a.levels)
    nm = name or "_x_".join(a.name for a in algs)
    srcs = [s for a in algs for s in a.sources]
    return Algorithm(nm, levels, U, V, W, srcs)


def brent_residual(alg: Algorithm) -> int:
    """Number of violated Brent equations (0 = the table is an exact algorithm).

    Sum_r U[(i,p),r] V[(p',j),r] W[(i',j'),r] must equal
    [i=i'][p=p'][j=j'] for every index combination, where (i,p) etc. are the
    grid blocks each composed row decodes to.  Evaluated in exact integer
    arithmetic after scaling every table by its power-of-two denominator.
    """
    mx, kx, nx = alg.dims
    Ug, Vg, Wg = grid_tables(alg)
    def scaled(T: List[List[Fraction]]):
        den = 1
        for row in T:
            for x in row:
                den = max(den, x.denominator)
        arr = np.array([[int(x * den) for x in row] for row in T], dtype=np.int64)
        return arr, den
    Ui, du = scaled(Ug)
    Vi, dv = scaled(Vg)
    Wi, dw = scaled(Wg)
    # T[a,b,c] = sum_r U[a,r] V[b,r] W[c,r] in int64 (values stay tiny)
    T = np.einsum("ar,br,cr->abc", Ui, Vi, Wi, optimize=True)
    scale = du * dv * dw
    want = np.zeros_like(T)
    for i in range(mx):
        for p in range(kx):
            for j in range(nx):
                want[i * kx + p, p * nx + j, i * nx + j] = scale
    return int(np.count_nonzero(T != want))


def grid_tables(alg: Algorithm):
    """Tables re-ordered so row = I*cols + J in the composed block grid."""
    mx, kx, nx = alg.dims
    out = []
    for which, (T, ncols) in enumerate(((alg.U, kx), (alg.V, nx), (alg.W, nx))):
        G: List[List[Fraction]] = [None] * len(T)  # type: ignore
        for row in range(len(T)):
            I, J = alg.block_of_row(which, row)
            G[I * ncols + J] = T[row]
        out.append(G)
    return out
