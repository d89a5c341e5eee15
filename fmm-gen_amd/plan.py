"""Job specification and partition arithmetic for one FMM call.

A job is what the reference CLI names with ``N a1 L1 ... aN LN pack_type``
(meta/control.py:120-142): algorithm ``a_i`` applied ``L_i`` times, levels in
command-line order.  Its jobname is ``<a1>-<L1>_..._<pack>``
(meta/control.py:47).

Partition (meta/common_mix.py:124-171, common/dgemm/bl_dgemm_util.c:316-353):
with Mx = prod mx, Kx = prod kx, Nx = prod nx over all levels,

    mr = m % (Mx*MR), kr = k % Kx, nr = n % (Nx*NR),  MR = 8, NR = 4
    md = m - mr,      kd = k - kr, nd = n - nr
    leaf block ms x ks x ns = md/Mx x kd/Kx x nd/Nx

The FMM covers C[0:md, 0:nd] with A[0:md, 0:kd] B[0:kd, 0:nd]; the fringe
is peeled with three plain GEMMs (k fringe into the core, right columns, bottom
rows).  MR and NR are the reference's micro-tile (common/include/bl_config.h:
16-17); they are kept as partition granularity so the FMM-covered region is the
same as the reference's for every shape.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence, Tuple

from .alg_dic import lookup
from .coeffs import Algorithm, compose, load_algorithm

MR = 8
NR = 4

PACK_TYPES = ("abc", "ab", "naive")


def jobname(algs: Sequence[str], levels: Sequence[int], pack_type: str) -> str:
    """meta/control.py:47 — e.g. ('222',), (1,), 'abc' -> '222-1_abc'."""
    return "_".join(f"{a}-{l}" for a, l in zip(algs, levels)) + "_" + pack_type


def build_algorithm(algs: Sequence[str], levels: Sequence[int]) -> Algorithm:
    """Compose the level list the way gen_abc_fmm does (meta/abc_gen_mix.py:409-431):
    each named algorithm repeated L times, in order."""
    if len(algs) != len(levels) or not algs:
        raise ValueError("need one level count per algorithm")
    per_level: List[Algorithm] = []
    for a, L in zip(algs, levels):
        if int(L) < 0:
            raise ValueError(f"negative level count {L} for {a}")
        path, dims = lookup(a)
        base = load_algorithm(path, dims, name=a)
        per_level += [base] * int(L)
    if not per_level:
        raise ValueError("all level counts are zero")
    return compose(per_level, name="_".join(f"{a}-{l}" for a, l in zip(algs, levels)))


@dataclass(frozen=True)
class Partition:
    m: int
    n: int
    k: int
    Mx: int
    Kx: int
    Nx: int
    md: int
    kd: int
    nd: int

    @property
    def ms(self) -> int:
        return self.md // self.Mx

    @property
    def ks(self) -> int:
        return self.kd // self.Kx

    @property
    def ns(self) -> int:
        return self.nd // self.Nx

    @property
    def fringe(self) -> Tuple[int, int, int]:
        return (self.m - self.md, self.k - self.kd, self.n - self.nd)


def partition(m: int, n: int, k: int, dims: Tuple[int, int, int]) -> Partition:
    Mx, Kx, Nx = dims
    md = m - m % (Mx * MR)
    kd = k - k % Kx
    nd = n - n % (Nx * NR)
    return Partition(m, n, k, Mx, Kx, Nx, md, kd, nd)


def fmm_flops(alg: Algorithm, m: int, n: int, k: int) -> int:
    """Multiply-add flops actually executed: R*2*ms*ks*ns on the FMM core plus
    2*(mnk - md*nd*kd) classical on the fringe."""
    p = partition(m, n, k, alg.dims)
    return alg.rank * 2 * p.ms * p.ks * p.ns + 2 * (m * n * k - p.md * p.nd * p.kd)
