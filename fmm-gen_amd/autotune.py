"""Per-problem choice of algorithm, level count and C-tile shape.

The cost model is the reference's (meta/model_gen.py:65-102: time = flops x
tau + memory operations x alpha, with per-product nnz counters from
meta/model_coefficient.py:56-100), restated for the unit structure of the
gfx950 kernels: a launch is a set of *units* (one product of an algorithm on
one BM x BN C tile over the whole leaf depth), each costing

    t_unit = t_fix + depth * (c_mfma * BM*BN + c_ab * ((nnzU_r-1)*BM + (nnzV_r-1)*BN))
             + c_c * nnzW_r * BM*BN

(the unit's MFMA work at the rate a plain-GEMM unit achieves, loads
included; the loads and adds of the extra operand-sum terms of product r --
the reference model's M_A_add / M_B_add counters, sum_r (nnzU_r - 1); the
read-modify-write of the C tiles the product feeds).  A problem's cost is the sum over its units (FMM core
products, k-fringe products, the two fringe GEMMs), with partial tiles
charged whole -- the kernels compute zero-padded tiles.  A persistent launch
spreads units over all CUs, so the choice that minimises a problem's summed
unit time minimises its share of the batch.

Seeds: c_mfma = 2 / (MFMA peak per CU), c_ab = element bytes / (HBM
bandwidth per CU), c_c = 2 x C element bytes / (HBM bandwidth per CU),
t_fix = 2 us.  With the extra-term feature a plain GEMM has no c_ab term, so
the fit separates the GEMM rate (c_mfma) from the FMM-only costs.  ``calibrate()`` refits the four constants on the device by
non-negative least squares over timed batches (units spread over all CUs).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .plan import MR, NR, build_algorithm

CANDIDATE_ALGS = {
    "gemm": None,
    "222L1": (["222"], [1]),
    "222L2": (["222"], [2]),
    "232L1": (["232"], [1]),
    "333L1": (["333"], [1]),
}

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0, "f16": 2500.0, "bf16x3": 2500.0 / 3.0}
ELEM_BYTES = {"f64": 8, "f32": 4, "bf16": 2, "f16": 2, "bf16x3": 4}
C_BYTES = {"f64": 8, "f32": 4, "bf16": 4, "f16": 4, "bf16x3": 4}
# C-tile shapes (BM, BN) per tile index; the kernels of these types.
TILES = {"f64": [(128, 128)], "f32": [(256, 128), (128, 256)], "bf16": [(256, 128), (128, 256)],
         "f16": [(256, 128), (128, 256)], "bf16x3": [(256, 128)]}
HBM_GBS = 8000.0
CUS = 256


@dataclass
class AlgCounts:
    """Per-product operand term counts and the partition of an algorithm."""
    dims: Tuple[int, int, int]
    nu: List[int]
    nv: List[int]
    nw: List[int]

    @property
    def rank(self) -> int:
        return len(self.nu)


def alg_counts(spec) -> AlgCounts:
    if spec is None:
        return AlgCounts((1, 1, 1), [1], [1], [1])
    alg = build_algorithm(*spec)
    U, V, W = alg.as_arrays()
    return AlgCounts(alg.dims, [int(x) for x in (U != 0).sum(0)], [int(x) for x in (V != 0).sum(0)],
                     [int(x) for x in (W != 0).sum(0)])


def _tiles(x: int, t: int) -> int:
    return (x + t - 1) // t


def unit_features(m: int, n: int, k: int, ac: AlgCounts, bm: int, bn: int) -> np.ndarray:
    """Summed model features of one problem: [units, sum depth*BM*BN,
    sum depth*((nu-1)*BM + (nv-1)*BN), sum nw*BM*BN], following the job split
    of a single call (fringe GEMMs, FMM core, k-fringe products)."""
    f = np.zeros(4)

    def add(units, depth, nu, nv, nw):
        f[0] += units
        f[1] += units * depth * bm * bn
        f[2] += units * depth * ((nu - 1) * bm + (nv - 1) * bn)
        f[3] += units * nw * bm * bn

    Mx, Kx, Nx = ac.dims
    if m <= 0 or n <= 0 or k <= 0:
        return f
    if ac.dims == (1, 1, 1) and ac.rank == 1:
        add(_tiles(m, bm) * _tiles(n, bn), k, 1, 1, 1)
        return f
    md, kd, nd = m - m % (Mx * MR), k - k % Kx, n - n % (Nx * NR)
    ms, ks, ns = md // Mx, kd // Kx, nd // Nx
    if n - nd > 0:
        add(_tiles(m, bm) * _tiles(n - nd, bn), k, 1, 1, 1)
    if m - md > 0 and nd > 0:
        add(_tiles(m - md, bm) * _tiles(nd, bn), k, 1, 1, 1)
    if ms <= 0 or ns <= 0:
        return f
    if ks <= 0:
        add(_tiles(md, bm) * _tiles(nd, bn), k, 1, 1, 1)
        return f
    tl = _tiles(ms, bm) * _tiles(ns, bn)
    for r in range(ac.rank):
        add(tl, ks, ac.nu[r], ac.nv[r], ac.nw[r])
    if k - kd > 0:
        add(tl * Mx * Nx, k - kd, 1, 1, 1)
    return f


@dataclass
class Model:
    dtype: str
    theta: np.ndarray = field(default=None)   # [t_fix, c_mfma, c_ab, c_c] in seconds
    calibrated: bool = False

    def __post_init__(self):
        if self.theta is None:
            per_cu_flops = PEAK_TFLOPS[self.dtype] * 1e12 / CUS
            bw = HBM_GBS * 1e9 / CUS
            self.theta = np.array([2e-6, 2.0 / per_cu_flops, ELEM_BYTES[self.dtype] / bw,
                                   2.0 * C_BYTES[self.dtype] / bw])

    def cost(self, feats: np.ndarray) -> float:
        return float(feats @ self.theta)


class Tuner:
    """Chooses (algorithm, tile) per problem by the model; plans are built
    once per algorithm."""

    def __init__(self, dtype: str = "f32", algs: Sequence[str] = tuple(CANDIDATE_ALGS), model: Optional[Model] = None):
        self.dtype = dtype
        self.algs = list(algs)
        self.model = model or Model(dtype)
        self.counts = {a: alg_counts(CANDIDATE_ALGS[a]) for a in self.algs}
        self._plans: Dict[str, object] = {}

    def candidates(self):
        for a in self.algs:
            for t in range(len(TILES[self.dtype])):
                yield a, t

    def features(self, m, n, k, alg, tile) -> np.ndarray:
        bm, bn = TILES[self.dtype][tile]
        return unit_features(m, n, k, self.counts[alg], bm, bn)

    def choose(self, m: int, n: int, k: int) -> Tuple[str, int, float]:
        best = None
        for a, t in self.candidates():
            c = self.model.cost(self.features(m, n, k, a, t))
            if best is None or c < best[2]:
                best = (a, t, c)
        return best

    def plan(self, alg: str, library=None):
        """The runtime Plan of a candidate algorithm (None for the GEMM)."""
        if CANDIDATE_ALGS[alg] is None:
            return None
        if alg not in self._plans:
            from . import runtime as rt
            self._plans[alg] = rt.Plan(build_algorithm(*CANDIDATE_ALGS[alg]), library)
        return self._plans[alg]

    def batch(self, operands, split: bool = False, choices=None, library=None):
        """A runtime Batch of (A, B, C) operand triples with the tuned (or the
        given) per-problem choices; returns (batch, choices)."""
        from . import runtime as rt
        if choices is None:
            choices = [self.choose(A.shape[0], B.shape[1], A.shape[1])[:2] for A, B, C in operands]
        probs = [(A, B, C, self.plan(a, library), t) for (A, B, C), (a, t) in zip(operands, choices)]
        return rt.Batch(probs, split=split, library=library), choices

    def calibrate(self, device_time, shapes: Sequence[Tuple[int, int, int]] = None, target_flops: float = 2e11):
        """Refit the model constants.  ``device_time(m, n, k, alg, tile, copies)``
        must return the seconds of one persistent launch of `copies` copies of
        the problem (one batch).  Returns the fitted constants."""
        from scipy.optimize import nnls
        shapes = shapes or [(512, 512, 512), (1024, 1024, 1024), (2048, 2048, 2048), (4096, 4096, 256),
                            (256, 4096, 4096), (3000, 1500, 700), (192, 192, 4096)]
        rows, ys = [], []
        for (m, n, k) in shapes:
            copies = max(1, int(target_flops / (2.0 * m * n * k)))
            for a, t in self.candidates():
                feats = self.features(m, n, k, a, t) * copies
                sec = device_time(m, n, k, a, t, copies)
                rows.append(feats / CUS)
                ys.append(sec)
        X = np.array(rows)
        y = np.array(ys)
        scale = X.max(axis=0)
        scale[scale == 0] = 1.0
        th, _ = nnls(X / scale, y)
        th = th / scale
        self.model.theta = np.where(th > 0, th, self.model.theta * 1e-3)
        self.model.calibrated = True
        return self.model.theta
