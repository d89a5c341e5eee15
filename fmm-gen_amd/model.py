"""Performance model of FMM variants (meta/model_coefficient.py:36-100 and
meta/model_gen.py:65-249), kept with the reference's constants, plus an
MI355X parameterisation that seeds the autotuner.

Counters per composed algorithm (model_coefficient.py:56-100):
    comp  = [R, sum_r (nnzU_r - 1), sum_r (nnzV_r - 1), nnzW]
    abc   = [nnzU, nnzV, nnzW, 0, 0, 0]
    ab    = [nnzU, nnzV, R,    0, 0, 3 nnzW]
    naive = [R,    R,    R,    nnzU + R, nnzV + R, 3 nnzW]
A model row is comp (4 entries) followed by one of the three (6 entries);
the plain GEMM row is [1,0,0,0,1,1,1,1,1,1] (model_gen.py:110).

Time model (model_gen.py:65-102): T = actual_flops * tau + mops * alpha with
    actual_flops = R*2*ms*ks*ns + Aadd*2*ms*ks + Badd*2*ks*ns + Cadd*2*ms*ns
    mops = NAmul*ms*ks*ceil(ns/nc) + NBmul*ns*ks + NCmul*2*ceil(ks/kc)*ms*ns
         + NAadd*ns*ks + NBadd*ns*ks + NCadd*ms*ns
and effective GFLOPS = 2mnk / T.  The reference's quirks are kept: the
leaf sizes are floor divisions (Python 2 '/', model_gen.py:81-83), N_A_add is
charged with the B panel size (model_gen.py:95), and the unused mops_A term
is dropped (model_gen.py:90-91).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import Dict, List, Sequence, Tuple

if __package__ in (None, ""):
    # run as a script (the reference's `python model_gen.py`): import through the package
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from __graft_entry__ import load_package  # noqa: E402
    load_package()
    from fmm_gen_amd.alg_dic import alg_dic, alg_set  # noqa: E402
    from fmm_gen_amd.coeffs import Algorithm, compose, load_algorithm  # noqa: E402
else:
    from .alg_dic import alg_dic, alg_set
    from .coeffs import Algorithm, compose, load_algorithm

GEMM_ROW = [1, 0, 0, 0, 1, 1, 1, 1, 1, 1]
PACKS = ("abc", "ab", "naive")


def gen_model_coefficient(alg: Algorithm) -> Tuple[List[int], List[int], List[int], List[int]]:
    """[comp, abc, ab, naive] counters of a composed algorithm."""
    comp = [0, 0, 0, 0]
    abc = [0] * 6
    ab = [0] * 6
    naive = [0] * 6
    R = alg.rank
    for r in range(R):
        nu = sum(1 for row in alg.U if row[r] != 0)
        nv = sum(1 for row in alg.V if row[r] != 0)
        nw = sum(1 for row in alg.W if row[r] != 0)
        abc[0] += nu
        ab[0] += nu
        naive[0] += 1
        naive[3] += nu + 1
        comp[1] += nu - 1
        abc[1] += nv
        ab[1] += nv
        naive[1] += 1
        naive[4] += nv + 1
        comp[2] += nv - 1
        abc[2] += nw
        ab[2] += 1
        ab[5] += 3 * nw
        naive[2] += 1
        naive[5] += 3 * nw
        comp[0] += 1
        comp[3] += nw
    return comp, abc, ab, naive


def model_row(alg: Algorithm, pack: str) -> List[int]:
    comp, abc, ab, naive = gen_model_coefficient(alg)
    return comp + {"abc": abc, "ab": ab, "naive": naive}[pack]


@dataclass(frozen=True)
class Machine:
    """Model constants.  tau: time per flop; alpha: time per element moved;
    mc/nc/kc: cache blocking of the packed panels."""
    name: str
    tau: float
    alpha: float
    mc: int = 96
    nc: int = 4096
    kc: int = 256
    overhead: float = 0.0   # fixed time per call (kernel launch)


# The reference's constants (model_gen.py:67-73): 3.54 GHz x 8 flop/cycle,
# 0.5 per element, one channel; the unit of time is the nanosecond, so the
# model's "GFLOPS" are flops per nanosecond.
REFERENCE = Machine("reference-avx-core", tau=1.0 / (8 * 3.54), alpha=0.5)

# MI355X, whole chip, in nanoseconds: FP64 MFMA 78.6 TFLOP/s -> tau in ns/flop;
# HBM3E 8 TB/s and 8-byte elements -> alpha in ns/element.  Blocking: the
# kernel's C tile is 128x128 and the whole leaf depth stays in one unit, so
# panels are re-read per 128-column tile (nc = 128) and C is updated once
# per product (kc = infinite depth -> one C pass).
MI355X_F64 = Machine("mi355x-f64", tau=1.0 / 78.6e3, alpha=8.0 / 8.0e3, mc=128, nc=128, kc=1 << 30,
                     overhead=8e3)
MI355X_F32 = Machine("mi355x-f32", tau=1.0 / 157.3e3, alpha=4.0 / 8.0e3, mc=128, nc=128, kc=1 << 30,
                     overhead=8e3)
MI355X_BF16 = Machine("mi355x-bf16", tau=1.0 / 2.5e6, alpha=2.0 / 8.0e3, mc=128, nc=128, kc=1 << 30,
                      overhead=8e3)


def get_model_gflops(m: int, n: int, k: int, dims: Sequence[int], level: int, coefficient: Sequence[float],
                     mach: Machine = REFERENCE) -> float:
    """Effective GFLOPS (flops per model time unit) of one variant; dims and
    level give the partition as in model_gen.py:65-102."""
    T = model_time(m, n, k, dims, level, coefficient, mach)
    return 2.0 * m * k * n / T


def model_time(m, n, k, dims, level, coefficient, mach: Machine = REFERENCE) -> float:
    (M_total_mul, M_A_add, M_B_add, M_C_add, N_A_mul, N_B_mul, N_C_mul, N_A_add, N_B_add,
     N_C_add) = coefficient
    Mx, Kx, Nx = dims[0] ** level, dims[1] ** level, dims[2] ** level
    ms, ks, ns = m // Mx, k // Kx, n // Nx
    actual = (M_total_mul * 2 * ms * ks * ns + M_A_add * 2 * ms * ks + M_B_add * 2 * ks * ns
              + M_C_add * 2 * ms * ns)
    mops_Bc = ns * ks
    mops_Ac = ms * ks * math.ceil(ns / mach.nc)
    mops_Cc = math.ceil(ks / mach.kc) * ms * ns
    mops_B = ns * ks
    mops_C = ms * ns
    mops = (N_A_mul * mops_Ac + N_B_mul * mops_Bc + N_C_mul * 2 * mops_Cc + N_A_add * mops_Bc
            + N_B_add * mops_B + N_C_add * mops_C)
    return actual * mach.tau + mops * mach.alpha + mach.overhead


def generate_coefficient_mat(algs: Sequence[str] = tuple(alg_set)) -> List[List[float]]:
    """Row 0 = GEMM; then for L in (1, 2) and pack in (abc, ab, naive) one row
    per algorithm, in the order of model_gen.py:104-130."""
    rows: Dict[int, List[float]] = {0: list(GEMM_ROW)}
    n = len(algs)
    for a_id, name in enumerate(algs):
        path, dims = alg_dic[name]
        base = load_algorithm(path, dims, name)
        for L in (1, 2):
            alg = compose([base] * L)
            comp, abc, ab, naive = gen_model_coefficient(alg)
            for p_id, cnt in enumerate((abc, ab, naive)):
                rows[1 + a_id + (3 * (L - 1) + p_id) * n] = comp + cnt
    return [rows[i] for i in range(1 + 6 * n)]


STEP = 240
SAMPLES = list(range(STEP, 12000 + 1, STEP))
HEADER = "dim,gemm," + ",".join(alg_set)


def generate_result_mat(coef: List[List[float]], test_type: str, mach: Machine = REFERENCE,
                        algs: Sequence[str] = tuple(alg_set)) -> List[List[float]]:
    out = []
    n = len(algs)
    for var_id, row in enumerate(coef):
        dims = (1, 1, 1) if var_id == 0 else alg_dic[algs[(var_id - 1) % n]][1]
        level = 1 if var_id - 1 < n * 3 else 2
        vals = []
        for s in SAMPLES:
            if test_type == "square":
                m = n_ = k = s
            elif test_type == "rankk":
                m, n_, k = 14400, 14400, s
            elif test_type == "fixk":
                m, n_, k = s, s, 1024
            else:
                raise ValueError(f"wrong test type {test_type!r}")
            vals.append(get_model_gflops(m, n_, k, dims, level, row, mach))
        out.append(vals)
    return out


MACHINES = {m.name: m for m in (REFERENCE, MI355X_F64, MI355X_F32, MI355X_BF16)}


def csv_suffix(mach: Machine) -> str:
    """The reference names its files ..._1core.csv (one AVX core); the MI355X
    parameterisations are named for the machine."""
    return "1core" if mach is REFERENCE else mach.name


def csv_files(out_dir: str, mach: Machine = REFERENCE) -> List[str]:
    """The reference's 18 CSVs: model_<shape>_<L>level_<pack>_<suffix>.csv,
    rows = sizes, columns = dim, gemm, 23 algorithms (model_gen.py:177-249)."""
    os.makedirs(out_dir, exist_ok=True)
    coef = generate_coefficient_mat()
    n = len(alg_set)
    written = []
    for shape in ("square", "rankk", "fixk"):
        res = generate_result_mat(coef, shape, mach)
        for i, (L, pack) in enumerate([(1, "abc"), (1, "ab"), (1, "naive"), (2, "abc"), (2, "ab"), (2, "naive")]):
            path = os.path.join(out_dir, f"model_{shape}_{L}level_{pack}_{csv_suffix(mach)}.csv")
            off = i * n
            with open(path, "w") as fh:
                fh.write(HEADER + "\n")
                for j, s in enumerate(SAMPLES):
                    cols = [s, res[0][j]] + [res[1 + off + a][j] for a in range(n)]
                    fh.write(",".join(f"{c:.2f}" for c in cols) + "\n")
            written.append(path)
    return written


def check_against_sweep(rows: Sequence[Dict[str, float]], mach: Machine = MI355X_F64) -> Dict[str, dict]:
    """Model against measured sweep points (rows of a sweep CSV, sweep.py):
    per variant '<alg>-<L>_<pack>' whose algorithm is in the registry, the
    median and range of measured / modelled effective GFLOPS, and the share
    of points where model and measurement agree on whether the variant beats
    the plain GEMM."""
    out: Dict[str, dict] = {}
    if not rows:
        return out
    labels = [c for c in rows[0] if c not in ("m", "n", "k", "gemm") and "/" not in c]
    for lab in labels:
        name, rest = lab.split("-", 1)
        level_s, pack = rest.split("_", 1)
        if name not in alg_dic or pack not in PACKS:
            continue
        path, dims = alg_dic[name]
        alg = compose([load_algorithm(path, dims, name)] * int(level_s))
        coef = model_row(alg, pack)
        ratios, agree = [], 0
        for r in rows:
            m, n, k = int(r["m"]), int(r["n"]), int(r["k"])
            mod = get_model_gflops(m, n, k, dims, int(level_s), coef, mach)
            mod_gemm = get_model_gflops(m, n, k, (1, 1, 1), 1, GEMM_ROW, mach)
            ratios.append(r[lab] / mod)
            agree += (r[lab] > r["gemm"]) == (mod > mod_gemm)
        ratios.sort()
        out[lab] = {"measured_over_model_median": round(ratios[len(ratios) // 2], 3),
                    "measured_over_model_min": round(ratios[0], 3),
                    "measured_over_model_max": round(ratios[-1], 3),
                    "winner_agreement": round(agree / len(rows), 3), "points": len(rows)}
    return out


def main(argv=None) -> None:
    """python model.py [out_dir] [--machine NAME] [--check sweep.csv ...]:
    write the 18 CSVs of a machine (the reference's model_gen.py main), and
    optionally compare the model with measured sweep CSVs."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description="performance model CSVs (meta/model_gen.py)")
    ap.add_argument("out_dir", nargs="?", default="model_csv")
    ap.add_argument("--machine", default=REFERENCE.name, choices=sorted(MACHINES))
    ap.add_argument("--check", nargs="*", default=[], help="sweep CSVs (sweep.py) to check the model against")
    a = ap.parse_args(argv)
    mach = MACHINES[a.machine]
    print("\n".join(csv_files(a.out_dir, mach)))
    for path in a.check:
        import csv
        with open(path) as fh:
            rows = [{k: float(v) for k, v in row.items()} for row in csv.DictReader(fh)]
        print(json.dumps({"sweep": path, "machine": mach.name, "check": check_against_sweep(rows, mach)}))


if __name__ == "__main__":
    main()
