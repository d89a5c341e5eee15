"""sweep.py -- the reference's performance sweeps, run locally on the GPU.

    python sweep.py [--shape square,rankk,fixk] [--algs 222,232,323,333] [--levels 1] [--extra 222:2]
                    [--packs abc,naive] [--step 240] [--reps 3] [--out DIR] [--dtype f64]

The same three sweeps as meta/sbatch_util.py:122-157 (driven there by
meta/run_sbatch_script.py:64-158 through SLURM jobs that each run a
generated test_<job>.x over a range and append to a MATLAB .m file):

    square  m = n = k = s,        s = 240 .. 12200 step 240
    rankk   m = n = 14400, k = s, s = 240 .. 14400 step 240
    fixk    m = n = s, k = 1024,  s = 240 .. 20000 step 240

Here one process runs every variant at every point on the current GPU, with
A, B, C resident in HBM (U[-1,1), C pre-filled, seeded), and writes per sweep
sweep_<shape>_<dtype>.csv: one row per point with m, n, k, the plain GEMM's
effective GFLOPS (2mnk / t) and, per variant <alg>-<L>_<pack>, its effective
GFLOPS and the FMM / GEMM ratio; GFLOPS are from the best of `reps` timed
calls (the reference driver's min-of-3, meta/misc_gen.py:196-227) and the
median is kept in sweep_<shape>_<dtype>_median.csv; the error of every
variant (max-norm against extended-precision dot products at sampled
entries, in units of u*k*max|A|*max|B|) goes to sweep_<shape>_<dtype>_err.csv.
A progress line per point goes to stdout.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import sys
import time
from typing import Dict, List, Sequence, Tuple

if __package__ in (None, ""):
    _root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, _root)
    from __graft_entry__ import load_package  # noqa: E402
    load_package()
    from fmm_gen_amd.plan import build_algorithm  # noqa: E402
else:
    from .plan import build_algorithm

SWEEPS = {
    # shape: (first, last, m(s), n(s), k(s))  (meta/sbatch_util.py:122-157)
    "square": (240, 12200, lambda s: s, lambda s: s, lambda s: s),
    "rankk": (240, 14400, lambda s: 14400, lambda s: 14400, lambda s: s),
    "fixk": (240, 20000, lambda s: s, lambda s: s, lambda s: 1024),
}


def points(shape: str, step: int = 240, limit: int = 0) -> List[Tuple[int, int, int]]:
    first, last, fm, fn, fk = SWEEPS[shape]
    out = []
    for s in range(first, last + 1, step):
        if limit and s > limit:
            break
        out.append((fm(s), fn(s), fk(s)))
    return out


def variants(algs: Sequence[str], levels: Sequence[int], extra: Sequence[str], packs: Sequence[str]):
    """(label, registry spec, pack) per variant, e.g. ('222-1_abc', (['222'], [1]), 'abc')."""
    specs = [(a, L) for a in algs for L in levels]
    for e in extra:
        a, L = e.split(":")
        specs.append((a, int(L)))
    out = []
    for a, L in specs:
        for p in packs:
            out.append((f"{a}-{L}_{p}", ([a], [L]), p))
    return out


def sampled_error(A, B, C0, C, nsamp: int = 64, seed: int = 1) -> float:
    """max |C - (C0 + A B)| / (u k max|A| max|B|) at sampled entries, exact
    values from extended-precision dot products."""
    import numpy as np
    import torch
    m, k = A.shape
    n = B.shape[1]
    g = np.random.default_rng(seed)
    ii = torch.as_tensor(g.integers(0, m, nsamp), device=A.device)
    jj = torch.as_tensor(g.integers(0, n, nsamp), device=A.device)
    ld = np.longdouble
    rows = A[ii, :].cpu().numpy().astype(ld)
    cols = B[:, jj].cpu().numpy().astype(ld)
    ex = C0[ii, jj].cpu().numpy().astype(ld) + (rows * cols.T).sum(axis=1)
    got = C[ii, jj].cpu().numpy().astype(ld)
    scale = 2.0 ** -53 * k * float(A.abs().max()) * float(B.abs().max())
    return float(np.abs(got - ex).max() / scale)


def time_calls(fn, reps: int) -> Tuple[float, float]:
    """(best, median) seconds of `reps` calls, each bracketed by HIP events,
    after one untimed call."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e-3)
    ts.sort()
    return ts[0], ts[len(ts) // 2]


def run_sweep(shape: str, vars_, dtype: str = "f64", step: int = 240, reps: int = 3, limit: int = 0,
              out_dir: str = ".", seed: int = 2024) -> Dict[str, str]:
    import torch
    from . import runtime as rt
    if dtype != "f64":
        raise ValueError("the sweeps time the FP64 drop-in path")
    plans = {}
    for label, spec, pack in vars_:
        key = tuple(map(tuple, spec))
        if key not in plans:
            plans[key] = rt.Plan(build_algorithm(*spec))
    labels = [v[0] for v in vars_]
    rows_best, rows_med, rows_err = [], [], []
    for (m, n, k) in points(shape, step, limit):
        t0 = time.perf_counter()
        g = torch.Generator(device="cuda").manual_seed(seed)
        A = rt.colmajor(m, k).uniform_(-1, 1, generator=g)
        B = rt.colmajor(k, n).uniform_(-1, 1, generator=g)
        C0 = rt.colmajor(m, n).uniform_(-1, 1, generator=g)
        C = rt.as_colmajor(C0)
        fl = 2.0 * m * n * k
        best, med, err = {}, {}, {}
        C.copy_(C0)
        rt.gemm(A, B, C)
        err["gemm"] = sampled_error(A, B, C0, C)
        b, md = time_calls(lambda: rt.gemm(A, B, C), reps)
        best["gemm"], med["gemm"] = fl / b / 1e9, fl / md / 1e9
        for label, spec, pack in vars_:
            plan = plans[tuple(map(tuple, spec))]
            ws = torch.empty(max(plan.workspace_bytes(m, n, k, pack), 1) + 256, dtype=torch.uint8, device="cuda")
            C.copy_(C0)
            plan.run(A, B, C, pack=pack, ws=ws)
            err[label] = sampled_error(A, B, C0, C)
            b, md = time_calls(lambda: plan.run(A, B, C, pack=pack, ws=ws), reps)
            best[label], med[label] = fl / b / 1e9, fl / md / 1e9
            del ws
        rows_best.append((m, n, k, best))
        rows_med.append((m, n, k, med))
        rows_err.append((m, n, k, err))
        top = max(labels, key=lambda l: best[l])
        print(json.dumps({"shape": shape, "m": m, "n": n, "k": k, "gemm_gf": round(best["gemm"], 1),
                          "best": top, "best_gf": round(best[top], 1),
                          "ratio": round(best[top] / best["gemm"], 3), "s": round(time.perf_counter() - t0, 2)}),
              flush=True)
        del A, B, C0, C
        torch.cuda.empty_cache()
    for p in plans.values():
        p.close()
    os.makedirs(out_dir, exist_ok=True)
    base = os.path.join(out_dir, f"sweep_{shape}_{dtype}")
    write_csv(base + ".csv", labels, rows_best, ratios=True)
    write_csv(base + "_median.csv", labels, rows_med, ratios=True)
    write_csv(base + "_err.csv", labels, rows_err, ratios=False, fmt="{:.3f}")
    return {"best": base + ".csv", "median": base + "_median.csv", "error": base + "_err.csv"}


def write_csv(path: str, labels: Sequence[str], rows, ratios: bool, fmt: str = "{:.1f}") -> None:
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        head = ["m", "n", "k", "gemm"] + list(labels)
        if ratios:
            head += [f"{l}/gemm" for l in labels]
        w.writerow(head)
        for m, n, k, vals in rows:
            r = [m, n, k, fmt.format(vals["gemm"])] + [fmt.format(vals[l]) for l in labels]
            if ratios:
                r += ["{:.3f}".format(vals[l] / vals["gemm"]) for l in labels]
            w.writerow(r)


def read_csv(path: str):
    """Rows of a sweep CSV as dicts of floats."""
    with open(path) as fh:
        return [{k: float(v) for k, v in row.items()} for row in csv.DictReader(fh)]


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--shape", default="square,rankk,fixk")
    ap.add_argument("--algs", default="222,232,323,333")
    ap.add_argument("--levels", default="1")
    ap.add_argument("--extra", default="222:2", help="further alg:level variants, comma separated")
    ap.add_argument("--packs", default="abc,naive")
    ap.add_argument("--step", type=int, default=240)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--limit", type=int, default=0, help="stop at this sweep value (0: the full range)")
    ap.add_argument("--dtype", default="f64")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default="sweep_results")
    a = ap.parse_args(argv)
    vars_ = variants(a.algs.split(","), [int(x) for x in a.levels.split(",")],
                     [e for e in a.extra.split(",") if e], a.packs.split(","))
    for shape in a.shape.split(","):
        out = run_sweep(shape, vars_, a.dtype, a.step, a.reps, a.limit, a.out, a.seed)
        print(json.dumps({"shape": shape, "files": out}), flush=True)


if __name__ == "__main__":
    if __package__ in (None, ""):
        from fmm_gen_amd import sweep as _s   # run through the package (relative imports)
        _s.main()
    else:
        main()
