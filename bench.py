"""Benchmark: effective GFLOPS (2mnk/t) of the best generated ABC FMM vs the
project's plain MFMA GEMM, on one MI355X.

One *step* is one C += A*B call (FP64, column-major, inputs resident in HBM)
through the C-ABI entry fmm_run with the fastest of the three pack types
(ABC, AB, Naive) on this workload; all three are reported.  Default workload: the one-level <3,3,3>
rank-23 algorithm (smirnov333-23-139), square m=n=k=9600, A, B, C ~ U[-1,1)
from a seed.  Other shapes of the same family: --workload.

Printed JSON line (one, rank 0): metric/value/unit as the headline metric,
plus the plain GEMM's rate on the same inputs, the sampled max-norm error of
both against an extended-precision dot product, a roofline object for the
FMM kernel and a CPU baseline (the oracle's C restatement of the reference
ABC algorithm, timed on the host cores on a bounded sample).

Further sections (``--sections``; all by default), reported beside the
headline in the same line:
  precision  8192^3 from FP64 U[-1,1) inputs: plain GEMM and the <2,2,2> L1,
             <2,2,2> L2 and <3,2,3> L1 FMMs with FP32, BF16, FP16 and
             split-BF16 products (FP32 accumulation), and FP64 refinement
             (fmm_refine); effective TFLOP/s and relative error against the
             FP64 product of the original inputs (sampled entries);
  lowp_margin  16384^3, BF16 and split-BF16: plain GEMM against <2,2,2> L1
             and L2 in the Naive pack type (median of 5 runs, errors);
  batched    1024 FP32 problems, m, n, k log-uniform in [192, 4096], one
             persistent launch with per-problem choices of the autotuner
             (calibrated on the device first) against all-GEMM batches and
             single calls;
  serving    Poisson arrivals of the GEMM request mix (60% square, 30%
             rank-k, 10% tall-skinny; 70% FP32, 30% FP64) under a 200 GB HBM
             budget: p50 / p99 latency and throughput.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (algs, levels, m, n, k)
    "333L1_sq9600": (["333"], [1], 9600, 9600, 9600),
    "222x323_sq9600": (["222", "323"], [1, 1], 9600, 9600, 9600),
    "333L1_rankk14400x480": (["333"], [1], 14400, 14400, 480),
    "222x323_rankk14400x480": (["222", "323"], [1, 1], 14400, 14400, 480),
    "222L1_sq2048": (["222"], [1], 2048, 2048, 2048),
    "222L1_sq9600": (["222"], [1], 9600, 9600, 9600),
    "222L2_sq9600": (["222"], [2], 9600, 9600, 9600),
    "222L1_rankk14400x480": (["222"], [1], 14400, 14400, 480),
}
DEFAULT = "333L1_sq9600"
PEAK_F64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# config 2: both algorithms at both shapes (reported beside the headline)
CONFIG2 = ["333L1_sq9600", "222x323_sq9600", "333L1_rankk14400x480", "222x323_rankk14400x480"]
# shapes where the FMM margin over GEMM is clear of run-to-run noise: a large
# square (two Strassen levels) and the deep end of the rank-k sweep
MARGIN = {"222L2_sq16384": (["222"], [2], 16384, 16384, 16384),
          "323L1_rankk14400x9600": (["323"], [1], 14400, 14400, 9600)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-gemm", action="store_true", help="skip the plain-GEMM comparison")
    ap.add_argument("--sections", default="headline,config2,margin,precision,lowp_margin,batched,serving",
                    help="comma list of headline, config2, margin, precision, lowp_margin, batched, serving")
    ap.add_argument("--runs", type=int, default=10, help="timed runs per variant in config2 / margin")
    ap.add_argument("--packs", default="abc,ab,naive", help="headline: pack types to measure")
    ap.add_argument("--gemm-only", action="store_true",
                    help="headline workload: only the plain GEMM (for per-kernel counter runs)")
    ap.add_argument("--prec-filter", default="",
                    help="precision section: only these dtype/alg/pack rows, e.g. bf16/323L1/abc,bf16/gemm/gemm")
    ap.add_argument("--qps", type=float, default=800.0, help="serving: arrival rate")
    ap.add_argument("--requests", type=int, default=1500, help="serving: number of requests")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 1 GiB three times (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration():
    """A plain copy of a known size, for scaling rocprofv3 FETCH_SIZE /
    WRITE_SIZE readings to bytes: each launch reads and writes 1 GiB."""
    import torch
    n = (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n * 8, "launches": 3}), flush=True)


def load_traffic(workload, pack="abc"):
    """HBM bytes per launch of the product kernel of this workload and pack
    type, from the committed rocprofv3 counter summary
    (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        return t.get(workload, {}).get(pack, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def sampled_error(A, B, C0, C, nsamp=256, seed=1):
    """Max |C - C_exact| / (u k max|A| max|B|) over sampled entries, with
    C_exact from extended-precision dot products of the same rows/columns."""
    import numpy as np
    import torch
    m, k = A.shape
    n = B.shape[1]
    g = np.random.default_rng(seed)
    ii = g.integers(0, m, nsamp)
    jj = g.integers(0, n, nsamp)
    It = torch.as_tensor(ii, device=A.device)
    Jt = torch.as_tensor(jj, device=A.device)
    rows = A[It, :].cpu().numpy().astype(np.longdouble)
    cols = B[:, Jt].cpu().numpy().astype(np.longdouble)
    c0 = C0[It, Jt].cpu().numpy().astype(np.longdouble)
    ex = c0 + (rows * cols.T).sum(axis=1)
    got = C[It, Jt].cpu().numpy().astype(np.longdouble)
    amax = float(A.abs().max())
    bmax = float(B.abs().max())
    return float(np.abs(got - ex).max() / (2.0 ** -53 * k * amax * bmax))


def cpu_baseline(F, algs, levels, m, n, k):
    """The oracle's C/OpenMP restatement of the reference ABC algorithm (the
    same tables, partition, product order and packing-sum order, built with
    the reference's -O3 -fopenmp -mavx), timed on this host on the headline
    workload itself, min of 3 runs (the reference driver's rule), with one
    thread per CPU this process may run on (BLISLAB_IC_NT overrides, as in
    the reference)."""
    import numpy as np
    from oracle import oracle
    alg = F.build_algorithm(algs, levels)
    try:
        ncpu = len(os.sched_getaffinity(0))
    except Exception:
        ncpu = os.cpu_count() or 1
    threads = int(os.environ.get("BLISLAB_IC_NT", ncpu))
    oracle.set_threads(threads)
    rng = np.random.default_rng(7)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    best = float("inf")
    for _ in range(3):
        C = C0.copy(order="F")
        t0 = time.perf_counter()
        oracle.fmm(alg, A, B, C, "abc")
        best = min(best, time.perf_counter() - t0)
        print(f"[cpu baseline] run {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    return {"value": round(2.0 * m * n * k / best / 1e9, 2), "unit": "GFLOPS", "cores": threads,
            "kind": "port",
            "sample": f"oracle C restatement (-O3 -fopenmp -mavx) of the reference ABC {'x'.join(algs)} "
                      f"levels {levels}, m={m} n={n} k={k} (the headline workload), FP64, seeded U[-1,1), "
                      f"C pre-filled, min of 3 runs, {threads} OpenMP threads = the CPUs available to "
                      f"the process (os.sched_getaffinity)"}


def measure_workload(F, rt, algs, levels, m, n, k, seed, runs, packs=("abc", "ab", "naive")):
    """GEMM and every pack type of one FP64 workload (inputs resident, U[-1,1),
    C pre-filled): effective GFLOPS 2mnk/t of the median and of the best of
    `runs` calls each timed with HIP events after two untimed calls, and the
    sampled max-norm error of one call against extended-precision dots."""
    import torch
    alg = F.build_algorithm(algs, levels)
    plan = rt.Plan(alg)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    A = rt.colmajor(m, k).uniform_(-1.0, 1.0, generator=gen)
    B = rt.colmajor(k, n).uniform_(-1.0, 1.0, generator=gen)
    C0 = rt.colmajor(m, n).uniform_(-1.0, 1.0, generator=gen)
    C = rt.as_colmajor(C0)
    fl = 2.0 * m * n * k
    out = {"m": m, "n": n, "k": k, "algorithm": "x".join(f"{a}-{l}" for a, l in zip(algs, levels)),
           "levels": len(levels), "runs": runs}

    def stats(fn):
        C.copy_(C0)
        fn()
        torch.cuda.synchronize()
        err = sampled_error(A, B, C0, C)
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(runs):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        ts.sort()
        med = ts[len(ts) // 2] if len(ts) % 2 else 0.5 * (ts[len(ts) // 2 - 1] + ts[len(ts) // 2])
        return {"gflops_median": round(fl / med / 1e6, 1), "gflops_best": round(fl / ts[0] / 1e6, 1),
                "ms_median": round(med, 4), "ms_best": round(ts[0], 4), "max_norm_err": round(err, 3)}
    out["gemm"] = stats(lambda: rt.gemm(A, B, C))
    for pk in packs:
        ws = torch.empty(plan.workspace_bytes(m, n, k, pk) + 256, dtype=torch.uint8, device="cuda")
        out[pk] = stats(lambda: plan.run(A, B, C, pack=pk, ws=ws))
        del ws
    best = max(packs, key=lambda pk: out[pk]["gflops_median"])
    out["best_pack"] = best
    out["speedup_median"] = round(out[best]["gflops_median"] / out["gemm"]["gflops_median"], 4)
    out["speedup_best"] = round(out[best]["gflops_best"] / out["gemm"]["gflops_best"], 4)
    out["err_metric"] = "max|C-C_exact|/(2^-53 k max|A| max|B|), 256 sampled entries"
    plan.close()
    del A, B, C0, C
    torch.cuda.empty_cache()
    return out


def rel_sampled(A, B, C0, C, nsamp=512, seed=3):
    """Relative error ||C - C_exact|| / ||C_exact|| over sampled entries, with
    C_exact from extended-precision dot products of the FP64 inputs."""
    import numpy as np
    import torch
    m, k = A.shape
    n = B.shape[1]
    g = np.random.default_rng(seed)
    ii = g.integers(0, m, nsamp)
    jj = g.integers(0, n, nsamp)
    It = torch.as_tensor(ii, device=A.device)
    Jt = torch.as_tensor(jj, device=A.device)
    rows = A[It, :].cpu().numpy().astype(np.longdouble)
    cols = B[:, Jt].cpu().numpy().astype(np.longdouble)
    ex = C0[It, Jt].cpu().numpy().astype(np.longdouble) + (rows * cols.T).sum(axis=1)
    got = C[It, Jt].cpu().numpy().astype(np.longdouble)
    return float(np.sqrt(((got - ex) ** 2).sum() / (ex ** 2).sum()))


def ev_ms(fn, reps=3):
    import torch
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def section_precision(F, rt, seed, only=()):
    """Config 'precision sweep': 8192^3, FP64 U[-1,1) inputs cast per dtype;
    each FMM in the ABC and Naive pack types.  `only`: dtype/alg/pack rows to
    run (all when empty; the refinement rows run only then)."""
    import torch
    N = 8192
    gen = torch.Generator(device="cuda").manual_seed(seed)
    A64 = rt.colmajor(N, N)
    B64 = rt.colmajor(N, N)
    A64.uniform_(-1.0, 1.0, generator=gen)
    B64.uniform_(-1.0, 1.0, generator=gen)
    Z64 = rt.colmajor(N, N).zero_()
    specs = {"gemm": None, "222L1": (["222"], [1]), "222L2": (["222"], [2]), "323L1": (["323"], [1])}
    plans = {a: (rt.Plan(F.build_algorithm(*sp)) if sp else None) for a, sp in specs.items()}
    fl = 2.0 * N ** 3
    rows = []
    C = rt.colmajor(N, N, dtype=torch.float32)
    for dt in ("f32", "bf16", "f16", "bf16x3"):
        if only and not any(o.startswith(dt + "/") for o in only):
            continue
        ti = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16, "bf16x3": torch.float32}[dt]
        A = rt.as_colmajor(A64.to(ti))
        B = rt.as_colmajor(B64.to(ti))
        split = dt == "bf16x3"
        for a, plan in plans.items():
            for pack in (("gemm",) if plan is None else ("abc", "naive")):
                if only and f"{dt}/{a}/{pack}" not in only:
                    continue
                ws = None if plan is None else torch.empty(plan.workspace_bytes(N, N, N, pack) + 256,
                                                           dtype=torch.uint8, device="cuda")

                def call():
                    if plan is None:
                        rt.gemm(A, B, C, split=split)
                    else:
                        plan.run(A, B, C, pack=pack, ws=ws, split=split)
                C.zero_()
                call()
                torch.cuda.synchronize()
                err = rel_sampled(A64, B64, Z64, C)
                ms = ev_ms(call)
                rows.append({"dtype": dt, "alg": a, "pack": pack, "ms": round(ms, 3),
                             "eff_tflops": round(fl / ms / 1e9, 1), "rel_err": float(f"{err:.3e}")})
                del ws
        del A, B
    if only:
        return {"shape": [N, N, N], "rows": rows, "filter": sorted(only)}
    C64 = rt.colmajor(N, N)
    for main, a, corr in (("bf16", "gemm", "f32"), ("bf16", "gemm", "bf16"), ("bf16", "222L1", "bf16"),
                          ("f16", "gemm", "f32"), ("f32", "222L1", "f32"), ("bf16x3", "222L1", "bf16x3")):
        def call():
            rt.refine(A64, B64, C64, main=main, main_plan=plans[a], corr=corr)
        C64.zero_()
        call()
        torch.cuda.synchronize()
        err = rel_sampled(A64, B64, Z64, C64)
        ms = ev_ms(call, 2)
        rows.append({"dtype": f"{main}+refine({corr})", "alg": a, "pack": "abc" if plans[a] else "gemm", "ms": round(ms, 3),
                     "eff_tflops": round(fl / ms / 1e9, 1), "rel_err": float(f"{err:.3e}")})
    # FMM against GEMM at equal stated error: for each error class, the
    # fastest FMM row and the fastest GEMM row (any dtype, refinement
    # included) whose measured error is within it
    equal = {}
    for bound in (1e-5, 1e-4, 1e-2):
        ok = [r for r in rows if r["rel_err"] <= bound]
        f = [r for r in ok if r["alg"] != "gemm"]
        g = [r for r in ok if r["alg"] == "gemm"]
        if f and g:
            bf = max(f, key=lambda r: r["eff_tflops"])
            bg = max(g, key=lambda r: r["eff_tflops"])
            equal[f"{bound:g}"] = {"fmm": f"{bf['dtype']} {bf['alg']}/{bf['pack']}", "fmm_tflops": bf["eff_tflops"],
                                   "fmm_err": bf["rel_err"], "gemm": bg["dtype"], "gemm_tflops": bg["eff_tflops"],
                                   "gemm_err": bg["rel_err"],
                                   "speedup": round(bf["eff_tflops"] / bg["eff_tflops"], 3)}
    # best FMM vs the same dtype's plain GEMM (both within that dtype's stated
    # accuracy, tests/tolerance.py)
    best = {}
    for dt in ("f32", "bf16", "f16", "bf16x3"):
        g = next(r for r in rows if r["dtype"] == dt and r["alg"] == "gemm")
        f = max((r for r in rows if r["dtype"] == dt and r["alg"] != "gemm"), key=lambda r: r["eff_tflops"])
        best[dt] = {"best_fmm": f"{f['alg']}/{f['pack']}", "fmm_tflops": f["eff_tflops"], "gemm_tflops": g["eff_tflops"],
                    "speedup": round(f["eff_tflops"] / g["eff_tflops"], 3)}
    return {"shape": [N, N, N], "inputs": "FP64 U[-1,1) cast per dtype, C = 0",
            "err_metric": "relative Frobenius error vs the FP64 product of the FP64 inputs, 512 sampled entries",
            "rows": rows, "best_vs_gemm": best, "equal_error": equal}


def section_lowp_margin(F, rt, seed, N=16384, runs=5):
    """The 16-bit types at 16384^3 (beside the 8192^3 precision sweep): plain
    GEMM against the <2,2,2> FMMs at one and two levels, Naive pack type (the
    fastest low-precision pack), BF16 and split-BF16; median of `runs`
    event-timed calls and the sampled relative error against the FP64
    product."""
    import numpy as np
    import torch
    gen = torch.Generator(device="cuda").manual_seed(seed)
    A64 = rt.colmajor(N, N)
    B64 = rt.colmajor(N, N)
    A64.uniform_(-1.0, 1.0, generator=gen)
    B64.uniform_(-1.0, 1.0, generator=gen)
    Z64 = rt.colmajor(N, N).zero_()
    plans = {"gemm": None, "222L1": rt.Plan(F.build_algorithm(["222"], [1])),
             "222L2": rt.Plan(F.build_algorithm(["222"], [2]))}
    fl = 2.0 * N ** 3
    C = rt.colmajor(N, N, dtype=torch.float32)
    rows = []
    for dt in ("bf16", "bf16x3"):
        ti = torch.bfloat16 if dt == "bf16" else torch.float32
        A = rt.as_colmajor(A64.to(ti))
        B = rt.as_colmajor(B64.to(ti))
        split = dt == "bf16x3"
        for a, plan in plans.items():
            pack = "gemm" if plan is None else "naive"
            ws = None if plan is None else torch.empty(plan.workspace_bytes(N, N, N, pack) + 256,
                                                       dtype=torch.uint8, device="cuda")

            def call():
                if plan is None:
                    rt.gemm(A, B, C, split=split)
                else:
                    plan.run(A, B, C, pack=pack, ws=ws, split=split)
            C.zero_()
            call()
            torch.cuda.synchronize()
            err = rel_sampled(A64, B64, Z64, C)
            ts = [ev_ms(call, 1) for _ in range(runs)]
            ms = float(np.median(ts))
            rows.append({"dtype": dt, "alg": a, "pack": pack, "ms_median": round(ms, 3), "ms_best": round(min(ts), 3),
                         "eff_tflops": round(fl / ms / 1e9, 1), "rel_err": float(f"{err:.3e}")})
            del ws
        del A, B
    best = {}
    for dt in ("bf16", "bf16x3"):
        g = next(r for r in rows if r["dtype"] == dt and r["alg"] == "gemm")
        f = max((r for r in rows if r["dtype"] == dt and r["alg"] != "gemm"), key=lambda r: r["eff_tflops"])
        best[dt] = {"best_fmm": f"{f['alg']}/{f['pack']}", "fmm_tflops": f["eff_tflops"], "gemm_tflops": g["eff_tflops"],
                    "speedup": round(f["eff_tflops"] / g["eff_tflops"], 3)}
    return {"shape": [N, N, N], "runs": runs, "inputs": "FP64 U[-1,1) cast per dtype, C = 0",
            "err_metric": "relative Frobenius error vs the FP64 product of the FP64 inputs, 512 sampled entries",
            "rows": rows, "best_vs_gemm": best}


def section_batched(F, rt, seed):
    """Config 'batched': 1024 FP32 problems in one persistent launch."""
    import numpy as np
    import torch
    from collections import Counter
    from fmm_gen_amd import autotune as AT
    tuner = AT.Tuner("f32")

    def device_time(m, n, k, alg, tile, copies):
        A = rt.colmajor(m, k, dtype=torch.float32).uniform_(-1, 1)
        B = rt.colmajor(k, n, dtype=torch.float32).uniform_(-1, 1)
        Cs = [rt.colmajor(m, n, dtype=torch.float32).zero_() for _ in range(copies)]
        b, _ = tuner.batch([(A, B, C) for C in Cs], choices=[(alg, tile)] * copies)
        t = ev_ms(b.run) / 1e3
        b.close()
        return t
    t0 = time.perf_counter()
    theta = tuner.calibrate(device_time)
    tcal = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    dims = np.exp(rng.uniform(np.log(192), np.log(4096), (1024, 3))).round().astype(int)
    ops = []
    for m, n, k in dims:
        A = rt.colmajor(int(m), int(k), dtype=torch.float32).uniform_(-1, 1)
        B = rt.colmajor(int(k), int(n), dtype=torch.float32).uniform_(-1, 1)
        C = rt.colmajor(int(m), int(n), dtype=torch.float32).uniform_(-1, 1)
        ops.append((A, B, C))
    fl = float(sum(2.0 * m * n * k for m, n, k in dims))
    tuned, choices = tuner.batch(ops)
    t_tuned = ev_ms(tuned.run)
    assert tuned.status() == 0
    gch = [("gemm", min(range(2), key=lambda t: tuner.model.cost(tuner.features(A.shape[0], B.shape[1], A.shape[1],
                                                                               "gemm", t)))) for A, B, C in ops]
    gb, _ = tuner.batch(ops, choices=gch)
    t_gemm = ev_ms(gb.run)

    def singles():
        for A, B, C in ops:
            rt.gemm(A, B, C)
    t_single = ev_ms(singles, 1)
    regret = tuner_regret(tuner, device_time, dims, seed)
    return {"problems": len(ops), "tflop": round(fl / 1e12, 3), "dtype": "f32",
            "tuned_ms": round(t_tuned, 3), "tuned_tflops": round(fl / t_tuned / 1e9, 2),
            "gemm_batch_ms": round(t_gemm, 3), "gemm_batch_tflops": round(fl / t_gemm / 1e9, 2),
            "gemm_single_calls_ms": round(t_single, 3),
            "choices": {f"{a}/tile{t}": c for (a, t), c in Counter(map(tuple, choices)).items()},
            "model_theta": [float(f"{x:.4g}") for x in theta], "calibration_s": round(tcal, 1),
            "tuner_regret": regret}


def tuner_regret(tuner, device_time, dims, seed, nsamp=24):
    """The tuner's pick against every candidate (algorithm x tile) on a seeded
    sample of the batch's problems: each candidate timed as one persistent
    launch of enough copies of the problem to fill the GPU (~2e11 flops);
    regret = t(pick) / t(best candidate) - 1."""
    import numpy as np
    rng = np.random.default_rng(seed + 1)
    idx = rng.choice(len(dims), nsamp, replace=False)
    per = []
    for i in idx:
        m, n, k = (int(x) for x in dims[i])
        copies = max(1, int(2e11 / (2.0 * m * n * k)))
        times = {(a, t): device_time(m, n, k, a, t, copies) for a, t in tuner.candidates()}
        pick = tuner.choose(m, n, k)[:2]
        best = min(times, key=times.get)
        per.append({"mnk": [m, n, k], "pick": f"{pick[0]}/tile{pick[1]}", "best": f"{best[0]}/tile{best[1]}",
                    "regret": round(times[pick] / times[best] - 1.0, 4)})
    r = np.array([p["regret"] for p in per])
    return {"sample": nsamp, "mean": round(float(r.mean()), 4), "max": round(float(r.max()), 4),
            "pick_is_best": int((r == 0).sum()),
            "best_is_fmm": int(sum(1 for p in per if not p["best"].startswith("gemm/"))), "per_problem": per}


def section_serving(F, rt, seed, qps, nreq, load_steps=(1600, 3200, 6400), load_requests=600):
    """Config 'serving': Poisson arrivals of the request mix at `qps`, then
    the same mix at higher offered loads: the saturation rate is the highest
    offered QPS the loop still serves at >= 95% of it."""
    from collections import Counter
    from fmm_gen_amd import serve as S
    ex = S.DeviceExecutor(use_graphs=True, seed=seed)
    srv = S.Server(ex, hbm_budget=200e9, lanes_per_dtype=2, max_batch=256)
    ex.warmup(srv.lanes)
    srv.serve(S.poisson_mix(nreq, qps=qps, seed=seed), timeout=600)
    st = srv.stats()
    st.update({"qps": qps, "hbm_budget_gb": 200, "graph_replays": ex.replays, "graph_captures": ex.captures,
               "algorithms": dict(Counter(f"{r.dtype}/{r.alg}" for r in srv.done))})
    load = []
    for q in load_steps:
        s2 = S.Server(ex, hbm_budget=200e9, lanes_per_dtype=2, max_batch=256)
        s2.lanes = srv.lanes     # the warmed-up lanes (captured graphs)
        s2.serve(S.poisson_mix(load_requests, qps=q, seed=seed + q), timeout=600)
        t = s2.stats()
        load.append({"offered_qps": q, "served_req_s": t["throughput_req_s"], "p50_ms": t["latency_p50_ms"],
                     "p99_ms": t["latency_p99_ms"], "mean_batch": t["mean_batch"], "failed": t["failed"]})
    ok = [x["offered_qps"] for x in load if x["served_req_s"] >= 0.95 * x["offered_qps"]]
    st["load_sweep"] = load
    st["saturation_qps"] = max(ok) if ok else None
    return st


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); sharding a call over "
              "several GPUs (by C tile position) is not built.", flush=True)
        sys.exit(2)
    import torch
    from __graft_entry__ import load_package
    F = load_package()
    from fmm_gen_amd import runtime as rt
    sections = set(args.sections.split(","))
    extra = {}
    if "precision" in sections:
        extra["precision_sweep"] = section_precision(F, rt, args.seed,
                                                     tuple(x for x in args.prec_filter.split(",") if x))
        torch.cuda.empty_cache()
    if "lowp_margin" in sections:
        extra["lowp_margin"] = section_lowp_margin(F, rt, args.seed)
        torch.cuda.empty_cache()
    if "batched" in sections:
        extra["batched"] = section_batched(F, rt, args.seed)
        torch.cuda.empty_cache()
    if "serving" in sections:
        extra["serving"] = section_serving(F, rt, args.seed, args.qps, args.requests)
        torch.cuda.empty_cache()
    if "config2" in sections:
        extra["config2"] = {}
        for w in CONFIG2:
            a, l, m_, n_, k_ = WORKLOADS[w]
            extra["config2"][w] = measure_workload(F, rt, a, l, m_, n_, k_, args.seed, args.runs)
            print(f"[config2] {w} done", file=sys.stderr, flush=True)
    if "margin" in sections:
        extra["margin"] = {}
        for w, (a, l, m_, n_, k_) in MARGIN.items():
            extra["margin"][w] = measure_workload(F, rt, a, l, m_, n_, k_, args.seed, args.runs,
                                                  packs=("abc", "naive"))
            print(f"[margin] {w} done", file=sys.stderr, flush=True)
    if "headline" not in sections:
        print(json.dumps(extra), flush=True)
        return

    algs, levels, m, n, k = WORKLOADS[args.workload]
    alg = F.build_algorithm(algs, levels)
    plan = rt.Plan(alg)
    gen = torch.Generator(device="cuda").manual_seed(args.seed)
    A = rt.colmajor(m, k)
    B = rt.colmajor(k, n)
    C0 = rt.colmajor(m, n)
    for X in (A, B, C0):
        X.uniform_(-1.0, 1.0, generator=gen)
    C = rt.colmajor(m, n)
    ws = torch.zeros(plan.workspace_bytes(m, n, k) // 4 + 16, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    if args.gemm_only:
        fl = 2.0 * m * n * k
        for _ in range(args.warmup):
            rt.gemm(A, B, C)
        torch.cuda.synchronize()
        ts = []
        for _ in range(args.steps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            rt.gemm(A, B, C)
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        print(json.dumps({"workload": args.workload, "gemm_only": True, "steps": args.steps,
                          "gflops_best": round(fl / min(ts) / 1e6, 1)}), flush=True)
        return
    # correctness sample of one call from C0, per pack type
    packs = tuple(args.packs.split(","))
    wsb = {pk: torch.empty(plan.workspace_bytes(m, n, k, pk) + 256, dtype=torch.uint8, device="cuda")
           for pk in packs}
    errs = {}
    for pk in packs:
        C.copy_(C0)
        if pk == "abc":
            plan.abc(A, B, C, ws=ws)
        else:
            plan.run(A, B, C, pack=pk, ws=wsb[pk])
        torch.cuda.synchronize()
        errs[pk] = sampled_error(A, B, C0, C)
    assert rt.ws_status(ws) == 0

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        for e0, e1 in ev:
            e0.record(stream)
            fn()
            e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per = [e0.elapsed_time(e1) for e0, e1 in ev]
        return wall, per

    def call(pk):
        if pk == "abc":
            return lambda: plan.abc(A, B, C, ws=ws)
        return lambda: plan.run(A, B, C, pack=pk, ws=wsb[pk])

    # every pack type once (short), then the best one timed over exactly K steps
    probe = {pk: ev_ms(call(pk), 2) for pk in packs}
    best = min(packs, key=lambda pk: probe[pk])
    wall, per = timed(call(best), args.steps, args.warmup)
    ms_step = wall * 1e3 / args.steps
    per_sorted = sorted(per)
    step_median = per_sorted[len(per) // 2] if len(per) % 2 else 0.5 * (per_sorted[len(per) // 2 - 1] +
                                                                          per_sorted[len(per) // 2])
    eff = 2.0 * m * n * k / (ms_step * 1e-3) / 1e9
    by_pack = {pk: {"ms": round(probe[pk], 3), "gflops": round(2.0 * m * n * k / probe[pk] / 1e6, 1),
                    "max_norm_err": round(errs[pk], 3)} for pk in packs}

    gemm = None
    if not args.no_gemm:
        C.copy_(C0)
        rt.gemm(A, B, C)
        torch.cuda.synchronize()
        err_gemm = sampled_error(A, B, C0, C)
        gwall, gper = timed(lambda: rt.gemm(A, B, C), args.steps, args.warmup)
        gms = gwall * 1e3 / args.steps
        gemm = {"gflops": round(2.0 * m * n * k / (gms * 1e-3) / 1e9, 1), "ms_per_step": round(gms, 4),
                "kernel_ms_mean": round(sum(gper) / len(gper), 4), "max_norm_err": round(err_gemm, 3)}

    # roofline of the dominant kernel of the reported FMM call -- its product
    # launch (f64d_kernel for Naive, abc_f64_kernel for ABC / AB), timed alone
    # with HIP events around the launch
    # (fmm_timing) over K calls: executed MFMA flops per launch over the mean
    # launch duration.  The same for the ABC pack type (the hot path).
    p = F.partition(m, n, k, alg.dims)
    f_mfma = F.fmm_flops(alg, m, n, k)

    def kernel_ms(pk):
        fn = call(pk)
        fn()
        rt.kernel_timing(True)
        vals = []
        for _ in range(args.steps):
            fn()
            vals.append(rt.last_kernel_ms())
        rt.kernel_timing(False)
        return sum(vals) / len(vals)
    kern_ms = kernel_ms(best)
    abc_ms = kern_ms if best == "abc" else (kernel_ms("abc") if "abc" in packs else None)
    achieved = f_mfma / (kern_ms * 1e-3) / 1e12
    nnzU, nnzV, nnzW = alg.nnz()
    if best == "naive":   # FP64: single-term products over S, T; M blocks stored once
        b_min = 8 * (alg.rank * p.ms * p.ks + alg.rank * p.ks * p.ns + alg.rank * p.ms * p.ns)
    elif best == "ab":
        b_min = 8 * (nnzU * p.ms * p.ks + nnzV * p.ks * p.ns + alg.rank * p.ms * p.ns)
    else:
        b_min = 8 * (nnzU * p.ms * p.ks + nnzV * p.ks * p.ns) + 8 * 2 * nnzW * p.ms * p.ns
    roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_F64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_F64_TFLOPS, 4), "traffic": load_traffic(args.workload, best),
            "kernel": ("f64d_kernel (single-term, direct-to-LDS)" if best == "naive" else
                       "abc_f64_kernel (operand sums)") + f", the product launch of the {best} pack type",
            "flops_per_launch": f_mfma,
            "algorithmic_bytes_per_launch": b_min, "kernel_ms_mean": round(kern_ms, 4),
            "abc_kernel": None if abc_ms is None else {
                "achieved": round(f_mfma / (abc_ms * 1e-3) / 1e12, 2),
                "frac": round(f_mfma / (abc_ms * 1e-3) / 1e12 / PEAK_F64_TFLOPS, 4),
                "kernel_ms_mean": round(abc_ms, 4), "traffic": load_traffic(args.workload, "abc")}}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(F, algs, levels, m, n, k)

    out = {
        "metric": "effective GFLOPS (2mnk/s) FMM vs plain GEMM at stated error",
        "value": round(eff, 1),
        "unit": "GFLOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded U[-1,1) inputs, C pre-filled)",
        "config": {"workload": args.workload, "algorithm": "x".join(f"{a}-{l}" for a, l in zip(algs, levels)),
                   "pack_type": best, "m": m, "n": n, "k": k, "rank": alg.rank, "dims": list(alg.dims),
                   "leaf": [p.ms, p.ks, p.ns], "parallelism": "single GPU"},
        "step_ms_median": round(step_median, 4),
        "step_ms_best": round(per_sorted[0], 4),
        "gflops_median": round(2.0 * m * n * k / step_median / 1e6, 1),
        "gflops_best": round(2.0 * m * n * k / per_sorted[0] / 1e6, 1),
        "max_norm_err": round(errs[best], 3),
        "by_pack_type": by_pack,
        "err_metric": "max|C-C_exact|/(2^-53 k max|A| max|B|) on 256 sampled entries",
        "plain_gemm": gemm,
        "speedup_vs_gemm": round(eff / gemm["gflops"], 4) if gemm else None,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
