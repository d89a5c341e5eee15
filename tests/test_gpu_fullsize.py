"""GPU parity at the full sizes of every benchmark configuration.

* FP64, <3,3,3> L1 and <2,2,2>x<3,2,3>, square 9600^3 and rank-k
  14400^2 x 480, U[-1,1) inputs with C pre-filled: the HIP ABC result is
  compared element by element with the oracle's C restatement of the
  reference ABC algorithm on the same inputs (both bounded by the stated
  per-level tolerance, so they differ by at most twice it), each is checked
  against extended-precision dot products at sampled entries, and the Naive
  pack type must equal ABC bit for bit at full size.
* 8192^3 precision sweep (FP32, BF16, FP16, split-BF16; <2,2,2> L1 / L2,
  <3,2,3> L1 and the plain GEMM; ABC and Naive): sampled relative error
  against the FP64 product of the stored inputs within the stated per-level
  bound plus the FP32 accumulation term of depth 8192, sampled entries
  against the oracle's restatement of the low-precision rounding points
  (oracle.fmm_lowp_at), and ABC == Naive bit for bit; then the FP64
  refinement variants against the FP64 product of the original inputs.
* The 1024-problem FP32 batch (m, n, k log-uniform in [192, 4096]) in one
  persistent launch, with every candidate algorithm present: a seeded sample
  of problems equals single calls bit for bit and lies within tolerance.
* A full-shape serving request stream: every result equals the same request
  computed alone, bit for bit.
Each case prints its errors (pytest -s shows them).
"""
import numpy as np
import pytest

from tolerance import U64, tol_for, tol_lowp_at, tol_lowp_vs_emu_at

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def sample_idx(m, n, nsamp, seed):
    g = np.random.default_rng(seed)
    return g.integers(0, m, nsamp), g.integers(0, n, nsamp)


def exact_at(A, B, C0, ii, jj):
    """C0 + A @ B at entries (ii, jj) in extended precision (numpy arrays)."""
    ld = np.longdouble
    rows = A[ii, :].astype(ld)
    cols = B[:, jj].astype(ld)
    return C0[ii, jj].astype(ld) + (rows * cols.T).sum(axis=1)


# --------------------------------------------------------------- FP64 config

FP64_CASES = [
    ("333L1_sq9600", ["333"], [1], 9600, 9600, 9600),
    ("222x323_sq9600", ["222", "323"], [1, 1], 9600, 9600, 9600),
    ("333L1_rankk14400x480", ["333"], [1], 14400, 14400, 480),
    ("222x323_rankk14400x480", ["222", "323"], [1, 1], 14400, 14400, 480),
]


@pytest.mark.parametrize("case", FP64_CASES, ids=[c[0] for c in FP64_CASES])
def test_fp64_full_size_vs_oracle(torch, fmm, rt, oracle, case):
    name, algs, levels, m, n, k = case
    alg = fmm.build_algorithm(algs, levels)
    nlev = len(levels)
    print(f"\n[{name}] generating inputs", flush=True)
    rng = np.random.default_rng(2024)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    # Fortran-ordered host arrays -> column-major device tensors (ld = rows)
    dA, dB, dC0 = (torch.from_numpy(X.T).to("cuda").t() for X in (A, B, C0))
    plan = rt.Plan(alg)
    ws = torch.zeros(plan.workspace_bytes(m, n, k) // 4 + 64, dtype=torch.int32, device="cuda")
    dC = rt.as_colmajor(dC0)
    print(f"[{name}] HIP ABC", flush=True)
    plan.abc(dA, dB, dC, ws=ws)
    torch.cuda.synchronize()
    assert rt.ws_status(ws) == 0
    # size-independent property: the Naive pack type applies the same C
    # updates in the same order, so it equals ABC bit for bit
    wsn = torch.empty(plan.workspace_bytes(m, n, k, "naive") + 256, dtype=torch.uint8, device="cuda")
    dCn = rt.as_colmajor(dC0)
    plan.run(dA, dB, dCn, pack="naive", ws=wsn)
    torch.cuda.synchronize()
    naive_equal = bool(torch.equal(dC, dCn))
    del dCn, wsn
    got = dC.cpu().numpy()
    print(f"[{name}] oracle ABC restatement ({oracle.get_threads()} threads)", flush=True)
    want = C0.copy(order="F")
    oracle.fmm(alg, A, B, want, "abc")
    scale = U64 * k * np.abs(A).max() * np.abs(B).max()
    tol = tol_for(nlev)
    diff = float(np.abs(got - want).max() / scale)
    ii, jj = sample_idx(m, n, 4096, 7)
    ex = exact_at(A, B, C0, ii, jj)
    e_gpu = float(np.abs(got[ii, jj] - ex).max() / scale)
    e_orc = float(np.abs(want[ii, jj] - ex).max() / scale)
    print(f"[{name}] levels {nlev}: e(GPU) {e_gpu:.3f}  e(oracle) {e_orc:.3f}  "
          f"max|GPU-oracle| {diff:.3f}  (units u*k*max|A|*max|B|; bound per result {tol})", flush=True)
    assert e_gpu <= tol and e_orc <= tol
    assert diff <= 2 * tol
    assert naive_equal
    plan.close()


# ------------------------------------------------------- precision sweep

def identity_alg():
    """The plain product as a one-level <1,1,1> rank-1 algorithm."""
    from fractions import Fraction
    from fmm_gen_amd.coeffs import Algorithm
    one = [[Fraction(1)]]
    return Algorithm("gemm", [(1, 1, 1)], one, one, one)


SWEEP_ALGS = {"gemm": None, "222L1": (["222"], [1]), "222L2": (["222"], [2]), "323L1": (["323"], [1])}
SWEEP_DTYPES = ["f32", "bf16", "f16", "bf16x3"]


@pytest.fixture(scope="module")
def sweep_inputs(torch, rt):
    N = 8192
    g = torch.Generator(device="cuda").manual_seed(2024)
    A64 = rt.colmajor(N, N).uniform_(-1.0, 1.0, generator=g)
    B64 = rt.colmajor(N, N).uniform_(-1.0, 1.0, generator=g)
    return N, A64, B64


@pytest.mark.parametrize("dtype", SWEEP_DTYPES)
def test_precision_sweep_8192(torch, fmm, rt, oracle, sweep_inputs, dtype):
    N, A64, B64 = sweep_inputs
    ti = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16, "bf16x3": torch.float32}[dtype]
    split = dtype == "bf16x3"
    A = rt.as_colmajor(A64.to(ti))
    B = rt.as_colmajor(B64.to(ti))
    ii, jj = sample_idx(N, N, 512, 3)
    It, Jt = torch.as_tensor(ii, device="cuda"), torch.as_tensor(jj, device="cuda")
    # the stored inputs, as float32 host arrays (for the restatement) and
    # their FP64 product at the sampled entries
    Ah = A.float().cpu().numpy()
    Bh = B.float().cpu().numpy()
    ex = exact_at(Ah.astype(np.float64), Bh.astype(np.float64), np.zeros((N, N), np.float32), ii, jj)
    for name, spec in SWEEP_ALGS.items():
        plan = rt.Plan(fmm.build_algorithm(*spec)) if spec else None
        nlev = 0 if spec is None else sum(spec[1])
        outs = {}
        for pack in (("gemm",) if plan is None else ("abc", "naive")):
            C = rt.colmajor(N, N, dtype=torch.float32).zero_()
            if plan is None:
                rt.gemm(A, B, C, split=split)
            else:
                plan.run(A, B, C, pack=pack, split=split)
            torch.cuda.synchronize()
            outs[pack] = C
        assert rt.ws_status() == 0
        C = outs.get("abc", outs.get("gemm"))
        got = C[It, Jt].cpu().numpy().astype(np.float64)
        rel = float(np.sqrt(((got - ex) ** 2).sum() / (ex ** 2).sum()))
        # the restatement at the sampled entries (the plain GEMM is the
        # one-product identity algorithm: same rounding points)
        alg = fmm.build_algorithm(*spec) if spec else identity_alg()
        emu = oracle.fmm_lowp_at(alg, Ah, Bh, np.zeros((N, N), np.float32), dtype, ii, jj).astype(np.float64)
        d_emu = float(np.sqrt(((got - emu) ** 2).sum() / (emu ** 2).sum()))
        same = True if plan is None else bool(torch.equal(outs["abc"], outs["naive"]))
        print(f"\n[{dtype} {name}] rel err vs FP64 of stored inputs {rel:.3e} (bound {tol_lowp_at(dtype, nlev, N):.2e}); "
              f"vs restatement {d_emu:.3e} (bound {tol_lowp_vs_emu_at(N):.2e}); abc == naive {same}", flush=True)
        assert rel <= tol_lowp_at(dtype, nlev, N)
        assert d_emu <= tol_lowp_vs_emu_at(N)
        assert same
        del outs, C
    del A, B


# Refined products against the FP64 product of the original FP64 inputs.
# Bounds: the correction removes the input rounding, so what remains is the
# FP32 accumulation of depth 8192 (plain and FP32 / split main products,
# measured 0.6e-6 .. 8e-6) or, for a BF16 FMM main product, the BF16
# rounding of its operand sums (<2,2,2> L1: measured 4.3e-3, the stated BF16
# one-level bound applies).
REFINE_CASES = [
    ("bf16", "gemm", "f32", 2e-5),
    ("bf16", "gemm", "bf16", 3e-5),
    ("f16", "gemm", "f32", 2e-5),
    ("f32", "222L1", "f32", 3e-5),
    ("bf16x3", "222L1", "bf16x3", 1e-4),
    ("bf16", "222L1", "bf16", 2.5e-2),
]


def test_refinement_8192(torch, fmm, rt, sweep_inputs):
    N, A64, B64 = sweep_inputs
    ii, jj = sample_idx(N, N, 512, 4)
    It, Jt = torch.as_tensor(ii, device="cuda"), torch.as_tensor(jj, device="cuda")
    Ah = A64[It, :].cpu().numpy()
    Bh = B64[:, Jt].cpu().numpy()
    ld = np.longdouble
    ex = (Ah.astype(ld) * Bh.T.astype(ld)).sum(axis=1)
    C = rt.colmajor(N, N)
    for main, a, corr, bound in REFINE_CASES:
        plan = rt.Plan(fmm.build_algorithm(*SWEEP_ALGS[a])) if SWEEP_ALGS[a] else None
        C.zero_()
        rt.refine(A64, B64, C, main=main, main_plan=plan, corr=corr)
        torch.cuda.synchronize()
        got = C[It, Jt].cpu().numpy().astype(ld)
        rel = float(np.sqrt(((got - ex) ** 2).sum() / (ex ** 2).sum()))
        print(f"\n[{main}+refine({corr}) {a}] rel err vs FP64 {rel:.3e} (bound {bound:.1e})", flush=True)
        assert rel <= bound
    assert rt.ws_status() == 0


# ------------------------------------------------------------- batched config

def test_batch_1024_problems(torch, fmm, rt):
    from fmm_gen_amd import autotune as AT
    rng = np.random.default_rng(2024)
    dims = np.exp(rng.uniform(np.log(192), np.log(4096), (1024, 3))).round().astype(int)
    g = torch.Generator(device="cuda").manual_seed(5)
    ops = []
    for m, n, k in dims:
        A = rt.colmajor(int(m), int(k), dtype=torch.float32).uniform_(-1, 1, generator=g)
        B = rt.colmajor(int(k), int(n), dtype=torch.float32).uniform_(-1, 1, generator=g)
        C = rt.colmajor(int(m), int(n), dtype=torch.float32).uniform_(-1, 1, generator=g)
        ops.append((A, B, C))
    tuner = AT.Tuner("f32")
    # every candidate algorithm and both tile shapes present in the launch
    cands = list(tuner.candidates())
    choices = [cands[i % len(cands)] for i in range(len(ops))]
    sample = sorted(set(rng.choice(len(ops), 48, replace=False).tolist()) | set(range(len(cands))))
    C0 = {i: ops[i][2].clone() for i in sample}
    b, _ = tuner.batch(ops, choices=choices)
    b.run()
    torch.cuda.synchronize()
    assert b.status() == 0
    worst = 0.0
    for i in sample:
        A, B, C = ops[i]
        a, t = choices[i]
        single = rt.as_colmajor(C0[i])
        plan = tuner.plan(a)
        if plan is None:
            rt.gemm(A, B, single)
        else:
            plan.run(A, B, single)
        torch.cuda.synchronize()
        assert torch.equal(single, C), (i, a, t, tuple(dims[i]))
        m, n, k = (int(x) for x in dims[i])
        si, sj = sample_idx(m, n, 64, i)
        ex = exact_at(A.cpu().numpy().astype(np.float64), B.cpu().numpy().astype(np.float64),
                      C0[i].cpu().numpy().astype(np.float64), si, sj)
        got = C.cpu().numpy()[si, sj].astype(np.longdouble)
        rel = float(np.sqrt(((got - ex) ** 2).sum() / (ex ** 2).sum()))
        nlev = 0 if AT.CANDIDATE_ALGS[a] is None else sum(AT.CANDIDATE_ALGS[a][1])
        assert rel <= tol_lowp_at("f32", nlev, k), (i, a, rel)
        worst = max(worst, rel / tol_lowp_at("f32", nlev, k))
    print(f"\n[batch 1024] {len(sample)} sampled problems equal single calls; worst error/bound {worst:.3f}; "
          f"{b.jobs} jobs, {b.units} units", flush=True)
    b.close()


# ------------------------------------------------------------- serving config

def test_serving_full_shapes_same_alone_and_in_mix(torch, fmm, rt):
    from fmm_gen_amd import serve as S
    reqs = S.poisson_mix(16, qps=50000.0, seed=9)
    kinds = {("sq" if r.m == r.n == r.k else "rk" if r.m == r.n else "ts", r.dtype) for r in reqs}
    rng = np.random.default_rng(9)
    for r in reqs:
        t = np.float32 if r.dtype == "f32" else np.float64
        r.host = tuple(rng.uniform(-1, 1, s).astype(t) for s in ((r.m, r.k), (r.k, r.n), (r.m, r.n)))
    ex = S.DeviceExecutor(keep_results=True)
    srv = S.Server(ex, hbm_budget=200e9, lanes_per_dtype=2, max_batch=16)
    ex.warmup(srv.lanes)
    srv.serve(reqs, timeout=300)
    assert len(srv.done) == len(reqs) and srv.stats()["failed"] == 0
    assert max(srv.batch_sizes) > 1
    for r in reqs:
        t = torch.float32 if r.dtype == "f32" else torch.float64
        dev = [rt.as_colmajor(torch.from_numpy(np.ascontiguousarray(h)).to("cuda", t)) for h in r.host]
        plan = ex.tuners[r.dtype].plan(r.alg)
        if plan is None:
            rt.gemm(*dev)
        else:
            plan.run(*dev)
        torch.cuda.synchronize()
        assert np.array_equal(dev[2].cpu().numpy(), r.result), (r.rid, r.dtype, r.alg, (r.m, r.n, r.k))
    print(f"\n[serving full shapes] {len(reqs)} requests, shape/dtype classes {sorted(kinds)}, "
          f"algorithms {sorted({r.alg for r in reqs})}, batches {srv.batch_sizes}", flush=True)
