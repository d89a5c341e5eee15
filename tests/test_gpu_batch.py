"""GPU tests of the batched persistent launch.

Every problem of a batch must get exactly (bit for bit) the result of a
single call with the same plan and dtype, whatever its C-tile shape and
whatever else runs in the launch; integer problems must be exact; a batch
captured in a HIP graph must give the eager result on every replay.
"""
import numpy as np
import pytest

from conftest import int_mats

pytestmark = pytest.mark.gpu

ALGS = [None, (["222"], [1]), (["222"], [2]), (["232"], [1]), (["333"], [1])]
DTYPES = ["f64", "f32", "bf16", "f16", "bf16x3"]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def tdt(torch, dtype):
    return {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16,
            "bf16x3": torch.float32}[dtype]


def make(torch, rt, dtype, A, B, C):
    ti = tdt(torch, dtype)
    tc = torch.float64 if dtype == "f64" else torch.float32

    def dev(X, t):
        d = rt.colmajor(X.shape[0], X.shape[1], dtype=t)
        d.copy_(torch.from_numpy(np.ascontiguousarray(X)).to(t))
        return d
    return dev(A, ti), dev(B, ti), dev(C, tc)


def shapes(rng, n):
    out = [(1, 1, 1), (7, 300, 1), (300, 5, 70), (256, 128, 64)]
    while len(out) < n:
        out.append(tuple(int(x) for x in rng.integers(1, 700, 3)))
    return out


@pytest.mark.parametrize("dtype", DTYPES)
def test_batch_equals_single_calls(torch, fmm, rt, dtype):
    rng = np.random.default_rng(21)
    plans = [None if a is None else rt.Plan(fmm.build_algorithm(*a)) for a in ALGS]
    split = dtype == "bf16x3"
    probs, singles, exact = [], [], []
    for i, (m, n, k) in enumerate(shapes(rng, 28)):
        A, B, C = int_mats(rng, m, n, k)
        if i % 2:   # float problems
            A = rng.uniform(-1, 1, (m, k))
            B = rng.uniform(-1, 1, (k, n))
            C = rng.uniform(-1, 1, (m, n))
        plan = plans[i % len(plans)]
        tile = int(rng.integers(0, 2))
        dA, dB, dC = make(torch, rt, dtype, A, B, C)
        probs.append((dA, dB, dC, plan, tile))
        sA, sB, sC = make(torch, rt, dtype, A, B, C)
        if plan is None:
            rt.gemm(sA, sB, sC, split=split)
        else:
            plan.run(sA, sB, sC, split=split)
        singles.append(sC)
        exact.append(None if i % 2 else C + A @ B)
    batch = rt.Batch(probs, split=split)
    batch.run()
    torch.cuda.synchronize()
    assert batch.status() == 0
    for i, ((_, _, dC, plan, tile), sC, ex) in enumerate(zip(probs, singles, exact)):
        got = dC.cpu().numpy()
        assert np.array_equal(got, sC.cpu().numpy()), (dtype, i, tile)
        if ex is not None:
            assert np.array_equal(got.astype(np.float64), ex), (dtype, i, tile)


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_batch_rerun_and_graph_replay(torch, fmm, rt, dtype):
    rng = np.random.default_rng(5)
    plans = [None if a is None else rt.Plan(fmm.build_algorithm(*a)) for a in ALGS]
    probs, C0s = [], []
    for i, (m, n, k) in enumerate(shapes(rng, 16)):
        A = rng.uniform(-1, 1, (m, k))
        B = rng.uniform(-1, 1, (k, n))
        C = rng.uniform(-1, 1, (m, n))
        dA, dB, dC = make(torch, rt, dtype, A, B, C)
        probs.append((dA, dB, dC, plans[i % len(plans)], i % 2))
        C0s.append(dC.clone())
    batch = rt.Batch(probs)

    def reset():
        for (_, _, dC, _, _), c0 in zip(probs, C0s):
            dC.copy_(c0)
    batch.run()
    torch.cuda.synchronize()
    eager = [p[2].clone() for p in probs]
    reset()
    batch.run()
    torch.cuda.synchronize()
    for p, e in zip(probs, eager):
        assert torch.equal(p[2], e)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    reset()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        batch.run()
    for rep in range(3):
        reset()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i, (p, e) in enumerate(zip(probs, eager)):
            assert torch.equal(p[2], e), (dtype, rep, i)
        assert batch.status() == 0


def test_f64_batch_variants(torch, fmm, rt):
    """FP64 batches pick their kernel by the largest operand-sum term count:
    only plain GEMMs -> the single-term kernel (variant 2), at most 2 terms
    (222) -> variant 0, more (333) -> variant 1; each problem still equals its
    single call bit for bit, and integer problems are exact."""
    rng = np.random.default_rng(23)
    p222 = rt.Plan(fmm.build_algorithm(["222"], [1]))
    p333 = rt.Plan(fmm.build_algorithm(["333"], [1]))
    for plan, want in ((None, 2), (p222, 0), (p333, 1)):
        probs, singles, exact = [], [], []
        for i, (m, n, k) in enumerate(shapes(rng, 8)):
            A, B, C = int_mats(rng, m, n, k)
            if i % 2:
                A = rng.uniform(-1, 1, (m, k))
                B = rng.uniform(-1, 1, (k, n))
            dA, dB, dC = make(torch, rt, "f64", A, B, C)
            probs.append((dA, dB, dC, plan, 0))
            sA, sB, sC = make(torch, rt, "f64", A, B, C)
            if plan is None:
                rt.gemm(sA, sB, sC)
            else:
                plan.run(sA, sB, sC)
            singles.append(sC)
            exact.append(None if i % 2 else C + A @ B)
        batch = rt.Batch(probs)
        assert batch.variant() == want
        batch.run()
        torch.cuda.synchronize()
        assert batch.status() == 0
        for i, ((_, _, dC, _, _), sC, ex) in enumerate(zip(probs, singles, exact)):
            got = dC.cpu().numpy()
            assert np.array_equal(got, sC.cpu().numpy()), (want, i)
            if ex is not None:
                assert np.array_equal(got, ex), (want, i)


def test_batch_errors_and_empty(torch, fmm, rt):
    A = rt.colmajor(8, 8, dtype=torch.float32)
    B = rt.colmajor(8, 8, dtype=torch.bfloat16)
    C = rt.colmajor(8, 8, dtype=torch.float32)
    with pytest.raises(rt.FmmError):
        rt.Batch([(A, A, C, None, 0), (B, B, C, None, 0)])   # mixed dtypes
    e = rt.Batch([])
    e.run()
    torch.cuda.synchronize()
    assert e.units == 0


def test_tuned_batch_matches_oracle(torch, fmm, rt, oracle):
    from fmm_gen_amd import autotune as AT
    rng = np.random.default_rng(9)
    tuner = AT.Tuner("f32")
    ops, host = [], []
    for _ in range(12):
        m, n, k = (int(x) for x in np.exp(rng.uniform(np.log(192), np.log(1200), 3)).round())
        A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
        B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
        C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
        ops.append(make(torch, rt, "f32", A, B, C))
        host.append((A, B, C))
    # every candidate appears at least once, the rest as tuned
    cands = list(tuner.candidates())
    choices = [cands[i] if i < len(cands) else tuner.choose(h[0].shape[0], h[1].shape[1], h[0].shape[1])[:2]
               for i, h in enumerate(host)]
    batch, choices = tuner.batch(ops, choices=choices)
    batch.run()
    torch.cuda.synchronize()
    for (A, B, C), (_, _, dC), (alg, tile) in zip(host, ops, choices):
        if AT.CANDIDATE_ALGS[alg] is None:
            emu = (C.astype(np.float64) + oracle._lowp_product(A, B, "f32")).astype(np.float32)
        else:
            emu = oracle.fmm_lowp(fmm.build_algorithm(*AT.CANDIDATE_ALGS[alg]), A, B, C, "f32")
        got = dC.cpu().numpy()
        assert np.linalg.norm(got - emu) / np.linalg.norm(emu) < 4e-6, (alg, tile, A.shape, B.shape)
