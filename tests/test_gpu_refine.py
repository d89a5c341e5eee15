"""GPU tests of the FP64 refinement entry (fmm_refine): exact on integer
inputs, equal to the numpy restatement of its split / products / FP64 update
(oracle.refine_lowp) up to FP32 accumulation differences, and never less
accurate than the main product alone."""
import numpy as np
import pytest

from conftest import int_mats
from tolerance import rel_frob

pytestmark = pytest.mark.gpu

CASES = [("bf16", None), ("bf16", (["222"], [1])), ("f16", None), ("f32", (["222"], [1])),
         ("bf16x3", (["222"], [1])), ("f32", (["333"], [1]))]
IDS = ["bf16-gemm", "bf16-222L1", "f16-gemm", "f32-222L1", "bf16x3-222L1", "f32-333L1"]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def dev64(torch, rt, X):
    d = rt.colmajor(X.shape[0], X.shape[1])
    d.copy_(torch.from_numpy(np.ascontiguousarray(X)))
    return d


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_refine_integer_exact(torch, fmm, rt, case):
    main, spec = case
    plan = rt.Plan(fmm.build_algorithm(*spec)) if spec else None
    rng = np.random.default_rng(1)
    for (m, n, k) in [(77, 65, 33), (2 * 8 * 20 + 3, 3 * 4 * 21 + 1, 3 * 30 + 2)]:
        A, B, C = int_mats(rng, m, n, k)
        dA, dB, dC = (dev64(torch, rt, X) for X in (A, B, C))
        rt.refine(dA, dB, dC, main=main, main_plan=plan, corr="f32")
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy(), C + A @ B), (case, m, n, k)


@pytest.mark.parametrize("corr", ["f32", "bf16x3", "bf16"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_refine_float(torch, fmm, rt, oracle, case, corr):
    main, spec = case
    if corr == "bf16" and main != "bf16":
        pytest.skip("a BF16 correction needs a BF16 main product (A1 exact in BF16)")
    alg = fmm.build_algorithm(*spec) if spec else None
    plan = rt.Plan(alg) if alg else None
    rng = np.random.default_rng(2)
    m, n, k = 3 * 8 * 20 + 5, 3 * 4 * 30 + 3, 3 * 40 + 1
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    C = rng.uniform(-1, 1, (m, n))
    dA, dB, dC = (dev64(torch, rt, X) for X in (A, B, C))
    rt.refine(dA, dB, dC, main=main, main_plan=plan, corr=corr)
    torch.cuda.synchronize()
    got = dC.cpu().numpy()
    exact = C + A @ B
    emu = oracle.refine_lowp(alg, main, None, corr, A, B, C)
    # the main product alone on the GPU (inputs rounded to its type, FP32
    # result), added to C in FP64
    inp = "f32" if main in ("f32", "bf16x3") else main
    ti = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}[inp]
    A1 = rt.as_colmajor(dA.to(ti))
    B1 = rt.as_colmajor(dB.to(ti))
    Ct = rt.colmajor(m, n, dtype=torch.float32).zero_()
    if plan is None:
        rt.gemm(A1, B1, Ct, split=main == "bf16x3")
    else:
        plan.abc(A1, B1, Ct, split=main == "bf16x3")
    torch.cuda.synchronize()
    alone = C + Ct.cpu().numpy().astype(np.float64)
    e_ref, e_alone, e_emu = rel_frob(got, A, B, C), rel_frob(alone, A, B, C), float(
        np.linalg.norm(got - emu) / np.linalg.norm(exact))
    print(f"{main} {spec} corr={corr}: alone {e_alone:.2e} refined {e_ref:.2e} vs restatement {e_emu:.2e}")
    assert e_emu <= 2e-6, (case, corr, e_emu)
    assert e_ref <= e_alone * 1.05 + 1e-12, (case, corr, e_ref, e_alone)
    if spec is None:   # a plain main product loses only input rounding, which the correction restores
        assert e_ref <= 5e-5, (case, corr, e_ref)
