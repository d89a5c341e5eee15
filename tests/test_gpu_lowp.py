"""GPU parity of the low-precision ABC FMM and plain GEMM (FP32 / BF16 /
FP16 inputs with FP32 C, and split-BF16 products of FP32 inputs).

* Integer known-answer cases are bit-exact in every type: every table here
  has coefficients +-1, so with inputs in [-4, 4] the operand sums are
  integers <= 32 (exact in BF16 and FP16) and every product and partial sum
  stays below 2^24 (exact in FP32).
* Float cases are checked twice: against the numpy restatement of the
  kernels' rounding points (oracle.fmm_lowp; only FP32 accumulation order
  differs) and against the FP64 product of the stored inputs, within the
  stated per-dtype, per-level tolerance (tests/tolerance.py).
"""
import numpy as np
import pytest

from conftest import int_mats
from tolerance import TOL_LOWP_VS_EMU, tol_lowp

pytestmark = pytest.mark.gpu

SPECS = [(["222"], [1]), (["222"], [2]), (["222", "323"], [1, 1]), (["333"], [1]), (["232"], [1])]
IDS = ["222L1", "222L2", "222x323", "333L1", "232L1"]
DTYPES = ["f32", "bf16", "f16", "bf16x3"]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def in_type(torch, dtype):
    return {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16, "bf16x3": torch.float32}[dtype]


def to_dev(torch, rt, X, tdtype, ld=None):
    t = rt.colmajor(X.shape[0], X.shape[1], ld=ld, dtype=tdtype, device="cuda")
    t.copy_(torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(tdtype))
    return t


def run(torch, rt, dtype, A, B, C, plan=None, ldpad=0):
    m, k = A.shape
    n = B.shape[1]
    ti = in_type(torch, dtype)
    dA = to_dev(torch, rt, A, ti, ld=m + ldpad if ldpad else None)
    dB = to_dev(torch, rt, B, ti, ld=k + ldpad if ldpad else None)
    dC = to_dev(torch, rt, C, torch.float32, ld=m + ldpad if ldpad else None)
    split = dtype == "bf16x3"
    if plan is None:
        rt.gemm(dA, dB, dC, split=split)
    else:
        # at least the ticket and error-flag slots (264 ints) even when the
        # shape needs no ordering, so the flag read below stays in bounds
        ws = torch.zeros(max(plan.workspace_bytes(m, n, k) // 4, 264) + 64, dtype=torch.int32, device="cuda")
        plan.abc(dA, dB, dC, ws=ws, split=split)
        torch.cuda.synchronize()
        flags = rt.ws_status(ws)
        assert flags == 0, f"launch error flags {flags} ({dtype}, {m}x{n}x{k})"
    torch.cuda.synchronize()
    return dC.cpu().numpy().astype(np.float32)


def stored(oracle, X, dtype):
    """The values an input of this dtype holds (float32 array)."""
    return oracle.round_op(np.asarray(X, dtype=np.float32), "f32" if dtype == "bf16x3" else dtype)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("dtype", DTYPES)
def test_lowp_gemm_kat_bit_exact(torch, rt, dtype):
    rng = np.random.default_rng(11)
    for (m, n, k) in [(1, 1, 1), (129, 130, 17), (300, 257, 513), (64, 8, 1000), (5, 700, 33), (256, 128, 64)]:
        A, B, C = int_mats(rng, m, n, k)
        got = run(torch, rt, dtype, A, B, C)
        assert np.array_equal(got, C + A @ B), (dtype, m, n, k, float(np.abs(got - (C + A @ B)).max()))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("spec", SPECS, ids=IDS)
def test_lowp_fmm_kat_bit_exact(torch, fmm, rt, spec, dtype):
    alg = fmm.build_algorithm(*spec)
    plan = rt.Plan(alg)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(5)
    for (m, n, k) in [(Mx * 8 * 40, Nx * 4 * 40, Kx * 70),              # no fringe, >1 tile
                      (Mx * 8 * 33 + 7, Nx * 4 * 33 + 3, Kx * 37 + 1),  # all fringes
                      (Mx * 8 - 3, 50, 20),                             # no FMM core
                      (Mx * 8 * 3, Nx * 4 * 3, max(Kx - 1, 1))]:        # depth < Kx
        A, B, C = int_mats(rng, m, n, k)
        want = C + A @ B
        got = run(torch, rt, dtype, A, B, C, plan=plan)
        assert np.array_equal(got, want), (spec, dtype, m, n, k, float(np.abs(got - want).max()))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("spec", [None] + SPECS, ids=["gemm"] + IDS)
def test_lowp_float_tolerance(torch, fmm, rt, oracle, spec, dtype):
    rng = np.random.default_rng(7)
    if spec is None:
        alg, plan, levels = None, None, 0
        m, n, k = 301, 283, 517
    else:
        alg = fmm.build_algorithm(*spec)
        plan = rt.Plan(alg)
        levels = sum(spec[1])
        Mx, Kx, Nx = alg.dims
        m, n, k = Mx * 8 * 30 + 5, Nx * 4 * 30 + 3, Kx * 60 + 1
    A = stored(oracle, rng.uniform(-1, 1, (m, k)), dtype)
    B = stored(oracle, rng.uniform(-1, 1, (k, n)), dtype)
    C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    got = run(torch, rt, dtype, A, B, C, plan=plan)
    exact = C.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
    if alg is None:
        emu = (C.astype(np.float64) + oracle._lowp_product(A, B, dtype)).astype(np.float32)
    else:
        emu = oracle.fmm_lowp(alg, A, B, C, dtype)
    e_emu = rel(got, emu)
    e_fp64 = rel(got, exact)
    print(f"{dtype} {'gemm' if spec is None else spec}: vs restatement {e_emu:.2e}, vs FP64 {e_fp64:.2e}")
    assert e_emu <= TOL_LOWP_VS_EMU, (dtype, spec, e_emu)
    assert e_fp64 <= tol_lowp(dtype, levels), (dtype, spec, e_fp64)


@pytest.mark.parametrize("dtype", DTYPES)
def test_lowp_strided_and_reproducible(torch, fmm, rt, oracle, dtype):
    alg = fmm.build_algorithm(["222"], [1])
    plan = rt.Plan(alg)
    rng = np.random.default_rng(3)
    m, n, k = 16 * 25 + 3, 8 * 27 + 1, 2 * 71 + 1
    A, B, C = int_mats(rng, m, n, k)
    for pad in (1, 3, 8):   # odd pads force the element-wise access paths
        got = run(torch, rt, dtype, A, B, C, plan=plan, ldpad=pad)
        assert np.array_equal(got, C + A @ B), (dtype, pad)
    A = stored(oracle, rng.uniform(-1, 1, (m, k)), dtype)
    B = stored(oracle, rng.uniform(-1, 1, (k, n)), dtype)
    C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    first = run(torch, rt, dtype, A, B, C, plan=plan)
    for _ in range(2):
        assert np.array_equal(run(torch, rt, dtype, A, B, C, plan=plan), first), dtype


def test_lowp_argument_errors(torch, fmm, rt):
    plan = rt.Plan(fmm.build_algorithm(["222"], [1]))
    A = rt.colmajor(64, 64, dtype=torch.bfloat16)
    B = rt.colmajor(64, 64, dtype=torch.bfloat16)
    C32 = rt.colmajor(64, 64, dtype=torch.float32)
    C64 = rt.colmajor(64, 64, dtype=torch.float64)
    with pytest.raises(rt.FmmError):
        plan.abc(A, B, C64)                                  # C must be FP32
    with pytest.raises(rt.FmmError):
        plan.abc(A, rt.colmajor(64, 64, dtype=torch.float16), C32)   # mixed inputs
    with pytest.raises(rt.FmmError):
        plan.abc(A, B, C32, split=True)                      # split needs FP32 inputs


@pytest.mark.parametrize("pack", ["ab", "naive"])
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("spec", [(["222"], [1]), (["333"], [1]), (["222", "323"], [1, 1])],
                         ids=["222L1", "333L1", "222x323"])
def test_lowp_packs_equal_abc(torch, fmm, rt, oracle, spec, dtype, pack):
    """AB and Naive with low-precision inputs: the stored operand sums are
    rounded like the ABC kernel's LDS sums, and AB's M_Add and Naive's fused
    C updates (the k-fringe in a second launch) apply the same FP32 updates
    in the same order, so the results equal ABC's bit for bit."""
    alg = fmm.build_algorithm(*spec)
    plan = rt.Plan(alg)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(31)
    ti = in_type(torch, dtype)
    split = dtype == "bf16x3"
    for (m, n, k) in [(Mx * 8 * 33 + 7, Nx * 4 * 33 + 3, Kx * 37 + 1), (Mx * 8 * 40, Nx * 4 * 40, Kx * 70)]:
        A = stored(oracle, rng.uniform(-1, 1, (m, k)), dtype)
        B = stored(oracle, rng.uniform(-1, 1, (k, n)), dtype)
        C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
        outs = {}
        for pk in ("abc", pack):
            dA = to_dev(torch, rt, A, ti)
            dB = to_dev(torch, rt, B, ti)
            dC = to_dev(torch, rt, C, torch.float32)
            plan.run(dA, dB, dC, pack=pk, split=split)
            torch.cuda.synchronize()
            outs[pk] = dC.cpu().numpy()
        assert np.array_equal(outs["abc"], outs[pack]), (spec, dtype, pack, m, n, k)
        Ai, Bi, Ci = int_mats(rng, m, n, k)
        dA, dB, dC = to_dev(torch, rt, Ai, ti), to_dev(torch, rt, Bi, ti), to_dev(torch, rt, Ci, torch.float32)
        plan.run(dA, dB, dC, pack=pack, split=split)
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy(), Ci + Ai @ Bi), (spec, dtype, pack, "int", m, n, k)


def sign_algorithm(fmm):
    """<1,2,1> with two one-term products whose operands carry -1:
    M1 = (-A0) B0, M2 = A1 (-B1), C = -M1 - M2 = A0 B0 + A1 B1."""
    from fractions import Fraction as Fr
    return fmm.Algorithm(name="sign121", levels=[(1, 2, 1)], U=[[Fr(-1), Fr(0)], [Fr(0), Fr(1)]],
                         V=[[Fr(1), Fr(0)], [Fr(0), Fr(-1)]], W=[[Fr(-1), Fr(-1)]])


@pytest.mark.parametrize("dtype", ["bf16", "f16", "bf16x3"])
def test_lowp_single_term_direct_paths(torch, fmm, rt, dtype):
    """The single-term 16-bit launches (direct-to-LDS kernel): plain GEMMs with
    aligned and odd leading dimensions (odd ones take the register-staged
    loads), edge tiles and partial last k stages; and products whose single
    operand terms carry -1 (the sign is applied to the accumulators), two of
    them on every C tile in product order."""
    rng = np.random.default_rng(17)
    for (m, n, k) in [(264, 136, 100), (520, 300, 96), (40, 520, 33)]:
        A, B, C = int_mats(rng, m, n, k)
        for pad in (0, 1, 3, 8):
            got = run(torch, rt, dtype, A, B, C, ldpad=pad)
            assert np.array_equal(got, C + A @ B), (dtype, m, n, k, pad)
    plan = rt.Plan(sign_algorithm(fmm))
    for (m, n, k) in [(512, 512, 256), (264, 300, 130)]:
        A, B, C = int_mats(rng, m, n, k)
        for pad in (0, 3):
            got = run(torch, rt, dtype, A, B, C, plan=plan, ldpad=pad)
            assert np.array_equal(got, C + A @ B), (dtype, "sign", m, n, k, pad)
