"""GPU parity of the HIP ABC FMM and plain GEMM against the oracle (FP64).

Integer known-answer cases must be bit-exact (all sums exact in FP64); float
cases must stay within the stated per-level tolerance (tests/tolerance.py)
and agree with the CPU restatement; repeated runs must be bitwise identical.
"""
import json
import os

import numpy as np
import pytest

from conftest import int_mats
from tolerance import TOL_GEMM, norm_err, tol_for

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SPECS = [(["222"], [1]), (["222"], [2]), (["222", "323"], [1, 1]), (["333"], [1]),
         (["232"], [1]), (["323"], [1])]
IDS = ["222L1", "222L2", "222x323", "333L1", "232L1", "323L1"]


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def to_dev(torch, rt, X, ld=None):
    t = rt.colmajor(X.shape[0], X.shape[1], ld=ld, device="cuda")
    t.copy_(torch.from_numpy(np.ascontiguousarray(X)))
    return t


def run_abc(torch, rt, plan, A, B, C, ldpad=0):
    m, k = A.shape
    n = B.shape[1]
    dA = to_dev(torch, rt, A, ld=m + ldpad if ldpad else None)
    dB = to_dev(torch, rt, B, ld=k + ldpad if ldpad else None)
    dC = to_dev(torch, rt, C, ld=m + ldpad if ldpad else None)
    plan.abc(dA, dB, dC)
    torch.cuda.synchronize()
    return dC.cpu().numpy()


def test_gemm_integer_and_float(torch, rt):
    rng = np.random.default_rng(1)
    for (m, n, k) in [(1, 1, 1), (129, 130, 17), (300, 257, 513), (64, 8, 1000), (5, 700, 33)]:
        A, B, C = int_mats(rng, m, n, k)
        dA, dB, dC = (to_dev(torch, rt, X) for X in (A, B, C))
        rt.gemm(dA, dB, dC)
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy(), C + A @ B), (m, n, k)
        A = rng.uniform(-1, 1, (m, k))
        B = rng.uniform(-1, 1, (k, n))
        C = rng.uniform(-1, 1, (m, n))
        dA, dB, dC = (to_dev(torch, rt, X) for X in (A, B, C))
        rt.gemm(dA, dB, dC)
        torch.cuda.synchronize()
        assert norm_err(dC.cpu().numpy(), A, B, C) <= TOL_GEMM, (m, n, k)


@pytest.mark.parametrize("spec", SPECS, ids=IDS)
def test_abc_kat_bit_exact(torch, fmm, rt, oracle, spec):
    alg = fmm.build_algorithm(*spec)
    plan = rt.Plan(alg)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(5)
    for (m, n, k) in [(Mx * 8 * 20, Nx * 4 * 40, Kx * 50),             # no fringe, >1 tile
                      (Mx * 8 * 20 + 7, Nx * 4 * 33 + 3, Kx * 37 + 1),   # all fringes
                      (Mx * 8 - 3, 50, 20),                              # no FMM core
                      (Mx * 8 * 3, Nx * 4 * 3, max(Kx - 1, 1))]:         # depth < Kx
        A, B, C = int_mats(rng, m, n, k)
        want = C + A @ B
        got = run_abc(torch, rt, plan, A, B, C)
        assert np.array_equal(got, want), (spec, m, n, k, float(np.abs(got - want).max()))
        ref = C.copy(order="F")
        oracle.fmm(alg, A, B, ref, "abc")
        assert np.array_equal(got, ref)


def test_abc_golden_kat(torch, fmm, rt):
    z = np.load(os.path.join(GOLDEN, "kat_abc.npz"), allow_pickle=False)
    index = json.loads(str(z["index"]))
    plans = {}
    for i, case in enumerate(index):
        key = (tuple(case["algs"]), tuple(case["levels"]))
        if key not in plans:
            plans[key] = rt.Plan(fmm.build_algorithm(list(key[0]), list(key[1])))
        A, B, C = (np.asfortranarray(z[f"{t}{i}"]) for t in ("A", "B", "C"))
        got = run_abc(torch, rt, plans[key], A, B, C)
        assert np.array_equal(got, z[f"out{i}"]), case


@pytest.mark.parametrize("spec", SPECS, ids=IDS)
def test_abc_float_tolerance_and_oracle(torch, fmm, rt, oracle, spec):
    alg = fmm.build_algorithm(*spec)
    plan = rt.Plan(alg)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(6)
    m, n, k = Mx * 8 * 12 + 5, Nx * 4 * 24 + 2, Kx * 96 + 1
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    C = rng.uniform(-1, 1, (m, n))
    got = run_abc(torch, rt, plan, A, B, C)
    tol = tol_for(len(alg.levels))
    assert norm_err(got, A, B, C) <= tol
    ref = np.asfortranarray(C.copy())
    oracle.fmm(alg, np.asfortranarray(A), np.asfortranarray(B), ref, "abc")
    # both within tolerance of the exact result, hence of each other
    assert norm_err(ref, A, B, C) <= tol


def test_abc_strided_leading_dims(torch, fmm, rt):
    alg = fmm.build_algorithm(["222"], [1])
    plan = rt.Plan(alg)
    rng = np.random.default_rng(7)
    A, B, C = int_mats(rng, 203, 161, 99)
    for pad in (1, 3, 8):
        got = run_abc(torch, rt, plan, A, B, C, ldpad=pad)
        assert np.array_equal(got, C + A @ B), pad


def test_abc_bitwise_reproducible(torch, fmm, rt):
    alg = fmm.build_algorithm(["222", "323"], [1, 1])
    plan = rt.Plan(alg)
    rng = np.random.default_rng(8)
    m, n, k = 1000, 900, 500
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    C = rng.uniform(-1, 1, (m, n))
    r1 = run_abc(torch, rt, plan, A, B, C)
    r2 = run_abc(torch, rt, plan, A, B, C)
    assert np.array_equal(r1, r2)


def test_workspace_reuse_and_status(torch, fmm, rt):
    alg = fmm.build_algorithm(["333"], [1])
    plan = rt.Plan(alg)
    rng = np.random.default_rng(9)
    m, n, k = 720, 600, 300
    nbytes = plan.workspace_bytes(m, n, k)
    assert nbytes > 0
    ws = torch.zeros(nbytes // 4 + 16, dtype=torch.int32, device="cuda")
    A, B, C = int_mats(rng, m, n, k)
    dA, dB = to_dev(torch, rt, A), to_dev(torch, rt, B)
    for _ in range(3):
        dC = to_dev(torch, rt, C)
        plan.abc(dA, dB, dC, ws=ws)
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy(), C + A @ B)
        assert rt.ws_status(ws) == 0


def test_host_dropin(torch, fmm, rt):
    alg = fmm.build_algorithm(["232"], [1])
    plan = rt.Plan(alg)
    rng = np.random.default_rng(10)
    A, B, C = int_mats(rng, 150, 90, 61)
    got = C.copy(order="F")
    plan.host_abc(A, B, got)
    assert np.array_equal(got, C + A @ B)


def test_config1_strassen_2048(torch, fmm, rt, oracle):
    """Strassen L1 ABC FP64, m=n=k=2048, A,B ~ U[0,1), C = 0."""
    alg = fmm.build_algorithm(["222"], [1])
    plan = rt.Plan(alg)
    rng = np.random.default_rng(2048)
    m = n = k = 2048
    A = np.asfortranarray(rng.random((m, k)))
    B = np.asfortranarray(rng.random((k, n)))
    C = np.zeros((m, n), order="F")
    got = run_abc(torch, rt, plan, A, B, C)
    assert norm_err(got, A, B, C) <= tol_for(1)
    again = run_abc(torch, rt, plan, A, B, C)
    assert np.array_equal(got, again)
    ref = C.copy(order="F")
    oracle.fmm(alg, A, B, ref, "abc")
    assert float(np.abs(got - ref).max()) <= 2 * tol_for(1) * 2.0 ** -53 * k


@pytest.mark.parametrize("handover", ["0", "1"])
def test_ordered_launch_both_handover_modes(torch, fmm, rt, oracle, handover, monkeypatch):
    """The ordered launches give the same results with the XCD-local C
    hand-over (release = wait for the writer's stores in the shared L2) and
    with the agent-scope release everywhere (FMM_LOCAL_HANDOVER=0): integer
    cases bit-exact, float cases bitwise equal between the modes.  Shapes with
    many products per C tile and more units than workgroups."""
    monkeypatch.setenv("FMM_LOCAL_HANDOVER", handover)
    rng = np.random.default_rng(31)
    for spec in ((["222", "323"], [1, 1]), (["333"], [1])):
        alg = fmm.build_algorithm(*spec)
        plan = rt.Plan(alg)
        Mx, Kx, Nx = alg.dims
        m, n, k = Mx * 8 * 40 + 3, Nx * 4 * 70 + 1, Kx * 90 + 2
        A, B, C = int_mats(rng, m, n, k)
        got = run_abc(torch, rt, plan, A, B, C)
        assert np.array_equal(got, C + A @ B), (spec, handover)
        Af = rng.uniform(-1, 1, (m, k))
        Bf = rng.uniform(-1, 1, (k, n))
        Cf = rng.uniform(-1, 1, (m, n))
        f_mode = run_abc(torch, rt, plan, Af, Bf, Cf)
        monkeypatch.setenv("FMM_LOCAL_HANDOVER", "1" if handover == "0" else "0")
        f_other = run_abc(torch, rt, plan, Af, Bf, Cf)
        monkeypatch.setenv("FMM_LOCAL_HANDOVER", handover)
        assert np.array_equal(f_mode, f_other), spec
        plan.close()
