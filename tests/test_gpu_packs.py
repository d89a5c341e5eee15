"""GPU parity of the AB and Naive pack types (FP64) against the oracle and
against the ABC path.

The three pack types apply the same C updates in the same order with the
same arithmetic (operand sums s = u0*x0, s = fma(u_t, x_t, s); products over
the whole depth in one accumulator; C_j = fma(w_j, M_r, C_j) in product
order, then the k-fringe products), so on any input their results must be
bitwise equal to ABC's.  Integer known-answer cases are also checked
bit-exact against numpy and against the oracle's restatements of the
reference's ab / naive generators.

FP64 Naive has two product launches, picked by leaf depth: ordered products
into M plus M_Add, and the C-stationary launch (every product of a tile
position applied to register-resident C tiles).  "naive" forces the first
(FMM_CS=0), "naive-cs" the second (FMM_CS=1, algorithms with <= 9 C blocks).
"""
import contextlib
import os

import numpy as np
import pytest

from conftest import int_mats
from tolerance import norm_err, tol_for

pytestmark = pytest.mark.gpu

SPECS = [(["222"], [1]), (["222"], [2]), (["222", "323"], [1, 1]), (["333"], [1]),
         (["232"], [1]), (["323"], [1])]
IDS = ["222L1", "222L2", "222x323", "333L1", "232L1", "323L1"]
PACKS = ["ab", "naive", "naive-cs"]
CS_MODE = {"naive": "0", "naive-cs": "1"}


def base_pack(pack):
    return "naive" if pack == "naive-cs" else pack


@contextlib.contextmanager
def cs_mode(pack):
    """FMM_CS for the duration of one call of test pack type `pack`."""
    old = os.environ.get("FMM_CS")
    if pack in CS_MODE:
        os.environ["FMM_CS"] = CS_MODE[pack]
    try:
        yield
    finally:
        if old is None:
            os.environ.pop("FMM_CS", None)
        else:
            os.environ["FMM_CS"] = old


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


def run(torch, rt, plan, pack, A, B, C, ldpad=0, ws=None):
    m, k = A.shape
    dA = to_dev(torch, rt, A, ld=m + ldpad if ldpad else None)
    dB = to_dev(torch, rt, B, ld=k + ldpad if ldpad else None)
    dC = to_dev(torch, rt, C, ld=m + ldpad if ldpad else None)
    print(f"pack={pack} m={m} n={B.shape[1]} k={k} ldpad={ldpad}", flush=True)
    with cs_mode(pack):
        plan.run(dA, dB, dC, base_pack(pack), ws=ws)
        torch.cuda.synchronize()
    return dC.cpu().numpy()


@pytest.mark.parametrize("pack", PACKS)
@pytest.mark.parametrize("spec", SPECS, ids=IDS)
def test_pack_kat_bit_exact(torch, fmm, rt, oracle, spec, pack):
    alg = fmm.build_algorithm(*spec)
    plan = rt.Plan(alg)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(15)
    for (m, n, k) in [(Mx * 8 * 20, Nx * 4 * 40, Kx * 50),             # no fringe, >1 tile
                      (Mx * 8 * 20 + 7, Nx * 4 * 33 + 3, Kx * 37 + 1),   # all fringes
                      (Mx * 8 - 3, 50, 20),                              # no FMM core
                      (Mx * 8 * 3, Nx * 4 * 3, max(Kx - 1, 1))]:         # depth < Kx
        A, B, C = int_mats(rng, m, n, k)
        want = C + A @ B
        got = run(torch, rt, plan, pack, A, B, C)
        assert np.array_equal(got, want), (spec, pack, m, n, k, float(np.abs(got - want).max()))
        ref = C.copy(order="F")
        oracle.fmm(alg, A, B, ref, base_pack(pack))
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("spec", SPECS, ids=IDS)
def test_packs_bitwise_equal_abc(torch, fmm, rt, spec):
    alg = fmm.build_algorithm(*spec)
    plan = rt.Plan(alg)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(16)
    for (m, n, k) in [(Mx * 8 * 12 + 5, Nx * 4 * 24 + 2, Kx * 96 + 1),
                      (Mx * 8 * 17, Nx * 4 * 35, Kx * 71)]:
        A = rng.uniform(-1, 1, (m, k))
        B = rng.uniform(-1, 1, (k, n))
        C = rng.uniform(-1, 1, (m, n))
        base = run(torch, rt, plan, "abc", A, B, C)
        assert norm_err(base, A, B, C) <= tol_for(len(alg.levels))
        for pack in PACKS:
            got = run(torch, rt, plan, pack, A, B, C)
            assert np.array_equal(got, base), (spec, pack, m, n, k, float(np.abs(got - base).max()))


@contextlib.contextmanager
def env(name, value):
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        yield
    finally:
        if old is None:
            os.environ.pop(name, None)
        else:
            os.environ[name] = old


@pytest.mark.parametrize("spec", SPECS, ids=IDS)
def test_abc_round_kernel_bitwise(torch, fmm, rt, oracle, spec):
    """The staggered-round ABC kernel (even leaf depth >= 32) against the
    chunked ABC kernels (FMM_ABC_ROUNDS=0): bitwise equal on random inputs,
    with edge tiles (leaf rows / columns not multiples of 128, whose
    out-of-range rows and columns are loaded clamped), k-fringes and leaf
    depths that end in a partial 32-deep slice (masked loads); exact on
    integer inputs against numpy and the oracle."""
    alg = fmm.build_algorithm(*spec)
    plan = rt.Plan(alg)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(21)
    for (m, n, k) in [(Mx * 8 * 20, Nx * 4 * 40, Kx * 64),
                      (Mx * 8 * 17 + 6, Nx * 4 * 29 + 1, Kx * 96 + 2),
                      (Mx * 136, Nx * 132, Kx * 32),
                      (Mx * 8 * 9 + 2, Nx * 4 * 11, Kx * 146),
                      (Mx * 8 * 5, Nx * 4 * 13 + 3, Kx * 50 + 1)]:
        A = rng.uniform(-1, 1, (m, k))
        B = rng.uniform(-1, 1, (k, n))
        C = rng.uniform(-1, 1, (m, n))
        got = run(torch, rt, plan, "abc", A, B, C)
        with env("FMM_ABC_ROUNDS", "0"):
            want = run(torch, rt, plan, "abc", A, B, C)
        assert np.array_equal(got, want), (spec, m, n, k, float(np.abs(got - want).max()))
        Ai, Bi, Ci = int_mats(rng, m, n, k)
        goti = run(torch, rt, plan, "abc", Ai, Bi, Ci)
        assert np.array_equal(goti, Ci + Ai @ Bi), (spec, m, n, k)
        ref = Ci.copy(order="F")
        oracle.fmm(alg, Ai, Bi, ref, "abc")
        assert np.array_equal(goti, ref)


CS_SPECS = [(["222"], [1]), (["232"], [1]), (["323"], [1]), (["333"], [1])]


def sign_algorithm(fmm):
    """<1,2,1> with two one-term products whose operands carry -1:
    M1 = (-A0) B0, M2 = A1 (-B1), C = -M1 - M2 = A0 B0 + A1 B1."""
    from fractions import Fraction as Fr
    return fmm.Algorithm(name="sign121", levels=[(1, 2, 1)], U=[[Fr(-1), Fr(0)], [Fr(0), Fr(1)]],
                         V=[[Fr(1), Fr(0)], [Fr(0), Fr(-1)]], W=[[Fr(-1), Fr(-1)]])


@pytest.mark.parametrize("spec", CS_SPECS, ids=["222L1", "232L1", "323L1", "333L1"])
def test_cs_direct_kernel_bitwise(torch, fmm, rt, oracle, spec):
    """The C-stationary kernel with direct-to-LDS loads against the register-
    staged one (FMM_CSD=0), bitwise on random inputs: leaf depths that end in
    a partial stage of a multiple of 4 (direct, fewer k-steps) or not
    (register-staged), k-fringe products, leaf rows / columns that leave edge
    positions, odd leading dimensions (A / B blocks register-staged while the
    pre-sums stay direct); exact on integer inputs against numpy and the
    oracle."""
    alg = fmm.build_algorithm(*spec)
    plan = rt.Plan(alg)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(22)
    for (m, n, k, pad) in [(Mx * 8 * 20, Nx * 4 * 40, Kx * 48, 0),
                           (Mx * 8 * 17 + 6, Nx * 4 * 29 + 1, Kx * 96 + 2, 0),
                           (Mx * 136, Nx * 132, Kx * 240, 0),
                           (Mx * 8 * 9 + 2, Nx * 4 * 11, Kx * 146, 1),
                           (Mx * 8 * 5, Nx * 4 * 13 + 3, Kx * 50 + 1, 3)]:
        A = rng.uniform(-1, 1, (m, k))
        B = rng.uniform(-1, 1, (k, n))
        C = rng.uniform(-1, 1, (m, n))
        got = run(torch, rt, plan, "naive-cs", A, B, C, ldpad=pad)
        with env("FMM_CSD", "0"):
            want = run(torch, rt, plan, "naive-cs", A, B, C, ldpad=pad)
        assert np.array_equal(got, want), (spec, m, n, k, pad, float(np.abs(got - want).max()))
        Ai, Bi, Ci = int_mats(rng, m, n, k)
        goti = run(torch, rt, plan, "naive-cs", Ai, Bi, Ci, ldpad=pad)
        assert np.array_equal(goti, Ci + Ai @ Bi), (spec, m, n, k, pad)
        ref = Ci.copy(order="F")
        oracle.fmm(alg, Ai, Bi, ref, "naive")
        assert np.array_equal(goti, ref)


def test_cs_direct_kernel_signs(torch, fmm, rt):
    """Single-term products whose operands carry -1 take the direct path with
    the sign folded into their W coefficients: bitwise equal to the
    register-staged kernel, which scales the operands."""
    plan = rt.Plan(sign_algorithm(fmm))
    rng = np.random.default_rng(23)
    for (m, n, k, pad) in [(512, 512, 256, 0), (264, 300, 130, 0), (200, 136, 96, 3)]:
        A = rng.uniform(-1, 1, (m, k))
        B = rng.uniform(-1, 1, (k, n))
        C = rng.uniform(-1, 1, (m, n))
        got = run(torch, rt, plan, "naive-cs", A, B, C, ldpad=pad)
        with env("FMM_CSD", "0"):
            want = run(torch, rt, plan, "naive-cs", A, B, C, ldpad=pad)
        assert np.array_equal(got, want), (m, n, k, pad)
        Ai, Bi, Ci = int_mats(rng, m, n, k)
        assert np.array_equal(run(torch, rt, plan, "naive-cs", Ai, Bi, Ci, ldpad=pad), Ci + Ai @ Bi)


@pytest.mark.parametrize("pack", ["ab", "naive"])
def test_madd_many_blocks_bitwise(torch, fmm, rt, oracle, pack):
    """M_Add of <2,2,2> x <3,2,3> (36 C blocks): the 4-wave kernel (each wave a
    quarter of the C blocks, its own product list) against the one-wave
    kernel (FMM_MADD_ALL4=0), bitwise on random inputs with edge rows and
    columns, odd leading dimensions and a k-fringe; exact on integer inputs
    against numpy and the oracle."""
    alg = fmm.build_algorithm(["222", "323"], [1, 1])
    plan = rt.Plan(alg)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(24)
    for (m, n, k, pad) in [(Mx * 8 * 17, Nx * 4 * 35, Kx * 71, 0),
                           (Mx * 8 * 12 + 5, Nx * 4 * 24 + 2, Kx * 40 + 3, 1)]:
        A = rng.uniform(-1, 1, (m, k))
        B = rng.uniform(-1, 1, (k, n))
        C = rng.uniform(-1, 1, (m, n))
        got = run(torch, rt, plan, pack, A, B, C, ldpad=pad)
        with env("FMM_MADD_ALL4", "0"):
            want = run(torch, rt, plan, pack, A, B, C, ldpad=pad)
        assert np.array_equal(got, want), (pack, m, n, k, pad, float(np.abs(got - want).max()))
        Ai, Bi, Ci = int_mats(rng, m, n, k)
        goti = run(torch, rt, plan, pack, Ai, Bi, Ci, ldpad=pad)
        assert np.array_equal(goti, Ci + Ai @ Bi), (pack, m, n, k, pad)
        ref = Ci.copy(order="F")
        oracle.fmm(alg, Ai, Bi, ref, pack)
        assert np.array_equal(goti, ref)


@pytest.mark.parametrize("pack", PACKS)
def test_pack_strided_leading_dims(torch, fmm, rt, pack):
    """Odd leading dimensions and odd leaf depth take the element-wise paths
    of the sum, product and M_Add kernels."""
    alg = fmm.build_algorithm(["222"], [1])
    plan = rt.Plan(alg)
    rng = np.random.default_rng(17)
    A, B, C = int_mats(rng, 203, 161, 99)
    for pad in (1, 3, 8):
        got = run(torch, rt, plan, pack, A, B, C, ldpad=pad)
        assert np.array_equal(got, C + A @ B), pad


@pytest.mark.parametrize("pack", PACKS)
def test_pack_workspace(torch, fmm, rt, pack):
    alg = fmm.build_algorithm(["333"], [1])
    plan = rt.Plan(alg)
    m, n, k = 720, 600, 301
    with cs_mode(pack):
        nbytes = plan.workspace_bytes(m, n, k, base_pack(pack))
    ms, ks, ns = 240, 100, 200
    Mblocks = alg.rank + 9            # products + k-fringe products
    header = -(-4 * 264 // 256) * 256   # status header (error word where an ABC workspace has it)
    # FP64 Naive: M blocks with the ordered product launch, none C-stationary
    want = header + (0 if pack == "naive-cs" else -(-8 * ms * ns * Mblocks // 256) * 256)
    if base_pack(pack) == "naive":
        ldt = -(-ks // 8) * 8             # T's leading dimension: a multiple of 8 elements
        want += -(-8 * ms * ks * alg.rank // 256) * 256 + -(-8 * ldt * ns * alg.rank // 256) * 256
    assert nbytes == want
    rng = np.random.default_rng(18)
    A, B, C = int_mats(rng, m, n, k)
    ws = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    for _ in range(2):   # reuse: nothing depends on the workspace's old contents
        got = run(torch, rt, plan, pack, A, B, C, ws=ws)
        assert np.array_equal(got, C + A @ B)
    small = torch.empty(nbytes // 8 - 1, dtype=torch.float64, device="cuda")
    with pytest.raises(rt.FmmError):
        run(torch, rt, plan, pack, A, B, C, ws=small)


@pytest.mark.parametrize("pack", PACKS)
def test_pack_host_dropin(torch, fmm, rt, pack):
    alg = fmm.build_algorithm(["232"], [1])
    plan = rt.Plan(alg)
    rng = np.random.default_rng(19)
    A, B, C = int_mats(rng, 150, 90, 61)
    got = C.copy(order="F")
    with cs_mode(pack):
        plan.host_run(A, B, got, base_pack(pack))
    assert np.array_equal(got, C + A @ B)


def test_naive_config1_strassen_2048(torch, fmm, rt, oracle):
    """Strassen L1 FP64 at m=n=k=2048, A, B ~ U[0,1), C = 0: Naive equals ABC
    bitwise and both are within the one-level bound of the exact product."""
    alg = fmm.build_algorithm(["222"], [1])
    plan = rt.Plan(alg)
    rng = np.random.default_rng(2048)
    m = n = k = 2048
    A = np.asfortranarray(rng.random((m, k)))
    B = np.asfortranarray(rng.random((k, n)))
    C = np.zeros((m, n), order="F")
    got = run(torch, rt, plan, "naive", A, B, C)
    assert np.array_equal(got, run(torch, rt, plan, "abc", A, B, C))
    assert norm_err(got, A, B, C) <= tol_for(1)
    ref = C.copy(order="F")
    oracle.fmm(alg, A, B, ref, "naive")
    assert float(np.abs(got - ref).max()) <= 2 * tol_for(1) * 2.0 ** -53 * k
