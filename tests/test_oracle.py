"""The CPU oracle, pinned (CPU only).

* exact small-integer known-answer tests: every restated FMM variant equals
  numpy's integer matmul bit for bit (all partial sums are exact in FP64);
* the blocked plain GEMM and the peeling are bitwise equal to the reference's
  own bl_dgemm / bl_dynamic_peeling compiled from its sources (oracle/_ref);
* the multi-C epilogue C_j = C_j + alpha_j*AB equals the reference's
  bl_dgemm_asm_8x4_mulstrassen bit for bit;
* committed golden fixtures (tests/golden) written from the reference build;
* FP64 error of the restated FMM stays under the stated per-level tolerance.
"""
import ctypes as ct
import os

import numpy as np
import pytest

from conftest import int_mats
from tolerance import norm_err, tol_for

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

ALGS = [(["222"], [1]), (["222"], [2]), (["222", "323"], [1, 1]), (["333"], [1]),
        (["232"], [1]), (["323"], [1]), (["grey_strassen"], [1]), (["322"], [1])]


def aligned_f(rows, cols, align=64):
    buf = np.zeros(rows * cols + align // 8, dtype=np.float64)
    off = (-buf.ctypes.data % align) // 8
    return buf[off:off + rows * cols].reshape((rows, cols), order="F")


@pytest.mark.parametrize("spec", ALGS, ids=lambda s: "x".join(f"{a}L{l}" for a, l in zip(*s)))
@pytest.mark.parametrize("pack", ["abc", "ab", "naive"])
def test_kat_bit_exact(fmm, oracle, spec, pack):
    alg = fmm.build_algorithm(*spec)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(hash((tuple(spec[0]), pack)) & 0xFFFF)
    shapes = [(Mx * 8 * 3, Nx * 4 * 5, Kx * 7),            # no fringe
              (Mx * 8 * 3 + 5, Nx * 4 * 2 + 3, Kx * 5 + 1),  # every fringe
              (Mx * 8 - 1, 17, 9),                           # no FMM core
              (37, 29, Kx - 1 if Kx > 1 else 1)]
    for (m, n, k) in shapes:
        A, B, C = int_mats(rng, m, n, k)
        want = C + A @ B
        got = C.copy(order="F")
        oracle.fmm(alg, A, B, got, pack)
        assert np.array_equal(got, want), (pack, m, n, k)


def test_blis_gemm_matches_reference_bitwise(oracle):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference runtime not built (needs the reference sources)")
    rng = np.random.default_rng(7)
    for (m, n, k) in [(37, 29, 300), (128, 64, 600), (8, 4, 256), (100, 100, 1), (96, 40, 513)]:
        A = np.asfortranarray(rng.random((m, k)))
        B = np.asfortranarray(rng.random((k, n)))
        ldc = ((m - 1) // 8 + 1) * 8
        Cr = aligned_f(ldc, n + 4)
        Co = aligned_f(ldc, n + 4)
        Cr[:m, :n] = Co[:m, :n] = rng.random((m, n))
        R.bl_dgemm(m, n, k, oracle._p(A), m, oracle._p(B), k, oracle._p(Cr), ldc)
        oracle.lib().oracle_dgemm_blis(m, n, k, oracle._p(A), m, oracle._p(B), k, oracle._p(Co), ldc)
        assert np.array_equal(Cr[:m, :n], Co[:m, :n]), (m, n, k)


def test_peeling_matches_reference_bitwise(oracle):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference runtime not built")
    rng = np.random.default_rng(8)
    for (m, n, k, d1, d2, d3) in [(70, 45, 35, 16, 2, 8), (100, 33, 300, 48, 4, 24), (20, 9, 7, 24, 3, 12)]:
        A = np.asfortranarray(rng.random((m, k)))
        B = np.asfortranarray(rng.random((k, n)))
        ldc = ((m - 1) // 8 + 1) * 8
        Cr = aligned_f(ldc, n + 4)
        Co = aligned_f(ldc, n + 4)
        R.bl_dynamic_peeling(m, n, k, oracle._p(A), m, oracle._p(B), k, oracle._p(Cr), ldc, d1, d2, d3)
        oracle.lib().oracle_peeling(m, n, k, oracle._p(A), m, oracle._p(B), k, oracle._p(Co), ldc, d1, d2, d3)
        assert np.array_equal(Cr[:m, :n], Co[:m, :n]), (m, n, k)


def test_multi_c_epilogue_matches_reference(oracle):
    """C_j = C_j + alpha_j * (A_panel B_panel) for 8x4 panels, the update the
    oracle's product_update applies (bl_dgemm_asm_8x4_mulstrassen)."""
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference runtime not built")
    rng = np.random.default_rng(9)
    k = 37
    packA = aligned_f(8 * k, 1)[:, 0]
    packB = aligned_f(4 * k, 1)[:, 0]
    packA[:] = rng.random(8 * k)
    packB[:] = rng.random(4 * k)
    alphas = np.array([1.0, -1.0, 0.5, -0.125], dtype=np.float64)
    Cs = [aligned_f(8, 4) for _ in alphas]
    for c in Cs:
        c[:] = rng.random((8, 4))
    want = []
    AB = np.zeros((8, 4))
    for p in range(k):  # sequential accumulation from 0, mul then add
        AB = AB + np.outer(packA[p * 8:(p + 1) * 8], packB[p * 4:(p + 1) * 4])
    for c, a in zip(Cs, alphas):
        want.append(c + a * AB)
    dp = ct.POINTER(ct.c_double)
    clist = (dp * len(Cs))(*[c.ctypes.data_as(dp) for c in Cs])
    aux = (ct.c_byte * 64)()
    R.bl_dgemm_asm_8x4_mulstrassen(k, oracle._p(packA), oracle._p(packB), len(Cs), 8,
                                   clist, alphas.ctypes.data_as(dp), ct.cast(aux, ct.c_void_p))
    for c, w in zip(Cs, want):
        assert np.array_equal(c, w)


def test_golden_fixtures(oracle):
    path = os.path.join(GOLDEN, "ref_dgemm.npz")
    if not os.path.exists(path):
        pytest.skip("golden fixtures not generated")
    z = np.load(path, allow_pickle=False)
    for i in range(int(z["count"])):
        A, B, C0, Cref = (z[f"{t}{i}"] for t in ("A", "B", "C", "out"))
        A, B = np.asfortranarray(A), np.asfortranarray(B)
        got = np.asfortranarray(C0.copy())
        oracle.dgemm(A, B, got)
        assert np.array_equal(got, Cref), i


@pytest.mark.parametrize("spec", ALGS[:6], ids=lambda s: "x".join(f"{a}L{l}" for a, l in zip(*s)))
def test_float_error_within_stated_tolerance(fmm, oracle, spec):
    alg = fmm.build_algorithm(*spec)
    rng = np.random.default_rng(11)
    m = n = 6 * 8 * 4
    k = 384
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    got = C0.copy(order="F")
    oracle.fmm(alg, A, B, got, "abc")
    assert norm_err(got, A, B, C0) <= tol_for(len(alg.levels))


# --- the numpy restatements of the low-precision kernels and the refinement ---

def test_round_bf16_known_answers(oracle):
    f = np.float32
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -(1.0 + 2 ** -8),
                  3.0e38, 2 ** -130, 0.1], dtype=f)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -1.0,
                     3.0e38, 2 ** -130, 0.10009765625], dtype=f)
    got = oracle.round_bf16(x)
    # ties go to even (1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6); 2^-7 + 2^-9 rounds down;
    # 3e38 -> the nearest BF16 (0x7f62 = 3.0040e38); 2^-130 is exact; 0.1 -> 0x3dcd
    want[5] = np.frombuffer(np.uint32(0x7F620000).tobytes(), dtype=f)[0]
    assert np.array_equal(got, want), (got, want)
    # every value it returns is representable: the low 16 bits are zero
    r = oracle.round_bf16(np.random.default_rng(0).standard_normal(10000).astype(f))
    assert not np.any(r.view(np.uint32) & 0xFFFF)
    # and it is the nearest one: no BF16 neighbour is closer
    y = np.random.default_rng(1).uniform(-4, 4, 10000).astype(f)
    ry = oracle.round_bf16(y).astype(np.float64)
    ulp = 2.0 ** (np.floor(np.log2(np.abs(ry))) - 7)
    assert np.all(np.abs(y - ry) <= ulp / 2 + 1e-30)


def test_round_op_f16_matches_numpy_half(oracle):
    x = np.random.default_rng(2).uniform(-10, 10, 1000).astype(np.float32)
    assert np.array_equal(oracle.round_op(x, "f16"), x.astype(np.float16).astype(np.float32))
    assert np.array_equal(oracle.round_op(x, "f32"), x)


def test_split_bf16_product_is_compensated(oracle):
    # S_hi T_hi + S_hi T_lo + S_lo T_hi recovers the FP32 product to ~2^-16
    rng = np.random.default_rng(3)
    S = rng.uniform(-1, 1, (64, 256)).astype(np.float32)
    T = rng.uniform(-1, 1, (256, 48)).astype(np.float32)
    exact = S.astype(np.float64) @ T.astype(np.float64)
    e3 = np.abs(oracle._lowp_product(S, T, "bf16x3") - exact).max() / np.abs(exact).max()
    e1 = np.abs(oracle._lowp_product(S, T, "bf16") - exact).max() / np.abs(exact).max()
    assert e3 < 2 ** -14 < e1


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16", "bf16x3"])
@pytest.mark.parametrize("spec", [(["222"], [1]), (["222"], [2]), (["333"], [1]), (["222", "323"], [1, 1])],
                         ids=["222L1", "222L2", "333L1", "222x323"])
def test_fmm_lowp_integer_exact(fmm, oracle, spec, dtype):
    # small integers: every operand sum and partial sum is exact in every type
    alg = fmm.build_algorithm(*spec)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(4)
    m, n, k = Mx * 8 * 3 + 5, Nx * 4 * 3 + 3, Kx * 7 + 1
    A = rng.integers(-4, 5, (m, k)).astype(np.float32)
    B = rng.integers(-4, 5, (k, n)).astype(np.float32)
    C = rng.integers(-4, 5, (m, n)).astype(np.float32)
    got = oracle.fmm_lowp(alg, A, B, C, dtype)
    assert np.array_equal(got, C.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64))


def test_fmm_lowp_planted_mistakes_are_caught(fmm, oracle):
    # a flipped W sign in the table must change the integer result
    alg = fmm.build_algorithm(["222"], [1])
    rng = np.random.default_rng(5)
    A = rng.integers(-4, 5, (32, 18)).astype(np.float32)
    B = rng.integers(-4, 5, (18, 16)).astype(np.float32)
    C = np.zeros((32, 16), np.float32)
    good = oracle.fmm_lowp(alg, A, B, C, "bf16")
    alg.W[0][0] = -alg.W[0][0]
    try:
        bad = oracle.fmm_lowp(alg, A, B, C, "bf16")
    finally:
        alg.W[0][0] = -alg.W[0][0]
    assert not np.array_equal(bad, good)


def test_refine_restatement_integer_and_float(fmm, oracle):
    rng = np.random.default_rng(6)
    m, n, k = 40, 24, 18
    A = rng.integers(-4, 5, (m, k)).astype(np.float64)
    B = rng.integers(-4, 5, (k, n)).astype(np.float64)
    C = rng.integers(-4, 5, (m, n)).astype(np.float64)
    for main, corr in (("bf16", "f32"), ("bf16", "bf16"), ("f16", "f32"), ("bf16x3", "bf16x3")):
        assert np.array_equal(oracle.refine_lowp(None, main, None, corr, A, B, C), C + A @ B)
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    exact = C + A @ B
    plain = C + oracle._lowp_product(oracle.round_op(A.astype(np.float32), "bf16"),
                                     oracle.round_op(B.astype(np.float32), "bf16"), "bf16")
    ref = oracle.refine_lowp(None, "bf16", None, "f32", A, B, C)
    e_plain = np.linalg.norm(plain - exact) / np.linalg.norm(exact)
    e_ref = np.linalg.norm(ref - exact) / np.linalg.norm(exact)
    assert e_ref < e_plain / 100       # the correction restores the input rounding
    # a dropped correction term (dB) is caught
    A1 = oracle.round_op(A.astype(np.float32), "bf16").astype(np.float64)
    half = C + A1 @ oracle.round_op(B.astype(np.float32), "bf16").astype(np.float64) + (A - A1) @ B
    assert np.linalg.norm(half - ref) / np.linalg.norm(exact) > 10 * e_ref


def test_dgemm_ref_integer_kat(oracle):
    """oracle_dgemm_ref (the j, i, p triple loop of common/dgemm/bl_dgemm_ref.c:40-47)
    on small-integer inputs equals numpy's exact integer product, with
    leading dimensions larger than the extents."""
    rng = np.random.default_rng(12)
    for (m, n, k) in [(1, 1, 1), (7, 5, 9), (33, 17, 65), (64, 3, 200)]:
        A = aligned_f(m + 3, k)
        B = aligned_f(k + 1, n)
        C = aligned_f(m + 5, n)
        A[:m, :] = rng.integers(-4, 5, (m, k))
        B[:k, :] = rng.integers(-4, 5, (k, n))
        C[:m, :] = rng.integers(-4, 5, (m, n))
        C[m:, :] = 7.0
        want = C.copy(order="F")
        want[:m, :] += (A[:m, :].astype(np.int64) @ B[:k, :].astype(np.int64)).astype(np.float64)
        oracle.lib().oracle_dgemm_ref(m, n, k, oracle._p(A), m + 3, oracle._p(B), k + 1, oracle._p(C), m + 5)
        assert np.array_equal(C, want), (m, n, k)      # padding rows untouched too
    # empty problems are no-ops (the reference's early return)
    C0 = C.copy(order="F")
    oracle.lib().oracle_dgemm_ref(0, n, k, oracle._p(A), 1, oracle._p(B), k + 1, oracle._p(C), m + 5)
    assert np.array_equal(C, C0)


def test_dgemm_ref_matches_reference_bitwise(oracle, rt):
    """oracle_dgemm_ref and the product library's CPU bl_dgemm_ref both equal
    the reference's own bl_dgemm_ref (built from its sources into oracle/_ref)
    bit for bit on random float inputs: same loop order, one rounding per
    multiply and per add."""
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference runtime not built (needs the reference sources)")
    rng = np.random.default_rng(13)
    for (m, n, k) in [(37, 29, 300), (8, 4, 256), (100, 100, 1), (5, 61, 513)]:
        A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
        B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
        C0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        outs = []
        for f in (R.bl_dgemm_ref, oracle.lib().oracle_dgemm_ref, rt.lib().bl_dgemm_ref):
            C = C0.copy(order="F")
            f(m, n, k, oracle._p(A), m, oracle._p(B), k, oracle._p(C), m)
            outs.append(C)
        assert np.array_equal(outs[0], outs[1]), ("oracle", m, n, k)
        assert np.array_equal(outs[0], outs[2]), ("product bl_dgemm_ref", m, n, k)
        assert not np.array_equal(outs[0], C0)


def test_dgemm_ref_catches_planted_order_change(oracle):
    """A dot product summed in another order (numpy's pairwise/BLAS sum) is not
    bitwise equal on random floats, so the bitwise check above has teeth."""
    rng = np.random.default_rng(14)
    m, n, k = 37, 29, 300
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    got = C.copy(order="F")
    oracle.lib().oracle_dgemm_ref(m, n, k, oracle._p(A), m, oracle._p(B), k, oracle._p(got), m)
    assert not np.array_equal(got, C + A @ B)
    assert np.abs(got - (C + A @ B)).max() < 1e-12


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16", "bf16x3"])
@pytest.mark.parametrize("spec", [(["222"], [1]), (["222"], [2]), (["323"], [1])],
                         ids=["222L1", "222L2", "323L1"])
def test_lowp_sampled_restatement_matches_full(fmm, oracle, spec, dtype):
    """fmm_lowp_at equals fmm_lowp at every sampled entry (core, every fringe
    region, k-fringe), up to one FP32 ulp where the FP64 dot products of the
    two forms round differently; a dropped or misplaced term would differ by
    O(1)."""
    alg = fmm.build_algorithm(*spec)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(15)
    m, n, k = Mx * 8 * 3 + 5, Nx * 4 * 4 + 3, Kx * 11 + 1
    A = oracle.round_op(rng.uniform(-1, 1, (m, k)).astype(np.float32), dtype)
    B = oracle.round_op(rng.uniform(-1, 1, (k, n)).astype(np.float32), dtype)
    C = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    full = oracle.fmm_lowp(alg, A, B, C, dtype)
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    got = oracle.fmm_lowp_at(alg, A, B, C, dtype, ii.ravel(), jj.ravel()).reshape(m, n)
    ulp = np.spacing(np.abs(full).astype(np.float32))
    assert np.all(np.abs(got - full) <= ulp), np.abs(got - full).max()
    # planted mistake: one U coefficient of product 0 with its sign flipped
    # changes the sampled result in the C blocks that product feeds
    import copy
    bad = copy.deepcopy(alg)
    row = next(i for i, r in enumerate(bad.U) if r[0] != 0)
    bad.U[row][0] = -bad.U[row][0]
    got_bad = oracle.fmm_lowp_at(bad, A, B, C, dtype, ii.ravel(), jj.ravel()).reshape(m, n)
    assert np.abs(got_bad - full).max() > 1e-2
