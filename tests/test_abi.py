"""The C-ABI library: it loads, exports every symbol include/fmm_amd.h declares,
and validates arguments before touching the device (CPU only)."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fmm_amd.h")
JOB_ONLY = {"bl_dgemm_strassen_abc", "bl_dgemm_strassen_ab", "bl_dgemm_strassen_naive", "fmm_job_abc",
            "fmm_job_ab", "fmm_job_naive", "fmm_job_plan", "fmm_job_pack"}


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(\w+)\s*\([^;{]*\)\s*;", text)) - {"defined"}


def test_header_declares_expected(rt):
    names = declared()
    assert set(rt.EXPORTS) <= names
    assert JOB_ONLY <= names


def test_runtime_exports_every_symbol(rt):
    lib = rt.lib()
    for name in declared() - JOB_ONLY:
        assert hasattr(lib, name), name


def test_plan_create_and_info(fmm, rt):
    alg = fmm.build_algorithm(["222", "323"], [1, 1])
    p = rt.Plan(alg)
    out = [ct.c_int() for _ in range(4)]
    assert rt.lib().fmm_plan_info(p.handle, *[ct.byref(o) for o in out]) == 0
    assert [o.value for o in out] == [6, 4, 6, 105]
    # 14400^2 x 480: one counter per C tile of the 2400x2400 leaf (19x19) + ticket + flags
    assert p.workspace_bytes(14400, 14400, 480) == (264 + 19 * 19) * 4
    # classical-depth problem smaller than the partition: no FMM core, no counters
    assert p.workspace_bytes(40, 40, 40) == 0
    p.close()


def test_plan_rejects_bad_tables(rt):
    lib = rt.lib()
    h = ct.c_void_p()
    dims = (ct.c_int * 3)(2, 2, 2)
    one = (ct.c_double * 1)(1.0)
    assert lib.fmm_plan_create(ct.byref(h), 0, dims, 1, one, one, one) == -1
    bad = (ct.c_int * 3)(2, 0, 2)
    assert lib.fmm_plan_create(ct.byref(h), 1, bad, 1, one, one, one) == -1
    assert lib.fmm_plan_create(ct.byref(h), 1, dims, 0, one, one, one) == -1


def test_argument_validation_before_device(fmm, rt):
    lib = rt.lib()
    p = rt.Plan(fmm.build_algorithm(["222"], [1]))
    # negative dims, short leading dimensions, unsupported dtype
    assert lib.fmm_abc(p.handle, 0, -1, 4, 4, None, 4, None, 4, None, 4, None, 0, None) == -1
    assert lib.fmm_abc(p.handle, 0, 8, 4, 4, None, 7, None, 4, None, 8, None, 0, None) == -1
    assert lib.fmm_abc(p.handle, 0, 8, 4, 4, None, 8, None, 3, None, 8, None, 0, None) == -1
    assert lib.fmm_abc(p.handle, 0, 8, 4, 4, None, 8, None, 4, None, 7, None, 0, None) == -1
    assert lib.fmm_abc(p.handle, 9, 8, 4, 4, None, 8, None, 4, None, 8, None, 0, None) == -4
    assert lib.fmm_gemm(0, 8, 4, -4, None, 8, None, 4, None, 8, None) == -1
    # empty problems are no-ops
    assert lib.fmm_abc(p.handle, 0, 0, 4, 4, None, 1, None, 4, None, 1, None, 0, None) == 0
    assert lib.fmm_gemm(0, 8, 0, 4, None, 8, None, 4, None, 8, None) == 0
    p.close()


def test_bl_dgemm_ref_is_textbook(rt):
    rng = np.random.default_rng(3)
    m, n, k = 7, 5, 9
    A = np.asfortranarray(rng.integers(-3, 4, (m, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-3, 4, (k, n)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-3, 4, (m, n)).astype(np.float64))
    want = C + A @ B
    dp = ct.POINTER(ct.c_double)
    rt.lib().bl_dgemm_ref(m, n, k, A.ctypes.data_as(dp), m, B.ctypes.data_as(dp), k, C.ctypes.data_as(dp), m)
    assert np.array_equal(C, want)


def test_batch_argument_validation_before_device(fmm, rt):
    lib = rt.lib()
    rt._declare_batch(lib)
    h = ct.c_void_p()
    probs = (rt._Problem * 1)(rt._Problem(8, 8, 8, None, 8, None, 8, None, 8, 3, 0))
    plans = (ct.c_void_p * 1)()
    assert lib.fmm_batch_create(ct.byref(h), 1, 1, probs, 0, plans) == -1      # plan index out of range
    probs[0].plan, probs[0].tile = -1, 2
    assert lib.fmm_batch_create(ct.byref(h), 1, 1, probs, 0, plans) == -1      # tile shape 0 or 1
    probs[0].tile = 0
    assert lib.fmm_batch_create(ct.byref(h), 9, 1, probs, 0, plans) == -4      # unknown dtype
    assert lib.fmm_batch_create(ct.byref(h), 1, -1, probs, 0, plans) == -1
    assert lib.fmm_batch_run(None, None) == -1


def test_noalloc_rejects_oversize_on_host(fmm, rt):
    """fmm_abc_noalloc checks its fixed scratch limits on the host, before any
    device call: a leaf block of more than 65,272 C tiles returns FMM_ETOOBIG."""
    lib = rt.lib()
    p = rt.Plan(fmm.build_algorithm(["222"], [1]))
    big = 2 * 33000        # leaf 33000 x 33000: 258^2 = 66,564 tiles of 128 x 128
    assert lib.fmm_abc_noalloc(p.handle, 0, big, big, 2, None, big, None, 2, None, big, None) == -6
    assert lib.fmm_abc_noalloc(p.handle, 0, -1, 4, 4, None, 4, None, 4, None, 4, None) == -1
    assert lib.fmm_abc_noalloc(p.handle, 9, 8, 4, 4, None, 8, None, 4, None, 8, None) == -4
    assert lib.fmm_abc_noalloc(p.handle, 0, 0, 4, 4, None, 1, None, 4, None, 1, None) == 0
    p.close()
