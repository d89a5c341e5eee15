"""ORACLE -- test infrastructure only.

ctypes access to the C restatement in ``fmm_oracle.c`` (built into
``oracle/_build/libfmm_oracle.so``) and, where it has been built, to the
reference's own CPU runtime compiled from its sources (``oracle/_ref``).
Only tests, the smoke check and the benchmark's CPU-baseline leg use this
module; nothing in the product path imports it.

All matrices are column-major float64 numpy arrays (Fortran order) so the
leading dimension is ``A.shape[0]`` unless a larger ``ld`` is passed.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfmm_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libblislab_ref.so")

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)

_lib = None
_ref = None


def build(ref: bool = False) -> None:
    """Compile the oracle (and the reference runtime when asked and present)."""
    targets = ["all"]
    if ref and os.path.isdir("/root/reference/common"):
        targets.append("ref")
    subprocess.check_call(["make", "-s", "-C", HERE] + targets)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        g = [ct.c_int] * 3 + [_dp, ct.c_int, _dp, ct.c_int, _dp, ct.c_int]
        for name in ("oracle_dgemm_ref", "oracle_dgemm_blis"):
            f = getattr(L, name)
            f.argtypes = g
            f.restype = None
        L.oracle_peeling.argtypes = g + [ct.c_int] * 3
        L.oracle_peeling.restype = None
        fm = [ct.c_int, _ip, ct.c_int, _dp, _dp, _dp] + g
        for name in ("oracle_fmm_abc", "oracle_fmm_ab", "oracle_fmm_naive"):
            f = getattr(L, name)
            f.argtypes = fm
            f.restype = ct.c_int
        L.oracle_set_threads.argtypes = [ct.c_int]
        L.oracle_get_threads.restype = ct.c_int
        _lib = L
    return _lib


def ref_lib():
    """The reference runtime (bl_dgemm, bl_dgemm_ref, bl_dynamic_peeling,
    bl_dgemm_asm_8x4_mulstrassen) or None when it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        L = ct.CDLL(REF_PATH)
        g = [ct.c_int] * 3 + [_dp, ct.c_int, _dp, ct.c_int, _dp, ct.c_int]
        for name in ("bl_dgemm", "bl_dgemm_ref"):
            getattr(L, name).argtypes = g
            getattr(L, name).restype = None
        L.bl_dynamic_peeling.argtypes = g + [ct.c_int] * 3
        L.bl_dynamic_peeling.restype = None
        L.bl_dgemm_asm_8x4_mulstrassen.argtypes = [
            ct.c_int, _dp, _dp, ct.c_ulonglong, ct.c_ulonglong,
            ct.POINTER(_dp), _dp, ct.c_void_p]
        L.bl_dgemm_asm_8x4_mulstrassen.restype = None
        _ref = L
    return _ref


def _p(a: np.ndarray):
    assert a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def _ld(a: np.ndarray) -> int:
    assert a.flags.f_contiguous or a.ndim == 1
    return a.shape[0]


def dgemm(A, B, C, kind: str = "blis") -> None:
    """C += A @ B in place (column-major arrays)."""
    m, k = A.shape
    n = B.shape[1]
    f = lib().oracle_dgemm_blis if kind == "blis" else lib().oracle_dgemm_ref
    f(m, n, k, _p(A), _ld(A), _p(B), _ld(B), _p(C), _ld(C))


def peeling(A, B, C, dim1, dim2, dim3) -> None:
    m, k = A.shape
    n = B.shape[1]
    lib().oracle_peeling(m, n, k, _p(A), _ld(A), _p(B), _ld(B), _p(C), _ld(C), dim1, dim2, dim3)


def fmm(alg, A, B, C, pack_type: str = "abc") -> None:
    """C += A @ B by the restated FMM of ``alg`` (an fmm_gen_amd Algorithm)."""
    U, V, W = (np.ascontiguousarray(t) for t in alg.as_arrays())
    dims = np.array([d for lv in alg.levels for d in lv], dtype=np.int32)
    m, k = A.shape
    n = B.shape[1]
    f = {"abc": lib().oracle_fmm_abc, "ab": lib().oracle_fmm_ab,
         "naive": lib().oracle_fmm_naive}[pack_type]
    rc = f(len(alg.levels), dims.ctypes.data_as(_ip), alg.rank, _p(U), _p(V), _p(W),
           m, n, k, _p(A), _ld(A), _p(B), _ld(B), _p(C), _ld(C))
    if rc != 0:
        raise RuntimeError(f"oracle_fmm_{pack_type} failed with {rc}")


def set_threads(nt: int) -> None:
    lib().oracle_set_threads(int(nt))


def get_threads() -> int:
    return int(lib().oracle_get_threads())
