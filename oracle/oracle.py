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


# --------------------------------------------------------------------------
# Low-precision ABC FMM: numpy restatement of the rounding points of the
# FP32 / BF16 / FP16 / split-BF16 kernels (no reference counterpart; the
# reference is FP64 only).  The structure is the reference's ABC algorithm
# (product order r = 0..R-1, then the k-fringe products, then the fringe
# GEMMs' contribution; meta/abc_gen_mix.py:448-472 and
# common/dgemm/bl_dgemm_util.c:316-353), and the rounding points are:
#   S = fl32(sum_t u_t A_t) summed in table order, then rounded to the
#   operand type (round to nearest even); likewise T;
#   P = S_op @ T_op evaluated in float64 (the kernel accumulates in FP32; the
#   difference is bounded by FP32 accumulation error only);
#   split-BF16: P = S_hi T_hi + S_hi T_lo + S_lo T_hi with S_hi = bf16(S),
#   S_lo = bf16(S - S_hi);
#   C_j = fl32(C_j + w_j * fl32(P)) for every C block of the product.

def round_bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).reshape(x.shape)


def round_op(x: np.ndarray, dtype: str) -> np.ndarray:
    x = np.asarray(x, dtype=np.float32)
    if dtype in ("f32", "bf16x3"):
        return x
    if dtype == "bf16":
        return round_bf16(x)
    if dtype == "f16":
        return x.astype(np.float16).astype(np.float32)
    raise ValueError(dtype)


def _lowp_product(S: np.ndarray, T: np.ndarray, dtype: str) -> np.ndarray:
    if dtype == "bf16x3":
        sh = round_bf16(S)
        sl = round_bf16(S - sh)
        th = round_bf16(T)
        tl = round_bf16(T - th)
        f = np.float64
        return sh.astype(f) @ th.astype(f) + sh.astype(f) @ tl.astype(f) + sl.astype(f) @ th.astype(f)
    So = round_op(S, dtype).astype(np.float64)
    To = round_op(T, dtype).astype(np.float64)
    return So @ To


def fmm_lowp(alg, A: np.ndarray, B: np.ndarray, C: np.ndarray, dtype: str) -> np.ndarray:
    """Returns C + A@B by the low-precision ABC FMM of ``alg`` (float32
    arrays; A and B already hold values of the input type)."""
    A = np.asarray(A, dtype=np.float32)
    B = np.asarray(B, dtype=np.float32)
    C = np.array(C, dtype=np.float32, copy=True)
    m, k = A.shape
    n = B.shape[1]
    Mx, Kx, Nx = alg.dims
    md, kd, nd = m - m % (Mx * 8), k - k % Kx, n - n % (Nx * 4)
    ms, ks, ns = md // Mx, kd // Kx, nd // Nx
    f32 = np.float32

    def upd(Cv, w, P):
        Cv[...] = (Cv + f32(w) * P.astype(f32)).astype(f32)

    # fringe GEMMs (one-term sums)
    if n - nd > 0:
        upd(C[:, nd:], 1.0, _lowp_product(A, B[:, nd:], dtype))
    if m - md > 0 and nd > 0:
        upd(C[md:, :nd], 1.0, _lowp_product(A[md:, :], B[:, :nd], dtype))
    if ms <= 0 or ns <= 0:
        return C
    if ks <= 0:
        upd(C[:md, :nd], 1.0, _lowp_product(A[:md, :], B[:, :nd], dtype))
        return C
    U, V, W = alg.as_arrays()

    def blk(X, rows, cols, I, J):
        return X[I * rows:(I + 1) * rows, J * cols:(J + 1) * cols]

    def opsum(X, rows, cols, table, which, r):
        s = None
        for row in range(table.shape[0]):
            u = table[row, r]
            if u == 0:
                continue
            I, J = alg.block_of_row(which, row)
            x = blk(X, rows, cols, I, J)
            s = (f32(u) * x).astype(f32) if s is None else (s + f32(u) * x).astype(f32)
        return s

    for r in range(alg.rank):
        S = opsum(A, ms, ks, U, 0, r)
        T = opsum(B, ks, ns, V, 1, r)
        P = _lowp_product(S, T, dtype)
        for row in range(W.shape[0]):
            w = W[row, r]
            if w == 0:
                continue
            I, J = alg.block_of_row(2, row)
            upd(blk(C, ms, ns, I, J), w, P)
    kr = k - kd
    if kr > 0:
        for I in range(Mx):
            for J in range(Nx):
                P = _lowp_product(A[I * ms:(I + 1) * ms, kd:], B[kd:, J * ns:(J + 1) * ns], dtype)
                upd(blk(C, ms, ns, I, J), 1.0, P)
    return C


def refine_lowp(alg_main, main: str, alg_corr, corr: str, A, B, C) -> np.ndarray:
    """numpy restatement of the FP64 refinement: A1 = fl_main(A),
    dA = fl32(A - A1), rounded on to BF16 for a BF16 correction (B likewise),
    C~ = A1 B1 by the low-precision FMM of
    alg_main (None: plain product), E = [dA | A1] [B1; dB] by alg_corr in
    `corr` arithmetic, result C + C~ + E in float64."""
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    C = np.asarray(C, np.float64)
    inp = "f32" if main in ("f32", "bf16x3") else main
    cin = "f32" if corr in ("f32", "bf16x3") else corr
    A1 = round_op(A.astype(np.float32), inp)
    B1 = round_op(B.astype(np.float32), inp)
    dA = round_op((A - A1.astype(np.float64)).astype(np.float32), cin)
    dB = round_op((B - B1.astype(np.float64)).astype(np.float32), cin)
    m, n = C.shape
    z = np.zeros((m, n), np.float32)
    Ct = (fmm_lowp(alg_main, A1, B1, z, main) if alg_main is not None
          else (z.astype(np.float64) + _lowp_product(A1, B1, main)).astype(np.float32))
    AA = np.concatenate([dA, A1], axis=1)
    BB = np.concatenate([B1, dB], axis=0)
    Et = (fmm_lowp(alg_corr, AA, BB, z, corr) if alg_corr is not None
          else (z.astype(np.float64) + _lowp_product(AA, BB, corr)).astype(np.float32))
    return C + (Ct.astype(np.float64) + Et.astype(np.float64))


def _lowp_dots(S: np.ndarray, T: np.ndarray, dtype: str) -> np.ndarray:
    """Row-wise products P[s] = S[s, :] . T[:, s] with the rounding points of
    _lowp_product (S: nsamp x depth, T: depth x nsamp, float32)."""
    f = np.float64
    if dtype == "bf16x3":
        sh = round_bf16(S)
        sl = round_bf16(S - sh)
        th = round_bf16(T)
        tl = round_bf16(T - th)
        return (np.einsum("sk,ks->s", sh.astype(f), th.astype(f)) + np.einsum("sk,ks->s", sh.astype(f), tl.astype(f))
                + np.einsum("sk,ks->s", sl.astype(f), th.astype(f)))
    return np.einsum("sk,ks->s", round_op(S, dtype).astype(f), round_op(T, dtype).astype(f))


def fmm_lowp_at(alg, A: np.ndarray, B: np.ndarray, C: np.ndarray, dtype: str, ii, jj) -> np.ndarray:
    """fmm_lowp evaluated only at the entries (ii[s], jj[s]) of the result:
    the same operand sums (table order, FP32, rounded to the operand type),
    the same product order and the same C updates fl32(C + w * fl32(P)),
    for one row of S_r and one column of T_r per sampled entry.  Lets a
    full-size result be checked entry by entry against the restatement."""
    A = np.asarray(A, dtype=np.float32)
    B = np.asarray(B, dtype=np.float32)
    ii = np.asarray(ii, dtype=np.int64)
    jj = np.asarray(jj, dtype=np.int64)
    f32 = np.float32
    out = np.asarray(C, dtype=np.float32)[ii, jj].copy()
    m, k = A.shape
    n = B.shape[1]
    Mx, Kx, Nx = alg.dims
    md, kd, nd = m - m % (Mx * 8), k - k % Kx, n - n % (Nx * 4)
    ms, ks, ns = md // Mx, kd // Kx, nd // Nx

    def upd(sel, w, P):
        out[sel] = (out[sel] + f32(w) * P[sel].astype(f32)).astype(f32)

    core = (ii < md) & (jj < nd)
    plain = ~core if (ms > 0 and ns > 0 and ks > 0) else np.ones_like(core)
    if plain.any():   # fringe GEMMs, or a core without FMM depth: one-term products over all of k
        P = _lowp_dots(A[ii], B[:, jj], dtype)
        upd(plain, 1.0, P)
    if not (ms > 0 and ns > 0 and ks > 0) or not core.any():
        return out
    U, V, W = alg.as_arrays()
    I, ip = ii // ms, ii % ms
    J, jp = jj // ns, jj % ns
    wrow = {alg.block_of_row(2, row): row for row in range(W.shape[0])}
    rows_w = np.array([wrow.get((int(a), int(b)), 0) for a, b in zip(I, J)])
    kr_ = np.arange(ks)
    for r in range(alg.rank):
        w_s = np.where(core, W[rows_w, r], 0.0)
        if not np.any(w_s):
            continue
        S = None
        for row in range(U.shape[0]):
            u = U[row, r]
            if u == 0:
                continue
            bi, bj = alg.block_of_row(0, row)
            x = A[(bi * ms + ip)[:, None], (bj * ks + kr_)[None, :]]
            S = (f32(u) * x).astype(f32) if S is None else (S + f32(u) * x).astype(f32)
        T = None
        for row in range(V.shape[0]):
            v = V[row, r]
            if v == 0:
                continue
            bi, bj = alg.block_of_row(1, row)
            x = B[(bi * ks + kr_)[:, None], (bj * ns + jp)[None, :]]
            T = (f32(v) * x).astype(f32) if T is None else (T + f32(v) * x).astype(f32)
        P = _lowp_dots(S, T, dtype)
        for w in np.unique(w_s[w_s != 0]):
            upd(w_s == w, w, P)
    kr = k - kd
    if kr > 0:
        P = _lowp_dots(A[ii, kd:], B[kd:, jj], dtype)
        upd(core, 1.0, P)
    return out
