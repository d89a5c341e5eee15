"""ctypes binding of the HIP C-ABI (include/fmm_amd.h) for torch device tensors.

Matrices are column-major, as in the reference: a torch tensor ``X`` of shape
(rows, cols) with ``X.stride(0) == 1`` and leading dimension ``X.stride(1)``.
``colmajor()`` allocates one.  PyTorch only supplies device memory and the
current stream here; all arithmetic runs in the library's own kernels.

There is no CPU fallback: if the library has not been built, or no GPU is
visible, the calls raise.
"""
from __future__ import annotations

import ctypes as ct
import os
from typing import Optional

import numpy as np

from .coeffs import Algorithm

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libfmm_amd.so")

FMM_F64, FMM_F32, FMM_BF16, FMM_F16, FMM_BF16X3 = 0, 1, 2, 3, 4
DTYPE_NAMES = {FMM_F64: "f64", FMM_F32: "f32", FMM_BF16: "bf16", FMM_F16: "f16", FMM_BF16X3: "bf16x3"}
PACKS = {"abc": 0, "ab": 1, "naive": 2}
STATUS = {0: "FMM_OK", -1: "FMM_EINVAL", -2: "FMM_EHIP", -3: "FMM_ENOMEM",
          -4: "FMM_EUNSUPPORTED", -5: "FMM_ESPIN", -6: "FMM_ETOOBIG"}

EXPORTS = ("fmm_plan_create", "fmm_plan_destroy", "fmm_plan_info", "fmm_workspace_bytes",
           "fmm_workspace_bytes_pack", "fmm_run", "fmm_abc", "fmm_ab", "fmm_naive", "fmm_gemm",
           "fmm_ws_status", "fmm_host_abc", "fmm_host_run", "bl_dgemm", "bl_dgemm_ref",
           "bl_finalize", "fmm_batch_create", "fmm_batch_set", "fmm_batch_run", "fmm_batch_variant",
           "fmm_batch_info", "fmm_batch_status", "fmm_batch_destroy", "fmm_batch_reserve",
           "fmm_batch_generation", "fmm_batch_last_status", "fmm_refine", "fmm_refine_workspace_bytes",
           "fmm_timing", "fmm_timing_last_ms", "fmm_abc_noalloc")


class FmmError(RuntimeError):
    pass


_lib = None


def _declare(L, has_job: bool = False):
    vp, i, dp = ct.c_void_p, ct.c_int, ct.POINTER(ct.c_double)
    L.fmm_plan_create.argtypes = [ct.POINTER(vp), i, ct.POINTER(i), i, dp, dp, dp]
    L.fmm_plan_create.restype = i
    L.fmm_plan_destroy.argtypes = [vp]
    L.fmm_plan_destroy.restype = i
    L.fmm_plan_info.argtypes = [vp] + [ct.POINTER(i)] * 4
    L.fmm_plan_info.restype = i
    L.fmm_workspace_bytes.argtypes = [vp, i, i, i]
    L.fmm_workspace_bytes.restype = ct.c_size_t
    L.fmm_workspace_bytes_pack.argtypes = [vp, i, i, i, i]
    L.fmm_workspace_bytes_pack.restype = ct.c_size_t
    for name in ("fmm_abc", "fmm_ab", "fmm_naive"):
        getattr(L, name).argtypes = [vp, i, i, i, i, vp, i, vp, i, vp, i, vp, ct.c_size_t, vp]
        getattr(L, name).restype = i
    L.fmm_abc_noalloc.argtypes = [vp, i, i, i, i, vp, i, vp, i, vp, i, vp]
    L.fmm_abc_noalloc.restype = i
    L.fmm_run.argtypes = [vp, i, i, i, i, i, vp, i, vp, i, vp, i, vp, ct.c_size_t, vp]
    L.fmm_run.restype = i
    L.fmm_host_run.argtypes = [vp, i, i, i, i, dp, i, dp, i, dp, i]
    L.fmm_host_run.restype = i
    L.fmm_gemm.argtypes = [i, i, i, i, vp, i, vp, i, vp, i, vp]
    L.fmm_gemm.restype = i
    L.fmm_ws_status.argtypes = [vp, ct.POINTER(i)]
    L.fmm_ws_status.restype = i
    L.fmm_host_abc.argtypes = [vp, i, i, i, dp, i, dp, i, dp, i]
    L.fmm_host_abc.restype = i
    for name in ("bl_dgemm", "bl_dgemm_ref"):
        getattr(L, name).argtypes = [i, i, i, dp, i, dp, i, dp, i]
        getattr(L, name).restype = None
    L.bl_finalize.argtypes = []
    L.bl_finalize.restype = None
    if has_job:
        L.fmm_job_plan.argtypes = []
        L.fmm_job_plan.restype = vp
        L.fmm_job_pack.argtypes = []
        L.fmm_job_pack.restype = i
        pack = {v: k for k, v in PACKS.items()}[L.fmm_job_pack()]
        getattr(L, f"bl_dgemm_strassen_{pack}").argtypes = [i, i, i, dp, i, dp, i, dp, i]
        getattr(L, f"bl_dgemm_strassen_{pack}").restype = None
        getattr(L, f"fmm_job_{pack}").argtypes = [i, i, i, vp, i, vp, i, vp, i, vp, ct.c_size_t, vp]
        getattr(L, f"fmm_job_{pack}").restype = i
        L.job_pack = pack
    return L


def load_library(path: str = LIB_PATH, has_job: bool = False):
    if not os.path.exists(path):
        raise FmmError(f"HIP library not built: {path} (run __graft_entry__.build())")
    return _declare(ct.CDLL(path), has_job)


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise FmmError(f"{what} failed: {STATUS.get(rc, rc)}")


def _stream_ptr(stream) -> Optional[int]:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def colmajor(rows: int, cols: int, ld: Optional[int] = None, dtype=None, device="cuda"):
    """Uninitialised column-major (rows x cols) tensor with leading dim ld."""
    import torch
    ld = max(rows, 1) if ld is None else ld
    if ld < max(rows, 1):
        raise ValueError("ld < rows")
    store = torch.empty((max(cols, 1), ld), dtype=dtype or torch.float64, device=device)
    return store.t()[:rows, :cols]


def as_colmajor(x):
    """Column-major copy of a 2-D tensor (same device and dtype)."""
    out = colmajor(x.shape[0], x.shape[1], dtype=x.dtype, device=x.device)
    out.copy_(x)
    return out


def _ld(x) -> int:
    if x.dim() != 2:
        raise ValueError("expected a 2-D tensor")
    if x.shape[0] > 1 and x.stride(0) != 1:
        raise ValueError("tensor is not column-major (stride(0) must be 1)")
    return max(int(x.stride(1)), x.shape[0], 1) if x.shape[1] > 1 else max(x.shape[0], 1)


def dtype_code(A, B, C, split: bool = False) -> int:
    """The C-ABI dtype of operands A, B, C: FP64 (C FP64), or FP32 / BF16 /
    FP16 inputs with an FP32 C; split=True asks for split-BF16 products on
    FP32 inputs."""
    import torch
    if A.dtype != B.dtype:
        raise FmmError(f"A and B must have one dtype (got {A.dtype}, {B.dtype})")
    if A.dtype == torch.float64:
        if C.dtype != torch.float64:
            raise FmmError("C must be float64 for float64 A, B")
        if split:
            raise FmmError("split products take float32 inputs")
        return FMM_F64
    if C.dtype != torch.float32:
        raise FmmError(f"C must be float32 for {A.dtype} A, B")
    code = {torch.float32: FMM_F32, torch.bfloat16: FMM_BF16, torch.float16: FMM_F16}.get(A.dtype)
    if code is None:
        raise FmmError(f"unsupported dtype {A.dtype}")
    if split:
        if code != FMM_F32:
            raise FmmError("split products take float32 inputs")
        return FMM_BF16X3
    return code


def _check_operands(A, B, C):
    for name, X in (("A", A), ("B", B), ("C", C)):
        if not X.is_cuda:
            raise FmmError(f"{name} must be a device tensor")
    m, k = A.shape
    k2, n = B.shape
    if k2 != k or tuple(C.shape) != (m, n):
        raise ValueError(f"shape mismatch: A{tuple(A.shape)} B{tuple(B.shape)} C{tuple(C.shape)}")
    return m, n, k


class Plan:
    """A run-time plan for one (possibly composed) algorithm."""

    def __init__(self, alg: Algorithm, library=None):
        self._lib = library or lib()
        U, V, W = (np.ascontiguousarray(t) for t in alg.as_arrays())
        dims = np.array([d for lv in alg.levels for d in lv], dtype=np.int32)
        h = ct.c_void_p()
        dp = ct.POINTER(ct.c_double)
        check(self._lib.fmm_plan_create(ct.byref(h), len(alg.levels),
                                        dims.ctypes.data_as(ct.POINTER(ct.c_int)), alg.rank,
                                        U.ctypes.data_as(dp), V.ctypes.data_as(dp),
                                        W.ctypes.data_as(dp)), "fmm_plan_create")
        self.handle = h
        self.alg = alg

    def close(self):
        if getattr(self, "handle", None):
            self._lib.fmm_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def workspace_bytes(self, m: int, n: int, k: int, pack: str = "abc") -> int:
        if pack == "abc":
            return int(self._lib.fmm_workspace_bytes(self.handle, m, n, k))
        return int(self._lib.fmm_workspace_bytes_pack(self.handle, PACKS[pack], m, n, k))

    def run(self, A, B, C, pack: str = "abc", ws=None, stream=None, split: bool = False) -> None:
        """C += A @ B on the GPU by the FMM of pack type `pack` ('abc', 'ab'
        or 'naive').  Device tensors: FP64 A, B, C; or FP32 / BF16 / FP16 A, B
        with FP32 C (ABC only); split=True: split-BF16 products of FP32 A, B."""
        m, n, k = _check_operands(A, B, C)
        dt = dtype_code(A, B, C, split)
        wptr, wbytes = (None, 0) if ws is None else (ws.data_ptr(), ws.numel() * ws.element_size())
        check(self._lib.fmm_run(self.handle, PACKS[pack], dt, m, n, k, A.data_ptr(), _ld(A),
                                B.data_ptr(), _ld(B), C.data_ptr(), _ld(C), wptr, wbytes,
                                _stream_ptr(stream)), f"fmm_run({pack}, {DTYPE_NAMES[dt]})")

    def abc(self, A, B, C, ws=None, stream=None, split: bool = False) -> None:
        """C += A @ B on the GPU by the ABC FMM (device tensors; dtypes as run())."""
        m, n, k = _check_operands(A, B, C)
        dt = dtype_code(A, B, C, split)
        wptr, wbytes = (None, 0) if ws is None else (ws.data_ptr(), ws.numel() * ws.element_size())
        check(self._lib.fmm_abc(self.handle, dt, m, n, k, A.data_ptr(), _ld(A), B.data_ptr(),
                                _ld(B), C.data_ptr(), _ld(C), wptr, wbytes, _stream_ptr(stream)),
              f"fmm_abc({DTYPE_NAMES[dt]})")

    def abc_noalloc(self, A, B, C, stream=None, split: bool = False) -> None:
        """C += A @ B by the ABC FMM with no device allocation (the library's
        static scratch array holds the workspace; calls overlapping on one
        GPU race)."""
        m, n, k = _check_operands(A, B, C)
        dt = dtype_code(A, B, C, split)
        check(self._lib.fmm_abc_noalloc(self.handle, dt, m, n, k, A.data_ptr(), _ld(A), B.data_ptr(),
                                        _ld(B), C.data_ptr(), _ld(C), _stream_ptr(stream)),
              f"fmm_abc_noalloc({DTYPE_NAMES[dt]})")

    def host_run(self, A: np.ndarray, B: np.ndarray, C: np.ndarray, pack: str = "abc") -> None:
        """Host-pointer drop-in on Fortran-ordered float64 numpy arrays."""
        dp = ct.POINTER(ct.c_double)
        m, k = A.shape
        n = B.shape[1]
        for X in (A, B, C):
            if X.dtype != np.float64 or not X.flags.f_contiguous:
                raise ValueError("host entries need Fortran-ordered float64 arrays")
        check(self._lib.fmm_host_run(self.handle, PACKS[pack], m, n, k, A.ctypes.data_as(dp),
                                     max(m, 1), B.ctypes.data_as(dp), max(k, 1),
                                     C.ctypes.data_as(dp), max(m, 1)), f"fmm_host_run({pack})")

    def host_abc(self, A: np.ndarray, B: np.ndarray, C: np.ndarray) -> None:
        self.host_run(A, B, C, "abc")


def gemm(A, B, C, stream=None, library=None, split: bool = False) -> None:
    """Plain C += A @ B on the same kernel family (identity plan); dtypes as
    Plan.run()."""
    L = library or lib()
    m, n, k = _check_operands(A, B, C)
    dt = dtype_code(A, B, C, split)
    check(L.fmm_gemm(dt, m, n, k, A.data_ptr(), _ld(A), B.data_ptr(), _ld(B), C.data_ptr(),
                     _ld(C), _stream_ptr(stream)), f"fmm_gemm({DTYPE_NAMES[dt]})")


def ws_status(ws=None, library=None) -> int:
    """Error flags of the last call that used workspace ``ws``; with
    ws=None, of every call since the last query that allocated its own
    workspace (cleared by the query).  Synchronises the device."""
    L = library or lib()
    flags = ct.c_int(0)
    rc = L.fmm_ws_status(None if ws is None else ws.data_ptr(), ct.byref(flags))
    if rc not in (0, -5):
        check(rc, "fmm_ws_status")
    return int(flags.value)


class _Problem(ct.Structure):
    _fields_ = [("m", ct.c_int), ("n", ct.c_int), ("k", ct.c_int),
                ("A", ct.c_void_p), ("lda", ct.c_int),
                ("B", ct.c_void_p), ("ldb", ct.c_int),
                ("C", ct.c_void_p), ("ldc", ct.c_int),
                ("plan", ct.c_int), ("tile", ct.c_int)]


def _declare_batch(L):
    if getattr(L, "_batch_declared", False):
        return
    vp, i = ct.c_void_p, ct.c_int
    L.fmm_batch_create.argtypes = [ct.POINTER(vp), i, i, ct.POINTER(_Problem), i, ct.POINTER(vp)]
    L.fmm_batch_create.restype = i
    L.fmm_batch_set.argtypes = [vp, i, ct.POINTER(_Problem), i, ct.POINTER(vp), vp]
    L.fmm_batch_set.restype = i
    L.fmm_batch_variant.argtypes = [vp]
    L.fmm_batch_variant.restype = i
    L.fmm_batch_reserve.argtypes = [vp, i, i]
    L.fmm_batch_reserve.restype = i
    L.fmm_batch_generation.argtypes = [vp]
    L.fmm_batch_generation.restype = i
    L.fmm_batch_run.argtypes = [vp, vp]
    L.fmm_batch_run.restype = i
    L.fmm_batch_info.argtypes = [vp, ct.POINTER(i), ct.POINTER(i)]
    L.fmm_batch_info.restype = i
    L.fmm_batch_status.argtypes = [vp, ct.POINTER(i)]
    L.fmm_batch_status.restype = i
    L.fmm_batch_last_status.argtypes = [vp, ct.POINTER(i)]
    L.fmm_batch_last_status.restype = i
    L.fmm_batch_destroy.argtypes = [vp]
    L.fmm_batch_destroy.restype = i
    L._batch_declared = True


class Batch:
    """Many independent C += A @ B problems run by ONE persistent launch.

    ``problems``: sequence of (A, B, C, plan, tile) with device tensors of one
    dtype family (as Plan.run), ``plan`` a Plan or None (plain GEMM) and
    ``tile`` 0 (256x128 C tiles) or 1 (128x256).  The job table lives on the
    device; ``run()`` is one persistent kernel on the stream (graph
    capturable); ``set()`` replaces the problems in place with the upload on
    a stream, so a graph captured from ``run()`` replays the new contents
    while ``variant()`` is unchanged.  Each problem's result equals a single
    call's bit for bit.
    """

    def __init__(self, problems, split: bool = False, library=None, dtype: Optional[int] = None):
        self._lib = library or lib()
        _declare_batch(self._lib)
        arr, plans, ph, dt = self._encode(problems, split)
        self.dtype = dt if dt is not None else (FMM_F64 if dtype is None else dtype)
        h = ct.c_void_p()
        check(self._lib.fmm_batch_create(ct.byref(h), self.dtype, len(problems), arr, len(plans), ph),
              "fmm_batch_create")
        self.handle = h
        self._info()

    def _encode(self, problems, split):
        keep, plans, index = [], [], {}
        arr = (_Problem * max(len(problems), 1))()
        dt = None
        for i, (A, B, C, plan, tile) in enumerate(problems):
            m, n, k = _check_operands(A, B, C)
            d = dtype_code(A, B, C, split)
            if dt is None:
                dt = d
            elif d != dt:
                raise FmmError("all problems of a batch must share one dtype")
            pi = -1
            if plan is not None:
                if id(plan) not in index:
                    index[id(plan)] = len(plans)
                    plans.append(plan)
                pi = index[id(plan)]
            arr[i] = _Problem(m, n, k, A.data_ptr(), _ld(A), B.data_ptr(), _ld(B), C.data_ptr(), _ld(C),
                              pi, int(tile))
            keep.append((A, B, C))
        ph = (ct.c_void_p * max(len(plans), 1))(*[p.handle for p in plans])
        self._keep = keep
        self._plans = plans
        return arr, plans, ph, dt

    def _info(self):
        jobs, units = ct.c_int(), ct.c_int()
        self._lib.fmm_batch_info(self.handle, ct.byref(jobs), ct.byref(units))
        self.jobs, self.units = jobs.value, units.value

    def set(self, problems, split: bool = False, stream=None) -> None:
        """Replace the problems (same dtype); the previous run must be done."""
        arr, plans, ph, dt = self._encode(problems, split)
        if dt is not None and dt != self.dtype:
            raise FmmError("a batch keeps its dtype")
        check(self._lib.fmm_batch_set(self.handle, len(problems), arr, len(plans), ph, _stream_ptr(stream)),
              "fmm_batch_set")
        self._info()

    def variant(self) -> int:
        return int(self._lib.fmm_batch_variant(self.handle))

    def generation(self) -> int:
        """Changes whenever the device buffers are reallocated (captured
        graphs of run() are then stale)."""
        return int(self._lib.fmm_batch_generation(self.handle))

    def reserve(self, jobs: int, ws_ints: int) -> None:
        check(self._lib.fmm_batch_reserve(self.handle, int(jobs), int(ws_ints)), "fmm_batch_reserve")

    def run(self, stream=None) -> None:
        check(self._lib.fmm_batch_run(self.handle, _stream_ptr(stream)), "fmm_batch_run")

    def status(self) -> int:
        """Error flags of the last run (synchronises the device)."""
        flags = ct.c_int(0)
        self._lib.fmm_batch_status(self.handle, ct.byref(flags))
        return int(flags.value)

    def last_status(self) -> int:
        """Error flags of the last completed run, without synchronising
        (valid once the caller has seen that run finish, e.g. by an event)."""
        flags = ct.c_int(0)
        self._lib.fmm_batch_last_status(self.handle, ct.byref(flags))
        return int(flags.value)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.fmm_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DTYPE_CODES = {v: k for k, v in DTYPE_NAMES.items()}


def refine(A, B, C, main: str = "bf16", main_plan: Optional[Plan] = None, corr: str = "f32",
           corr_plan: Optional[Plan] = None, stream=None, library=None) -> None:
    """C += A @ B for FP64 device tensors: the main product A1 @ B1 in
    `main` precision (FMM of main_plan, or plain GEMM), one correction
    [dA | A1] @ [B1; dB] in `corr` precision ('f32' or 'bf16x3'), both added
    to C in FP64 (see fmm_refine in include/fmm_amd.h)."""
    import torch
    L = library or lib()
    if not getattr(L, "_refine_declared", False):
        vp, i = ct.c_void_p, ct.c_int
        L.fmm_refine.argtypes = [vp, i, vp, i, i, i, i, vp, i, vp, i, vp, i, vp, ct.c_size_t, vp]
        L.fmm_refine.restype = i
        L.fmm_refine_workspace_bytes.argtypes = [vp, i, vp, i, i, i, i]
        L.fmm_refine_workspace_bytes.restype = ct.c_size_t
        L._refine_declared = True
    m, n, k = _check_operands(A, B, C)
    for X in (A, B, C):
        if X.dtype != torch.float64:
            raise FmmError("refine takes float64 A, B, C")
    mp = main_plan.handle if main_plan is not None else None
    cp = corr_plan.handle if corr_plan is not None else None
    check(L.fmm_refine(mp, DTYPE_CODES[main], cp, DTYPE_CODES[corr], m, n, k, A.data_ptr(), _ld(A),
                       B.data_ptr(), _ld(B), C.data_ptr(), _ld(C), None, 0, _stream_ptr(stream)),
          f"fmm_refine({main}, {corr})")


def kernel_timing(enable: bool, library=None) -> None:
    """Bracket every product-kernel launch of this thread with HIP events."""
    L = library or lib()
    L.fmm_timing.argtypes = [ct.c_int]
    L.fmm_timing.restype = ct.c_int
    check(L.fmm_timing(1 if enable else 0), "fmm_timing")


def last_kernel_ms(library=None) -> float:
    """Duration of the last timed product-kernel launch (waits for it)."""
    L = library or lib()
    L.fmm_timing_last_ms.argtypes = [ct.POINTER(ct.c_float)]
    L.fmm_timing_last_ms.restype = ct.c_int
    v = ct.c_float(0.0)
    check(L.fmm_timing_last_ms(ct.byref(v)), "fmm_timing_last_ms")
    return float(v.value)
