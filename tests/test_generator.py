"""The control.py generator: job naming, emitted HIP source, per-job library
exports (CPU), and the generated library + driver on the GPU."""
import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, int_mats

JOBS = os.path.join(ROOT, "fmm-gen_amd", "lib", "jobs")


def test_cli_usage_errors():
    from fmm_gen_amd import control
    with pytest.raises(SystemExit):
        control.parse_args(["control.py"])
    with pytest.raises(SystemExit):
        control.parse_args(["control.py", "2", "222", "1"])
    assert control.parse_args(["control.py", "1", "222", "1"]) == (["222"], [1], "abc", "../")
    assert control.parse_args(["control.py", "2", "222", "1", "323", "1", "abc", "/x"]) == \
        (["222", "323"], [1, 1], "abc", "/x")


def test_emitted_source(tmp_path, fmm):
    from fmm_gen_amd import codegen
    alg = fmm.build_algorithm(["222"], [1])
    src = codegen.emit_source(alg, "222-1_abc", "abc")
    assert "Largest number of A or B terms of one product: 2" in src
    assert "void bl_dgemm_strassen_abc(" in src
    # Strassen's first product in the reference table: (A00 + A11)(B00 + B11)
    assert "M0 = (A(0,0) + A(1,1)) * (B(0,0) + B(1,1)); C(0,0) += M0; C(1,1) += M0" in src
    eqs = codegen.equations(alg)
    assert len(eqs) == 7


@pytest.mark.parametrize("job", ["222-1_abc", "222-1_323-1_abc", "333-1_abc", "222-1_ab", "333-1_naive"])
def test_built_job_exports(job):
    lib = os.path.join(JOBS, job, "lib", f"lib{job}.so")
    if not os.path.exists(lib):
        pytest.skip("jobs not built (run __graft_entry__.build())")
    pack = job.rsplit("_", 1)[1]
    L = ct.CDLL(lib)
    for sym in (f"bl_dgemm_strassen_{pack}", f"fmm_job_{pack}", "fmm_job_plan", "fmm_job_pack", "fmm_run",
                "fmm_abc", "bl_dgemm"):
        assert hasattr(L, sym)
    for other in {"abc", "ab", "naive"} - {pack}:
        assert not hasattr(L, f"bl_dgemm_strassen_{other}")
    L.fmm_job_pack.restype = ct.c_int
    assert L.fmm_job_pack() == {"abc": 0, "ab": 1, "naive": 2}[pack]
    assert os.access(os.path.join(JOBS, job, "test", f"test_{job}_st.x"), os.X_OK)
    assert os.access(os.path.join(JOBS, job, "test", f"test_{job}_mt.x"), os.X_OK)


@pytest.mark.gpu
@pytest.mark.parametrize("job,spec", [("222-1_abc", (["222"], [1])),
                                      ("222-1_323-1_abc", (["222", "323"], [1, 1])),
                                      ("333-1_abc", (["333"], [1])),
                                      ("222-1_ab", (["222"], [1])),
                                      ("333-1_naive", (["333"], [1]))])
def test_job_library_host_dropin(job, spec, fmm, rt):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = rt.load_library(os.path.join(JOBS, job, "lib", f"lib{job}.so"), has_job=True)
    alg = fmm.build_algorithm(*spec)
    Mx, Kx, Nx = alg.dims
    rng = np.random.default_rng(4)
    m, n, k = Mx * 8 * 9 + 5, Nx * 4 * 11 + 3, Kx * 13 + 2
    A, B, C = int_mats(rng, m, n, k)
    want = C + A @ B
    # reference contract: ldc = roundup(m, 8), C with n+4 columns
    ldc = ((m - 1) // 8 + 1) * 8
    Cp = np.zeros((ldc, n + 4), order="F")
    Cp[:m, :n] = C
    dp = ct.POINTER(ct.c_double)
    entry = getattr(L, f"bl_dgemm_strassen_{L.job_pack}")
    entry(m, n, k, A.ctypes.data_as(dp), m, B.ctypes.data_as(dp), k, Cp.ctypes.data_as(dp), ldc)
    assert np.array_equal(Cp[:m, :n], want)
    assert not Cp[m:, :].any() and not Cp[:, n:].any()   # nothing outside m x n written


@pytest.mark.gpu
@pytest.mark.parametrize("job", ["222-1_abc", "333-1_naive"])
def test_job_driver_prints_reference_line(job):
    exe = os.path.join(JOBS, job, "test", f"test_{job}_mt.x")
    res = subprocess.run([exe, "1000", "1000", "1000"], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, FMM_SEED="1"))
    assert res.returncode == 0, res.stderr
    lines = [l for l in res.stdout.splitlines() if l.strip()]
    assert not any("!=" in l for l in lines), lines[:5]
    cols = re.split(r"\s+", lines[-1].strip())
    assert cols[:3] == ["1000", "1000", "1000"] and len(cols) == 5
    assert float(cols[3]) > 0 and float(cols[4]) > 0
    res = subprocess.run([exe, "10"], capture_output=True, text=True, timeout=60)
    assert "require 3 arguments" in res.stdout


@pytest.mark.gpu
def test_custom_table_job_equals_runtime_plan(fmm, rt):
    """The job library abc_gen.py built from table files (Strassen then
    smirnov333-23-139) gives the runtime plan's result for the same composition
    (registry names 222, 333) bit for bit, on float inputs, through the
    reference's host entry and the device entry."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from __graft_entry__ import CUSTOM_DIR
    L = rt.load_library(os.path.join(CUSTOM_DIR, "libfmm_custom.so"), has_job=True)
    assert L.job_pack == "abc"
    plan = rt.Plan(fmm.build_algorithm(["222", "333"], [1, 1]))
    rng = np.random.default_rng(6)
    m, n, k = 6 * 8 * 9 + 5, 6 * 4 * 13 + 3, 6 * 40 + 2
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    host = C.copy(order="F")
    dp = ct.POINTER(ct.c_double)
    L.bl_dgemm_strassen_abc(m, n, k, A.ctypes.data_as(dp), m, B.ctypes.data_as(dp), k, host.ctypes.data_as(dp), m)
    dA, dB, dC, dR = (torch.from_numpy(X.T).to("cuda").t() for X in (A, B, C, C))
    assert L.fmm_job_abc(m, n, k, dA.data_ptr(), m, dB.data_ptr(), k, dC.data_ptr(), m, None, 0, None) == 0
    plan.abc(dA, dB, dR)
    torch.cuda.synchronize()
    assert np.array_equal(host, dR.cpu().numpy())
    assert torch.equal(dC, dR)
