"""HIP source emitter for one generated job (the role of meta/abc_gen_mix.py,
meta/misc_gen.py and meta/control.py:42-118 in the reference).

For a job such as ``222-1_abc`` it writes, under ``<gen_path>/<job>/``:

  include/        the kernel and runtime headers (copied, as the reference
                  copies common/ into every job directory)
  kernels/fmm_<job>.hip
                  the generated source: the composed [U,V,W] tables as
                  constexpr data, the product equations as comments, the
                  kernels instantiated for this job (source chunk = the job's
                  largest per-product term count) and the reference's
                  per-job entry bl_dgemm_strassen_<pack>
  test/test_<job>.cpp
                  the driver: ``test_<job>_{st,mt}.x m n k`` prints
                  "m n k fmm_GF gemm_GF" like meta/misc_gen.py:147-261
  makefile_<job>  the hipcc recipe (for rebuilding by hand)

and builds ``lib/lib<job>.so`` and ``test/test_<job>_{st,mt}.x``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from typing import List, Sequence

from . import build as B
from .coeffs import Algorithm

# every kernel / runtime header: a job directory builds on its own (makefile)
HEADERS = sorted(f for f in os.listdir(B.CSRC) if f.endswith(".hpp"))


def _fmt(x) -> str:
    v = float(x)
    if v == int(v):
        return f"{int(v)}.0"
    return repr(v)


def _block_name(alg: Algorithm, which: int, row: int) -> str:
    I, J = alg.block_of_row(which, row)
    return f"{'ABC'[which]}({I},{J})"


def equations(alg: Algorithm) -> List[str]:
    """One comment line per product, e.g.
    M0 = (A(0,0) + A(1,1)) * (B(0,0) + B(1,1)); C(0,0) += M0; C(1,1) += M0"""
    out = []
    for r in range(alg.rank):
        def side(T, which):
            parts = []
            for a, row in enumerate(T):
                c = row[r]
                if c == 0:
                    continue
                nm = _block_name(alg, which, a)
                if c == 1:
                    parts.append(("+", nm))
                elif c == -1:
                    parts.append(("-", nm))
                else:
                    parts.append(("+", f"({float(c):g})*{nm}"))
            s = "".join(f" {sg} {t}" for sg, t in parts).strip()
            return s[2:] if s.startswith("+ ") else s
        cs = []
        for a, row in enumerate(alg.W):
            c = row[r]
            if c == 0:
                continue
            nm = _block_name(alg, 2, a)
            cs.append(f"{nm} {'+=' if c == 1 else '-=' if c == -1 else f'+= ({float(c):g})*'} M{r}")
        out.append(f"M{r} = ({side(alg.U, 0)}) * ({side(alg.V, 1)}); " + "; ".join(cs))
    return out


def _table(name: str, T) -> str:
    flat = [_fmt(x) for row in T for x in row]
    lines = []
    for i in range(0, len(flat), 16):
        lines.append("    " + ", ".join(flat[i:i + 16]) + ",")
    return f"constexpr double {name}[{len(flat)}] = {{\n" + "\n".join(lines) + "\n};"


def max_terms(alg: Algorithm) -> int:
    best = 1
    for r in range(alg.rank):
        best = max(best, sum(1 for row in alg.U if row[r] != 0), sum(1 for row in alg.V if row[r] != 0))
    return best


PACK_CONST = {"abc": "FMM_PACK_ABC", "ab": "FMM_PACK_AB", "naive": "FMM_PACK_NAIVE"}
PACK_REF = {"abc": "meta/abc_gen_mix.py:448", "ab": "meta/ab_gen_mix.py:341",
            "naive": "meta/naive_gen_mix.py:244"}


def emit_source(alg: Algorithm, job: str, pack_type: str) -> str:
    if pack_type not in PACK_CONST:
        raise ValueError(f"pack_type {pack_type!r}: one of {sorted(PACK_CONST)}")
    pk = PACK_CONST[pack_type]
    Mx, Kx, Nx = alg.dims
    ch = max_terms(alg)
    dims = ", ".join(str(d) for lv in alg.levels for d in lv)
    eq = "\n".join("// " + e for e in equations(alg))
    src = f"""// Generated by fmm-gen_amd control.py for job {job}: do not edit.
//
// Algorithm {alg.name}: <{Mx},{Kx},{Nx}>, rank {alg.rank}, {len(alg.levels)} level(s),
// nnz(U,V,W) = {alg.nnz()}.  Products, in execution order (blocks are
// (row-block, column-block) of the composed grid):
{eq}
//
// Largest number of A or B terms of one product: {ch} (the kernel variant is
// chosen from it: 32-deep slices for at most 2 terms, else 16-deep).
// Pack type: {pack_type}.
#include "fmm_capi.hpp"

namespace {{
constexpr int kLevels = {len(alg.levels)};
constexpr int kDims[{3 * len(alg.levels)}] = {{{dims}}};
constexpr int kRank = {alg.rank};
// Composed tables, reference row order (level 0 most significant), row-major
// [rows x rank].
{_table("kU", alg.U)}
{_table("kV", alg.V)}
{_table("kW", alg.W)}

fmm_plan_t job_plan() {{
  static fmm_plan_t plan = [] {{
    fmm_plan_t p = nullptr;
    if (fmm_plan_create(&p, kLevels, kDims, kRank, kU, kV, kW) != FMM_OK) p = nullptr;
    return p;
  }}();
  return plan;
}}
}}  // namespace

extern "C" {{

fmm_plan_t fmm_job_plan(void) {{ return job_plan(); }}

int fmm_job_pack(void) {{ return {pk}; }}

int fmm_job_{pack_type}(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                void* ws, size_t ws_bytes, void* stream) {{
  fmm_plan_t p = job_plan();
  if (!p) return FMM_ENOMEM;
  return fmm_run(p, {pk}, FMM_F64, m, n, k, A, lda, B, ldb, C, ldc, ws, ws_bytes, stream);
}}

// The reference's generated entry ({PACK_REF[pack_type]}): host pointers,
// C += A*B, no return value.
void bl_dgemm_strassen_{pack_type}(int m, int n, int k, double* XA, int lda, double* XB, int ldb, double* XC,
                                   int ldc) {{
  fmm_plan_t p = job_plan();
  int rc = p ? fmm_host_run(p, {pk}, m, n, k, XA, lda, XB, ldb, XC, ldc) : FMM_ENOMEM;
  fmm::report("bl_dgemm_strassen_{pack_type}", rc);
}}

}}  // extern "C"
"""
    return src


DRIVER = r"""// Generated by fmm-gen_amd control.py: test driver of job @JOB@.
//
// Usage: test_@JOB@_{st,mt}.x m n k
// Same contract as the reference driver (meta/misc_gen.py:147-261): A (lda=m)
// and B (ldb=k) with entries drand48() in [0,1), C and C_ref = 0 with
// ldc = roundup(m, 8) and n+4 columns, the FMM applied 3 times and the plain
// GEMM applied 3 times, the best time of each, a print-only 1e-10 comparison,
// then "m n k fmm_GF gemm_GF" with GF = 2mnk/1e9/t.  The seed is FMM_SEED
// when set (the reference always seeds with time(NULL)).
//
// The GFLOPS printed are those of the device-pointer entries with A, B, C
// resident on the GPU (hipEvent timing); the host-pointer drop-in
// bl_dgemm_strassen_@PACK@, whose time includes the PCIe copies, is run once
// and reported on a second line when FMM_HOST_RATE is set.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "fmm_amd.h"

#define CHECK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
  fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); exit(1); } } while (0)

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char** argv) {
  if (argc != 4) {
    printf("Error: require 3 arguments, but only %d provided.\n", argc - 1);
    return 0;
  }
  int m = atoi(argv[1]), n = atoi(argv[2]), k = atoi(argv[3]);
  if (m <= 0 || n <= 0 || k <= 0) {
    printf("Error: m, n, k must be positive.\n");
    return 0;
  }
  const char* seed = getenv("FMM_SEED");
  srand48(seed ? atol(seed) : (long)time(NULL));
  int lda = m, ldb = k, ldc = ((m - 1) / 8 + 1) * 8;
  size_t na = (size_t)lda * k, nb = (size_t)ldb * n, nc = (size_t)ldc * (n + 4);
  double* A = (double*)malloc(na * sizeof(double));
  double* B = (double*)malloc(nb * sizeof(double));
  double* C = (double*)calloc(nc, sizeof(double));
  double* Cref = (double*)calloc(nc, sizeof(double));
  for (size_t i = 0; i < na; ++i) A[i] = drand48();
  for (size_t i = 0; i < nb; ++i) B[i] = drand48();
  double *dA, *dB, *dC, *dCref;
  CHECK(hipMalloc(&dA, na * sizeof(double)));
  CHECK(hipMalloc(&dB, nb * sizeof(double)));
  CHECK(hipMalloc(&dC, nc * sizeof(double)));
  CHECK(hipMalloc(&dCref, nc * sizeof(double)));
  CHECK(hipMemcpy(dA, A, na * sizeof(double), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(dB, B, nb * sizeof(double), hipMemcpyHostToDevice));
  CHECK(hipMemset(dC, 0, nc * sizeof(double)));
  CHECK(hipMemset(dCref, 0, nc * sizeof(double)));
  // the job's workspace, taken once outside the timed calls
  size_t ws_bytes = fmm_workspace_bytes_pack(fmm_job_plan(), fmm_job_pack(), m, n, k);
  void* ws = NULL;
  if (ws_bytes) CHECK(hipMalloc(&ws, ws_bytes));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const int nrepeats = 3;
  float best_fmm = 1e30f, best_gemm = 1e30f, t;
  for (int r = 0; r < nrepeats; ++r) {
    CHECK(hipEventRecord(e0, 0));
    int rc = fmm_job_@PACK@(m, n, k, dA, lda, dB, ldb, dC, ldc, ws, ws_bytes, NULL);
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    if (rc != FMM_OK) { fprintf(stderr, "fmm_job_@PACK@: error %d\n", rc); return 1; }
    CHECK(hipEventElapsedTime(&t, e0, e1));
    if (t < best_fmm) best_fmm = t;
  }
  for (int r = 0; r < nrepeats; ++r) {
    CHECK(hipEventRecord(e0, 0));
    int rc = fmm_gemm(FMM_F64, m, n, k, dA, lda, dB, ldb, dCref, ldc, NULL);
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    if (rc != FMM_OK) { fprintf(stderr, "fmm_gemm: error %d\n", rc); return 1; }
    CHECK(hipEventElapsedTime(&t, e0, e1));
    if (t < best_gemm) best_gemm = t;
  }
  CHECK(hipMemcpy(C, dC, nc * sizeof(double), hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(Cref, dCref, nc * sizeof(double), hipMemcpyDeviceToHost));
  // print-only comparison, first mismatch per row (bl_compare_error)
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j)
      if (fabs(C[(size_t)j * ldc + i] - Cref[(size_t)j * ldc + i]) > 1e-10) {
        printf("C[ %d ][ %d ] != C_ref, %E, %E\n", i, j, C[(size_t)j * ldc + i], Cref[(size_t)j * ldc + i]);
        break;
      }
  double gf = (m * (double)n / 1e9) * (2.0 * k);
  printf("%5d\t %5d\t %5d\t %5.2lf\t %5.2lf\n", m, n, k, gf / (best_fmm * 1e-3), gf / (best_gemm * 1e-3));
  if (getenv("FMM_HOST_RATE")) {
    double* Ch = (double*)calloc(nc, sizeof(double));
    double t0 = now_s();
    bl_dgemm_strassen_@PACK@(m, n, k, A, lda, B, ldb, Ch, ldc);
    double th = now_s() - t0;
    printf("host-pointer bl_dgemm_strassen_@PACK@ (PCIe-inclusive): %5.2lf GFLOPS\n", gf / th);
    free(Ch);
  }
  fflush(stdout);
  CHECK(hipFree(dA));
  CHECK(hipFree(dB));
  CHECK(hipFree(dC));
  CHECK(hipFree(dCref));
  if (ws) CHECK(hipFree(ws));
  free(A);
  free(B);
  free(C);
  free(Cref);
  bl_finalize();
  return 0;
}
"""


def emit_driver(job: str, pack_type: str) -> str:
    return DRIVER.replace("@JOB@", job).replace("@PACK@", pack_type)


def emit_makefile(job: str) -> str:
    flags = " ".join(B.HIPFLAGS)
    return f"""# Generated by fmm-gen_amd control.py for job {job}.
HIPCC ?= hipcc
FLAGS := {flags} -Iinclude

all: lib/lib{job}.so test/test_{job}_st.x test/test_{job}_mt.x

lib/lib{job}.so: kernels/fmm_{job}.hip
\t@mkdir -p lib
\t$(HIPCC) $(FLAGS) -shared -o $@ $<

test/test_{job}_st.x: test/test_{job}.cpp lib/lib{job}.so
\t$(HIPCC) $(FLAGS) -o $@ $< -Llib -l{job} -Wl,-rpath,'$$ORIGIN/../lib'

test/test_{job}_mt.x: test/test_{job}_st.x
\tcp $< $@

clean:
\trm -f lib/lib{job}.so test/test_{job}_st.x test/test_{job}_mt.x
"""


def generate_job(alg: Algorithm, job: str, pack_type: str, job_dir: str, compile: bool = True) -> dict:
    for sub in ("include", "kernels", "test", "lib"):
        os.makedirs(os.path.join(job_dir, sub), exist_ok=True)
    for h in HEADERS:
        shutil.copy(os.path.join(B.CSRC, h), os.path.join(job_dir, "include", h))
    shutil.copy(os.path.join(B.INCLUDE, "fmm_amd.h"), os.path.join(job_dir, "include", "fmm_amd.h"))
    src = os.path.join(job_dir, "kernels", f"fmm_{job}.hip")
    with open(src, "w") as fh:
        fh.write(emit_source(alg, job, pack_type))
    drv = os.path.join(job_dir, "test", f"test_{job}.cpp")
    with open(drv, "w") as fh:
        fh.write(emit_driver(job, pack_type))
    with open(os.path.join(job_dir, f"makefile_{job}"), "w") as fh:
        fh.write(emit_makefile(job))
    out = {"source": src, "driver": drv}
    if compile:
        inc = [os.path.join(job_dir, "include")]
        lib = os.path.join(job_dir, "lib", f"lib{job}.so")
        B.build_shared(src, lib, includes=inc)
        st = os.path.join(job_dir, "test", f"test_{job}_st.x")
        cmd = [B.hipcc()] + B.HIPFLAGS + ["-I", inc[0], "-o", st, drv, "-L", os.path.dirname(lib),
                                           f"-l{job}", "-Wl,-rpath,$ORIGIN/../lib"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("driver build failed:\n" + res.stderr[-4000:])
        mt = os.path.join(job_dir, "test", f"test_{job}_mt.x")
        shutil.copy(st, mt)
        out.update({"library": lib, "driver_st": st, "driver_mt": mt})
    return out
