"""Build of the HIP libraries (gfx950) with hipcc, in-tree.

``build_runtime()`` compiles ``csrc/fmm_runtime.hip`` into
``lib/libfmm_amd.so``.  Per-job libraries are built by ``control.py``
through ``build_shared()`` with the same flags.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from typing import List, Optional

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "lib")
INCLUDE = os.path.join(os.path.dirname(PKG_DIR), "include")
ARCH = "gfx950"

# -packed-fp32-ops off: keep the compiler from forming v_pk_*_f32 on its own.
# amdgpu-mfma-vgpr-form: MFMA accumulators in VGPRs; with the default AGPR
# form this compiler copies every accumulator AGPR<->VGPR around each loop
# iteration of an MFMA chain, which costs more than the MFMAs themselves.
# amdgpu-sched-strategy=max-memory-clause: the machine scheduler variant that
# measured best for the FP64 single-term kernel (GEMM 9600^3 63.5 -> 65.8
# TFLOP/s, Naive 333 L1 64.4 -> 66.3) and neutral for the others; the default
# and the other strategies (max-ilp, iterative-ilp, iterative-minreg) were
# timed beside it.
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-mllvm", "-amdgpu-mfma-vgpr-form=1",
            "-mllvm", "-amdgpu-sched-strategy=max-memory-clause"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_shared(src: str, out: str, extra: Optional[List[str]] = None, includes: Optional[List[str]] = None) -> str:
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cmd = [hipcc()] + HIPFLAGS + ["-shared"]
    for inc in [CSRC, INCLUDE] + list(includes or []):
        cmd += ["-I", inc]
    cmd += list(extra or []) + ["-o", out + ".tmp", src]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stderr[-4000:])
    os.replace(out + ".tmp", out)
    return out


def build_runtime() -> str:
    return build_shared(os.path.join(CSRC, "fmm_runtime.hip"), os.path.join(LIB_DIR, "libfmm_amd.so"))


if __name__ == "__main__":
    print(build_runtime())
