"""Write the committed golden fixtures (run in the build container).

ref_dgemm.npz: seeded float64 inputs and the output of the reference's own
    bl_dgemm (common/dgemm/my_dgemm.c, compiled from its sources into
    oracle/_ref by oracle/Makefile) -- C += A*B with its KC-blocked
    summation order.  Pins the oracle's blocked GEMM bit for bit on machines
    without the reference sources.
kat_abc.npz: small-integer known-answer cases (A, B, C0 in [-4, 4]) with
    C0 + A@B from exact integer arithmetic, per algorithm configuration and
    fringe residue.  Any correct FMM reproduces them bit for bit.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def aligned_f(rows, cols, align=64):
    buf = np.zeros(rows * cols + align // 8, dtype=np.float64)
    off = (-buf.ctypes.data % align) // 8
    return buf[off:off + rows * cols].reshape((rows, cols), order="F")


def ref_dgemm():
    from oracle import oracle
    oracle.build(ref=True)
    R = oracle.ref_lib()
    assert R is not None, "reference runtime did not build"
    rng = np.random.default_rng(20260101)
    out = {}
    cases = [(17, 9, 300), (24, 12, 257), (9, 5, 1), (16, 8, 520)]
    for i, (m, n, k) in enumerate(cases):
        A = np.asfortranarray(rng.random((m, k)))
        B = np.asfortranarray(rng.random((k, n)))
        C0 = rng.uniform(-1, 1, (m, n))
        ldc = ((m - 1) // 8 + 1) * 8
        C = aligned_f(ldc, n + 4)
        C[:m, :n] = C0
        R.bl_dgemm(m, n, k, oracle._p(A), m, oracle._p(B), k, oracle._p(C), ldc)
        out.update({f"A{i}": A, f"B{i}": B, f"C{i}": C0, f"out{i}": np.array(C[:m, :n])})
    out["count"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "ref_dgemm.npz"), **out)


KAT_SPECS = [(["222"], [1]), (["222"], [2]), (["222", "323"], [1, 1]), (["333"], [1]),
             (["232"], [1]), (["323"], [1])]


def kat_abc():
    from __graft_entry__ import load_package
    load_package()
    import fmm_gen_amd as F
    rng = np.random.default_rng(20260102)
    out, index = {}, []
    for spec in KAT_SPECS:
        alg = F.build_algorithm(*spec)
        Mx, Kx, Nx = alg.dims
        for (m, n, k) in [(Mx * 8 * 2 + 3, Nx * 4 * 3 + 1, Kx * 6 + 1), (Mx * 8 * 2, Nx * 4 * 4, Kx * 4)]:
            i = len(index)
            A = rng.integers(-4, 5, (m, k)).astype(np.float64)
            B = rng.integers(-4, 5, (k, n)).astype(np.float64)
            C0 = rng.integers(-4, 5, (m, n)).astype(np.float64)
            want = (C0.astype(np.int64) + A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
            out.update({f"A{i}": A, f"B{i}": B, f"C{i}": C0, f"out{i}": want})
            index.append({"algs": spec[0], "levels": spec[1], "m": m, "n": n, "k": k})
    out["index"] = np.array(json.dumps(index))
    np.savez_compressed(os.path.join(HERE, "kat_abc.npz"), **out)


if __name__ == "__main__":
    ref_dgemm()
    kat_abc()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
