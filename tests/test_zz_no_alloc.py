"""fmm_abc_noalloc (the allocation-free ABC entry point) against fmm_abc.

Two calls in a row through the C-ABI, the second with a smaller problem,
each compared bit for bit with fmm_abc on the same inputs (same kernels, same
product order).  The plans are fresh, so the first calls also copy the
product tables into the library's scratch array; a call with an FP32 problem
follows.  After the first call has set the device up, the calls leave the
device's free memory unchanged (they allocate nothing).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def operands(torch, rt, m, n, k, seed, dtype):
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = rt.colmajor(m, k, dtype=dtype).uniform_(-1, 1, generator=g)
    B = rt.colmajor(k, n, dtype=dtype).uniform_(-1, 1, generator=g)
    C = rt.colmajor(m, n, dtype=torch.float64 if dtype == torch.float64 else torch.float32)
    C.uniform_(-1, 1, generator=g)
    return A, B, C


def test_noalloc_equals_abc(torch, fmm, rt):
    alg = fmm.build_algorithm(["222", "323"], [1, 1])      # R = 105: ordered launch, counters in scratch
    na = rt.Plan(alg)            # used only by the allocation-free entry point
    ref = rt.Plan(alg)           # the existing entry point, fmm_abc
    free = []
    for i, (m, n, k, dt) in enumerate([(3000, 2600, 1900, torch.float64), (1000, 700, 333, torch.float64),
                                       (2100, 1500, 1200, torch.float32)]):
        print(f"\n[noalloc] call {i}: {m}x{n}x{k} {dt}", flush=True)
        A, B, C0 = operands(torch, rt, m, n, k, 40 + i, dt)
        C1 = rt.as_colmajor(C0)
        C2 = rt.as_colmajor(C0)
        torch.cuda.synchronize()
        f0 = torch.cuda.mem_get_info()[0]
        na.abc_noalloc(A, B, C1)
        torch.cuda.synchronize()
        free.append((f0, torch.cuda.mem_get_info()[0]))
        ref.abc(A, B, C2)
        torch.cuda.synchronize()
        assert torch.equal(C1, C2), (i, (C1 - C2).abs().max().item())
        assert not torch.equal(C1, C0)
    assert rt.ws_status() == 0
    # after the first call's set-up, no call changes the free device memory
    for before, after in free[1:]:
        assert before == after, free
    na.close()
    ref.close()
