"""Error flags of the product-order waits reach every kind of caller (GPU).

The order-wait bound is forced to 0 ms from the host (FMM_SPIN_LIMIT_MS=0),
so every wait that is not satisfied at once gives up and sets its flag; the
kernels still run to the end (the C updates are then unordered, so C is not
checked).  Shapes with one C tile per leaf block and many products make waits
certain: every product of the launch updates the same tile, and a queue's
workgroups start their units together.  Then, with the bound back to its
default, the same calls report no flag.
"""
import numpy as np
import pytest

from conftest import int_mats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def dev_ops(torch, rt, m, n, k, dtype=None):
    rng = np.random.default_rng(0)
    A, B, C = int_mats(rng, m, n, k)
    t = dtype or torch.float64
    return tuple(rt.as_colmajor(torch.from_numpy(np.ascontiguousarray(X)).to("cuda", t)) for X in (A, B, C))


def test_owned_workspace_flags_reach_status(torch, rt, fmm, monkeypatch):
    plan = rt.Plan(fmm.build_algorithm(["222", "323"], [1, 1]))    # R = 105, leaf 128 x 64 x 128
    A, B, C = dev_ops(torch, rt, 6 * 128, 6 * 128, 4 * 64)
    rt.ws_status()                              # clears anything earlier tests left
    monkeypatch.setenv("FMM_SPIN_LIMIT_MS", "0")
    plan.abc(A, B, C)                           # ws=None: the library allocates it
    assert rt.ws_status() & 1
    assert rt.ws_status() == 0                  # the query cleared it
    # a caller-supplied workspace keeps the flag itself
    ws = torch.zeros(plan.workspace_bytes(A.shape[0], B.shape[1], A.shape[1]) // 4 + 64, dtype=torch.int32, device="cuda")
    plan.abc(A, B, C, ws=ws)
    assert rt.ws_status(ws) & 1
    assert rt.ws_status() == 0                  # and nothing went to the device word
    # the fused Naive launch (FP32) and its workspace's status header
    A32, B32, C32 = dev_ops(torch, rt, 6 * 128, 6 * 128, 4 * 64, torch.float32)
    plan.run(A32, B32, C32, pack="naive")
    assert rt.ws_status() & 1
    monkeypatch.setenv("FMM_SPIN_LIMIT_MS", "5000")
    plan.abc(A, B, C)
    plan.abc(A, B, C, ws=ws)
    plan.run(A32, B32, C32, pack="naive")
    assert rt.ws_status() == 0 and rt.ws_status(ws) == 0
    plan.close()


def test_host_entry_returns_espin(torch, rt, fmm, monkeypatch):
    plan = rt.Plan(fmm.build_algorithm(["222", "323"], [1, 1]))
    rng = np.random.default_rng(1)
    A, B, C = int_mats(rng, 6 * 128, 6 * 128, 4 * 64)
    rt.ws_status()
    monkeypatch.setenv("FMM_SPIN_LIMIT_MS", "0")
    with pytest.raises(rt.FmmError, match="FMM_ESPIN"):
        plan.host_abc(A, B, C.copy(order="F"))
    monkeypatch.setenv("FMM_SPIN_LIMIT_MS", "5000")
    got = C.copy(order="F")
    plan.host_abc(A, B, got)
    assert np.array_equal(got, C + A @ B)
    plan.close()


def test_refine_flags_reach_its_status_header(torch, rt, fmm, monkeypatch):
    plan = rt.Plan(fmm.build_algorithm(["222", "323"], [1, 1]))
    A, B, C = dev_ops(torch, rt, 6 * 128, 6 * 128, 4 * 64)
    rt.ws_status()
    monkeypatch.setenv("FMM_SPIN_LIMIT_MS", "0")
    rt.refine(A, B, C, main="f32", main_plan=plan, corr="f32")
    assert rt.ws_status() & 1
    monkeypatch.setenv("FMM_SPIN_LIMIT_MS", "5000")
    rt.refine(A, B, C, main="f32", main_plan=plan, corr="f32")
    assert rt.ws_status() == 0
    plan.close()


def test_batch_last_status_without_sync(torch, rt, fmm, monkeypatch):
    plan = rt.Plan(fmm.build_algorithm(["222", "323"], [1, 1]))
    A, B, C = dev_ops(torch, rt, 6 * 128, 6 * 128, 4 * 64)
    monkeypatch.setenv("FMM_SPIN_LIMIT_MS", "0")
    b = rt.Batch([(A, B, C, plan, 0)])
    b.run()
    ev = torch.cuda.Event()
    ev.record()
    ev.synchronize()
    assert b.last_status() & 1
    monkeypatch.setenv("FMM_SPIN_LIMIT_MS", "5000")
    b.set([(A, B, C, plan, 0)])
    b.run()
    ev.record()
    ev.synchronize()
    assert b.last_status() == 0 and b.status() == 0
    b.close()
    plan.close()
