"""GPU test of the serving loop: requests served in a mix (continuous
batching, HBM-budget admission, graph replays) get the same result as the
same request computed alone, bit for bit, and within the stated tolerance
of the FP64 product."""
import numpy as np
import pytest

from tolerance import norm_err, rel_frob, tol_for, tol_lowp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_serving_same_alone_and_in_mix(torch, fmm, rt):
    from fmm_gen_amd import autotune as AT
    from fmm_gen_amd import serve as S
    reqs = S.poisson_mix(48, qps=20000.0, seed=5, scale=16)
    rng = np.random.default_rng(5)
    for r in reqs:
        r.host = (rng.uniform(-1, 1, (r.m, r.k)), rng.uniform(-1, 1, (r.k, r.n)), rng.uniform(-1, 1, (r.m, r.n)))
    ex = S.DeviceExecutor(keep_results=True)
    # tiny problems: let the tuner take FMM choices too, so mixes hold both
    for t in ex.tuners.values():
        t.model.theta = np.array([0.0, 1.0, 0.0, 0.0])
    budget = 4 * max(r.nbytes for r in reqs)
    srv = S.Server(ex, hbm_budget=budget, lanes_per_dtype=2, max_batch=16)
    ex.warmup(srv.lanes)
    srv.serve(reqs, timeout=120)
    assert len(srv.done) == len(reqs)
    assert srv.peak_in_use <= budget
    assert max(srv.batch_sizes) > 1
    assert ex.replays + ex.captures >= srv.launches and ex.replays > 0
    algs = set()
    for r in reqs:
        algs.add(r.alg)
        t = torch.float32 if r.dtype == "f32" else torch.float64
        dev = []
        for h in r.host:
            d = rt.colmajor(h.shape[0], h.shape[1], dtype=t)
            d.copy_(torch.from_numpy(np.ascontiguousarray(h)).to(t))
            dev.append(d)
        plan = ex.tuners[r.dtype].plan(r.alg)
        if plan is None:
            rt.gemm(*dev)
        else:
            plan.run(*dev)
        torch.cuda.synchronize()
        alone = dev[2].cpu().numpy()
        assert np.array_equal(alone, r.result), (r.rid, r.dtype, r.alg, r.tile, (r.m, r.n, r.k))
        A, B, C = r.host
        levels = 0 if AT.CANDIDATE_ALGS[r.alg] is None else sum(AT.CANDIDATE_ALGS[r.alg][1])
        if r.dtype == "f64":
            assert norm_err(r.result, A, B, C) <= tol_for(levels) if levels else 16.0
        else:
            A32, B32, C32 = (x.astype(np.float32).astype(np.float64) for x in (A, B, C))
            assert rel_frob(r.result, A32, B32, C32) <= tol_lowp("f32", levels)
    assert len(algs) > 1
