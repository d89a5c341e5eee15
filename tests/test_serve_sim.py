"""CPU tests of the serving scheduler (fmm_gen_amd/serve.py) with a simulated
executor: admission by HBM budget in arrival order, continuous batching over
overlapping lanes, and every request completed with its own result."""
import numpy as np

from fmm_gen_amd import serve as S


class Tracking(S.SimExecutor):
    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.max_running = 0
        self.max_allocated = 0
        self.lanes = []

    def init_lane(self, ln):
        super().init_lane(ln)
        self.lanes.append(ln)

    def alloc(self, r):
        super().alloc(r)
        self.max_allocated = max(self.max_allocated, self.allocated)

    def launch(self, ln, reqs):
        super().launch(ln, reqs)
        self.max_running = max(self.max_running, 1 + sum(1 for x in self.lanes if x.busy))


def mix(n=60, seed=3):
    reqs = S.poisson_mix(n, qps=4000.0, seed=seed, scale=64)
    rng = np.random.default_rng(seed)
    for r in reqs:
        r.host = (rng.uniform(-1, 1, (r.m, r.k)), rng.uniform(-1, 1, (r.k, r.n)), rng.uniform(-1, 1, (r.m, r.n)))
    return reqs


def test_mix_shapes_and_dtypes():
    reqs = S.poisson_mix(4000, qps=100.0, seed=1)
    sq = sum(1 for r in reqs if r.m == r.n == r.k)
    rk = sum(1 for r in reqs if r.m == r.n and r.k in (256, 512) and r.m in (4096, 8192))
    ts = sum(1 for r in reqs if (r.m, r.n, r.k) == (16384, 512, 512))
    f32 = sum(1 for r in reqs if r.dtype == "f32")
    assert abs(sq / 4000 - 0.6) < 0.03 and abs(rk / 4000 - 0.3) < 0.03 and abs(ts / 4000 - 0.1) < 0.02
    assert abs(f32 / 4000 - 0.7) < 0.03
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 0.01) < 0.001          # Poisson at 100 QPS


def test_budget_fifo_and_completion():
    reqs = mix()
    budget = 3 * max(r.nbytes for r in reqs)
    ex = Tracking(clock=__import__("time").perf_counter, flops_per_s=2e9, launch_s=2e-4)
    srv = S.Server(ex, hbm_budget=budget, lanes_per_dtype=2, max_batch=8)
    srv.serve(reqs, timeout=60)
    assert len(srv.done) == len(reqs)
    assert ex.max_allocated <= budget and srv.peak_in_use <= budget
    assert ex.allocated == 0 and srv.in_use == 0
    # admission follows arrival order
    adm = [r.t_admit for r in sorted(reqs, key=lambda r: r.arrival)]
    assert all(a <= b for a, b in zip(adm, adm[1:]))
    for r in reqs:
        A, B, C = r.host
        assert np.allclose(r.result, C + A @ B)
        assert r.arrival <= r.t_admit <= r.t_launch <= r.t_done
    st = srv.stats()
    assert st["requests"] == len(reqs) and st["latency_p99_ms"] >= st["latency_p50_ms"] > 0


def test_continuous_batching_overlaps_lanes():
    reqs = mix(80, seed=4)
    for r in reqs:
        r.arrival *= 0.01          # a burst: many requests wait at once
    ex = Tracking(clock=__import__("time").perf_counter, flops_per_s=1e9, launch_s=1e-3)
    srv = S.Server(ex, hbm_budget=1e12, lanes_per_dtype=2, max_batch=16)
    srv.serve(reqs, timeout=60)
    assert len(srv.done) == len(reqs)
    assert max(srv.batch_sizes) > 1          # requests share launches
    assert ex.max_running >= 2               # launches overlap on the device


def test_launch_error_flags_reach_requests():
    """The executor's per-launch error flags (a batch's product-order wait
    that timed out) mark exactly the requests of that launch as failed."""
    class Faulty(S.SimExecutor):
        def __init__(self, *a, **k):
            super().__init__(*a, **k)
            self.n = 0

        def launch(self, ln, reqs):
            super().launch(ln, reqs)
            self.n += 1
            ln.state["error"] = 1 if self.n == 2 else 0     # the second launch reports a flag

    reqs = mix(30, seed=5)
    ex = Faulty(clock=__import__("time").perf_counter, flops_per_s=1e9, launch_s=1e-4)
    srv = S.Server(ex, hbm_budget=1e12, lanes_per_dtype=1, max_batch=4)
    launched = []
    orig = ex.launch
    ex.launch = lambda ln, rs: (launched.append(list(rs)), orig(ln, rs))
    srv.serve(reqs, timeout=60)
    bad = {r.rid for r in launched[1]}
    assert {r.rid for r in srv.done if r.error} == bad
    assert srv.stats()["failed"] == len(bad) > 0
