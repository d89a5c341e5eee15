"""Serving loop: continuous batching of GEMM requests onto persistent launches.

Requests (C += A @ B of mixed shapes, FP32 or FP64) arrive over time.  The
server

* admits them in arrival order while the HBM their operands need fits the
  budget (operands stay resident on the device from admission until the
  result is done, then the memory returns to the budget);
* keeps, per dtype, a few *lanes*: a stream, one reusable device job table
  (runtime.Batch) and one HIP graph per kernel variant captured from that
  table's launch -- since the launch arguments depend only on the table's
  buffers, refilling the table (Batch.set, stream ordered) and replaying the
  graph runs any new set of problems;
* whenever a lane is idle, launches every admitted request of its dtype not
  yet running (up to a cap) as ONE persistent launch, each request with the
  algorithm and C-tile shape the autotuner picks for its shape;
* completes the requests of a lane when its event has fired, and records
  each request's latency (completion - arrival).

Requests that arrive while launches run are admitted as memory frees up and
go out in the next launch on another lane, overlapping on the device with
the running ones.  A request's result does not depend on what it shares a
launch with (each problem of a batch gets a single call's result bit for
bit).  Device work goes through an executor, so the scheduling policy is
tested on the CPU with a simulated one.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Deque, Dict, List, Optional

import numpy as np

ELEM = {"f32": 4, "f64": 8}


def request_bytes(m: int, n: int, k: int, dtype: str) -> int:
    return (m * k + k * n + m * n) * ELEM[dtype]


@dataclass
class Request:
    rid: int
    m: int
    n: int
    k: int
    dtype: str                      # "f32" or "f64"
    arrival: float                  # seconds on the server clock
    host: Optional[tuple] = None    # optional (A, B, C) numpy inputs
    nbytes: int = 0
    alg: Optional[str] = None
    tile: int = 0
    t_admit: Optional[float] = None
    t_launch: Optional[float] = None
    t_done: Optional[float] = None
    dev: Optional[tuple] = None     # device (A, B, C) while resident
    result: Optional[np.ndarray] = None
    error: int = 0                  # error flags of the launch that ran it (0: result valid)

    def __post_init__(self):
        self.nbytes = request_bytes(self.m, self.n, self.k, self.dtype)

    @property
    def flops(self) -> float:
        return 2.0 * self.m * self.n * self.k

    @property
    def latency(self) -> Optional[float]:
        return None if self.t_done is None else self.t_done - self.arrival


@dataclass
class Lane:
    dtype: str
    index: int
    busy: bool = False
    reqs: List[Request] = field(default_factory=list)
    state: dict = field(default_factory=dict)   # executor-owned (stream, batch, graphs, event)


class Server:
    def __init__(self, executor, hbm_budget: float = 200e9, lanes_per_dtype: int = 2, max_batch: int = 256,
                 dtypes=("f32", "f64"), clock: Callable[[], float] = time.perf_counter):
        self.ex = executor
        self.budget = float(hbm_budget)
        self.max_batch = max_batch
        self.clock = clock
        self.lanes = [Lane(dt, i) for dt in dtypes for i in range(lanes_per_dtype)]
        for ln in self.lanes:
            self.ex.init_lane(ln)
        self.pending: Deque[Request] = collections.deque()     # arrived, not admitted
        self.ready: Dict[str, Deque[Request]] = {dt: collections.deque() for dt in dtypes}
        self.in_use = 0.0
        self.peak_in_use = 0.0
        self.done: List[Request] = []
        self.launches = 0
        self.batch_sizes: List[int] = []
        self.t0 = None

    # --- scheduling ------------------------------------------------------
    def submit(self, r: Request) -> None:
        if r.nbytes > self.budget:
            raise ValueError(f"request {r.rid} needs {r.nbytes} bytes, more than the HBM budget")
        self.pending.append(r)

    def _complete(self, now: float) -> int:
        n = 0
        for ln in self.lanes:
            if ln.busy and self.ex.finished(ln):
                err = self.ex.status(ln)
                for r in ln.reqs:
                    r.t_done = now
                    r.error = err
                    self.ex.collect(r)
                    self.ex.free(r)
                    self.in_use -= r.nbytes
                    self.done.append(r)
                n += len(ln.reqs)
                ln.reqs = []
                ln.busy = False
        return n

    def _admit(self, now: float) -> int:
        n = 0
        while self.pending and self.in_use + self.pending[0].nbytes <= self.budget:
            r = self.pending.popleft()
            self.ex.alloc(r)
            r.t_admit = now
            self.in_use += r.nbytes
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            self.ready[r.dtype].append(r)
            n += 1
        return n

    def _launch(self, now: float) -> int:
        n = 0
        for ln in self.lanes:
            q = self.ready[ln.dtype]
            if ln.busy or not q:
                continue
            reqs = []
            while q and len(reqs) < self.max_batch:
                reqs.append(q.popleft())
            for r in reqs:
                r.t_launch = now
            self.ex.launch(ln, reqs)
            ln.reqs = reqs
            ln.busy = True
            self.launches += 1
            self.batch_sizes.append(len(reqs))
            n += 1
        return n

    def step(self) -> bool:
        """One scheduling pass; True if anything happened."""
        now = self.clock() - self.t0
        a = self._complete(now)
        b = self._admit(now)
        c = self._launch(now)
        return bool(a or b or c)

    def busy(self) -> bool:
        return bool(self.pending) or any(self.ready.values()) or any(ln.busy for ln in self.lanes)

    def serve(self, requests: List[Request], idle_sleep: float = 20e-6, timeout: float = 600.0) -> None:
        """Feed `requests` at their arrival times (seconds from now) and run
        until every one is done."""
        order = sorted(requests, key=lambda r: r.arrival)
        i = 0
        self.t0 = self.clock()
        while i < len(order) or self.busy():
            now = self.clock() - self.t0
            if now > timeout:
                raise TimeoutError("serving loop did not drain")
            while i < len(order) and order[i].arrival <= now:
                self.submit(order[i])
                i += 1
            if not self.step():
                self.ex.idle(idle_sleep)

    # --- results -------------------------------------------------------------
    def stats(self) -> dict:
        lat = np.array([r.latency for r in self.done])
        if len(lat) == 0:
            return {"requests": 0}
        first = min(r.arrival for r in self.done)
        last = max(r.t_done for r in self.done)
        span = max(last - first, 1e-12)
        fl = sum(r.flops for r in self.done)
        return {
            "requests": len(self.done),
            "duration_s": round(span, 4),
            "latency_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "latency_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "latency_mean_ms": round(float(lat.mean()) * 1e3, 3),
            "throughput_req_s": round(len(self.done) / span, 2),
            "throughput_tflops": round(fl / span / 1e12, 3),
            "launches": self.launches,
            "mean_batch": round(float(np.mean(self.batch_sizes)), 2) if self.batch_sizes else 0.0,
            "peak_hbm_gb": round(self.peak_in_use / 1e9, 3),
            "failed": sum(1 for r in self.done if r.error),
        }


# --- workload --------------------------------------------------------------

def poisson_mix(nreq: int, qps: float, seed: int = 0, scale: int = 1) -> List[Request]:
    """The serving mix: 60% square n in {1024, 2048, 4096}; 30% rank-k m = n in
    {4096, 8192}, k in {256, 512}; 10% tall-skinny 16384 x 512 x 512; 70% FP32,
    30% FP64; Poisson arrivals at `qps`.  `scale` divides every dimension
    (for small test runs)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(nreq):
        t += rng.exponential(1.0 / qps)
        u = rng.random()
        if u < 0.6:
            n = int(rng.choice([1024, 2048, 4096]))
            m, k = n, n
        elif u < 0.9:
            m = int(rng.choice([4096, 8192]))
            n, k = m, int(rng.choice([256, 512]))
        else:
            m, n, k = 16384, 512, 512
        dt = "f32" if rng.random() < 0.7 else "f64"
        out.append(Request(i, max(m // scale, 1), max(n // scale, 1), max(k // scale, 1), dt, t))
    return out


# --- executors ---------------------------------------------------------------

class SimExecutor:
    """CPU stand-in for the device: a launch finishes `cost(reqs)` seconds
    after it starts, launches of different lanes overlap, and results are
    computed with numpy (for scheduler tests)."""

    def __init__(self, clock: Callable[[], float], flops_per_s: float = 1e12, launch_s: float = 1e-4):
        self.clock = clock
        self.rate = flops_per_s
        self.launch_s = launch_s
        self.allocated = 0

    def init_lane(self, ln):
        ln.state["end"] = None

    def alloc(self, r):
        self.allocated += r.nbytes

    def free(self, r):
        self.allocated -= r.nbytes

    def launch(self, ln, reqs):
        ln.state["end"] = self.clock() + self.launch_s + sum(r.flops for r in reqs) / self.rate

    def finished(self, ln):
        return self.clock() >= ln.state["end"]

    def status(self, ln):
        return ln.state.get("error", 0)

    def collect(self, r):
        if r.host is not None:
            A, B, C = r.host
            r.result = C + A @ B

    def idle(self, s):
        time.sleep(s)


class DeviceExecutor:
    """The GPU: operands in HBM from admission, one stream + Batch + graphs
    per lane, the autotuner's choice per request."""

    def __init__(self, use_graphs: bool = True, seed: int = 0, keep_results: bool = False, library=None):
        import torch
        from . import autotune as AT
        from . import runtime as rt
        self.torch, self.rt = torch, rt
        self.use_graphs = use_graphs
        self.keep_results = keep_results
        self.lib = library
        self.tuners = {"f32": AT.Tuner("f32"), "f64": AT.Tuner("f64")}
        self.gen = torch.Generator(device="cuda").manual_seed(seed)
        self.admit_stream = torch.cuda.Stream()
        self.captures = 0
        self.replays = 0
        # device table capacity per lane: enough for a few thousand jobs and
        # order counters, so refills rarely reallocate (and drop graphs)
        self.reserve_jobs = 4096
        self.reserve_ws = 1 << 18

    def tdtype(self, dt):
        return self.torch.float32 if dt == "f32" else self.torch.float64

    def init_lane(self, ln):
        torch = self.torch
        st = ln.state
        st["stream"] = torch.cuda.Stream()
        st["event"] = torch.cuda.Event()
        # an empty table of the lane's dtype; filled per launch by set()
        z = self.rt.colmajor(1, 1, dtype=self.tdtype(ln.dtype))
        st["batch"] = self.rt.Batch([(z, z, self.rt.colmajor(1, 1, dtype=self.tdtype(ln.dtype)), None, 0)],
                                    library=self.lib)
        st["batch"].reserve(self.reserve_jobs, self.reserve_ws)
        st["graphs"] = {}
        st["gen"] = st["batch"].generation()

    def alloc(self, r):
        torch, rt = self.torch, self.rt
        t = self.tdtype(r.dtype)
        with torch.cuda.stream(self.admit_stream):
            A = rt.colmajor(r.m, r.k, dtype=t)
            B = rt.colmajor(r.k, r.n, dtype=t)
            C = rt.colmajor(r.m, r.n, dtype=t)
            if r.host is not None:
                for X, h in zip((A, B, C), r.host):
                    X.copy_(torch.from_numpy(np.ascontiguousarray(h)).to(t), non_blocking=False)
            else:
                A.uniform_(-1.0, 1.0, generator=self.gen)
                B.uniform_(-1.0, 1.0, generator=self.gen)
                C.uniform_(-1.0, 1.0, generator=self.gen)
            ev = torch.cuda.Event()
            ev.record(self.admit_stream)
        r.dev = (A, B, C, ev)
        tuner = self.tuners[r.dtype]
        r.alg, r.tile, _ = tuner.choose(r.m, r.n, r.k)

    def launch(self, ln, reqs):
        torch = self.torch
        st = ln.state
        s = st["stream"]
        tuner = self.tuners[ln.dtype]
        probs = []
        for r in reqs:
            A, B, C, ev = r.dev
            s.wait_event(ev)
            probs.append((A, B, C, tuner.plan(r.alg, self.lib), r.tile))
        b = st["batch"]
        with torch.cuda.stream(s):
            b.set(probs, stream=s)
            if b.generation() != st["gen"]:
                st["graphs"] = {}
                st["gen"] = b.generation()
            if self.use_graphs:
                v = b.variant()
                g = st["graphs"].get(v)
                if g is None:
                    # first launch of this variant on this lane: run it
                    # eagerly, then capture the launch for later replays
                    b.run(stream=s)
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=s):
                        b.run(stream=s)
                    st["graphs"][v] = g
                    self.captures += 1
                else:
                    g.replay()
                    self.replays += 1
            else:
                b.run(stream=s)
            st["event"].record(s)

    def finished(self, ln):
        return ln.state["event"].query()

    def status(self, ln):
        """Error flags of the lane's completed launch (the batch copies them
        to host-mapped memory at the end of every run, so no sync)."""
        return ln.state["batch"].last_status()

    def collect(self, r):
        if self.keep_results:
            r.result = r.dev[2].cpu().numpy()

    def free(self, r):
        r.dev = None

    def idle(self, s):
        time.sleep(s)

    def warmup(self, lanes) -> None:
        """Capture each lane's graphs for the kernel variants the tuner can
        need (capture synchronises the device, so it is done before serving)."""
        torch, rt = self.torch, self.rt
        for ln in lanes:
            combos = [[("gemm", 0)], [("222L1", 0)], [("333L1", 0)]] if ln.dtype == "f64" else \
                [[("gemm", 0)], [("gemm", 1)], [("gemm", 0), ("gemm", 1)]]
            for combo in combos:
                reqs = []
                for j, (alg, tile) in enumerate(combo):
                    r = Request(-1 - j, 64, 64, 64, ln.dtype, 0.0)
                    t = self.tdtype(ln.dtype)
                    ev = torch.cuda.Event()
                    ops = tuple(rt.colmajor(64, 64, dtype=t).zero_() for _ in range(3))
                    ev.record()
                    r.dev = ops + (ev,)
                    r.alg, r.tile = alg, tile
                    reqs.append(r)
                self.launch(ln, reqs)
                torch.cuda.synchronize()
        for tuner in self.tuners.values():
            for a in tuner.algs:
                tuner.plan(a, self.lib)
