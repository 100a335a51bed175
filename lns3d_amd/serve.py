"""A serving loop for linear-stability requests with continuous batching.

A request is (grid nx x ny, beta, omega, Mach, Re, steps, seed): form the
coefficients and factors for it on the device, then march `steps` implicit
steps of the linearized equations from its seeded initial disturbance.  The
default step keeps complex-float factors with one complex-double defect
correction per step (relative residual of the AF system about 1e-13 after
it on these problems).

Continuous batching: requests are admitted as they arrive while device
memory (a budget in bytes, checked with lns3d_ctx_bytes) allows, each onto
its own HIP stream, and all admitted requests progress on the device
together; a request that arrives while others run starts at once if memory
allows.  Contexts are pooled by (nx, ny, precision) and reused for later
requests of the same grid (new physics via lns3d_ctx_set_params), so no
allocation or free sits on the path of a request once the pool is warm.  A
request's steps are a graph of `chunk` steps captured once per pooled
context and launched steps/chunk times on the request's stream.

Results do not depend on what else runs: every request's arithmetic is its
own (own buffers, deterministic kernels), which the GPU test checks bit for
bit.
"""
from __future__ import annotations

import dataclasses
import time
from typing import List, Optional

import numpy as np

from .config import Params, disturbance_envelope, make_disturbance_device, make_grid, make_mean
from .native import Graph, Step, ctx_bytes, MODE_FP64, MODE_REFINED

GRIDS = ((129, 65), (257, 129), (513, 257), (1025, 257))
# parameters that reach kernels by value rather than through the device
# coefficients (the physics: beta, omega, Mach, Re, ... does not)
STRUCTURAL = ("nx", "ny", "optx", "opty", "xper", "yper", "wall", "wallt", "top", "left", "right",
              "linear", "mattyp", "loctime", "alfa", "periodic_exact")


@dataclasses.dataclass
class Request:
    rid: int
    nx: int
    ny: int
    beta: float
    omega: float
    Ma: float
    Re: float
    steps: int
    seed: int
    arrival: float = 0.0          # seconds after the start of the run

    def params(self) -> Params:
        return Params(nx=self.nx, ny=self.ny, kz=self.beta, omega=self.omega, Ma=self.Ma, Re=self.Re)


def draw_requests(count: int, rate: Optional[float], seed: int = 0, steps=(20, 100), chunk=10,
                  grids=GRIDS, grid_weights=(0.4, 0.3, 0.2, 0.1)) -> List[Request]:
    """seeded requests: beta U[0,1], omega U[0.01,0.2], Mach U[0.3,0.8],
    Re U[500,3000], grid from `grids`, steps a multiple of `chunk` in the
    range; Poisson arrivals at `rate` per second (None: all at t = 0)"""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    w = np.asarray(grid_weights, float)
    w = w / w.sum()
    for k in range(count):
        nx, ny = grids[int(rng.choice(len(grids), p=w))]
        nst = int(rng.integers(steps[0] // chunk, steps[1] // chunk + 1)) * chunk
        if rate:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(rid=k, nx=nx, ny=ny, beta=float(rng.uniform(0, 1)),
                           omega=float(rng.uniform(0.01, 0.2)), Ma=float(rng.uniform(0.3, 0.8)),
                           Re=float(rng.uniform(500, 3000)), steps=nst,
                           seed=int(rng.integers(1 << 31)), arrival=t if rate else 0.0))
    return out


class _Slot:
    """a pooled context with its stream, state buffers and chunk graph"""

    def __init__(self, p: Params, precision: int, mode: int, chunk: int, corrections: int):
        import torch
        self.key = (p.nx, p.ny, precision)
        self.stream = torch.cuda.Stream()
        self.st = Step(p, precision=precision)
        self.st.set_stream(self.stream)
        P = p.nx * p.ny
        self.v = torch.zeros(5 * P, dtype=torch.complex128, device="cuda")
        self.dv = torch.zeros_like(self.v)
        self.bytes = ctx_bytes(p, precision) + 2 * 80 * P
        self.mode, self.chunk, self.corrections = mode, chunk, corrections
        self.graph = None
        self.graph_key = None
        self.req = None
        self.done_event = None
        self.admitted = 0.0

    def start(self, req: Request, grid_cache):
        import torch
        p = req.params()
        x, y, met, envelope = grid_cache(p)
        vm = make_mean(p, y)
        self.st.set_params(p)
        self.st.setup(met, vm)
        self.st.factor()
        with torch.cuda.stream(self.stream):
            self.v.copy_(make_disturbance_device(envelope, req.seed))
        self.st.itrbc3d(self.v)
        # kernel arguments passed by value (grid, scheme and boundary
        # selectors) are frozen in the graph: capture again when they change
        key = tuple(getattr(p, f) for f in STRUCTURAL)
        if self.graph is None or key != self.graph_key:
            if self.graph is not None:
                self.graph.close()
            self.graph = Graph([self.st], [self.v], [self.dv], nsteps=self.chunk, mode=self.mode,
                               corrections=self.corrections)
            self.graph_key = key
        for _ in range(req.steps // self.chunk):
            self.graph.launch(self.stream)
        self.done_event = torch.cuda.Event()
        self.done_event.record(self.stream)
        self.req = req

    def finished(self) -> bool:
        return self.req is not None and self.done_event.query()

    def close(self):
        if self.graph is not None:
            self.graph.close()
        self.st.close()


class Server:
    """budget_bytes: device memory the pooled contexts may hold"""

    def __init__(self, budget_bytes: int, precision: int = 1, mode: int = MODE_REFINED,
                 chunk: int = 10, corrections: int = 1, keep_states: bool = False):
        self.budget = budget_bytes
        self.precision, self.mode, self.chunk, self.corrections = precision, mode, chunk, corrections
        self.keep_states = keep_states
        self.pool: List[_Slot] = []
        self._grids = {}

    def _grid(self, p):
        """grid, metrics and the disturbance envelope on the device, per grid"""
        import torch
        key = (p.nx, p.ny)
        if key not in self._grids:
            x, y, met = make_grid(p)
            env = disturbance_envelope(x).reshape(-1, order="F")
            self._grids[key] = (x, y, met, torch.from_numpy(env.astype(np.complex128)).cuda())
        return self._grids[key]

    def _held(self):
        return sum(s.bytes for s in self.pool)

    def _slot_for(self, req: Request) -> Optional[_Slot]:
        key = (req.nx, req.ny, self.precision)
        for s in self.pool:
            if s.req is None and s.key == key:
                return s
        need = ctx_bytes(req.params(), self.precision) + 2 * 80 * req.nx * req.ny
        # make room by dropping idle contexts of other grids
        while self._held() + need > self.budget:
            idle = [s for s in self.pool if s.req is None]
            if not idle:
                return None
            victim = max(idle, key=lambda s: s.bytes)
            victim.close()
            self.pool.remove(victim)
        s = _Slot(req.params(), self.precision, self.mode, self.chunk, self.corrections)
        self.pool.append(s)
        return s

    def run(self, requests: List[Request], poll_s: float = 2e-4):
        """serve the requests in real time (arrival offsets from now); returns
        one record per request"""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        out = {}
        t0 = time.perf_counter()
        active: List[_Slot] = []
        while pending or active:
            now = time.perf_counter() - t0
            # retire finished requests
            for s in list(active):
                if s.finished():
                    r = s.req
                    rec = dict(rid=r.rid, nx=r.nx, ny=r.ny, steps=r.steps, arrival=r.arrival,
                               admitted=s.admitted, done=now, latency=now - r.arrival,
                               vmax=float(s.v.abs().max()))
                    if self.keep_states:
                        rec["state"] = s.v.cpu().numpy()
                    out[r.rid] = rec
                    s.req = None
                    active.remove(s)
            # admit arrived requests in order while memory allows
            while pending and pending[0].arrival <= now:
                s = self._slot_for(pending[0])
                if s is None:
                    break
                r = pending.pop(0)
                s.admitted = time.perf_counter() - t0
                s.start(r, self._grid)
                active.append(s)
            if not active and pending:
                time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
            else:
                time.sleep(poll_s)
        return [out[k] for k in sorted(out)]

    def close(self):
        for s in self.pool:
            s.close()
        self.pool = []


def summarize(records, t_total=None):
    lat = np.array([r["latency"] for r in records])
    span = t_total if t_total is not None else (max(r["done"] for r in records) -
                                                min(r["arrival"] for r in records))
    return dict(requests=len(records), solves_per_s=len(records) / span,
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3))


def serial_fp64(requests: List[Request]):
    """the same requests one at a time with complex-double factors and steps
    (no overlap): per-request service times.  As in the server, one context
    per grid is created beforehand and reused, and the disturbance is drawn
    on the device."""
    import torch
    times = []
    ctx = {}
    srv = Server(0)                       # only for its per-grid cache
    for r in requests:
        key = (r.nx, r.ny)
        if key not in ctx:
            ctx[key] = Step(r.params(), precision=0)
            srv._grid(r.params())
    torch.cuda.synchronize()
    for r in requests:
        p = r.params()
        x, y, met, envelope = srv._grid(p)
        st = ctx[(r.nx, r.ny)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st.set_params(p)
        st.setup(met, make_mean(p, y))
        st.factor()
        v = make_disturbance_device(envelope, r.seed)
        st.itrbc3d(v)
        for _ in range(r.steps):
            st.step(v)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    for st in ctx.values():
        st.close()
    t = np.array(times)
    return dict(requests=len(times), solves_per_s=len(times) / t.sum(),
                p50_ms=float(np.percentile(t, 50) * 1e3), p99_ms=float(np.percentile(t, 99) * 1e3))


def sweep(count=48, seed=0, budget_gb=64.0, precision=1, mode=MODE_REFINED, load=0.8):
    """the serving benchmark: peak throughput with every request waiting at
    t = 0 (closed loop), then Poisson arrivals at `load` x peak; returns both
    summaries and the serial complex-double reference"""
    budget = int(budget_gb * 2**30)
    srv = Server(budget, precision=precision, mode=mode)
    # warm the pool and the graphs, then the closed-loop peak
    srv.run(draw_requests(min(count, 8), None, seed + 1))
    reqs = draw_requests(count, None, seed)
    t0 = time.perf_counter()
    recs = srv.run(reqs)
    peak = summarize(recs, time.perf_counter() - t0)
    rate = load * peak["solves_per_s"]
    reqs = draw_requests(count, rate, seed + 2)
    t0 = time.perf_counter()
    recs = srv.run(reqs)
    open_loop = summarize(recs, time.perf_counter() - t0)
    srv.close()
    return dict(peak=peak, rate=rate, open_loop=open_loop)
