"""A serving loop for receptivity requests with continuous batching.

A request is (grid nx x ny, beta, omega, Mach, Re, seed): a forced
receptivity solve in the frequency domain.  An acoustic wave of frequency
omega enters through the inflow sponge (ispg = 2: the outflow sponge spg and
the inflow sponge spg2 toward the incoming wave, src/lrhs3d.f90:222-282,
profiles as src/sponge.f90:20 forms them), and the linearized equations are
marched in pseudo-time with local time steps (loctime = 1,
src/temporal.f90:145-160, CFL 20 per 128 cells along x) until the residual norm the reference prints per
step (src/resstat3d.f90:61-89, totres) has fallen to `tol` (1e-9) times its
value at the first step, or `max_steps` have been taken.  The reference has
no stopping test; this one is the ratio it reports as jtotrs.

Such a march converges only when every mode of the operator decays in
pseudo-time.  The tanh boundary layer carries growing modes at Re of order
1000 (the residual then stalls near 1e-3 and rises), so requests draw Re from
[50, 150], where the march reaches 1e-9.

The default step keeps complex-float factors and forms the complex-double
residual every step (the pseudo-time iteration refines the answer: its fixed
point is the complex-double steady state).  Each request's residual norm is
recorded on the device at every step (lns3d_ctx_set_residual_output) and
copied back per chunk of steps without blocking the host.

Continuous batching: requests are admitted as they arrive while device
memory (a budget in bytes, checked with lns3d_ctx_bytes) allows, each onto
its own HIP stream, and all admitted requests progress on the device
together.  Contexts are pooled by (nx, ny, precision) and reused for later
requests of the same grid (new physics via lns3d_ctx_set_params).  A
request's steps are a graph of `chunk` steps captured once per pooled
context; two chunks stay queued per request so that the device never waits
for the host's convergence test.  A request whose set-up fails (a refused
parameter, a singular factor) or whose residual stops being finite is
recorded as failed and its slot freed; the others continue.

Batched per grid (Server(batched=True)): the requests of one grid share a
staged graph, one launch per stage over all of them (lns3d_graph_create
mode 3, residual norms recorded per request), captured again when a request
joins or leaves (about 1-2 ms of host time); the grids' cohorts run on
streams of their own.  The chunk discipline is the same, so a request takes
the same steps and reaches the same state either way.  Measured on the
serving mixes here, the two choices serve at the same rate: with a graph
per request, four hardware queues already overlap the small requests'
latency-bound launches, and both are then held by the line solves' two
workgroups per CU.  The staged graph makes one launch per stage for all of
a grid's requests instead of one per request (2.5x faster under a kernel
tracer, which adds a cost per launch).

Results do not depend on what else runs: every request's arithmetic is its
own (own buffers, deterministic kernels and reductions), which the GPU test
checks bit for bit.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time
from typing import List, Optional

import numpy as np

from .config import Params, disturbance_envelope, make_disturbance_device, make_grid, make_mean
from .native import (Graph, Step, ctx_bytes, set_line_variant, LNS3DError, MODE_FP64, MODE_FP32, MODE_REFINED,
                     MODE_BATCHED)

GRIDS = ((129, 65), (257, 129), (513, 257), (1025, 257))
# Pseudo-time CFL of a request (local time steps): 20 per 128 cells along x.
# Steps to 1e-9 measured for three drawn requests per grid (sum over the
# three): 129x65 -- CFL 10: 6999, 20: 5322, 40: 9081; 257x129 -- 20: 21433,
# 40: 12886, 80: 14034; 513x257 -- 40: 50995, 80: 30285, 160: 25464;
# 1025x257 -- 80: 33284, 160: 21923, 320: 31435.
def request_cfl(nx: int) -> float:
    return 20.0 * (nx - 1) / 128.0


# parameters that reach kernels by value rather than through the device
# coefficients (the physics: beta, omega, Mach, Re, ... does not)
STRUCTURAL = ("nx", "ny", "optx", "opty", "xper", "yper", "wall", "wallt", "top", "left", "right",
              "linear", "mattyp", "loctime", "alfa", "periodic_exact")
MODES = {"fp32_outer": (1, MODE_FP32), "refined": (1, MODE_REFINED), "fp64": (0, MODE_FP64)}


@dataclasses.dataclass
class Request:
    rid: int
    nx: int
    ny: int
    beta: float
    omega: float
    Ma: float
    Re: float
    seed: int
    arrival: float = 0.0          # seconds after the start of the run
    tol: float = 1e-9
    max_steps: int = 30000

    def params(self) -> Params:
        return Params(nx=self.nx, ny=self.ny, kz=self.beta, omega=self.omega, Ma=self.Ma, Re=self.Re,
                      loctime=1, cflmax=request_cfl(self.nx))


def sponge_profiles(x):
    """ispg = 2 profiles on the physical x (nx, ny): spg = ((x - 70)/30)^3
    for x >= 70 (outflow), spg2 = ((x - 20)/(0 - 20))^3 for x <= 20 (inflow,
    toward the incoming wave), the forms of src/sponge.f90:20-120 with As =
    As2 = 1, Ns = Ns2 = 3"""
    spg = np.where(x >= 70.0, ((x - 70.0) / 30.0) ** 3, 0.0)
    spg2 = np.where(x <= 20.0, ((x - 20.0) / (0.0 - 20.0)) ** 3, 0.0)
    return np.asfortranarray(spg), np.asfortranarray(spg2)


def draw_requests(count: int, rate: Optional[float], seed: int = 0, grids=GRIDS,
                  grid_weights=(0.4, 0.3, 0.2, 0.1), Re=(50.0, 150.0), tol=1e-9,
                  max_steps=30000) -> List[Request]:
    """seeded requests: beta U[0,1], omega U[0.01,0.2], Mach U[0.3,0.8],
    Re U[Re], grid from `grids`; Poisson arrivals at `rate` per second (None:
    all at t = 0)"""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    w = np.asarray(grid_weights, float)
    w = w / w.sum()
    for k in range(count):
        nx, ny = grids[int(rng.choice(len(grids), p=w))]
        if rate:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(rid=k, nx=nx, ny=ny, beta=float(rng.uniform(0, 1)),
                           omega=float(rng.uniform(0.01, 0.2)), Ma=float(rng.uniform(0.3, 0.8)),
                           Re=float(rng.uniform(*Re)), seed=int(rng.integers(1 << 31)),
                           arrival=t if rate else 0.0, tol=tol, max_steps=max_steps))
    return out


def with_arrivals(requests: List[Request], rate: float, seed: int = 0) -> List[Request]:
    """the same requests with Poisson arrival times at `rate` per second,
    drawn from their own seeded stream (so a request set keeps its
    parameters whatever the arrival process)"""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for r in requests:
        t += float(rng.exponential(1.0 / rate))
        out.append(dataclasses.replace(r, arrival=t))
    return out


def fifo_latencies(service_s, arrivals):
    """latencies of one-at-a-time service in arrival order: request k, with
    service time service_s[k], arrives at arrivals[k] and starts when it has
    arrived and the one before it is done"""
    order = np.argsort(arrivals, kind="stable")
    free, lat = 0.0, np.zeros(len(order))
    for k in order:
        start = max(arrivals[k], free)
        free = start + service_s[k]
        lat[k] = free - arrivals[k]
    return lat


class _Slot:
    """a pooled context with its stream, state buffers, per-step residual
    norms and chunk graph"""

    def __init__(self, p: Params, precision: int, mode: int, chunk: int, corrections: int):
        import torch
        self.key = (p.nx, p.ny, precision)
        self.stream = torch.cuda.Stream()
        self.st = Step(p, precision=precision)
        self.st.set_stream(self.stream)
        P = p.nx * p.ny
        self.v = torch.zeros(5 * P, dtype=torch.complex128, device="cuda")
        self.dv = torch.zeros_like(self.v)
        self.norms = torch.zeros(chunk, dtype=torch.float64, device="cuda")
        self.hnorms = [torch.zeros(chunk, dtype=torch.float64).pin_memory() for _ in range(2)]
        self.st.set_residual_output(self.norms, chunk)
        self.bytes = ctx_bytes(p, precision) + 2 * 80 * P + 8 * (14 * P + p.nx)
        self.mode, self.chunk, self.corrections = mode, chunk, corrections
        self.graph = None
        self.graph_key = None
        self.req = None
        self.admitted = 0.0

    def start(self, req: Request, grid_cache, own_graph: bool = True):
        """set-up and, with its own graph, the first two chunks (the caller
        handles failures); without, a cohort launches its chunks"""
        import torch
        p = req.params()
        x, y, met, envelope = grid_cache(p)
        vm = make_mean(p, y)
        spg, spg2 = sponge_profiles(x)
        self.req = req
        self.st.set_params(p)
        self.st.setup(met, vm)
        self.st.set_sponge(2, spg, spg2, x, None, 0.0)
        self.st.factor()
        with torch.cuda.stream(self.stream):
            self.v.copy_(make_disturbance_device(envelope, req.seed))
        self.st.itrbc3d(self.v)
        self.r0 = None
        self.steps = 0            # steps whose norms were examined
        self.launched = 0         # chunks launched
        self.converged_at = None
        self.final_ratio = None
        self.failed = None
        self.inflight = collections.deque()
        self.queued = 0           # chunks of a cohort's launches not yet examined
        self.done = False
        if not own_graph:
            return
        # kernel arguments passed by value (grid, scheme and boundary
        # selectors) are frozen in the graph: capture again when they change
        key = tuple(getattr(p, f) for f in STRUCTURAL)
        if self.graph is None or key != self.graph_key:
            if self.graph is not None:
                self.graph.close()
                self.graph = None
            self.st.set_residual_output(self.norms, self.chunk)
            self.graph = Graph([self.st], [self.v], [self.dv], nsteps=self.chunk, mode=self.mode,
                               corrections=self.corrections)
            self.graph_key = key
        self._launch()
        self._launch()

    def _launch(self):
        import torch
        self.graph.launch(self.stream)
        buf = self.hnorms[self.launched % 2]
        with torch.cuda.stream(self.stream):
            buf.copy_(self.norms, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(self.stream)
        self.inflight.append((ev, buf))
        self.launched += 1

    def examine(self, nrm) -> bool:
        """the convergence test over one chunk's norms; True once done"""
        if self.r0 is None:
            self.r0 = float(nrm[0])
        for k in range(self.chunk):
            ratio = float(nrm[k]) / self.r0 if self.r0 > 0 else 0.0
            if not math.isfinite(ratio):
                self.failed = "non-finite residual"
                return True
            if ratio <= self.req.tol:
                self.converged_at = self.steps + k + 1
                self.final_ratio = ratio
                return True
        self.steps += self.chunk
        self.final_ratio = float(nrm[-1]) / self.r0 if self.r0 > 0 else 0.0
        return self.steps >= self.req.max_steps

    def wants_chunk(self) -> bool:
        return not self.done and self.launched * self.chunk < self.req.max_steps

    def poll(self) -> bool:
        """examine the chunks that have finished; True once the request is done"""
        while self.inflight and self.inflight[0][0].query():
            _, buf = self.inflight.popleft()
            if self.examine(buf.numpy()):
                return True
            if self.wants_chunk():
                self._launch()
        return False

    def drained(self) -> bool:
        """every chunk still queued for a finished request has run"""
        while self.inflight and self.inflight[0][0].query():
            self.inflight.popleft()
        return not self.inflight and self.queued == 0

    def drain(self):
        self.stream.synchronize()
        self.inflight.clear()

    def close(self):
        if self.graph is not None:
            self.graph.close()
        self.st.close()


class _Cohort:
    """the requests of one grid and precision stepped together: a staged
    graph (one launch per stage over every member, lns3d_graph_create mode
    3) of `chunk` steps, captured again whenever the members change.

    Two launches stay queued.  A member takes part in every launch from the
    first after it joins until the one queued when its norms showed it done,
    which is the chunk count its own graph would give: its chunk k + 1 is
    queued before chunk k is examined, and no later chunk follows.  Replaced
    graphs are destroyed when the serving run ends (freeing their tables
    waits for the whole device)."""

    def __init__(self, key, chunk: int, corrections: int):
        import torch
        self.key = key
        self.chunk, self.corrections = chunk, corrections
        self.stream = torch.cuda.Stream()
        self.members: List[_Slot] = []      # in the next launch
        self.graph = None
        self.graph_members = ()
        self.old = []                       # replaced graphs
        self.inflight = collections.deque() # (event, [(slot, host buffer)])
        self.rebuilds = 0
        self.rebuild_s = 0.0                # host time capturing graphs
        self.failed: List[_Slot] = []       # members whose graph could not be made

    def join(self, s: _Slot):
        self.members.append(s)
        if not self.inflight:
            self._fill()

    def _launch(self):
        import torch
        members = tuple(self.members)
        if members != self.graph_members:
            t0 = time.perf_counter()
            self.retire_graph()
            for m in members:
                m.st.set_residual_output(m.norms, self.chunk)
            self.graph = Graph([m.st for m in members], [m.v for m in members], [m.dv for m in members],
                               nsteps=self.chunk, mode=MODE_BATCHED, corrections=self.corrections)
            self.graph_members = members
            self.rebuilds += 1
            self.rebuild_s += time.perf_counter() - t0
        for m in members:
            self.stream.wait_stream(m.stream)    # a new member's set-up
        self.graph.launch(self.stream)
        entries = []
        with torch.cuda.stream(self.stream):
            for m in members:
                buf = m.hnorms[m.launched % 2]
                buf.copy_(m.norms, non_blocking=True)
                m.launched += 1
                m.queued += 1
                entries.append((m, buf))
        ev = torch.cuda.Event()
        ev.record(self.stream)
        self.inflight.append((ev, entries))

    def _fill(self):
        self.members = [m for m in self.members if m.wants_chunk()]
        while self.members and len(self.inflight) < 2:
            try:
                self._launch()
            except LNS3DError as e:
                # no graph for these members: they fail, the other grids go on
                for m in self.members:
                    m.failed, m.done = f"batched graph: {e}", True
                    self.failed.append(m)
                self.members = []

    def poll(self) -> List[_Slot]:
        """examine the launches that have finished; returns the members
        found done"""
        out, self.failed = self.failed, []
        while self.inflight and self.inflight[0][0].query():
            _, entries = self.inflight.popleft()
            for m, buf in entries:
                m.queued -= 1
                if not m.done and m.examine(buf.numpy()):
                    m.done = True
                    out.append(m)
            self._fill()
        return out

    def retire_graph(self):
        """the current graph is not launched again"""
        if self.graph is not None:
            self.old.append(self.graph)
        self.graph, self.graph_members = None, ()

    def release(self):
        """destroy replaced graphs once nothing of the cohort runs (their
        tables are freed, which waits for the device)"""
        self.stream.synchronize()
        for g in self.old:
            g.close()
        self.old = []

    def close(self):
        self.retire_graph()
        self.release()


class Server:
    """budget_bytes: device memory the pooled contexts may hold.  batched:
    the requests of one grid share a staged graph (one launch per stage over
    all of them); otherwise each request has its own graph and stream.  The
    complex-double mode is never batched."""

    def __init__(self, budget_bytes: int, mode: str = "fp32_outer", chunk: int = 50,
                 corrections: int = 1, keep_states: bool = False, batched: bool = False,
                 small_line_variant: Optional[int] = None):
        self.budget = budget_bytes
        # small_line_variant: the line-solve variant (lns3d_tune_set) for
        # the contexts of grids of at most 257 rows in either direction, in
        # both precisions, pinned for the server's life and dropped at
        # close (e.g. 4, three lines per workgroup: more lines in flight
        # when many small requests run together, longer chains per line)
        self.small_line_variant = small_line_variant
        self._pinned = set()
        self.precision, self.gmode = MODES[mode]
        self.mode_name, self.chunk, self.corrections = mode, chunk, corrections
        self.keep_states = keep_states
        self.batched = batched and self.gmode != MODE_FP64
        self.pool: List[_Slot] = []
        self.cohorts = {}
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

    def _need(self, req):
        P = req.nx * req.ny
        return ctx_bytes(req.params(), self.precision) + 2 * 80 * P + 8 * (14 * P + req.nx)

    def _slot_for(self, req: Request) -> Optional[_Slot]:
        key = (req.nx, req.ny, self.precision)
        for s in self.pool:
            if s.req is None and s.key == key:
                return s
        need = self._need(req)
        # make room by dropping idle contexts of other grids
        while self._held() + need > self.budget:
            idle = [s for s in self.pool if s.req is None]
            if not idle:
                return None
            victim = max(idle, key=lambda s: s.bytes)
            self._drop(victim)
        if self.small_line_variant is not None and max(req.nx, req.ny) <= 257:
            for shape in ((req.nx, req.ny), (req.ny, req.nx)):
                for prec in (0, 1):
                    key3 = shape + (prec,)
                    if key3 not in self._pinned:
                        set_line_variant(shape[0], shape[1], prec, self.small_line_variant)
                        self._pinned.add(key3)
        s = _Slot(req.params(), self.precision, self.gmode, self.chunk, self.corrections)
        self.pool.append(s)
        return s

    def _drop(self, s: _Slot):
        """close a pooled context.  A cohort graph that holds it is never
        launched again (its members differ from the cohort's from now on) and
        is destroyed with the cohort's other replaced graphs."""
        c = self.cohorts.get(s.key)
        if c is not None and s in c.graph_members:
            c.retire_graph()
        s.close()
        self.pool.remove(s)

    def _retire(self, s: _Slot, now: float, out: dict):
        r = s.req
        rec = dict(rid=r.rid, nx=r.nx, ny=r.ny, arrival=r.arrival, admitted=s.admitted, done=now,
                   latency=now - r.arrival, steps=s.converged_at or s.steps,
                   converged=s.converged_at is not None, ratio=s.final_ratio, error=s.failed)
        if s.failed is None:
            rec["vmax"] = float(s.v.abs().max())
            if self.keep_states:
                rec["state"] = s.v.cpu().numpy()
        out[r.rid] = rec
        s.req = None

    def run(self, requests: List[Request], poll_s: float = 2e-4):
        """serve the requests in real time (arrival offsets from now); returns
        one record per request (error set for a request that failed)"""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        out = {}
        t0 = time.perf_counter()
        active: List[_Slot] = []
        retiring: List[_Slot] = []    # done; a queued chunk may still run on its stream
        while pending or active or retiring:
            now = time.perf_counter() - t0
            if self.batched:
                for c in self.cohorts.values():
                    for s in c.poll():
                        active.remove(s)
                        retiring.append(s)
            else:
                for s in list(active):
                    if s.poll():
                        active.remove(s)
                        retiring.append(s)
            for s in list(retiring):
                if s.drained():
                    self._retire(s, time.perf_counter() - t0, out)
                    retiring.remove(s)
            # admit arrived requests in order while memory allows
            while pending and pending[0].arrival <= now:
                try:
                    s = self._slot_for(pending[0])
                except LNS3DError as e:          # a context of this grid cannot be created
                    r = pending.pop(0)
                    out[r.rid] = dict(rid=r.rid, nx=r.nx, ny=r.ny, arrival=r.arrival, error=str(e),
                                      done=time.perf_counter() - t0, converged=False)
                    continue
                if s is None:
                    break
                r = pending.pop(0)
                s.admitted = time.perf_counter() - t0
                try:
                    s.start(r, self._grid, own_graph=not self.batched)
                except LNS3DError as e:
                    # this request alone fails: record it, drop its slot (its
                    # context may be half set up), keep serving the others
                    s.drain()
                    out[r.rid] = dict(rid=r.rid, nx=r.nx, ny=r.ny, arrival=r.arrival, error=str(e),
                                      done=time.perf_counter() - t0, converged=False)
                    s.req = None
                    self._drop(s)
                    continue
                if self.batched:
                    if s.key not in self.cohorts:
                        self.cohorts[s.key] = _Cohort(s.key, self.chunk,
                                                      0 if self.gmode == MODE_FP32 else self.corrections)
                    self.cohorts[s.key].join(s)
                active.append(s)
            if not active and not retiring and pending:
                time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
            else:
                time.sleep(poll_s)
        for c in self.cohorts.values():
            c.release()
        return [out[k] for k in sorted(out)]

    def close(self):
        for c in self.cohorts.values():
            c.close()
        self.cohorts = {}
        for s in self.pool:
            s.close()
        self.pool = []
        for n, lines, prec in self._pinned:
            set_line_variant(n, lines, prec, -1)
        self._pinned = set()


def summarize(records, t_total=None):
    ok = [r for r in records if r.get("error") is None]
    lat = np.array([r["latency"] for r in ok]) if ok else np.zeros(1)
    span = t_total if t_total is not None else (max(r["done"] for r in records) -
                                                min(r["arrival"] for r in records))
    return dict(requests=len(records), failed=len(records) - len(ok),
                converged=sum(1 for r in ok if r["converged"]), solves_per_s=len(ok) / span,
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_steps=float(np.mean([r["steps"] for r in ok])) if ok else 0.0)


def serial(requests: List[Request], mode: str = "fp64", chunk: int = 50):
    """the same requests one at a time (no overlap), with the same
    convergence test: per-request service times.  As in the server, one
    context per grid is created beforehand and reused, and the disturbance
    is drawn on the device."""
    srv = Server(1 << 62, mode=mode, chunk=chunk)
    times, recs = [], []
    try:
        warm = {}
        for r in requests:                       # one context and graph per grid
            warm.setdefault((r.nx, r.ny), dataclasses.replace(r, arrival=0.0, max_steps=chunk))
        srv.run(list(warm.values()))
        for r in requests:
            t0 = time.perf_counter()
            rec = srv.run([dataclasses.replace(r, arrival=0.0)], poll_s=5e-5)[0]
            times.append(time.perf_counter() - t0)
            recs.append(rec)
    finally:
        srv.close()
    t = np.array(times)
    ok = [r for r in recs if r.get("error") is None]
    return dict(requests=len(times), solves_per_s=len(times) / t.sum(),
                p50_ms=float(np.percentile(t, 50) * 1e3), p99_ms=float(np.percentile(t, 99) * 1e3),
                converged=sum(1 for r in ok if r["converged"]),
                mean_steps=float(np.mean([r["steps"] for r in ok])) if ok else 0.0,
                service_s=[float(x) for x in t]), recs


def peak(count=24, seed=0, budget_gb=64.0, mode="fp32_outer", chunk=50, batched=True, small_line_variant=None,
         grids=GRIDS, grid_weights=(0.4, 0.3, 0.2, 0.1)):
    """closed-loop peak throughput of the requests of `sweep` (after a
    warm-up run), with the server's batching and variant choices; adds the
    batched server's graph captures (count and host seconds)"""
    srv = Server(int(budget_gb * 2**30), mode=mode, chunk=chunk, batched=batched,
                 small_line_variant=small_line_variant)
    srv.run(draw_requests(min(count, 4), None, seed + 1, grids=grids, grid_weights=grid_weights, max_steps=200))
    t0 = time.perf_counter()
    recs = srv.run(draw_requests(count, None, seed, grids=grids, grid_weights=grid_weights))
    out = summarize(recs, time.perf_counter() - t0)
    out["graph_captures"] = sum(c.rebuilds for c in srv.cohorts.values())
    out["capture_s"] = sum(c.rebuild_s for c in srv.cohorts.values())
    srv.close()
    return out


def sweep(count=24, seed=0, budget_gb=64.0, mode="fp32_outer", load=0.8, chunk=50, batched=False):
    """the serving benchmark: peak throughput with every request waiting at
    t = 0 (closed loop), then the same requests arriving as a Poisson
    process at `load` x peak; returns both summaries and the arrival times
    (in request order)"""
    budget = int(budget_gb * 2**30)
    srv = Server(budget, mode=mode, chunk=chunk, batched=batched)
    # warm the pool and the graphs, then the closed-loop peak
    srv.run(draw_requests(min(count, 4), None, seed + 1, max_steps=200))
    reqs = draw_requests(count, None, seed)
    t0 = time.perf_counter()
    recs = srv.run(reqs)
    peak = summarize(recs, time.perf_counter() - t0)
    rate = load * peak["solves_per_s"]
    reqs = with_arrivals(reqs, rate, seed + 2)
    t0 = time.perf_counter()
    recs = srv.run(reqs)
    open_loop = summarize(recs, time.perf_counter() - t0)
    srv.close()
    return dict(peak=peak, rate=rate, open_loop=open_loop, arrivals=[r.arrival for r in reqs])
