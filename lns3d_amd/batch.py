"""Batches of uneven implicit-step instances: many (grid, beta, Mach, Re)
problems stepped together with one graph launch.

Each instance is a device-resident context (coefficients, factors, state in
HBM).  Before the contexts are built, the line-solve kernel of every line
shape in the batch (xi lines: ny lines of nx rows; eta lines: nx lines of ny
rows) is chosen by the autotuner (lns3d_tune_lines: every variant timed on
synthetic lines of that shape).  The steps of all instances are then captured
into one graph, so one launch runs nsteps x count steps: with
MODE_BATCHED every stage of a step (residual, line solves grouped by solve
kernel, corrections, update) is one kernel launch over all instances; with
the other modes each instance runs on its own branch.
"""
from __future__ import annotations

import math
import time

import numpy as np

from .config import Params, make_grid, make_mean, make_disturbance
from .native import Graph, Step, tune_lines, set_line_variant, MODE_FP64, MODE_FP32, MODE_REFINED, MODE_BATCHED


def draw_instances(count: int = 64, seed: int = 0, nx_range=(65, 513), ny_range=(33, 257)):
    """instance parameters: nx, ny log-uniform on the ranges, beta U[0,1],
    Mach U[0.3,0.8], Re U[500,3000]; the tanh boundary-layer baseflow"""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx = int(round(math.exp(rng.uniform(math.log(nx_range[0]), math.log(nx_range[1])))))
        ny = int(round(math.exp(rng.uniform(math.log(ny_range[0]), math.log(ny_range[1])))))
        out.append(Params(nx=nx, ny=ny, kz=float(rng.uniform(0.0, 1.0)),
                          Ma=float(rng.uniform(0.3, 0.8)), Re=float(rng.uniform(500.0, 3000.0))))
    return out


def tune_shapes(params_list, precisions=(0, 1)):
    """autotune the solve kernel of every line shape of the batch; returns
    {(n, lines, precision): (variant, times_ms)}"""
    done = {}
    for p in params_list:
        for n, lines in ((p.nx, p.ny), (p.ny, p.nx)):
            for prec in precisions:
                key = (n, lines, prec)
                if key not in done:
                    done[key] = tune_lines(n, lines, prec)
    return done


# line-solve variants (lns3d_tune_lines numbering) the batch-level tuner
# tries for every line shape of a staged batch: one chain per line with
# three or four lines per workgroup.  A shape alone is solved fastest by
# the variant that shortens the chain (twisted lines); a staged launch over
# thousands of lines is bound by throughput, where more lines per workgroup
# share the issue of one wave.
BATCH_CANDIDATES = (("plain3", 4), ("plain4", 1))


def pin_variants(params_list, variant, tuned=None, precisions=(0, 1)):
    """use one solve variant for every line shape of the batch; None: the
    per-shape choice again (the records of `tuned`, a tune_shapes result,
    else the built-in rule)"""
    for p in params_list:
        for n, lines in ((p.nx, p.ny), (p.ny, p.nx)):
            for prec in precisions:
                v = variant
                if v is None:
                    v = tuned[(n, lines, prec)][0] if tuned and (n, lines, prec) in tuned else -1
                set_line_variant(n, lines, prec, v)


def tune_batch(params_list, precision=1, nsteps=2, corrections=1, seed=0, tuned=None):
    """the batch-level tuner of a staged batch: times a short staged graph
    with the per-shape variants (`tuned`, a tune_shapes result) and with
    each candidate of BATCH_CANDIDATES for all lines; pins the fastest and
    returns (name, {name: seconds})"""
    import torch
    times = {}
    for name, variant in (("per-shape", None),) + BATCH_CANDIDATES:
        pin_variants(params_list, variant, tuned)
        b = Batch(params_list, precision=precision, seed=seed)
        g = b.capture(nsteps, MODE_BATCHED, corrections)
        g.launch()
        g.synchronize()
        b.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.launch()
        g.synchronize()
        times[name] = time.perf_counter() - t0
        g.close()
        b.close()
    best = min(times, key=times.get)
    pin_variants(params_list, dict(BATCH_CANDIDATES).get(best), tuned)
    return best, times


class Batch:
    """count instances with their own contexts and states; `capture` builds
    the graph of nsteps steps of all of them"""

    def __init__(self, params_list, precision: int = 1, seed: int = 0):
        import torch
        self.params = list(params_list)
        self.precision = precision
        self.steps, self.v, self.dv, self.v0 = [], [], [], []
        for k, p in enumerate(self.params):
            x, y, met = make_grid(p)
            vm = make_mean(p, y)
            st = Step(p, precision=precision)
            st.setup(met, vm)
            st.factor()
            v0 = make_disturbance(p, x, seed=seed + k)
            v = torch.from_numpy(np.ascontiguousarray(v0.reshape(-1, order="F"))).cuda()
            st.itrbc3d(v)
            self.steps.append(st)
            self.v.append(v)
            self.dv.append(torch.zeros_like(v))
        torch.cuda.synchronize()
        self.v0 = [v.clone() for v in self.v]

    @property
    def points(self) -> int:
        return sum(p.nx * p.ny for p in self.params)

    def reset(self):
        for v, v0 in zip(self.v, self.v0):
            v.copy_(v0)

    def capture(self, nsteps: int, mode: int = MODE_REFINED, corrections: int = 1) -> Graph:
        return Graph(self.steps, self.v, self.dv, nsteps=nsteps, mode=mode, corrections=corrections)

    def run_eager(self, k: int, nsteps: int, mode: int = MODE_REFINED, corrections: int = 1):
        """the same steps for instance k, launched one by one (the check of
        the graph's result)"""
        st, v, dv = self.steps[k], self.v[k], self.dv[k]
        for _ in range(nsteps):
            if mode == MODE_FP64:
                st.step(v, dv)
            elif mode == MODE_FP32:
                st.step_fp32(v, dv)
            else:
                st.step_refined_fixed(v, dv, corrections)

    def close(self):
        for st in self.steps:
            st.close()
        self.steps = []


def run_batch(count=64, nsteps=20, seed=0, precision=1, mode=MODE_BATCHED, corrections=1,
              repeats=3, tune=True):
    """the batch benchmark: tune, build, capture, replay; returns a dict of
    timings and throughput (instance-steps/s and grid-point-steps/s)"""
    import torch
    params = draw_instances(count, seed)
    t0 = time.perf_counter()
    tuned = tune_shapes(params, (0, 1) if precision else (0,)) if tune else {}
    choice, choice_times = None, {}
    if tune and mode == MODE_BATCHED:
        choice, choice_times = tune_batch(params, precision, corrections=corrections, seed=seed, tuned=tuned)
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = Batch(params, precision=precision, seed=seed)
    t_build = time.perf_counter() - t0
    g = b.capture(nsteps, mode, corrections)
    g.launch()
    g.synchronize()
    times = []
    for _ in range(repeats):
        b.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.launch()
        g.synchronize()
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    out = dict(instances=count, steps_per_instance=nsteps, points=b.points,
               seconds=t, instance_steps_per_s=count * nsteps / t,
               point_steps_per_s=b.points * nsteps / t, tune_seconds=t_tune,
               build_seconds=t_build,
               variants={f"{k[0]}x{k[1]}p{k[2]}": v[0] for k, v in list(tuned.items())[:8]})
    if choice is not None:
        out["batch_variant"] = choice
        out["batch_variant_seconds"] = {k: round(v, 5) for k, v in choice_times.items()}
        pin_variants(params, None, tuned)
    g.close()
    b.close()
    return out
