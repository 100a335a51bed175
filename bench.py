"""Benchmark: implicit approximate-factorization steps per second on one MI355X.

Workload: one implicit step of LNS3D's ImpDrv3D (lrhs3D + rbe3D + rhsBC3D,
xi line solve, eta line solve, update, itrBC3D) on the 1025x257 grid of the
benchmark baseflow (parallel boundary layer, Ma 0.5, Re 1000, beta 0.3),
complex double, all arrays resident in HBM.  The LHS is formed and factored
once before timing (the reference factors once per run, impdrv3d.f90:248).

Prints one JSON line.  value = complex-double steps/s; the complex-float
factor + refinement path is reported beside it.  roofline: the dominant
kernel (the xi-line solve) measured with HIP events on its stream;
cpu_baseline: the reference's own Fortran step (oracle/_ref), or the C
restatement when that library is absent, timed on the host on a bounded
sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NX, NY, BETA = 1025, 257, 0.3
# algorithmic bytes per grid point of one line solve (factors 2000 + rhs in
# 80 + out 80; complex-float factors halve the factor bytes)
SOLVE_BYTES_PT = {0: 2160, 1: 1160}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-refine", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--no-sweep", action="store_true", help="skip the serving sweep leg")
    ap.add_argument("--requests", type=int, default=16, help="requests in the serving sweep")
    ap.add_argument("--kernels-only", action="store_true",
                    help="launch 5 xi solves, 5 eta solves, 5 steps and a calibration copy of "
                         "one solve's algorithmic bytes, then exit (for rocprofv3 --pmc runs)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_solve_xi_1025x257.json")


def measured_traffic():
    """HBM bytes per xi-solve launch from the committed rocprofv3 --pmc runs
    (FETCH_SIZE + WRITE_SIZE of k_solve_bc, each scaled by the counter's
    reading for a copy of known size); None when that file is absent"""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def kernels_only(st, P):
    import torch
    r = torch.randn(P * 5, dtype=torch.complex128, device="cuda")
    for which in ("x", "y"):
        pen = _dir_solver(st, which)
        for _ in range(5):
            pen(r)
    v = r.clone()
    dv = torch.zeros_like(v)
    st.itrbc3d(v)
    for _ in range(5):
        st.step(v, dv)
    # calibration: a plain device copy of one xi solve's algorithmic bytes
    nb = SOLVE_BYTES_PT[0] * P
    a = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"kernels_only": True, "copy_bytes": nb}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the multi-GPU decomposition is described, not built)"}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from lns3d_amd.native import Step, lib

    p = Params(nx=NX, ny=NY, kz=BETA)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v0 = make_disturbance(p, x, seed=0)
    P = NX * NY

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).cuda()

    st = Step(p, precision=0 if args.no_refine else 1)
    st.setup(met, vm)
    st.factor()
    if args.kernels_only:
        kernels_only(st, P)
        return
    v = dev(v0)
    st.itrbc3d(v)
    dv = torch.zeros_like(v)
    torch.cuda.synchronize()

    # ---- complex-double steps (value) ----
    for _ in range(args.warmup):
        st.step(v, dv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.step(v, dv)
    torch.cuda.synchronize()
    t_fp64 = time.perf_counter() - t0
    steps_s = args.steps / t_fp64

    # ---- dominant kernel: xi-line solve, HIP events on the default stream ----
    L = lib()
    r = torch.randn(P * 5, dtype=torch.complex128, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nrep = 20
    sol_ms = {}
    for name, fn in (("xi", "x"), ("eta", "y")):
        pen = _dir_solver(st, fn)
        pen(r)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(nrep):
            pen(r)
        e1.record()
        torch.cuda.synchronize()
        sol_ms[name] = e0.elapsed_time(e1) / nrep
    bytes_launch = SOLVE_BYTES_PT[0] * P
    achieved = bytes_launch / (sol_ms["xi"] * 1e-3) / 1e9

    out = {
        "metric": "implicit steps/s at 1025x257 (FP64 vs FP32+refine); sweep p50/p99 latency",
        "value": round(steps_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_fp64 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded disturbance on the analytic tanh-grid boundary-layer baseflow)",
        "config": {"workload": "LNS3D implicit AF step: lrhs3D + xi/eta block-penta line solves + "
                               "update, complex double, factors formed once",
                   "nx": NX, "ny": NY, "beta": BETA, "Ma": p.Ma, "Re": p.Re, "cfl": p.cflmax},
        "roofline": {"bound": "hbm", "kernel": "k_solve_tw<double,6,1,false,4,true> (xi lines: 257 lines x 1025 rows, twisted, column halves)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "bytes_per_launch": bytes_launch,
                     "avg_launch_ms": round(sol_ms["xi"], 4),
                     "eta_solve_ms": round(sol_ms["eta"], 4)},
    }

    # ---- complex-float factors with refinement to the complex-double step ----
    if not args.no_refine:
        vr = dev(v0)
        st.itrbc3d(vr)
        vref = vr.clone()
        dvr = torch.zeros_like(vr)
        dv64 = torch.zeros_like(vr)
        st.step(vref, dv64)
        its, res = st.step_refined(vr, dvr, tol=1e-10, max_iters=8)
        torch.cuda.synchronize()
        err = float((dvr - dv64).abs().max() / dv64.abs().max())
        vr = dev(v0)
        st.itrbc3d(vr)
        for _ in range(args.warmup):
            st.step_refined(vr, dvr, tol=1e-10, max_iters=8)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        iters = []
        nref = max(5, args.steps // 5)
        for _ in range(nref):
            it, _ = st.step_refined(vr, dvr, tol=1e-10, max_iters=8)
            iters.append(it)
        torch.cuda.synchronize()
        t_ref = time.perf_counter() - t0
        vf = dev(v0)
        st.itrbc3d(vf)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st.step_fp32(vf, dvr)
        torch.cuda.synchronize()
        t_f32 = time.perf_counter() - t0
        # one correction per step, no host synchronisation (the graph form)
        vx = dev(v0)
        st.itrbc3d(vx)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st.step_refined_fixed(vx, dvr, corrections=1)
        torch.cuda.synchronize()
        t_fix = time.perf_counter() - t0
        st.step_refined_fixed(vx, dvr, corrections=1, check_residual=True)
        res_fix = st.last_residual()
        out["fp32_refine"] = {"steps_s": round(nref / t_ref, 3), "iters_mean": float(np.mean(iters)),
                              "tol": 1e-10, "dv_rel_err_vs_fp64_step": err,
                              "first_step_iters": its, "first_step_resid": res,
                              "one_correction_steps_s": round(args.steps / t_fix, 3),
                              "one_correction_resid": res_fix,
                              "fp32_no_refine_steps_s": round(args.steps / t_f32, 3),
                              "ratio_to_fp64": round((args.steps / t_fix) / steps_s, 3)}
    st.close()

    # ---- uneven batch: 64 instances x 20 refined steps in one graph launch ----
    if not args.no_batch:
        from lns3d_amd.batch import run_batch
        b = run_batch(count=64, nsteps=20, seed=0)
        out["batch"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in b.items()}

    # ---- serving sweep: receptivity requests marched to 1e-9 of their first
    # residual; closed-loop peak, then Poisson arrivals at 0.8 x peak, with
    # complex-float factors and the complex-double residual each step (the
    # pseudo-time iteration as the refinement), beside the same requests
    # solved one at a time with complex-double factors and with complex-float
    # factors (time to solution of each path) ----
    if not args.no_sweep:
        from lns3d_amd.serve import sweep, serial, draw_requests
        nreq = args.requests
        ser64, _ = serial(draw_requests(nreq, None, 0), mode="fp64")
        ser32, _ = serial(draw_requests(nreq, None, 0), mode="fp32_outer")
        sw = sweep(count=nreq, seed=0, mode="fp32_outer")
        rnd = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()}
        out["sweep"] = {"requests": nreq,
                        "request": "receptivity (ispg 2 sponge forcing, loctime 1, CFL 80) to 1e-9 of "
                                   "the first residual; grids 129x65..1025x257, Re U[50,150]",
                        "mode": "FP32 factors, FP64 residual each step (outer refinement)",
                        "peak": rnd(sw["peak"]), "arrival_rate_s": round(sw["rate"], 3),
                        "poisson": rnd(sw["open_loop"]),
                        "serial_fp64": rnd(ser64), "serial_fp32_outer": rnd(ser32),
                        "time_to_solution_ratio_fp64_over_fp32":
                            round(ser32["solves_per_s"] / ser64["solves_per_s"], 3)}

    # ---- CPU baseline ----
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, met, vm, v0, args.cpu_steps)
    print(json.dumps(out), flush=True)


def _dir_solver(st, which):
    """in-place xi or eta solve of a device rhs through the C ABI"""
    import ctypes
    from lns3d_amd.native import lib, check, ptr
    L = lib()
    if not hasattr(L, "lns3d_solve_dir"):
        raise RuntimeError("lns3d_solve_dir missing from the library")
    L.lns3d_solve_dir.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.lns3d_solve_dir.restype = ctypes.c_int
    d = 1 if which == "x" else 2
    return lambda r: check(L.lns3d_solve_dir(st.h, d, ptr(r)), "lns3d_solve_dir")


class _quiet_stdout:
    """silence fd 1 (the reference's Fortran writes progress lines there)"""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *exc):
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)


def cpu_baseline(p, met, vm, v0, nsteps):
    with _quiet_stdout():
        out = _cpu_baseline(p, met, vm, v0, nsteps)
        from oracle.oracle import Reference
        if Reference.available():
            Reference().lib.ref_flush()
    return out


def _cpu_baseline(p, met, vm, v0, nsteps):
    """The reference's Fortran step (oracle/_ref/libref.so, OpenMP threads on
    the host) or, where that library was not built, the C restatement (one
    thread).  The LU is formed once outside the timed region, as in the
    reference; the sample is nsteps full steps at the benchmark size."""
    from oracle.oracle import Oracle, Reference
    from lns3d_amd.config import make_grid
    x, y, met2 = make_grid(p)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if Reference.available():
        R = Reference()
        R.setup(p, x, y, met2, vm)
        R.genmtrx()
        R.set_time_step()
        v = np.asfortranarray(v0.copy())
        R.itrbc3d(v)
        R.step(v, True)                        # forms and factors the LHS
        t0 = time.perf_counter()
        for _ in range(nsteps):
            R.step(v, False)
        dt = time.perf_counter() - t0
        return {"value": round(nsteps / dt, 4), "unit": "steps/s", "cores": cores,
                "kind": "reference",
                "sample": f"{nsteps} implicit steps at {p.nx}x{p.ny} after one factoring step "
                          "(flang -O2 -fopenmp build of the reference sources)"}
    O = Oracle()
    c = O.genmtrx(p, vm, met)
    dtl = O.set_time_step(p, vm, met)
    v = np.asfortranarray(v0.copy())
    O.itrbc3d(p, v)
    f = O.factor(p, c, dtl)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        O.step(p, c, dtl, f, v)
    dt = time.perf_counter() - t0
    return {"value": round(nsteps / dt, 4), "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"{nsteps} implicit steps at {p.nx}x{p.ny} (C restatement, one thread)"}


if __name__ == "__main__":
    main()
