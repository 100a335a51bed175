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

AFFINITY_AT_START = len(os.sched_getaffinity(0))

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
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-refine", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--no-sweep", action="store_true", help="skip the serving sweep leg")
    ap.add_argument("--requests", type=int, default=16, help="requests in the serving sweep")
    ap.add_argument("--no-dropin", action="store_true", help="skip the host-array drop-in timings")
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

    # ---- the drop-in line solves with host arrays (the reference's call
    # sites): the assembled blocks and the right-hand side cross PCIe ----
    if not args.no_dropin:
        out["dropin"] = dropin_rates(st, p)
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
        from lns3d_amd.serve import sweep, peak, serial, draw_requests, fifo_latencies
        nreq = args.requests
        ser64, _ = serial(draw_requests(nreq, None, 0), mode="fp64")
        ser32, _ = serial(draw_requests(nreq, None, 0), mode="fp32_outer")
        sw = sweep(count=nreq, seed=0, mode="fp32_outer")
        pk_b = peak(count=nreq, seed=0, mode="fp32_outer", batched=True)
        # many small requests: 2 nreq on the two smaller grids, with a graph
        # per request and batched per grid with three lines per workgroup
        small = dict(count=2 * nreq, seed=0, mode="fp32_outer", grids=((129, 65), (257, 129)),
                     grid_weights=(0.6, 0.4))
        pk_s0 = peak(batched=False, **small)
        pk_s1 = peak(batched=True, small_line_variant=4, **small)
        rnd = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()
                         if k != "service_s"}

        def fifo(ser):
            # the same requests served one at a time in arrival order under
            # the server's Poisson arrivals (measured service times)
            lat = fifo_latencies(np.array(ser["service_s"]), np.array(sw["arrivals"]))
            return {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                    "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}

        out["sweep"] = {"requests": nreq,
                        "request": "receptivity (ispg 2 sponge forcing, loctime 1, CFL 20 per 128 cells in x) to 1e-9 of "
                                   "the first residual; grids 129x65..1025x257, Re U[50,150]; one request set for "
                                   "every leg",
                        "mode": "FP32 factors, FP64 residual each step (outer refinement)",
                        "peak": rnd(sw["peak"]), "arrival_rate_s": round(sw["rate"], 3),
                        "peak_batched_per_grid": rnd(pk_b),
                        "small_grids_peak": {"requests": 2 * nreq, "grids": "129x65 (0.6), 257x129 (0.4)",
                                             "graph_per_request": rnd(pk_s0),
                                             "batched_per_grid_three_lines_per_workgroup": rnd(pk_s1)},
                        "poisson": rnd(sw["open_loop"]),
                        "serial_fp64": dict(rnd(ser64), same_arrivals=fifo(ser64)),
                        "serial_fp32_outer": dict(rnd(ser32), same_arrivals=fifo(ser32)),
                        "time_to_solution_ratio_fp64_over_fp32":
                            round(ser32["solves_per_s"] / ser64["solves_per_s"], 3)}

    # ---- CPU baseline ----
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, met, vm, v0, args.cpu_steps)
    print(json.dumps(out), flush=True)


def dropin_rates(st, p, nsolve=5):
    """lns3d_cpenta2bc (xi lines) and lns3d_cpenta1bc (eta lines) called as
    the reference calls cpenta2bc / cpenta1bc, with host arrays: code 0 copies
    the assembled blocks (2000 B per point) and factors, code 1 copies the
    right-hand side in and the solution out (80 B per point each way) around
    the device solve.  Times per call (median of nsolve code-1 calls)"""
    import torch
    from lns3d_amd.native import lib, check
    L = lib()
    nx, ny = p.nx, p.ny
    out = {}
    rng = np.random.default_rng(5)
    for name, d, fn in (("cpenta2bc_xi", 1, L.lns3d_cpenta2bc), ("cpenta1bc_eta", 2, L.lns3d_cpenta1bc)):
        m_dev = torch.zeros(125 * nx * ny, dtype=torch.complex128, device="cuda")
        st.assemble(d, m_dev)
        torch.cuda.synchronize()
        mat = m_dev.cpu().numpy()
        del m_dev
        rhs = (rng.standard_normal(5 * nx * ny) + 1j * rng.standard_normal(5 * nx * ny)).astype(np.complex128)
        t0 = time.perf_counter()
        check(fn(nx, ny, 5, mat.ctypes.data, rhs.ctypes.data, 0), name)
        t_factor = time.perf_counter() - t0
        ts = []
        for _ in range(nsolve):
            t0 = time.perf_counter()
            check(fn(nx, ny, 5, mat.ctypes.data, rhs.ctypes.data, 1), name)
            ts.append(time.perf_counter() - t0)
        check(L.lns3d_release_factors(), "lns3d_release_factors")
        out[name] = {"factor_call_ms": round(t_factor * 1e3, 3),
                     "solve_call_ms": round(float(np.median(ts)) * 1e3, 3),
                     "host_bytes_factor": 2000 * nx * ny, "host_bytes_solve": 160 * nx * ny}
        del mat
    return out


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


def host_description():
    """CPU model, logical CPUs and the NUMA layout of the host"""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    numa = {}
    base = "/sys/devices/system/node"
    try:
        for d in sorted(os.listdir(base)):
            if d.startswith("node") and d[4:].isdigit():
                with open(os.path.join(base, d, "cpulist")) as f:
                    numa[d] = f.read().strip()
    except OSError:
        pass
    # the process's CPUs as it started (the OpenMP runtime binds the calling
    # thread to one place when it starts, with OMP_PROC_BIND=close)
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "numa_nodes": numa,
            "affinity_cpus": AFFINITY_AT_START}


def _cpu_baseline(p, met, vm, v0, nsteps, warm=3):
    """The reference's Fortran step (oracle/_ref/libref.so: its own lrhs3D,
    rhsBC3D, lhs*, cpenta* and itrBC3D compiled with flang -O2 -fopenmp,
    the transposes of its driver in parallel as src/impdrv3d.f90:257-304) on
    the host: OMP_NUM_THREADS threads (OMP_PROC_BIND=close, OMP_PLACES=cores
    unless set) and one thread.  The first step forms and factors the LHS
    (timed apart); then `warm` steps, then `nsteps` steps timed one by one:
    median and p90, and the median of each phase.  Where the library was not
    built, the C restatement on one thread."""
    import ctypes
    # thread placement of the reference's OpenMP runtime, read when it starts
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle.oracle import Oracle, Reference
    from lns3d_amd.config import make_grid
    x, y, met2 = make_grid(p)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if Reference.available():
        R = Reference()
        R.setup(p, x, y, met2, vm)
        R.genmtrx()
        R.set_time_step()
        ph = (ctypes.c_double * 5)()
        res = {}
        for nthr in (cores, 1):
            R.lib.ref_set_threads(ctypes.c_int(nthr))
            v = np.asfortranarray(v0.copy())
            R.itrbc3d(v)
            t0 = time.perf_counter()
            R.step(v, True)                    # forms and factors the LHS
            t_first = time.perf_counter() - t0
            R.lib.ref_phase_times(ph)
            factor_s = ph[4]
            for _ in range(warm):
                R.step(v, False)
            ts, phases = [], []
            n = nsteps if nthr == cores else max(5, nsteps // 4)
            for _ in range(n):
                t0 = time.perf_counter()
                R.step(v, False)
                ts.append(time.perf_counter() - t0)
                R.lib.ref_phase_times(ph)
                phases.append(list(ph)[:4])
            med = float(np.median(ts))
            pm = np.median(np.array(phases), axis=0)
            res[nthr] = {"steps_s": round(1.0 / med, 4), "median_ms": round(1e3 * med, 3),
                         "p90_ms": round(1e3 * float(np.percentile(ts, 90)), 3), "steps_timed": n,
                         "phases_ms": {"rhs": round(1e3 * pm[0], 3), "xi_solve": round(1e3 * pm[1], 3),
                                       "eta_solve": round(1e3 * pm[2], 3), "update": round(1e3 * pm[3], 3)},
                         "factor_ms": round(1e3 * factor_s, 2), "first_step_ms": round(1e3 * t_first, 2)}
        R.lib.ref_set_threads(ctypes.c_int(cores))
        top = res[cores]
        return {"value": top["steps_s"], "unit": "steps/s", "cores": cores, "kind": "reference",
                "sample": f"{warm} warm-up then {top['steps_timed']} implicit steps at {p.nx}x{p.ny}, each "
                          "timed, median (flang -O2 -fopenmp build of the reference sources)",
                "median_ms": top["median_ms"], "p90_ms": top["p90_ms"], "phases_ms": top["phases_ms"],
                "factor_ms": top["factor_ms"], "one_thread": res[1],
                "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
                "host": host_description()}
    O = Oracle()
    c = O.genmtrx(p, vm, met)
    dtl = O.set_time_step(p, vm, met)
    v = np.asfortranarray(v0.copy())
    O.itrbc3d(p, v)
    f = O.factor(p, c, dtl)
    ts = []
    for _ in range(max(3, nsteps // 4)):
        t0 = time.perf_counter()
        O.step(p, c, dtl, f, v)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    return {"value": round(1.0 / med, 4), "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"{len(ts)} implicit steps at {p.nx}x{p.ny} (C restatement, one thread), median",
            "host": host_description()}


if __name__ == "__main__":
    main()
