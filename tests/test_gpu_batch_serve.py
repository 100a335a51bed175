"""Batches, graphs, the refined step without host synchronisation, the
autotuner and the serving loop, on the GPU.

* A graph of several uneven instances (each on its own branch, or every
  stage one launch over all instances) gives the same states, bit for bit,
  as the same steps launched one by one; it is replayed several times with
  eager work and synchronisation between the replays, and every replay is
  checked.  The staged graph groups the line solves by solve kernel: a batch
  whose shapes are pinned to different kernels is checked too.
* The refined step with one correction agrees with the complex-double step
  to 1e-10 relative, and its residual check reports <= 1e-10.
* The autotuner returns a variant whose plans still solve like the oracle.
* Serving: receptivity requests converge to 1e-9 of their first residual;
  a request gives the same state alone and inside a mix of overlapping
  requests (bit for bit); requests batched per grid in staged graphs give
  the states of a graph per request (bit for bit); the complex-float mode
  reaches the complex-double march's steady state; a refused request fails
  alone; a memory budget that holds one context at a time still completes
  every request.
* The staged graph records each instance's residual norms with the bits of
  the per-instance steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


SHAPES = [dict(nx=33, ny=17, kz=0.3), dict(nx=47, ny=21, kz=0.8, Ma=0.7),
          dict(nx=65, ny=33, kz=0.1, Re=2500.0)]


@pytest.mark.parametrize("mode,pin", [(0, None), (2, None), (3, None), (3, "mixed")])
def test_graph_matches_eager(torch_cuda, mode, pin):
    from lns3d_amd.batch import Batch
    from lns3d_amd.config import Params
    from lns3d_amd.native import lib
    torch = torch_cuda
    print(f"graph vs eager mode={mode} pin={pin}", flush=True)
    params = [Params(**kw) for kw in SHAPES]
    pinned = []
    if pin:
        # xi lines: twisted pairs, twisted halves, plain; eta lines: three,
        # two and one lines per workgroup (both precisions of a shape in
        # the same family)
        for p, (vx, vy) in zip(params, [(5, 4), (6, 3), (0, 0)]):
            for prec in (0, 1):
                pinned += [(p.nx, p.ny, prec), (p.ny, p.nx, prec)]
                lib().lns3d_tune_set(p.nx, p.ny, prec, vx if prec else (6 if vx == 5 else vx))
                lib().lns3d_tune_set(p.ny, p.nx, prec, vy)
    b = Batch(params, precision=1, seed=3)
    for key in pinned:
        lib().lns3d_tune_set(*key, -1)
    nsteps = 3
    # eager reference
    for k in range(len(params)):
        b.run_eager(k, nsteps, mode=mode)
    torch.cuda.synchronize()
    ref = [v.clone() for v in b.v]
    g = b.capture(nsteps, mode=mode)
    # replays on two caller streams, then the graph's own (a staged graph
    # copies its tables on its first launch's stream; later launches on
    # another stream wait for that copy)
    streams = [torch.cuda.Stream(), torch.cuda.Stream(), None]
    for rep in range(3):
        b.reset()
        junk = torch.randn(1000, device="cuda").sum()      # eager work between replays
        torch.cuda.synchronize()
        g.launch(streams[rep])
        if streams[rep] is None:
            g.synchronize()
        else:
            streams[rep].synchronize()
        for k in range(len(params)):
            assert torch.equal(b.v[k], ref[k]), (rep, k)
        del junk
    g.close()
    b.close()


def test_staged_norms_match_branches(torch_cuda):
    """the staged graph's complex-float steps (corrections = 0) with a
    residual output per instance: the states and every recorded norm equal,
    bit for bit, those of the same steps on one branch per instance, on two
    consecutive launches (the second continues the slots)"""
    from lns3d_amd.batch import Batch
    from lns3d_amd.config import Params
    from lns3d_amd.native import Graph, MODE_FP32, MODE_BATCHED
    torch = torch_cuda
    params = [Params(**kw) for kw in SHAPES]
    b = Batch(params, precision=1, seed=5)
    nsteps, slots = 3, 8
    out = {}
    for mode in (MODE_FP32, MODE_BATCHED):
        print(f"norms mode={mode}", flush=True)
        b.reset()
        norms = [torch.full((slots,), -1.0, dtype=torch.float64, device="cuda") for _ in params]
        for st, n in zip(b.steps, norms):
            st.set_residual_output(n, slots)
        g = Graph(b.steps, b.v, b.dv, nsteps=nsteps, mode=mode, corrections=0)
        g.launch()
        g.synchronize()
        g2 = Graph(b.steps, b.v, b.dv, nsteps=nsteps, mode=mode, corrections=0)   # slots 3..5
        g2.launch()
        g2.synchronize()
        out[mode] = ([v.clone() for v in b.v], [n.clone() for n in norms])
        for st in b.steps:
            st.set_residual_output(None, 0)
        g.close()
        g2.close()
    (va, na), (vb, nb) = out[MODE_FP32], out[MODE_BATCHED]
    for k in range(len(params)):
        assert torch.equal(va[k], vb[k]), k
        assert torch.equal(na[k], nb[k]), (k, na[k], nb[k])
        assert bool((na[k][:2 * nsteps] > 0).all()) and bool((na[k][2 * nsteps:] == -1.0).all()), na[k]
    b.close()


@pytest.mark.parametrize("corrections", [1, 0])
def test_staged_periodic_matches_branches(torch_cuda, corrections):
    """a staged graph with periodic members (xper; yper with one and with
    four lines per workgroup of the batched periodic solve, the latter above
    512 lines) and a non-periodic one: three steps give the states of the
    same steps on one branch per instance, bit for bit"""
    from lns3d_amd.batch import Batch
    from lns3d_amd.config import Params
    from lns3d_amd.native import Graph, MODE_FP32, MODE_REFINED, MODE_BATCHED
    torch = torch_cuda
    params = [Params(nx=33, ny=17, kz=0.3, xper=1, periodic_exact=1), Params(nx=47, ny=21, kz=0.8, yper=1),
              Params(nx=600, ny=24, kz=0.5, yper=1, periodic_exact=1), Params(nx=65, ny=33, kz=0.1)]
    print(f"staged periodic corrections={corrections}", flush=True)
    b = Batch(params, precision=1, seed=9)
    out = {}
    for mode in (MODE_REFINED if corrections else MODE_FP32, MODE_BATCHED):
        b.reset()
        g = Graph(b.steps, b.v, b.dv, nsteps=3, mode=mode, corrections=corrections)
        g.launch()
        g.synchronize()
        out[mode] = [v.clone() for v in b.v]
        g.close()
    va, vb = out.values()
    for k in range(len(params)):
        assert bool(torch.isfinite(vb[k]).all()), k
        assert torch.equal(va[k], vb[k]), k
    b.close()


@pytest.mark.parametrize("impl,niter", [(1, 2), (2, 1), (2, 2), (3, 1)])
def test_graph_time_schemes_match_eager(torch_cuda, impl, niter):
    """complex-double graphs (mode 0) of steps with corrector iterations,
    impl = 2 (after its first two steps: the history exists and the factors
    hold alfa = 2/3) and impl = 3: two replays in a row equal the same steps
    launched one by one, bit for bit"""
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from lns3d_amd.native import Step, Graph, MODE_FP64
    torch = torch_cuda
    print(f"graph impl={impl} niter={niter}", flush=True)
    p = Params(nx=33, ny=17, kz=0.3)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v0 = make_disturbance(p, x, seed=4)

    def make():
        st = Step(p)
        st.setup(met, vm)
        st.set_time_scheme(impl)
        st.set_corrector_iterations(niter)
        st.factor()
        v = torch.from_numpy(np.ascontiguousarray(v0.reshape(-1, order="F"))).cuda()
        st.itrbc3d(v)
        dv = torch.zeros_like(v)
        for _ in range(2):
            st.step(v, dv)
        return st, v, dv

    sa, va, dva = make()
    sb, vb, dvb = make()
    assert torch.equal(va, vb)
    nsteps = 2
    g = Graph([sb], [vb], [dvb], nsteps=nsteps, mode=MODE_FP64)
    for rep in range(2):
        for _ in range(nsteps):
            sa.step(va, dva)
        g.launch()
        g.synchronize()
        assert bool(torch.isfinite(vb).all())
        assert torch.equal(va, vb), rep
    g.close()
    sa.close()
    sb.close()


def test_refined_fixed_matches_fp64(torch_cuda):
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from lns3d_amd.native import Step
    torch = torch_cuda
    p = Params(nx=129, ny=65)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v0 = make_disturbance(p, x, seed=2)
    st = Step(p, precision=1)
    st.setup(met, vm)
    st.factor()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).cuda()
    v64, v32 = dev(v0), dev(v0)
    st.itrbc3d(v64)
    st.itrbc3d(v32)
    d64, d32 = torch.zeros_like(v64), torch.zeros_like(v64)
    for k in range(3):
        st.step(v64, d64)
        st.step_refined_fixed(v32, d32, corrections=1, check_residual=True)
        r = st.last_residual()
        assert r <= 1e-10, (k, r)
        e = float((d32 - d64).abs().max() / d64.abs().max())
        assert e <= 1e-10, (k, e)
    st.close()


def test_tuner_choice_solves(torch_cuda, orc):
    from lns3d_amd.native import Penta, tune_lines, TUNE_VARIANTS
    from tests.penta_cases import random_lines, relerr
    torch = torch_cuda
    nsys, n = 70, 40
    var, times = tune_lines(n, nsys, 0)
    print("tuned", var, times, flush=True)
    assert len(times) == TUNE_VARIANTS and all(t > 0 for t in times)   # every variant applies at n = 40
    assert 0 <= var < TUNE_VARIANTS and times[var] == min(times)
    mat, rhs = random_lines(nsys, n, 2, seed=5)
    m, xr = mat.copy(order="F"), rhs.copy(order="F")
    orc.cpenta_bc(2, m, xr, 2)
    h = Penta(0, 2, nsys, n)
    dm = torch.from_numpy(np.ascontiguousarray(mat.reshape(-1, order="F"))).cuda()
    dr = torch.from_numpy(np.ascontiguousarray(rhs.reshape(-1, order="F"))).cuda()
    h.factor(dm)
    h.solve(dr)
    assert relerr(dr.cpu().numpy().reshape(rhs.shape, order="F"), xr) < 1e-12
    h.close()


def _reqs(count, rate, seed, grids, weights, **kw):
    from lns3d_amd.serve import draw_requests
    return draw_requests(count, rate, seed, grids=grids, grid_weights=weights, **kw)


def test_serve_alone_and_in_mix(torch_cuda):
    """receptivity requests over 129x65 .. 1025x257 arriving while others
    run: each converges to 1e-9 of its first residual, and the one on the
    largest grid gives the same state, bit for bit, alone and in the mix"""
    from lns3d_amd.serve import Server, GRIDS
    print("serve: Poisson mix over 129x65 .. 1025x257", flush=True)
    mix = _reqs(6, 40.0, 11, GRIDS, (1, 1, 1, 1))
    mix[-1] = dataclass_replace(mix[-1], nx=1025, ny=257)
    srv = Server(64 * 2**30, keep_states=True)
    recs = srv.run(mix)
    for r in recs:
        print(f"  rid {r['rid']} {r['nx']}x{r['ny']}: {r['steps']} steps, ratio {r['ratio']:.2e}, "
              f"latency {r['latency']*1e3:.0f} ms", flush=True)
    assert len(recs) == len(mix) and all(r["error"] is None and r["converged"] for r in recs)
    target = mix[-1]
    alone = srv.run([dataclass_replace(target, arrival=0.0)])[0]
    srv.close()
    got = [r for r in recs if r["rid"] == target.rid][0]
    assert alone["steps"] == got["steps"]
    assert np.array_equal(alone["state"], got["state"])
    assert np.all(np.isfinite(got["state"]))


def test_serve_batched_matches_unbatched(torch_cuda):
    """requests of two grids arriving while others run: served with one
    staged graph per grid (rebuilt as requests come and go) they reach the
    same states, bit for bit, in the same number of steps, as with a graph
    and stream per request"""
    from lns3d_amd.serve import Server
    grids = ((129, 65), (257, 129))
    reqs = _reqs(6, 60.0, 13, grids, (1, 1), max_steps=3000)
    out = {}
    for batched in (False, True):
        print(f"serve: batched={batched}", flush=True)
        srv = Server(64 * 2**30, keep_states=True, batched=batched)
        out[batched] = srv.run(reqs)
        if batched:
            reb = {k: (c.rebuilds, round(c.rebuild_s, 3)) for k, c in srv.cohorts.items()}
            print(f"  graph rebuilds (count, s) per grid: {reb}", flush=True)
        srv.close()
    for a, b in zip(out[False], out[True]):
        print(f"  rid {a['rid']} {a['nx']}x{a['ny']}: steps {a['steps']} / {b['steps']}", flush=True)
        assert a["error"] is None and b["error"] is None
        assert a["steps"] == b["steps"] and a["converged"] == b["converged"]
        assert np.array_equal(a["state"], b["state"]), a["rid"]


def test_serve_small_line_variant(torch_cuda):
    """three lines per workgroup on the small grids: batched and per request
    agree bit for bit, and the server drops its variant records at close"""
    from lns3d_amd.serve import Server
    reqs = _reqs(4, None, 17, ((129, 65), (257, 129)), (1, 1), max_steps=500)
    out = {}
    for batched in (False, True):
        srv = Server(64 * 2**30, keep_states=True, batched=batched, small_line_variant=4)
        out[batched] = srv.run(reqs)
        pinned = set(srv._pinned)
        srv.close()
        assert pinned and not srv._pinned
    for a, b in zip(out[False], out[True]):
        assert a["error"] is None and a["steps"] == b["steps"]
        assert np.array_equal(a["state"], b["state"]), a["rid"]


def test_serve_matches_complex_double_march(torch_cuda):
    """per grid, a request served with complex-float factors (the default
    mode) converges to the steady state of the complex-double march (the
    complex-double step is the oracle-pinned one)"""
    from lns3d_amd.serve import Server, GRIDS
    reqs = [dataclass_replace(r, rid=k, nx=g[0], ny=g[1], arrival=0.0)
            for k, (r, g) in enumerate(zip(_reqs(4, None, 5, GRIDS, (1, 1, 1, 1)), GRIDS))]
    out = {}
    for mode in ("fp32_outer", "fp64"):
        print(f"serve: one request per grid, {mode}", flush=True)
        srv = Server(64 * 2**30, mode=mode, keep_states=True)
        out[mode] = srv.run(reqs)
        srv.close()
    for a, b in zip(out["fp32_outer"], out["fp64"]):
        assert a["converged"] and b["converged"]
        e = float(np.abs(a["state"] - b["state"]).max() / np.abs(b["state"]).max())
        print(f"  {a['nx']}x{a['ny']}: steps {a['steps']} (fp64 {b['steps']}), state rel diff {e:.2e}",
              flush=True)
        assert e < 1e-6


@pytest.mark.parametrize("batched", [False, True])
def test_serve_isolates_a_failing_request(torch_cuda, batched):
    """a request the library refuses fails alone: the others complete with
    the states they reach without it"""
    from lns3d_amd.serve import Server
    grids, w = ((129, 65), (257, 129)), (1, 1)
    reqs = _reqs(4, None, 7, grids, w, max_steps=400)
    bad = dataclass_replace(reqs[1], Re=-1.0)            # refused by lns3d_ctx_set_params
    srv = Server(64 * 2**30, keep_states=True, batched=batched)
    recs = srv.run([reqs[0], bad, reqs[2], reqs[3]])
    assert recs[1]["error"] and "Re must be" in recs[1]["error"]
    solo = {r.rid: srv.run([r])[0] for r in (reqs[0], reqs[2], reqs[3])}
    srv.close()
    for r in (recs[0], recs[2], recs[3]):
        assert r["error"] is None
        assert np.array_equal(r["state"], solo[r["rid"]]["state"])


def test_serve_batched_graph_failure(torch_cuda, monkeypatch):
    """a staged graph that cannot be made (forced here on the first capture
    for the 257x129 grid) fails that cohort's requests with the error; the
    other grid's requests complete with the states they reach without them"""
    from lns3d_amd import serve
    from lns3d_amd.native import LNS3DError
    grids, w = ((129, 65), (257, 129)), (1, 1)
    reqs = _reqs(4, None, 19, grids, w, max_steps=300)
    real = serve.Graph
    tripped = []

    def flaky(steps, *a, **k):
        if not tripped and steps[0].params.nx == 257:
            tripped.append(1)
            raise LNS3DError("forced failure", 2)
        return real(steps, *a, **k)

    monkeypatch.setattr(serve, "Graph", flaky)
    srv = serve.Server(64 * 2**30, keep_states=True, batched=True)
    recs = srv.run(reqs)
    monkeypatch.setattr(serve, "Graph", real)
    solo = {r.rid: srv.run([r])[0] for r in reqs if r.nx == 129}
    srv.close()
    assert tripped
    for r in recs:
        if r["nx"] == 257:
            assert r["error"] and "forced failure" in r["error"], r
        else:
            assert r["error"] is None
            assert np.array_equal(r["state"], solo[r["rid"]]["state"])


def test_serve_tight_budget(torch_cuda):
    """the pool never holds more than the budget; with room for one context
    of the single grid in use, requests run one after another"""
    from lns3d_amd.serve import Server
    reqs = _reqs(4, None, 4, ((129, 65),), (1,), max_steps=200)
    srv = Server(0, keep_states=False)
    one = srv._need(reqs[0])
    srv.budget = one + 1024
    recs = srv.run(reqs)
    assert len(recs) == 4 and all(r["error"] is None for r in recs)
    assert srv._held() <= srv.budget and len(srv.pool) == 1
    assert all(b["admitted"] >= a["done"] - 1e-3 for a, b in zip(recs, recs[1:]))
    srv.close()


def dataclass_replace(r, **kw):
    import dataclasses
    return dataclasses.replace(r, **kw)
