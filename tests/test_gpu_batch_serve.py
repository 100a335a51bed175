"""Batches, graphs, the refined step without host synchronisation, the
autotuner and the serving loop, on the GPU.

* A graph of several uneven instances (each on its own branch) gives the
  same states, bit for bit, as the same steps launched one by one; it is
  replayed several times with eager work and synchronisation between the
  replays, and every replay is checked.
* The refined step with one correction agrees with the complex-double step
  to 1e-10 relative, and its residual check reports <= 1e-10.
* The autotuner returns a variant whose plans still solve like the oracle.
* Serving: a request gives the same state alone and inside a mix of
  overlapping requests (bit for bit), and a memory budget that holds one
  context at a time still completes every request."""
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


@pytest.mark.parametrize("mode", [0, 2])
def test_graph_matches_eager(torch_cuda, mode):
    from lns3d_amd.batch import Batch
    from lns3d_amd.config import Params
    torch = torch_cuda
    print(f"graph vs eager mode={mode}", flush=True)
    params = [Params(**kw) for kw in SHAPES]
    b = Batch(params, precision=1, seed=3)
    nsteps = 3
    # eager reference
    for k in range(len(params)):
        b.run_eager(k, nsteps, mode=mode)
    torch.cuda.synchronize()
    ref = [v.clone() for v in b.v]
    g = b.capture(nsteps, mode=mode)
    for rep in range(3):
        b.reset()
        junk = torch.randn(1000, device="cuda").sum()      # eager work between replays
        torch.cuda.synchronize()
        g.launch()
        g.synchronize()
        for k in range(len(params)):
            assert torch.equal(b.v[k], ref[k]), (rep, k)
        del junk
    g.close()
    b.close()


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


GRIDS = ((33, 17), (65, 33))


def _reqs(count, rate, seed):
    from lns3d_amd.serve import draw_requests
    return draw_requests(count, rate, seed, steps=(10, 30), chunk=10, grids=GRIDS, grid_weights=(1, 1))


def test_serve_alone_and_in_mix(torch_cuda):
    from lns3d_amd.serve import Server
    print("serve alone vs mix", flush=True)
    mix = _reqs(8, 400.0, 11)
    target = mix[3]
    srv = Server(8 * 2**30, keep_states=True)
    alone = srv.run([dataclass_replace(target, arrival=0.0)])[0]["state"]
    recs = srv.run(mix)
    srv.close()
    got = [r for r in recs if r["rid"] == target.rid][0]["state"]
    assert len(recs) == len(mix)
    assert np.array_equal(alone, got)
    assert np.all(np.isfinite(got))


def test_serve_tight_budget(torch_cuda):
    """the pool never holds more than the budget; with room for one context
    of the single grid in use, requests run one after another"""
    import dataclasses
    from lns3d_amd.native import ctx_bytes
    from lns3d_amd.serve import Server
    reqs = _reqs(5, None, 4)
    one = max(ctx_bytes(r.params(), 1) + 160 * r.nx * r.ny for r in reqs)
    srv = Server(one + 1024, keep_states=False)
    recs = srv.run(reqs)
    assert len(recs) == 5
    assert srv._held() <= srv.budget
    srv.close()
    same = [dataclasses.replace(r, nx=65, ny=33) for r in reqs]
    one = ctx_bytes(same[0].params(), 1) + 160 * 65 * 33
    srv = Server(one + 1024, keep_states=False)
    recs = srv.run(same)
    assert len(recs) == 5 and len(srv.pool) == 1
    assert all(b["admitted"] >= a["done"] - 1e-3 for a, b in zip(recs, recs[1:]))
    srv.close()


def dataclass_replace(r, **kw):
    import dataclasses
    return dataclasses.replace(r, **kw)
