"""The mixed-precision and batched forms at their benchmark sizes.

* 1025x257, beta 0.3: the refined step (complex-float factors as the
  preconditioner of a defect correction of the complex-double AF system, run
  to a relative residual of 1e-10) against the oracle's complex-double step
  on the same input, per variable (rho, u, v, w, T) in the max norm, over
  three steps; the refinement count of each step is printed.
* 64 uneven instances (65x33 to 513x257, beta, Mach, Re drawn from the
  seed), line-solve kernels chosen by the autotuner, 20 refined steps of
  every instance in one graph launch with one launch per stage over all
  instances: the same states, bit for bit, as the graph with one branch per
  instance; every instance against 20 complex-double steps of the same
  instance launched one by one (1e-10), and the three smallest instances
  against 20 of the oracle's steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("rho", "u", "v", "w", "T")


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to("cuda")


def _host(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def per_variable(a, b):
    """max-norm relative difference of each of the five variables"""
    return [float(np.abs(a[k] - b[k]).max() / max(np.abs(b[k]).max(), 1e-300)) for k in range(5)]


def test_refined_1025x257_per_variable(torch_cuda, orc):
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from lns3d_amd.native import Step
    torch = torch_cuda
    print("refined step 1025x257 beta 0.3 vs oracle", flush=True)
    p = Params(nx=1025, ny=257, kz=0.3)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v0 = make_disturbance(p, x, seed=0)
    c = orc.genmtrx(p, vm, met)
    dtl = orc.set_time_step(p, vm, met)
    f = orc.factor(p, c, dtl)
    vo = v0.copy(order="F")
    orc.itrbc3d(p, vo)
    st = Step(p, precision=1)
    st.setup(met, vm)
    st.factor()
    v = _dev(torch, v0)
    st.itrbc3d(v)
    dv = torch.zeros_like(v)
    for k in range(3):
        dvo = orc.step(p, c, dtl, f, vo)
        its, res = st.step_refined(v, dv, tol=1e-10, max_iters=10)
        e_dv = per_variable(_host(dv, v0.shape), dvo)
        e_v = per_variable(_host(v, v0.shape), vo)
        print(f"  step {k}: {its} solves, AF residual {res:.2e}; dv error "
              + " ".join(f"{n} {e:.1e}" for n, e in zip(NAMES, e_dv))
              + "; v error " + " ".join(f"{n} {e:.1e}" for n, e in zip(NAMES, e_v)), flush=True)
        assert res <= 1e-10
        assert max(e_dv) <= 1e-10, e_dv
        assert max(e_v) <= 1e-10, e_v
    st.close()


def test_uneven_batch_64_instances(torch_cuda, orc):
    from lns3d_amd.batch import Batch, draw_instances, tune_shapes
    from lns3d_amd.config import make_grid, make_mean, make_disturbance
    from lns3d_amd.native import MODE_REFINED, MODE_BATCHED
    torch = torch_cuda
    params = draw_instances(64, 0)
    print("batch of 64: tuning", flush=True)
    tune_shapes(params, (0, 1))
    b = Batch(params, precision=1, seed=0)
    print(f"batch of 64: {b.points} points, capture", flush=True)
    nsteps = 20
    # one launch per stage (the benchmark's graph), then one branch per
    # instance: the same states bit for bit
    g = b.capture(nsteps, mode=MODE_BATCHED, corrections=1)
    g.launch()
    g.synchronize()
    got = [v.clone() for v in b.v]
    g.close()
    b.reset()
    g = b.capture(nsteps, mode=MODE_REFINED, corrections=1)
    g.launch()
    g.synchronize()
    for k in range(len(params)):
        assert torch.equal(b.v[k], got[k]), (k, params[k])
    # the same instances, 20 complex-double steps each, one by one
    worst = 0.0
    for k, st in enumerate(b.steps):
        v, dv = b.v0[k].clone(), torch.zeros_like(b.v0[k])
        for _ in range(nsteps):
            st.step(v, dv)
        e = float((got[k] - v).abs().max() / v.abs().max())
        worst = max(worst, e)
        assert e <= 1e-10, (k, params[k], e)
    print(f"  graph vs complex-double steps: worst {worst:.2e}", flush=True)
    # the three smallest instances against the oracle
    order = sorted(range(len(params)), key=lambda k: params[k].nx * params[k].ny)[:3]
    for k in order:
        p = params[k]
        x, y, met = make_grid(p)
        vm = make_mean(p, y)
        c = orc.genmtrx(p, vm, met)
        dtl = orc.set_time_step(p, vm, met)
        f = orc.factor(p, c, dtl)
        vo = make_disturbance(p, x, seed=k)
        orc.itrbc3d(p, vo)
        assert np.abs(vo - _host(b.v0[k], vo.shape)).max() <= 1e-15 * np.abs(vo).max()
        for _ in range(nsteps):
            orc.step(p, c, dtl, f, vo)
        e = per_variable(_host(got[k], vo.shape), vo)
        print(f"  instance {k} {p.nx}x{p.ny} vs oracle: " + " ".join(f"{n} {x:.1e}" for n, x in zip(NAMES, e)),
              flush=True)
        assert max(e) <= 1e-10, (k, e)
    g.close()
    b.close()
