"""Complex-float line solves and the refined step.

The complex-float records come from the same complex-double LU, so one
complex-float solve reproduces the complex-double solve to single-precision
rounding; the refined step (complex-float solves as the preconditioner of a
defect correction on the complex-double AF operator) reproduces the
complex-double step to the requested tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to("cuda")


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("nsys,n", [(3, 9), (65, 129), (257, 300)])
def test_fp32_solve_close_to_fp64(torch_cuda, layout, nsys, n):
    from lns3d_amd.native import Penta
    from tests.penta_cases import random_lines, relerr
    print(f"fp32 solve layout={layout} nsys={nsys} n={n}", flush=True)
    torch = torch_cuda
    mat, rhs = random_lines(nsys, n, layout, seed=5 + n)
    dm = _dev(torch, mat)
    h64, h32 = Penta(0, layout, nsys, n, 0), Penta(0, layout, nsys, n, 1)
    h64.factor(dm)
    h32.factor(dm)
    r64, r32 = _dev(torch, rhs), _dev(torch, rhs)
    h64.solve(r64)
    h32.solve(r32)
    e = relerr(r32.cpu().numpy(), r64.cpu().numpy())
    print(f"  rel diff {e:.2e}", flush=True)
    assert e < 1e-4


@pytest.mark.parametrize("nx,ny", [(33, 17), (129, 65)])
def test_refined_step_matches_fp64(torch_cuda, nx, ny):
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from lns3d_amd.native import Step
    print(f"refined step {nx}x{ny}", flush=True)
    torch = torch_cuda
    p = Params(nx=nx, ny=ny, kz=0.3)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v = make_disturbance(p, x, seed=2)
    st = Step(p, precision=1)
    st.setup(met, vm)
    st.factor()
    va = _dev(torch, v)
    st.itrbc3d(va)
    vb = va.clone()
    dva, dvb = torch.zeros_like(va), torch.zeros_like(va)
    st.step(va, dva)
    its, res = st.step_refined(vb, dvb, tol=1e-12, max_iters=10)
    e = float((dvb - dva).abs().max() / dva.abs().max())
    print(f"  iters {its} resid {res:.2e} dv rel diff {e:.2e}", flush=True)
    assert res <= 1e-12 and e < 1e-10


@pytest.mark.parametrize("nx,ny,xper,yper", [(129, 65, 1, 0), (600, 40, 0, 1), (65, 600, 1, 0)])
def test_refined_step_periodic_matches_fp64(torch_cuda, nx, ny, xper, yper):
    """periodic lines: complex-float records rounded from the complex-double
    bordered LU (one line per workgroup up to 512 lines, four above), the
    operator without a row on the duplicate periodic row; the refined step's
    increment equals the complex-double step's to 1e-10 (both with the exact
    cyclic factor: the refinement converges to the cyclic system's solution,
    which the shipped factor misses on short lines)"""
    import dataclasses
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from lns3d_amd.native import Step
    print(f"refined step {nx}x{ny} xper={xper} yper={yper}", flush=True)
    torch = torch_cuda
    p = Params(nx=nx, ny=ny, kz=0.3)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v = make_disturbance(p, x, seed=2)
    p = dataclasses.replace(p, xper=xper, yper=yper, periodic_exact=1)
    st = Step(p, precision=1)
    st.setup(met, vm)
    st.factor()
    va = _dev(torch, v)
    st.itrbc3d(va)
    vb = va.clone()
    dva, dvb = torch.zeros_like(va), torch.zeros_like(va)
    st.step(va, dva)
    its, res = st.step_refined(vb, dvb, tol=1e-12, max_iters=10)
    # unique points: the duplicate periodic column of the increment (xper)
    # comes from its own eta line solved on a copied right-hand side, where
    # the refinement has no residual row; v takes column 0 there (itrBC3D)
    d, d0 = (dvb - dva).reshape(ny, nx, 5), dva.reshape(ny, nx, 5)
    if xper:
        d, d0 = d[:, :-1], d0[:, :-1]
    e = float(d.abs().max() / d0.abs().max())
    ev = float((vb - va).abs().max() / va.abs().max())
    print(f"  iters {its} resid {res:.2e} dv rel diff {e:.2e} v rel diff {ev:.2e}", flush=True)
    assert res <= 1e-12 and e < 1e-10 and ev < 1e-12
    st.close()
