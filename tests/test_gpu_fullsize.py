"""Line solves at the benchmark's full sizes against the oracle, on the LHS
factors the oracle assembles from the benchmark baseflow (1025x257, beta
0.3) and a seeded complex N(0,1) right-hand side:
  * eta lines, cpenta1bc(nsys=1025, n=257)  (src/impdrv3d.f90:327)
  * xi lines,  cpenta2bc(n=1025, nsys=257)
  * periodic xi lines on a 1024x512 grid, cpenta2p(np=1024, nsys=512)
Tolerance 1e-12 relative (max norm).  Size-independent properties at the
same sizes: linearity of the solve in the right-hand side, and the residual
of the solution under the assembled operator (block-penta matvec in numpy)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


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


def _problem(orc, **kw):
    from lns3d_amd.config import Params, make_grid, make_mean
    p = Params(**kw)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    c = orc.genmtrx(p, vm, met)
    dtl = orc.set_time_step(p, vm, met)
    return p, c, dtl


def _rhs(shape, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def _matvec_lines(mat, x, layout, periodic):
    """y = A x for the assembled lines (slot k of row i at column i-2+k,
    wrapped rows for the non-periodic closures), vectorised over lines"""
    from tests.penta_cases import _bc_col
    if layout == 1:                                  # (5,5,5,nsys,n), (5,nsys,n)
        m, v = mat, x
    else:                                            # (5,5,5,n,nsys) -> line index first
        m, v = mat.transpose(0, 1, 2, 4, 3), x.transpose(0, 2, 1)
    n = m.shape[4] - (1 if periodic else 0)
    y = np.zeros_like(v)
    for i in range(n):
        for k in range(5):
            col = (i - 2 + k) % n if periodic else _bc_col(n, i, k)
            y[:, :, i] += np.einsum("rcs,cs->rs", m[k, :, :, :, i], v[:, :, col])
    if periodic:
        y[:, :, n] = y[:, :, 0]
    return y if layout == 1 else y.transpose(0, 2, 1)


@pytest.mark.parametrize("which", ["eta_cpenta1bc", "xi_cpenta2bc"])
def test_benchmark_lines(torch_cuda, orc, which):
    from lns3d_amd.native import Penta
    from tests.penta_cases import relerr
    torch = torch_cuda
    print(f"full size {which}", flush=True)
    p, c, dtl = _problem(orc, nx=1025, ny=257, kz=0.3)
    if which == "eta_cpenta1bc":
        mat = orc.lhs(p, 2, c, dtl)                  # (5,5,5,nx,ny): nsys = nx, n = ny
        layout, nsys, n = 1, 1025, 257
    else:
        mat = orc.lhs(p, 1, c, dtl)                  # (5,5,5,nx,ny): n = nx, nsys = ny
        layout, nsys, n = 2, 257, 1025
    rshape = (5, nsys, n) if layout == 1 else (5, n, nsys)
    b = _rhs(rshape, 1)
    h = Penta(0, layout, nsys, n)
    h.factor(_dev(torch, mat))
    db = _dev(torch, b)
    h.solve(db)
    x = _host(db, rshape)
    mref, xref = mat.copy(order="F"), b.copy(order="F")
    orc.cpenta_bc(layout, mref, xref, 2)
    assert relerr(x, xref) < TOL
    # the solution satisfies the assembled system
    res = _matvec_lines(mat, x, layout, False) - b
    assert np.abs(res).max() < 1e-10 * np.abs(b).max()
    # linearity: solve(a b + b2) = a x + x2
    b2 = _rhs(rshape, 2)
    a = 0.7 - 0.2j
    d3 = _dev(torch, np.asfortranarray(a * b + b2))
    h.solve(d3)
    d2 = _dev(torch, b2)
    h.solve(d2)
    assert relerr(_host(d3, rshape), a * x + _host(d2, rshape)) < 1e-12
    h.close()


def test_periodic_1024x512(torch_cuda, orc):
    from lns3d_amd.native import Penta
    from tests.penta_cases import relerr
    torch = torch_cuda
    print("full size periodic cpenta2p(1024, 512)", flush=True)
    p, c, dtl = _problem(orc, nx=1024, ny=512, kz=0.3, xper=1)
    mat = orc.lhs(p, 1, c, dtl)                      # (5,5,5,np=1024,nsys=512)
    np_, nsys = 1024, 512
    b = _rhs((5, np_, nsys), 3)
    b[:, np_ - 1, :] = b[:, 0, :]
    h = Penta(1, 2, nsys, np_)
    h.factor(_dev(torch, mat))
    db = _dev(torch, b)
    h.solve(db)
    x = _host(db, b.shape)
    shp = (2, 5, 5, np_, nsys)
    mref, xref = mat.copy(order="F"), b.copy(order="F")
    orc.cpenta_p(2, mref, xref, np.zeros(shp, complex, order="F"),
                 np.zeros(shp, complex, order="F"), 2, exact=0)
    assert relerr(x, xref) < TOL
    assert np.array_equal(x[:, np_ - 1, :], x[:, 0, :])
    h.close()


def test_benchmark_step(torch_cuda, orc):
    """two implicit steps of the benchmark problem (1025x257, beta 0.3)"""
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from lns3d_amd.native import Step
    from tests.penta_cases import relerr
    torch = torch_cuda
    print("full size step 1025x257", flush=True)
    p = Params(nx=1025, ny=257, kz=0.3)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v0 = make_disturbance(p, x, seed=0)
    c = orc.genmtrx(p, vm, met)
    dtl = orc.set_time_step(p, vm, met)
    vo = v0.copy(order="F")
    orc.itrbc3d(p, vo)
    f = orc.factor(p, c, dtl)
    st = Step(p)
    st.setup(met, vm)
    st.factor()
    v = _dev(torch, v0)
    st.itrbc3d(v)
    dv = torch.zeros_like(v)
    for k in range(2):
        dvo = orc.step(p, c, dtl, f, vo)
        st.step(v, dv)
        assert relerr(_host(dv, v0.shape), dvo) < TOL, k
    assert relerr(_host(v, v0.shape), vo) < TOL
    st.close()
