"""GPU parity of the line solvers against the oracle (C restatement of
src/cpenta1bc.f, cpenta2bc.f, cpenta1p.f, cpenta2p.f).  Tolerance: 1e-12
relative in the max norm for complex-double factors."""
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


@pytest.fixture(scope="module")
def orc():
    from oracle.oracle import Oracle
    return Oracle()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to("cuda")


def _host(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


# sizes cover: the shortest lines; lines whose last rows (the boundary-row
# blocks) fall in one chunk or straddle two chunks of the streaming solve;
# many lines (four per workgroup) with a line count that is not a multiple
# of four
# Lines of 16 or more rows with fewer than 512 lines use the twisted factor
# (two chains per line meeting in the middle): the sizes below include its
# shortest lines, odd and even lengths and chunk boundaries on either chain.
BC_SIZES = [(1, 7), (3, 8), (5, 15), (5, 16), (2, 17), (3, 18), (2, 24), (3, 31), (4, 40),
            (1, 101), (65, 129), (2, 33), (3, 35), (513, 14), (515, 17), (516, 9)]


def twisted(nsys, n):
    return n >= 16 and nsys < 512


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("nsys,n", BC_SIZES)
def test_bc_random(torch_cuda, orc, layout, nsys, n):
    from lns3d_amd.native import Penta
    from tests.penta_cases import random_lines, relerr
    print(f"bc layout={layout} nsys={nsys} n={n}", flush=True)
    mat, rhs = random_lines(nsys, n, layout, seed=nsys * 100 + n)
    m_ref, x_ref = mat.copy(order="F"), rhs.copy(order="F")
    orc.cpenta_bc(layout, m_ref, x_ref, 2)
    h = Penta(0, layout, nsys, n)
    dm, dr = _dev(torch_cuda, mat), _dev(torch_cuda, rhs)
    h.factor(dm)
    h.solve(dr)
    x = _host(dr, rhs.shape)
    assert relerr(x, x_ref) < TOL
    # second solve with the same factors: repeatable
    dr2 = _dev(torch_cuda, rhs)
    h.solve(dr2)
    assert np.array_equal(_host(dr2, rhs.shape), x)
    if not twisted(nsys, n):
        # the reference-order kernel on the same factors
        dr3 = _dev(torch_cuda, rhs)
        h.solve_reference_order(dr3)
        assert relerr(_host(dr3, rhs.shape), x_ref) < TOL
    h.close()


@pytest.mark.parametrize("exact", [0, 1])
@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("nsys,np_", [(2, 8), (4, 17), (33, 65)])
def test_periodic_random(torch_cuda, orc, layout, nsys, np_, exact):
    from lns3d_amd.native import Penta
    from tests.penta_cases import random_lines, relerr
    print(f"periodic layout={layout} nsys={nsys} np={np_} exact={exact}", flush=True)
    mat, rhs = random_lines(nsys, np_, layout, seed=7 * nsys + np_)
    m_ref, x_ref = mat.copy(order="F"), rhs.copy(order="F")
    per = np.zeros((2, 5, 5, nsys, np_) if layout == 1 else (2, 5, 5, np_, nsys), complex, order="F")
    per2 = per.copy(order="F")
    orc.cpenta_p(layout, m_ref, x_ref, per, per2, 2, exact)
    h = Penta(1, layout, nsys, np_, 0, exact)
    dm, dr = _dev(torch_cuda, mat), _dev(torch_cuda, rhs)
    h.factor(dm)
    h.solve(dr)
    assert relerr(_host(dr, rhs.shape), x_ref) < TOL
    h.close()


@pytest.mark.parametrize("exact", [0, 1])
@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("nsys,np_", [(4, 17), (33, 65), (520, 24)])
def test_periodic_complex_float(torch_cuda, orc, layout, nsys, np_, exact):
    """periodic lines with complex-float records (rounded from the handle's
    complex-double bordered LU; one line per workgroup, four above 512
    lines): one solve agrees with the oracle's cpenta1p/cpenta2p of the same
    factor (shipped or exact) to single-precision accuracy"""
    from lns3d_amd.native import Penta
    from tests.penta_cases import random_lines, relerr
    print(f"periodic complex float layout={layout} nsys={nsys} np={np_} exact={exact}", flush=True)
    mat, rhs = random_lines(nsys, np_, layout, seed=11 * nsys + np_)
    m_ref, x_ref = mat.copy(order="F"), rhs.copy(order="F")
    per = np.zeros((2, 5, 5, nsys, np_) if layout == 1 else (2, 5, 5, np_, nsys), complex, order="F")
    per2 = per.copy(order="F")
    orc.cpenta_p(layout, m_ref, x_ref, per, per2, 2, exact)
    h = Penta(1, layout, nsys, np_, 1, exact)
    dm, dr = _dev(torch_cuda, mat), _dev(torch_cuda, rhs)
    h.factor(dm)
    h.solve(dr)
    e = relerr(_host(dr, rhs.shape), x_ref)
    print(f"  rel diff {e:.2e}", flush=True)
    assert e < 1e-5
    h.close()


# complex-float twisted lines take chain steps in pairs (12-row chunks): the
# sizes include the shortest twisted lines (no pairs), pair regions ending
# inside a chunk, several chunks, and many lines (no pairs: plain variants)
FP32_SIZES = [(3, 35), (2, 61), (515, 17), (2, 16), (3, 17), (2, 18), (2, 29), (3, 40), (3, 63),
              (1, 101), (65, 129)]


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("nsys,n", FP32_SIZES)
def test_bc_complex_float_factors(torch_cuda, orc, layout, nsys, n):
    """complex-float records (the refinement path's preconditioner): one
    solve agrees with the complex-double oracle to single-precision accuracy
    (1e-5 relative for these well-conditioned lines)"""
    from lns3d_amd.native import Penta
    from tests.penta_cases import random_lines, relerr
    print(f"bc-fp32 layout={layout} nsys={nsys} n={n}", flush=True)
    mat, rhs = random_lines(nsys, n, layout, seed=31 * nsys + n)
    m_ref, x_ref = mat.copy(order="F"), rhs.copy(order="F")
    orc.cpenta_bc(layout, m_ref, x_ref, 2)
    h = Penta(0, layout, nsys, n, 1)
    dm, dr = _dev(torch_cuda, mat), _dev(torch_cuda, rhs)
    h.factor(dm)
    h.solve(dr)
    assert relerr(_host(dr, rhs.shape), x_ref) < 1e-5
    h.close()


# every solve variant pinned in turn (lns3d_tune_set) on shapes no other test
# uses: the shortest twisted lines, odd and even lengths, chunk boundaries in
# both chunk sizes, and many lines
VARIANT_SIZES = [(2, 16), (3, 37), (1, 103), (5, 64), (33, 97)]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nsys,n", VARIANT_SIZES)
def test_every_variant(torch_cuda, orc, nsys, n, prec):
    from lns3d_amd.native import Penta, set_line_variant, TUNE_VARIANTS, VARIANT_NAMES
    from tests.penta_cases import random_lines, relerr
    layout = 1
    mat, rhs = random_lines(nsys, n, layout, seed=17 * nsys + n)
    m_ref, x_ref = mat.copy(order="F"), rhs.copy(order="F")
    orc.cpenta_bc(layout, m_ref, x_ref, 2)
    tol = TOL if prec == 0 else 1e-5
    try:
        for v in range(TUNE_VARIANTS):
            print(f"variant {v} ({VARIANT_NAMES[v]}) nsys={nsys} n={n} prec={prec}", flush=True)
            set_line_variant(n, nsys, prec, v)
            h = Penta(0, layout, nsys, n, prec)
            dm, dr = _dev(torch_cuda, mat), _dev(torch_cuda, rhs)
            h.factor(dm)
            h.solve(dr)
            assert relerr(_host(dr, rhs.shape), x_ref) < tol, VARIANT_NAMES[v]
            h.close()
    finally:
        set_line_variant(n, nsys, prec, -1)
