"""The allocation-free residual norm (lns3d_residual_norm_noalloc) against
the context's lns3d_residual_norm, through the C ABI: the same kernels and
the same order of summation, so the same bits -- on two calls in a row, the
second on a smaller grid (the static partials array keeps the first call's
contents, which the second must not read)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_noalloc_norm_matches_context_norm(torch_cuda):
    from lns3d_amd.config import Params, make_grid, make_mean
    from lns3d_amd.native import Step, lib, ptr
    torch = torch_cuda
    L = lib()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    for nx, ny in ((1025, 257), (65, 33)):
        print(f"no-alloc norm {nx}x{ny}", flush=True)
        p = Params(nx=nx, ny=ny, kz=0.3)
        x, y, met = make_grid(p)
        st = Step(p)
        st.setup(met, make_mean(p, y))
        _, dtl = st.get_coeffs()
        P = nx * ny
        r = torch.randn(5 * P, dtype=torch.complex128, device="cuda", generator=gen)
        d = torch.from_numpy(np.ascontiguousarray(dtl.reshape(-1, order="F"))).cuda()
        ref, got = (torch.full((1,), -1.0, dtype=torch.float64, device="cuda") for _ in range(2))
        st.residual_norm(r, ref)
        torch.cuda.synchronize()
        assert L.lns3d_residual_norm_noalloc(ptr(r), ptr(d), P, ptr(got), None) == 0
        torch.cuda.synchronize()
        assert float(ref) > 0 and torch.equal(got, ref), (float(got), float(ref))
        # a plain numpy sum agrees to rounding
        rh = r.cpu().numpy().reshape(P, 5)
        want = np.sqrt(np.sum(np.abs(rh) ** 2 / dtl.reshape(-1, order="F")[:, None] ** 2))
        assert abs(float(got) - want) <= 1e-12 * want
        st.close()
    assert L.lns3d_residual_norm_noalloc(ptr(r), ptr(d), 0, ptr(got), None) == 1
    assert L.lns3d_residual_norm_noalloc(ptr(r), ptr(d), (1 << 28) + 1, ptr(got), None) == 1
    assert b"LNS3D_NOALLOC_MAX_POINTS" in L.lns3d_last_error()
