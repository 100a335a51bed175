"""GPU parity of the symmetry boundaries, left/right = 7 (the symmetric
derivative closures of src/grad.f90, grad2.f90, cgrad.f90, cgrad2.f90, the
xi-factor rows of src/lhsbs1f3d.f90 with its transposed Ahi(6,ny,nx) read,
rhsBC3D/itrBC3D leaving those rows alone), against the reference's own
vectors (tests/golden/ref_sym_<case>.npz, written by
tests/golden/gen_golden.py from the compiled reference): coefficients, dtl,
itrBC3D, lrhs3D, both assembled factors and two steps at 1e-12 relative (max
norm); the refined step (complex-float factors, the operator regenerated in
double) reaches the complex-double increment.  A grid whose symmetric rows
would read off the grid in the reference (ny > nx) is refused."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("case", ["left7", "both7", "mse"])
def test_symmetry_against_reference(case):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lns3d_amd.config import Params
    from lns3d_amd.native import Step
    z = np.load(os.path.join(GOLD, f"ref_sym_{case}.npz"), allow_pickle=False)
    p = Params(**json.loads(str(z["params"])))
    print(f"symmetry {case}: left={p.left} right={p.right} top={p.top} {p.nx}x{p.ny}", flush=True)
    met, vm = np.asfortranarray(z["met"]), np.asfortranarray(z["vm"])

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to("cuda")

    def host(t, shape):
        return t.cpu().numpy().reshape(shape, order="F")

    st = Step(p)
    st.setup(met, vm)
    cg, dtg = st.get_coeffs()
    for k in ("Ah", "Bh", "Dh", "Dhi", "Vh11", "Vh12", "Vh22"):
        assert rel(cg[k], z["c_" + k]) < TOL, k
    assert rel(cg["ABhi"][:4], z["c_ABhi"][:4]) < TOL
    assert rel(dtg, z["dtl"]) < TOL
    if p.top == 1:
        st.set_top_forcing(z["x"], None, float(z["x0"]))
    v0 = z["v0"]
    vd = dev(v0)
    st.itrbc3d(vd)
    assert rel(host(vd, v0.shape), z["itrbc"]) < TOL
    vd = dev(v0)
    rd = torch.zeros_like(vd)
    st.lrhs3d(vd, rd)
    assert rel(host(rd, v0.shape), z["lrhs"]) < TOL
    for d in (1, 2):
        m_dev = torch.zeros(125 * p.nx * p.ny, dtype=torch.complex128, device="cuda")
        st.assemble(d, m_dev)
        assert rel(host(m_dev, z[f"lhs{d}"].shape), z[f"lhs{d}"]) < TOL, d
    st.factor()
    vd = dev(z["itrbc"])
    dvd = torch.zeros_like(vd)
    for k in (1, 2):
        st.step(vd, dvd)
        e = rel(host(dvd, v0.shape), z[f"dv{k}"])
        print(f"  step {k}: dv rel {e:.2e}", flush=True)
        assert e < TOL, k
    assert rel(host(vd, v0.shape), z["v2"]) < TOL
    st.close()

    # complex-float factors with refinement reach the same increment
    st32 = Step(p, precision=1)
    st32.setup(met, vm)
    if p.top == 1:
        st32.set_top_forcing(z["x"], None, float(z["x0"]))
    st32.factor()
    vb = dev(z["itrbc"])
    dvb = torch.zeros_like(vb)
    its, res = st32.step_refined(vb, dvb, tol=1e-12, max_iters=10)
    e = rel(host(dvb, v0.shape), z["dv1"])
    print(f"  refined: iters {its} resid {res:.2e} dv rel diff {e:.2e}", flush=True)
    assert res <= 1e-12 and e < 1e-10
    st32.close()


def test_symmetry_refuses_tall_grid():
    from lns3d_amd.config import Params
    from lns3d_amd.native import Step, LNS3DError
    with pytest.raises(LNS3DError):
        Step(Params(nx=9, ny=13, left=7))
    with pytest.raises(LNS3DError):
        Step(Params(nx=12, ny=13, right=7))
    Step(Params(nx=13, ny=13, left=7, right=7)).close()
