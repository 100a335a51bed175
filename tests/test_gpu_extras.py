"""GPU parity of the &extras deck options and the inviscid deck against the
reference's own vectors (tests/golden/ref_extra_<case>.npz, written by
tests/golden/gen_golden.py from the compiled reference on a 13 x 9 grid whose
wall normals are not along y):

* Re = 0 with mattyp = 0 (src/input.f90:31-36, Navier = .false.): the rotated
  wall rows of src/rhsbc3d.f90:108-118, src/lhsbc1f3d.f90, src/lhsbc2f3d.f90,
  the tangency condition of src/itrbc3d.f90:120-132 and bcfix after the eta
  solve (src/impdrv3d.f90:333-336), over the whole wall and over is:ie;
* the wall extent is:ie (src/lns3d.f90:73) of the viscous and the bump wall;
* useCalcd (src/lhs1f3d.f90:27-37: the D terms in the eta factor, the sponge
  in the xi factor, alternating with istep - 1 + iter) and updateLHS
  (src/impdrv3d.f90:248,310: the factors formed again each step);
* idamp = 1 (src/damper.f90:26-29, src/genmtrx.f90:87-96, 877-881: the
  outflow column's mean xi second derivatives and Vh11 zeroed);
* right = 9 (src/rhsbc3d.f90:317-320, src/itrbc3d.f90:408-410: the right
  column extrapolated from the two inside it), periodic in y.

lrhs3D, itrBC3D, the assembled factors (where they do not depend on the step)
and two or three implicit steps, 1e-12 relative in the max norm."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["inviscid", "inviscid_isie", "wall_isie", "bump_isie", "calcd", "calcd_update", "calcd_sym", "idamp",
         "right9"]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_extras_against_reference(name):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lns3d_amd.config import Params
    from lns3d_amd.native import Step
    z = np.load(os.path.join(GOLD, f"ref_extra_{name}.npz"), allow_pickle=False)
    p = Params(**json.loads(str(z["params"])))
    print(f"extras {name}: {p.nx}x{p.ny} Re={p.Re} is:ie={p.wall_is}:{p.wall_ie} "
          f"calcd={p.use_calcd} update={p.update_lhs}", flush=True)
    met, vm = np.asfortranarray(z["met"]), np.asfortranarray(z["vm"])
    v0 = z["v0"]

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to("cuda")

    def host(t, shape):
        return t.cpu().numpy().reshape(shape, order="F")

    st = Step(p)
    st.setup(met, vm)
    if "spg" in z.files:
        st.set_sponge(1, z["spg"])
    if "uw" in z.files:
        st.set_wall(z["uw"])
    vd = dev(v0)
    rd = torch.zeros_like(vd)
    st.lrhs3d(vd, rd)
    assert rel(host(rd, v0.shape), z["lrhs"]) < TOL
    st.itrbc3d(vd)
    assert rel(host(vd, v0.shape), z["itrbc"]) < TOL
    st.factor()
    for d in (1, 2):
        if f"lhs{d}" in z.files:
            m_dev = torch.zeros(125 * p.nx * p.ny, dtype=torch.complex128, device="cuda")
            st.assemble(d, m_dev)
            assert rel(host(m_dev, z[f"lhs{d}"].shape), z[f"lhs{d}"]) < TOL, d
    dvd = torch.zeros_like(vd)
    nsteps = 3 if p.update_lhs else 2
    for k in range(1, nsteps + 1):
        st.step(vd, dvd)
        e_dv, e_v = rel(host(dvd, v0.shape), z[f"dv{k}"]), rel(host(vd, v0.shape), z[f"v{k}"])
        print(f"  step {k}: dv {e_dv:.2e} v {e_v:.2e}", flush=True)
        assert e_dv < TOL and e_v < TOL, k
    st.close()


@pytest.mark.parametrize("name", ["calcd_update", "inviscid", "inviscid_isie"])
def test_complex_float_refined_against_reference(name):
    """complex-float factors under the refined step (defect correction to
    1e-10 of the complex-double AF system, whose operator regenerates the
    blocks in double): useCalcd with updateLHS forms both precisions' factors
    again as the D terms change factor; the inviscid deck's operator applies
    the rotated wall rows of the assembled factors (src/lhsbc1f3d.f90,
    src/lhsbc2f3d.f90) and bcfix follows the solve.  The reference's steps
    to 1e-9"""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lns3d_amd.config import Params
    from lns3d_amd.native import Step
    z = np.load(os.path.join(GOLD, f"ref_extra_{name}.npz"), allow_pickle=False)
    p = Params(**json.loads(str(z["params"])))
    met, vm, v0 = np.asfortranarray(z["met"]), np.asfortranarray(z["vm"]), z["v0"]
    print(f"complex float {name}: {p.nx}x{p.ny} Re={p.Re} is:ie={p.wall_is}:{p.wall_ie}", flush=True)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to("cuda")

    st = Step(p, precision=1)
    st.setup(met, vm)
    if "spg" in z.files:
        st.set_sponge(1, z["spg"])
    if "uw" in z.files:
        st.set_wall(z["uw"])
    vd = dev(v0)
    st.itrbc3d(vd)
    st.factor()
    dvd = torch.zeros_like(vd)
    for k in range(1, (4 if p.update_lhs else 3)):
        iters, res = st.step_refined(vd, dvd, tol=1e-10)
        dv = dvd.cpu().numpy().reshape(v0.shape, order="F")
        e = rel(dv, z[f"dv{k}"])
        print(f"  step {k}: {iters} solves, residual {res:.1e}, dv {e:.2e}", flush=True)
        assert e < 1e-9, k
    st.close()


@pytest.mark.parametrize("corrections", [1, 0])
def test_staged_inviscid_matches_branches(corrections):
    """a staged graph (one launch per stage over all contexts) of two
    inviscid contexts (whole wall, is:ie) and one viscous one (is:ie): the
    inviscid members' rotated wall rows in the residual, the operator and
    itrBC3D, and bcfix between the last correction and the update; three
    steps give the states of the same steps on one branch per context, bit
    for bit"""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lns3d_amd.config import Params
    from lns3d_amd.native import Step, Graph, MODE_FP32, MODE_REFINED, MODE_BATCHED
    names = ["inviscid", "inviscid_isie", "wall_isie"]
    print(f"staged inviscid corrections={corrections}", flush=True)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to("cuda")

    steps, v0s = [], []
    for name in names:
        z = np.load(os.path.join(GOLD, f"ref_extra_{name}.npz"), allow_pickle=False)
        p = Params(**json.loads(str(z["params"])))
        st = Step(p, precision=1)
        st.setup(np.asfortranarray(z["met"]), np.asfortranarray(z["vm"]))
        if "spg" in z.files:
            st.set_sponge(1, z["spg"])
        if "uw" in z.files:
            st.set_wall(z["uw"])
        v = dev(z["v0"])
        st.itrbc3d(v)
        st.factor()
        steps.append(st)
        v0s.append(v)
    out = {}
    for mode in (MODE_REFINED if corrections else MODE_FP32, MODE_BATCHED):
        vs = [v.clone() for v in v0s]
        dvs = [torch.zeros_like(v) for v in v0s]
        g = Graph(steps, vs, dvs, nsteps=3, mode=mode, corrections=corrections)
        g.launch()
        g.synchronize()
        out[mode] = vs
        g.close()
    a, b = out.values()
    for k, name in enumerate(names):
        assert bool(torch.isfinite(b[k]).all()), name
        assert torch.equal(a[k], b[k]), name
    for st in steps:
        st.close()


def test_inviscid_refusals():
    """the inviscid deck needs the wall normals (lns3d_setup with the metric)
    before its factors and its boundary rows"""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lns3d_amd.config import Params
    from lns3d_amd.native import Step, LNS3DError
    z = np.load(os.path.join(GOLD, "ref_extra_inviscid.npz"), allow_pickle=False)
    p = Params(**json.loads(str(z["params"])))
    st = Step(p)
    c = {k[2:]: z[k] for k in z.files if k.startswith("c_")}
    st.set_coeffs(c, z["dtl"])
    with pytest.raises(LNS3DError):
        st.factor()
    vd = torch.zeros(5 * p.nx * p.ny, dtype=torch.complex128, device="cuda")
    with pytest.raises(LNS3DError):
        st.itrbc3d(vd)
    st.close()
