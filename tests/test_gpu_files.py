"""Files through a device step: the metric, mean and restart files of a
small case are written and read back through the library's readers and
writers (lns3d_amd/csrc/fileio.cpp), the step runs on the device from the
read data, and its increment agrees with the oracle's step on the original
arrays to 1e-12 (relative, max norm); the stepped state written as a restart
reads back bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_step_from_files(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lns3d_amd import files
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from lns3d_amd.native import Step
    from oracle.oracle import Oracle
    p = Params(nx=33, ny=17, kz=0.3)
    print(f"files -> step {p.nx}x{p.ny}", flush=True)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v = make_disturbance(p, x, seed=5)
    O = Oracle()
    O.itrbc3d(p, v)                       # a state that satisfies the boundary conditions
    h = files.FieldHeader(0, 0.0, p.nx, p.ny, 1, 5, p.Re, p.Ma, p.Pr, p.gamma, 716.5)
    files.write_grid(tmp_path / "grid.dat", x, y)
    files.write_metric(tmp_path / "metric.dat", met)
    files.write_mean(tmp_path / "mean.dat", h, vm)
    files.write_restart(tmp_path / "output.R.0", h, v)
    gx, gy = files.read_grid(tmp_path / "grid.dat")
    assert np.array_equal(gx, x) and np.array_equal(gy, y)
    met2 = files.read_metric(tmp_path / "metric.dat", p.nx, p.ny)
    _, vm2 = files.read_mean(tmp_path / "mean.dat")
    h2, v2 = files.read_restart(tmp_path / "output.R.0")
    assert h2 == h

    c = O.genmtrx(p, vm, met)
    dtl = O.set_time_step(p, vm, met)
    f = O.factor(p, c, dtl)
    dvo = O.step(p, c, dtl, f, v.copy(order="F"))

    st = Step(p)
    st.setup(met2, vm2)
    st.factor()
    vd = torch.from_numpy(np.ascontiguousarray(v2.reshape(-1, order="F"))).to("cuda:0")
    dvd = torch.zeros_like(vd)
    st.step(vd, dvd)
    torch.cuda.synchronize()
    dv = dvd.cpu().numpy().reshape(v.shape, order="F")
    err = float(np.abs(dv - dvo).max() / np.abs(dvo).max())
    assert err < 1e-12, err
    vnew = vd.cpu().numpy().reshape(v.shape, order="F")
    files.write_restart(tmp_path / "output.R.1", files.FieldHeader(1, 1.0, p.nx, p.ny, 1, 5, p.Re, p.Ma,
                                                                    p.Pr, p.gamma, 716.5), vnew)
    h3, v3 = files.read_restart(tmp_path / "output.R.1")
    assert h3.lstep == 1 and np.array_equal(v3, vnew)
    st.close()
