"""The Fortran host layer (liblns3d_fortran.so, built from
lns3d_amd/fortran/lns3d_hip.f90 by flang) called the way Fortran code calls
it: every argument by reference.  The drop-in cpenta1bc_/cpenta2bc_/
cpenta1p_/cpenta2p_ externals solve as the oracle does (1e-12), a solve
(code 1) before any factor reports LNS3D_ENOTFACTORED through the status
query, and the module's host-array driver (init, put, advance, get) runs the
implicit step as the oracle does."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLIB = os.path.join(ROOT, "lns3d_amd", "liblns3d_fortran.so")
TOL = 1e-12


@pytest.fixture(scope="module")
def flib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(FLIB):
        pytest.fail("liblns3d_fortran.so not built (flang missing at build time?)")
    from lns3d_amd.native import lib
    lib()                                   # the C library first (RTLD_GLOBAL not needed)
    return ctypes.CDLL(FLIB)


def I(v):
    return ctypes.byref(ctypes.c_int(v))


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def status(flib):
    f = flib.lns3d_hip_status_f_
    f.restype = ctypes.c_int
    return f()


@pytest.mark.parametrize("layout", [1, 2])
def test_dropin_bc(flib, orc, layout):
    from tests.penta_cases import random_lines, relerr
    nsys, n = 6, 21
    print(f"fortran drop-in bc layout={layout}", flush=True)
    mat, rhs = random_lines(nsys, n, layout, seed=40 + layout)
    mref, xref = mat.copy(order="F"), rhs.copy(order="F")
    orc.cpenta_bc(layout, mref, xref, 2)
    m, x = mat.copy(order="F"), rhs.copy(order="F")
    if layout == 1:
        flib.cpenta1bc_(I(nsys), I(n), I(5), P(m), P(x), I(0))
        assert status(flib) == 0
        flib.cpenta1bc_(I(nsys), I(n), I(5), P(m), P(x), I(1))
    else:
        flib.cpenta2bc_(I(n), I(nsys), I(5), P(m), P(x), I(2))
    assert status(flib) == 0
    assert relerr(x, xref) < TOL
    assert np.array_equal(m, mat)           # factors stay on the device


@pytest.mark.parametrize("layout", [1, 2])
def test_dropin_periodic(flib, orc, layout):
    from tests.penta_cases import random_lines, relerr
    nsys, np_ = 4, 18
    print(f"fortran drop-in periodic layout={layout}", flush=True)
    mat, rhs = random_lines(nsys, np_, layout, seed=60 + layout)
    shp = (2, 5, 5) + mat.shape[3:]
    mref, xref = mat.copy(order="F"), rhs.copy(order="F")
    orc.cpenta_p(layout, mref, xref, np.zeros(shp, complex, order="F"),
                 np.zeros(shp, complex, order="F"), 2, exact=0)
    m, x = mat.copy(order="F"), rhs.copy(order="F")
    per, per2 = np.zeros(shp, complex, order="F"), np.zeros(shp, complex, order="F")
    if layout == 1:
        flib.cpenta1p_(I(nsys), I(np_), I(5), P(m), P(x), P(per), P(per2), I(2))
    else:
        flib.cpenta2p_(I(np_), I(nsys), I(5), P(m), P(x), P(per), P(per2), I(2))
    assert status(flib) == 0
    assert relerr(x, xref) < TOL


def test_dropin_solve_before_factor(flib):
    from lns3d_amd.native import lib
    mat = np.zeros((5, 5, 5, 3, 9), complex, order="F")
    rhs = np.zeros((5, 3, 9), complex, order="F")
    flib.cpenta1bc_(I(3), I(9), I(5), P(mat), P(rhs), I(1))
    assert status(flib) == 3                # LNS3D_ENOTFACTORED
    assert b"before factor" in lib().lns3d_last_error()


class FParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("nx", "ny", "optx", "opty", "xper", "yper", "wall", "wallt", "top", "left",
                 "right", "linear", "mattyp", "loctime")] + \
               [(n, ctypes.c_double) for n in
                ("Ma", "Re", "Pr", "gamma", "kz", "omega", "cflmax", "dtmax", "alfa")] + \
               [("periodic_exact", ctypes.c_int), ("precision", ctypes.c_int)] + \
               [(n, ctypes.c_int) for n in ("wall_is", "wall_ie", "idamp", "use_calcd", "update_lhs")]


def test_module_driver_steps(flib, orc):
    """lns3d_hip_init/put/advance/get (module procedures, flang's external
    names) against three oracle steps at 33x17"""
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from tests.penta_cases import relerr
    p = Params(nx=33, ny=17, wall=1, wallt=1)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v0 = make_disturbance(p, x, seed=5, amp=1.0)
    fp = FParams()
    for name, _ in FParams._fields_:
        setattr(fp, name, 0 if name == "precision" else getattr(p, name, 0))
    fn = lambda s: getattr(flib, "_QMlns3d_hipPlns3d_hip_" + s)
    print("fortran module driver 33x17", flush=True)
    fn("init")(ctypes.byref(fp), P(met), P(vm))
    assert status(flib) == 0
    v = v0.copy(order="F")
    fn("put")(P(v))
    dv = np.zeros_like(v, order="F")
    fn("advance")(I(3), P(dv))
    assert status(flib) == 0
    fn("get")(P(v))
    fn("free")()
    c = orc.genmtrx(p, vm, met)
    dtl = orc.set_time_step(p, vm, met)
    vo = v0.copy(order="F")
    orc.itrbc3d(p, vo)
    f = orc.factor(p, c, dtl)
    for _ in range(3):
        dvo = orc.step(p, c, dtl, f, vo)
    assert relerr(dv, dvo) < TOL
    assert relerr(v, vo) < TOL


def test_module_driver_scheme(flib, orc):
    """lns3d_hip_scheme(impl = 2, niter = 2) then three steps through the
    module driver, against the oracle's two-step backward corrector"""
    import dataclasses
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    from tests.penta_cases import relerr
    p = Params(nx=33, ny=17, wall=1, wallt=1)
    x, y, met = make_grid(p)
    vm = make_mean(p, y)
    v0 = make_disturbance(p, x, seed=6, amp=1.0)
    fp = FParams()
    for name, _ in FParams._fields_:
        setattr(fp, name, 0 if name == "precision" else getattr(p, name, 0))
    fn = lambda s: getattr(flib, "_QMlns3d_hipPlns3d_hip_" + s)
    print("fortran module driver 33x17 impl 2 niter 2", flush=True)
    fn("init")(ctypes.byref(fp), P(met), P(vm))
    assert status(flib) == 0
    fn("scheme")(I(2), I(2))
    assert status(flib) == 0
    v = v0.copy(order="F")
    fn("put")(P(v))
    dv = np.zeros_like(v, order="F")
    fn("advance")(I(3), P(dv))
    assert status(flib) == 0
    fn("get")(P(v))
    fn("scheme")(I(3), I(2))            # impl 3 takes one iteration: refused at the step
    assert status(flib) == 0
    fn("advance")(I(1), P(dv))
    assert status(flib) != 0
    fn("free")()
    c = orc.genmtrx(p, vm, met)
    dtl = orc.set_time_step(p, vm, met)
    vo = v0.copy(order="F")
    orc.itrbc3d(p, vo)
    f = orc.factor(p, c, dtl)
    v2old = None
    for k in range(3):
        if k == 1:
            f = orc.factor(dataclasses.replace(p, alfa=6.6666666666666666666e-1), c, dtl)
        start = vo.copy(order="F")
        pk = p if k == 0 else dataclasses.replace(p, alfa=6.6666666666666666666e-1)
        dvo = orc.step_iter(pk, c, dtl, f, vo, 2, v2old)
        v2old = start
    assert relerr(v, vo) < TOL
    assert np.abs(dv - dvo).max() < TOL * np.abs(vo).max()


@pytest.mark.parametrize("case", ["sponge_3", "top1", "smooth_base", "inflow_left4", "bump_t1",
                                  "sutherland"])
def test_module_driver_deck_extras(flib, case):
    """the Fortran module driver with a deck's extras (lns3d_hip_sponge,
    lns3d_hip_top, lns3d_hip_smoother, lns3d_hip_inflow, then
    lns3d_hip_factor) against one of the reference's steps
    (tests/golden/ref_<case>.npz)"""
    import json
    import os
    from lns3d_amd.config import Params
    from tests.penta_cases import relerr
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    z = np.load(os.path.join(gold, f"ref_{case}.npz"), allow_pickle=False)
    p = Params(**json.loads(str(z["params"])))
    print(f"fortran module driver {case} {p.nx}x{p.ny}", flush=True)
    fp = FParams()
    for name, _ in FParams._fields_:
        setattr(fp, name, 0 if name == "precision" else getattr(p, name, 0))
    fn = lambda s: getattr(flib, "_QMlns3d_hipPlns3d_hip_" + s)
    D = lambda v: ctypes.byref(ctypes.c_double(v))
    met, vm = np.asfortranarray(z["met"]), np.asfortranarray(z["vm"])
    if case == "sutherland":
        fn("sutherland")(D(float(z["S0"])))
    fn("init")(ctypes.byref(fp), P(met), P(vm))
    assert status(flib) == 0
    if case == "sutherland":
        pass
    elif case.startswith("bump"):
        fn("wall")(P(np.asfortranarray(z["uw"])))
    elif case.startswith("sponge"):
        x = np.asfortranarray(z["x"])
        fn("sponge")(I(int(z["ispg"])), P(np.asfortranarray(z["spg"])), P(np.asfortranarray(z["spg2"])),
                     P(x), P(np.ascontiguousarray(z["wamp"])), D(float(z["x0"])))
    elif case == "top1":
        fn("top")(P(np.asfortranarray(z["x"])), P(np.ones(p.nx)), D(float(z["x0"])))
    elif case.startswith("smooth"):
        fn("smoother")(D(float(z["eps_e"])), P(np.asfortranarray(z["buff"])))
    else:
        la = np.array([complex(z["lalpha"]).real, complex(z["lalpha"]).imag])
        fn("inflow")(P(np.asfortranarray(z["x"])), D(float(z["x0"])), P(np.asfortranarray(z["eig"])), P(la))
    assert status(flib) == 0
    fn("factor")()
    assert status(flib) == 0
    v = np.asfortranarray(z["itrbc"].copy())
    fn("put")(P(v))
    dv = np.zeros_like(v, order="F")
    fn("advance")(I(1), P(dv))
    assert status(flib) == 0
    fn("get")(P(v))
    fn("free")()
    assert relerr(dv, z["dv1"]) < TOL
