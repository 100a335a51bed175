"""CPU-side checks of the product library and its host layer (no GPU calls):
the C-ABI library loads and exports every function include/lns3d_amd.h
declares; the ctypes signature table covers the header; the parameter
struct has the header's layout; host-side input generation; the benchmark's
refusal of more than one GPU; the F3 finding against the compiled reference
where that was built."""
import ctypes
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lns3d_amd.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    return sorted(set(re.findall(r"\b(lns3d_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from lns3d_amd import native
    if not os.path.exists(native.LIB_PATH):
        from lns3d_amd.build import build
        build()
    return ctypes.CDLL(native.LIB_PATH)


def test_library_exports_header(lib):
    names = header_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_signature_table_covers_header():
    from lns3d_amd.native import SIGNATURES
    names = header_functions()
    assert sorted(SIGNATURES) == names


def test_error_query_without_gpu(lib):
    lib.lns3d_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.lns3d_last_error(), bytes)


def test_params_struct_layout():
    """CParams mirrors lns3d_params: 14 ints, 9 doubles, 7 ints"""
    from lns3d_amd.native import CParams
    names = [f[0] for f in CParams._fields_]
    src = open(HEADER).read()
    body = re.search(r"typedef struct\s*{(.*?)}\s*lns3d_params;", src, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    decl = []
    for typ, rest in re.findall(r"\b(int|double)\s+([^;]+);", body):
        decl += [(typ, v.strip()) for v in rest.split(",")]
    assert [d[1] for d in decl] == names
    for (typ, name), (fname, ftype) in zip(decl, CParams._fields_):
        assert (ftype is ctypes.c_int) == (typ == "int"), name


def test_grid_and_mean():
    from lns3d_amd.config import Params, make_grid, make_mean, make_disturbance
    p = Params(nx=33, ny=17)
    x, y, met = make_grid(p)
    assert x.shape == (33, 17) and met.shape == (33, 17, 10)
    assert x.flags["F_CONTIGUOUS"] and met.flags["F_CONTIGUOUS"]
    assert y[0, 0] == 0.0 and abs(y[0, -1] - p.ymax) < 1e-12
    assert np.all(np.diff(y[0]) > 0)
    # the analytic metric n2 = d(eta)/dy against a centred difference on a
    # fine grid
    pf = Params(nx=9, ny=513)
    _, yf, metf = make_grid(pf)
    deta = 1.0 / (pf.ny - 1)
    fd = 2 * deta / (yf[0, 2:] - yf[0, :-2])
    assert np.abs(fd / metf[0, 1:-1, 3] - 1).max() < 2e-3
    vm = make_mean(p, y)
    assert np.allclose(vm[0] * vm[4], 1.0)
    assert vm[1][0, 0] == 0.0
    v = make_disturbance(p, x, seed=1)
    assert v.shape == (5, 33, 17) and np.iscomplexobj(v)
    assert np.array_equal(v, make_disturbance(p, x, seed=1))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_native_fails_loudly_without_library(tmp_path, monkeypatch):
    from lns3d_amd import native
    monkeypatch.setattr(native, "_lib", None)
    monkeypatch.setattr(native, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(native.LNS3DError):
        native.lib()


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref.so")),
                    reason="compiled reference not built (reference sources absent)")
def test_f3_shipped_index_slip():
    """cpenta1bc as shipped (line 675 uses the wrong row index in the forward
    sweep of the last row) does not solve the system; with that one index
    corrected it agrees with a dense solve, and so does the oracle"""
    from oracle.oracle import Oracle, Reference
    from tests.penta_cases import random_lines, relerr, dense_bc
    R, O = Reference(), Oracle()
    nsys, n = 2, 10
    mat, rhs = random_lines(nsys, n, 1, seed=77)
    xs = []
    for shipped in (False, True):
        m = mat.copy(order="F")
        # the shipped routine writes past the rhs array: give it room
        buf = np.zeros(rhs.size * 8, complex)
        buf[:rhs.size] = rhs.reshape(-1, order="F")
        R._call("cpenta1bc_shipped_" if shipped else "cpenta1bc_", nsys, n, 5, m, buf, 2)
        xs.append(buf[:rhs.size].reshape(rhs.shape, order="F"))
    mo, xo = mat.copy(order="F"), rhs.copy(order="F")
    O.cpenta_bc(1, mo, xo, 2)
    for s in range(nsys):
        xd = np.linalg.solve(dense_bc(mat[:, :, :, s, :]), rhs[:, s, :].T.reshape(-1)).reshape(n, 5).T
        assert relerr(xs[0][:, s, :], xd) < 1e-13
        assert relerr(xo[:, s, :], xd) < 1e-13
        assert relerr(xs[1][:, s, :], xd) > 1e-6


def test_request_and_instance_draws():
    """serving requests and batch instances: seeded, in range, Poisson
    arrivals with the requested mean rate"""
    from lns3d_amd.serve import draw_requests
    from lns3d_amd.batch import draw_instances
    a = draw_requests(400, 50.0, seed=3)
    b = draw_requests(400, 50.0, seed=3)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 50.0) < 0.15 / 50.0
    assert all(0 <= r.beta <= 1 and 0.01 <= r.omega <= 0.2 and 50 <= r.Re <= 150 for r in a)
    assert all(r.arrival == 0.0 for r in draw_requests(5, None, seed=1))
    inst = draw_instances(64, seed=0)
    assert len(inst) == 64
    assert all(65 <= p.nx <= 513 and 33 <= p.ny <= 257 for p in inst)
    assert all(0.3 <= p.Ma <= 0.8 and 500 <= p.Re <= 3000 for p in inst)
    assert len({(p.nx, p.ny) for p in inst}) > 32


def test_serving_arrivals_and_fifo():
    """Poisson arrivals on a fixed request set keep its parameters, and the
    one-at-a-time baseline queues requests in arrival order"""
    from lns3d_amd.serve import draw_requests, with_arrivals, fifo_latencies
    reqs = draw_requests(50, None, seed=4)
    a = with_arrivals(reqs, 20.0, seed=9)
    assert [(r.nx, r.ny, r.beta, r.Re) for r in a] == [(r.nx, r.ny, r.beta, r.Re) for r in reqs]
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0) and abs(np.diff(np.r_[0.0, t]).mean() - 1 / 20.0) < 0.3 / 20.0
    # arrivals 0, 1, 1.5 with service 2, 1, 0.25: the second waits until 2,
    # the third until 3
    lat = fifo_latencies(np.array([2.0, 1.0, 0.25]), np.array([0.0, 1.0, 1.5]))
    assert np.allclose(lat, [2.0, 2.0, 1.75])
    # out of order arrivals are served in arrival order
    lat = fifo_latencies(np.array([1.0, 1.0]), np.array([5.0, 0.0]))
    assert np.allclose(lat, [1.0, 1.0])


@pytest.mark.parametrize("change,code,text", [
    (dict(Re=-1.0), 1, b"Re must be >= 0"),
    (dict(Ma=float("nan")), 1, b"must be finite"),
    (dict(kz=float("inf")), 1, b"must be finite"),
    (dict(Re=0.0, mattyp=1), 1, b"mattyp = 0 only"),
    (dict(wall_is=5, wall_ie=3), 1, b"wall extent"),
    (dict(wall_is=2, wall_ie=34), 1, b"wall extent"),
    (dict(wall_is=-1), 1, b"wall extent"),
    (dict(idamp=3), 1, b"idamp"),
    (dict(use_calcd=2), 1, b"use_calcd"),
])
def test_params_refused_before_any_device_call(lib, change, code, text):
    """lns3d_ctx_create checks the deck before it touches the device: values
    the reference defines differently, or that the library does not
    implement, come back as LNS3D_EINVAL (1) or LNS3D_EUNSUPPORTED (4) with
    a message, never as a context that computes something else (Re = 0 with
    constant viscosity is the reference's inviscid deck, src/input.f90:31-36)"""
    import dataclasses
    from lns3d_amd.config import Params
    from lns3d_amd.native import CParams
    change = dict(change)
    precision = change.pop("precision", 0)
    p = dataclasses.replace(Params(nx=33, ny=17), **change)
    cp = CParams.from_params(p, precision)
    out = ctypes.c_void_p()
    lib.lns3d_ctx_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.lns3d_last_error.restype = ctypes.c_char_p
    rc = lib.lns3d_ctx_create(ctypes.byref(cp), ctypes.byref(out))
    assert rc == code
    assert not out.value
    assert text in lib.lns3d_last_error()


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref_hip.so")),
                    reason="relinked harness not built (reference sources absent)")
def test_relinked_harness_takes_the_dropins():
    """oracle/_ref/libref_hip.so defines none of the routines the drop-ins
    replace: it imports those it calls, and liblns3d_fortran.so defines
    every one"""
    names = ["lrhs3d_", "lhs1f3d_", "lhs2f3d_", "cgrad_", "cgrad2_", "cpenta1bc_", "cpenta2bc_",
             "cpenta1p_", "cpenta2p_"]
    nm = lambda lib, opt: set(l.split()[-1] for l in subprocess.run(
        ["nm", "-D", opt, lib], capture_output=True, text=True, check=True).stdout.splitlines())
    hip = os.path.join(ROOT, "oracle", "_ref", "libref_hip.so")
    flib = os.path.join(ROOT, "lns3d_amd", "liblns3d_fortran.so")
    assert not set(names) & nm(hip, "--defined-only")
    # every one the reference's driver calls (cpenta2bc is not among them:
    # its ImpDrv3D calls cpenta1bc for both sweeps, src/impdrv3d.f90:35-43)
    assert set(names) - {"cpenta2bc_"} <= nm(hip, "--undefined-only")
    assert set(names) <= nm(flib, "--defined-only")
