"""The reference's data files (restart, mean, grid, metric) through the
library's readers and writers (lns3d_amd/csrc/fileio.cpp).

Pins, without a GPU:
* files written by the Fortran runtime itself (tests/golden/fortran_files.f90,
  built with flang here; the record structure of src/impdrv3d.f90:393-399,
  src/genmean.f90:28-33, src/gengrid.f90:17-32, src/gengrid.f90:76-80) decode
  to the closed forms the program wrote, and the library's writers reproduce
  those files byte for byte;
* the reference's own fixtures (test/mse-recep: complex.R.0, mean.R.0,
  metric.dat), when the reference tree is present: headers agree with the
  input decks next to them, the mean flow's spanwise velocity is zero (a 2D
  mean flow: pins the component-fastest layout), the ten metric arrays obey
  the chain rule for second derivatives (xi_xx = xi_x d(xi_x)/dxi +
  eta_x d(xi_x)/deta, ...; a swapped pair fails by O(1)), and rewriting each
  file gives the same bytes;
* records longer than a subrecord are joined, and malformed files are
  rejected with an error."""
import hashlib
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from lns3d_amd import files
from lns3d_amd.native import LNS3DError

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.environ.get("LNS3D_REFERENCE", "/root/reference"), "test", "mse-recep")


def _flang():
    for c in (shutil.which("flang"), "/opt/rocm/lib/llvm/bin/flang"):
        if c and os.path.exists(c):
            return c
    return None


@pytest.fixture(scope="module")
def fortran_dir(tmp_path_factory):
    fl = _flang()
    if fl is None:
        pytest.skip("flang not available")
    d = tmp_path_factory.mktemp("ffiles")
    exe = str(d / "wf")
    subprocess.run([fl, "-O1", "-o", exe, os.path.join(HERE, "golden", "fortran_files.f90")], check=True)
    subprocess.run([exe], cwd=str(d), check=True)
    return d


NX, NY = 7, 5


def _expected():
    e, i, j = np.meshgrid(np.arange(1, 6), np.arange(1, NX + 1), np.arange(1, NY + 1), indexing="ij")
    v = (e + 10 * i + 100 * j) - 1j * (e * i + j) / 8.0
    vold = -0.5 * e + 1j * (i + 0.25 * j)
    vm = e * 1000.0 + i * 10.0 + j + 0.5
    ii, jj = np.meshgrid(np.arange(1, NX + 1), np.arange(1, NY + 1), indexing="ij")
    x = 1.5 * ii + 0.25 * jj
    y = 0.75 * jj - 0.125 * ii
    mt = np.stack([100.0 * k + ii + jj / 16.0 for k in range(1, 11)], axis=2)
    return v, vold, vm, x, y, mt


def _sha(p):
    return hashlib.sha256(open(p, "rb").read()).hexdigest()


def test_fortran_written_files_decode(fortran_dir):
    v, vold, vm, x, y, mt = _expected()
    h, gv, gvo = files.read_restart(fortran_dir / "restart.R", with_vold=True)
    assert (h.lstep, h.nx, h.ny, h.nz, h.ndof) == (1234, NX, NY, 1, 5)
    assert (h.time, h.Re, h.Ma, h.Pr, h.gamma, h.cv) == (5.25, 2400.0, 0.1, 0.72, 1.4, 716.5)
    assert np.array_equal(gv, v) and np.array_equal(gvo, vold)
    hm, gm = files.read_mean(fortran_dir / "mean.dat")
    assert hm == h and np.array_equal(gm, vm)
    gx, gy = files.read_grid(fortran_dir / "grid.dat")
    assert np.array_equal(gx, x) and np.array_equal(gy, y)
    assert np.array_equal(files.read_metric(fortran_dir / "metric.dat", NX, NY), mt)


def test_writers_match_fortran_bytes(fortran_dir, tmp_path):
    v, vold, vm, x, y, mt = _expected()
    h = files.read_header(fortran_dir / "restart.R")
    files.write_restart(tmp_path / "restart.R", h, v, vold)
    files.write_mean(tmp_path / "mean.dat", h, vm)
    files.write_grid(tmp_path / "grid.dat", x, y)
    files.write_metric(tmp_path / "metric.dat", mt)
    for f in ("restart.R", "mean.dat", "grid.dat", "metric.dat"):
        assert _sha(tmp_path / f) == _sha(fortran_dir / f), f


def test_restart_without_vold(tmp_path):
    v = _expected()[0]
    h = files.FieldHeader(7, 0.5, NX, NY, 1, 5, 1000.0, 0.5, 0.72, 1.4, 716.5)
    files.write_restart(tmp_path / "r", h, v)
    h2, v2, vo = files.read_restart(tmp_path / "r", with_vold=True)
    assert h2 == h and np.array_equal(v2, v) and vo is None
    # framing: 68-byte header record, then 16 * 5 * nx * ny bytes
    b = open(tmp_path / "r", "rb").read()
    n = 16 * 5 * NX * NY
    assert len(b) == 8 + 68 + 8 + n
    assert struct.unpack("<i", b[:4])[0] == 68 and struct.unpack("<i", b[72:76])[0] == 68
    assert struct.unpack("<i", b[76:80])[0] == n and struct.unpack("<i", b[-4:])[0] == n


def test_subrecords_joined(tmp_path):
    # one record of 40 doubles split after the third into two subrecords: a
    # negative leading marker continues the record, a negative trailing
    # marker continues one
    m = np.arange(1.0, 1.0 + 2 * 2 * 10)     # nx = ny = 2 metric: 40 doubles
    raw = m.tobytes()
    a, b = raw[:24], raw[24:]
    blob = (struct.pack("<i", -len(a)) + a + struct.pack("<i", len(a)) +
            struct.pack("<i", len(b)) + b + struct.pack("<i", -len(b)))
    p = tmp_path / "metric.dat"
    p.write_bytes(blob)
    got = files.read_metric(p, 2, 2)
    assert np.array_equal(got.reshape(-1, order="F"), m)


def test_malformed_files_rejected(tmp_path):
    v = _expected()[0]
    h = files.FieldHeader(1, 0.0, NX, NY, 1, 5, 1.0, 0.5, 0.72, 1.4, 1.0)
    files.write_restart(tmp_path / "r", h, v)
    b = open(tmp_path / "r", "rb").read()
    (tmp_path / "trunc").write_bytes(b[:-10])
    with pytest.raises(LNS3DError):
        files.read_restart(tmp_path / "trunc")
    bad = bytearray(b)
    bad[-4:] = struct.pack("<i", 12345)                   # trailing marker differs
    (tmp_path / "bad").write_bytes(bytes(bad))
    with pytest.raises(LNS3DError):
        files.read_restart(tmp_path / "bad")
    with pytest.raises(LNS3DError):                        # metric of the wrong size
        files.read_metric(tmp_path / "r", NX, NY)
    with pytest.raises(LNS3DError):
        files.read_header(tmp_path / "missing")


# ---- the reference's fixtures (this container only)

need_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference fixtures not present")


def _deck(name):
    """Ma, Re, Pr of an lns3d input deck (first, third and fifth lines)"""
    lines = [l.split("!")[0].strip() for l in open(os.path.join(REF, name))]
    return float(lines[0]), float(lines[2]), float(lines[4])


@need_ref
def test_reference_restart_and_mean(tmp_path):
    for fname, deck in (("complex.R.0", "complex.inp"), ("mean.R.0", "mean.inp")):
        h = files.read_header(os.path.join(REF, fname))
        Ma, Re, Pr = _deck(deck)
        assert (h.nx, h.ny, h.nz, h.ndof) == (384, 127, 1, 5)
        assert (h.Ma, h.Re, h.Pr) == (Ma, Re, Pr)
    h, v = files.read_restart(os.path.join(REF, "complex.R.0"))
    assert np.all(np.isfinite(v.real)) and np.all(np.isfinite(v.imag))
    # no-slip, isothermal wall (complex.inp: wall = 0, wallt = 0): u, v, w, T vanish at j = 1
    assert np.abs(v[1:, :, 0]).max() == 0.0 and np.abs(v[0, :, 0]).max() > 0.0
    files.write_restart(tmp_path / "c", h, v)
    assert _sha(tmp_path / "c") == _sha(os.path.join(REF, "complex.R.0"))
    hm, vm = files.read_mean(os.path.join(REF, "mean.R.0"))
    assert np.abs(vm[3]).max() == 0.0                   # spanwise velocity of a 2D mean flow
    assert 0.9 < np.median(vm[0]) < 1.1 and 0.9 < np.median(vm[4]) < 1.1   # density, temperature ~ 1
    files.write_mean(tmp_path / "m", hm, vm)
    assert _sha(tmp_path / "m") == _sha(os.path.join(REF, "mean.R.0"))


@need_ref
def test_reference_metric_chain_rule(tmp_path):
    nx, ny = 384, 127
    m = files.read_metric(os.path.join(REF, "metric.dat"), nx, ny)
    m1, m2, n1, n2, m11, m12, m22, n11, n12, n22 = (m[:, :, k] for k in range(10))
    dxi, deta = 1.0 / (nx - 1), 1.0 / (ny - 1)      # src/gengrid.f90: uniform computational grid

    def dxi_(a):
        return (a[2:, 1:-1] - a[:-2, 1:-1]) / (2 * dxi)

    def deta_(a):
        return (a[1:-1, 2:] - a[1:-1, :-2]) / (2 * deta)

    def inner(a):
        return a[1:-1, 1:-1]

    pairs = {
        "m11": (m11, inner(m1) * dxi_(m1) + inner(n1) * deta_(m1)),
        "m12": (m12, inner(m2) * dxi_(m1) + inner(n2) * deta_(m1)),
        "m22": (m22, inner(m2) * dxi_(m2) + inner(n2) * deta_(m2)),
        "n11": (n11, inner(m1) * dxi_(n1) + inner(n1) * deta_(n1)),
        "n12": (n12, inner(m2) * dxi_(n1) + inner(n2) * deta_(n1)),
        "n22": (n22, inner(m2) * dxi_(n2) + inner(n2) * deta_(n2)),
    }
    for k, (a, b) in pairs.items():
        a = inner(a)
        assert np.abs(a - b).max() < 5e-3 * np.abs(a).max(), k
    # the check tells arrays apart: m22 in the place of m11 fails by O(1)
    assert np.abs(inner(m22) - pairs["m11"][1]).max() > 0.5 * np.abs(inner(m22)).max()
    files.write_metric(tmp_path / "metric.dat", m)
    assert _sha(tmp_path / "metric.dat") == _sha(os.path.join(REF, "metric.dat"))
