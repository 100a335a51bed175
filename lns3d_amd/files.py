"""The reference's data files through the library's readers and writers
(lns3d_read_* / lns3d_write_* in include/lns3d_amd.h): restart fields
(src/impdrv3d.f90:393-399, src/genini3d.f90:26-57), mean fields
(src/genmean.f90:28-33), grids (src/gengrid.f90:17-32) and mesh metrics
(src/gengrid.f90:76-80).  Arrays are numpy arrays in Fortran order with the
reference's shapes."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from .native import CI, check, lib


class _Header(ctypes.Structure):
    _fields_ = [("lstep", ctypes.c_int), ("time", ctypes.c_double), ("nx", ctypes.c_int),
                ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("ndof", ctypes.c_int),
                ("Re", ctypes.c_double), ("Ma", ctypes.c_double), ("Pr", ctypes.c_double),
                ("gamma", ctypes.c_double), ("cv", ctypes.c_double)]


@dataclass
class FieldHeader:
    """first record of restart and mean files"""
    lstep: int
    time: float
    nx: int
    ny: int
    nz: int
    ndof: int
    Re: float
    Ma: float
    Pr: float
    gamma: float
    cv: float

    def _c(self) -> _Header:
        return _Header(self.lstep, self.time, self.nx, self.ny, self.nz, self.ndof, self.Re, self.Ma,
                       self.Pr, self.gamma, self.cv)

    @classmethod
    def _from(cls, h: _Header) -> "FieldHeader":
        return cls(*(getattr(h, f) for f, _ in _Header._fields_))


def _path(p) -> bytes:
    return str(p).encode()


def _shape(h: FieldHeader):
    return (h.ndof, h.nx, h.ny) if h.nz == 1 else (h.ndof, h.nx, h.ny, h.nz)


def _farr(a, dtype):
    a = np.asarray(a, dtype=dtype)
    return np.asfortranarray(a)


def read_header(path) -> FieldHeader:
    h = _Header()
    check(lib().lns3d_read_header(_path(path), ctypes.byref(h)), "lns3d_read_header")
    return FieldHeader._from(h)


def read_restart(path, with_vold: bool = False):
    """(header, v[, vold]) of a restart file; v complex (ndof,nx,ny).  With
    with_vold, vold is None when the file holds no previous field."""
    hd = read_header(path)
    h = _Header()
    v = np.empty(_shape(hd), dtype=np.complex128, order="F")
    vold = np.empty_like(v, order="F") if with_vold else None
    got = CI(0)
    check(lib().lns3d_read_restart(_path(path), ctypes.byref(h), v.ctypes.data,
                                   vold.ctypes.data if vold is not None else None, ctypes.byref(got)),
          "lns3d_read_restart")
    if with_vold:
        return FieldHeader._from(h), v, (vold if got.value else None)
    return FieldHeader._from(h), v


def write_restart(path, header: FieldHeader, v, vold=None) -> None:
    v = _farr(v, np.complex128)
    vo = _farr(vold, np.complex128) if vold is not None else None
    n = header.ndof * header.nx * header.ny * header.nz
    if v.size != n or (vo is not None and vo.size != n):
        raise ValueError("field size does not match the header")
    h = header._c()
    check(lib().lns3d_write_restart(_path(path), ctypes.byref(h), v.ctypes.data,
                                    vo.ctypes.data if vo is not None else None), "lns3d_write_restart")


def read_mean(path):
    """(header, vm) of a mean-field file; vm real (ndof,nx,ny)"""
    hd = read_header(path)
    h = _Header()
    vm = np.empty(_shape(hd), dtype=np.float64, order="F")
    check(lib().lns3d_read_mean(_path(path), ctypes.byref(h), vm.ctypes.data), "lns3d_read_mean")
    return FieldHeader._from(h), vm


def write_mean(path, header: FieldHeader, vm) -> None:
    vm = _farr(vm, np.float64)
    if vm.size != header.ndof * header.nx * header.ny * header.nz:
        raise ValueError("field size does not match the header")
    h = header._c()
    check(lib().lns3d_write_mean(_path(path), ctypes.byref(h), vm.ctypes.data), "lns3d_write_mean")


def read_grid(path):
    """(x, y) of a grid file, each (nx,ny) (or (nx,ny,nz) when nz > 1)"""
    nx, ny, nz = CI(0), CI(0), CI(0)
    check(lib().lns3d_read_grid(_path(path), ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz), None, None),
          "lns3d_read_grid")
    shp = (nx.value, ny.value) if nz.value == 1 else (nx.value, ny.value, nz.value)
    x = np.empty(shp, dtype=np.float64, order="F")
    y = np.empty(shp, dtype=np.float64, order="F")
    check(lib().lns3d_read_grid(_path(path), ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz),
                                x.ctypes.data, y.ctypes.data), "lns3d_read_grid")
    return x, y


def write_grid(path, x, y) -> None:
    x, y = _farr(x, np.float64), _farr(y, np.float64)
    if x.shape != y.shape or x.ndim not in (2, 3):
        raise ValueError("x and y must have one (nx,ny[,nz]) shape")
    nz = x.shape[2] if x.ndim == 3 else 1
    check(lib().lns3d_write_grid(_path(path), x.shape[0], x.shape[1], nz, x.ctypes.data, y.ctypes.data),
          "lns3d_write_grid")


def read_metric(path, nx: int, ny: int):
    """metric (nx,ny,10): m1, m2, n1, n2, m11, m12, m22, n11, n12, n22"""
    m = np.empty((nx, ny, 10), dtype=np.float64, order="F")
    check(lib().lns3d_read_metric(_path(path), nx, ny, m.ctypes.data), "lns3d_read_metric")
    return m


def write_metric(path, metric) -> None:
    m = _farr(metric, np.float64)
    if m.ndim != 3 or m.shape[2] != 10:
        raise ValueError("metric must be (nx,ny,10)")
    check(lib().lns3d_write_metric(_path(path), m.shape[0], m.shape[1], m.ctypes.data), "lns3d_write_metric")
