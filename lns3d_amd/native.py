"""ctypes binding of liblns3d_amd.so (the C ABI in include/lns3d_amd.h).

The library is the product: every compute entry point here calls into it,
and loading fails loudly when it has not been built (there is no fallback).
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "liblns3d_amd.so")
FLIB_PATH = os.path.join(_PKG, "liblns3d_fortran.so")

VP = ctypes.c_void_p
CI = ctypes.c_int
CD = ctypes.c_double

# exported symbols and their argument types (kept in sync with the header)
SIGNATURES = {
    "lns3d_last_error": ([], ctypes.c_char_p),
    "lns3d_cpenta1bc": ([CI, CI, CI, VP, VP, CI], CI),
    "lns3d_cpenta2bc": ([CI, CI, CI, VP, VP, CI], CI),
    "lns3d_cpenta1p": ([CI, CI, CI, VP, VP, VP, VP, CI], CI),
    "lns3d_cpenta2p": ([CI, CI, CI, VP, VP, VP, VP, CI], CI),
    "lns3d_set_periodic_exact": ([CI], CI),
    "lns3d_release_factors": ([], CI),
    "lns3d_penta_create": ([CI, CI, CI, CI, CI, CI, ctypes.POINTER(VP)], CI),
    "lns3d_penta_factor": ([VP, VP, VP], CI),
    "lns3d_penta_solve": ([VP, VP, VP], CI),
    "lns3d_penta_solve_reference_order": ([VP, VP, VP], CI),
    "lns3d_penta_destroy": ([VP], CI),
}

_lib = None


class LNS3DError(RuntimeError):
    pass


def lib():
    """Load the native library (once)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LNS3DError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                             "or python lns3d_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().lns3d_last_error().decode(errors="replace")
        raise LNS3DError(f"{what}: error {rc}: {msg}")


def ptr(t) -> int:
    """device or host address of a torch tensor / numpy array"""
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


class Penta:
    """Device-resident factor-once / solve-many handle (lns3d_penta_*)."""

    def __init__(self, kind: int, layout: int, nsys: int, nrows: int, precision: int = 0,
                 exact: int = 0):
        h = VP()
        check(lib().lns3d_penta_create(kind, layout, nsys, nrows, precision, exact,
                                       ctypes.byref(h)), "lns3d_penta_create")
        self.h = h

    def factor(self, mat_dev, stream=None):
        check(lib().lns3d_penta_factor(self.h, ptr(mat_dev), stream), "lns3d_penta_factor")

    def solve(self, rhs_dev, stream=None):
        check(lib().lns3d_penta_solve(self.h, ptr(rhs_dev), stream), "lns3d_penta_solve")

    def solve_reference_order(self, rhs_dev, stream=None):
        check(lib().lns3d_penta_solve_reference_order(self.h, ptr(rhs_dev), stream),
              "lns3d_penta_solve_reference_order")

    def close(self):
        if self.h:
            lib().lns3d_penta_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
