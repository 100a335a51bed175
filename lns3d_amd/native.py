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
    "lns3d_ctx_create": ([VP, ctypes.POINTER(VP)], CI),
    "lns3d_ctx_destroy": ([VP], CI),
    "lns3d_ctx_set_stream": ([VP, VP], CI),
    "lns3d_ctx_set_params": ([VP, VP], CI),
    "lns3d_ctx_synchronize": ([VP], CI),
    "lns3d_step_refined_fixed": ([VP, VP, VP, CI, CI], CI),
    "lns3d_last_residual": ([VP, ctypes.POINTER(CD)], CI),
    "lns3d_ctx_bytes": ([VP, ctypes.POINTER(ctypes.c_size_t)], CI),
    "lns3d_graph_create": ([VP, VP, VP, CI, CI, CI, CI, ctypes.POINTER(VP)], CI),
    "lns3d_graph_launch": ([VP, VP], CI),
    "lns3d_graph_synchronize": ([VP], CI),
    "lns3d_graph_destroy": ([VP], CI),
    "lns3d_tune_lines": ([CI, ctypes.c_long, CI, ctypes.POINTER(CI), ctypes.POINTER(ctypes.c_float)], CI),
    "lns3d_tune_set": ([CI, ctypes.c_long, CI, CI], CI),
    "lns3d_state_upload": ([VP, VP], CI),
    "lns3d_state_download": ([VP, VP], CI),
    "lns3d_advance": ([VP, CI, VP], CI),
    "lns3d_setup": ([VP, VP, VP], CI),
    "lns3d_set_coeffs": ([VP] * 10, CI),
    "lns3d_get_coeffs": ([VP] * 10, CI),
    "lns3d_factor": ([VP], CI),
    "lns3d_assemble": ([VP, CI, VP], CI),
    "lns3d_lrhs3d": ([VP, VP, VP], CI),
    "lns3d_itrbc3d": ([VP, VP], CI),
    "lns3d_solve": ([VP, VP], CI),
    "lns3d_solve_dir": ([VP, CI, VP], CI),
    "lns3d_step": ([VP, VP, VP], CI),
    "lns3d_step_refined": ([VP, VP, VP, CD, CI, ctypes.POINTER(CI), ctypes.POINTER(CD)], CI),
    "lns3d_step_fp32": ([VP, VP, VP], CI),
    "lns3d_set_sponge": ([VP, CI, VP, VP, VP, VP, CD], CI),
    "lns3d_set_top_forcing": ([VP, VP, VP, CD], CI),
    "lns3d_set_time_scheme": ([VP, CI], CI),
    "lns3d_set_time_history": ([VP, VP, ctypes.c_long], CI),
    "lns3d_set_corrector_iterations": ([VP, CI], CI),
    "lns3d_set_inflow": ([VP, VP, ctypes.c_double, VP, VP], CI),
    "lns3d_set_smoother": ([VP, ctypes.c_double, VP], CI),
    "lns3d_set_wall": ([VP, VP], CI),
    "lns3d_set_sutherland": ([VP, ctypes.c_double], CI),
    "lns3d_step_forced": ([VP, VP, VP, VP, VP], CI),
    "lns3d_forced_residual": ([VP, VP, VP, VP, VP, ctypes.POINTER(ctypes.c_double)], CI),
    "lns3d_residual_norm": ([VP, VP, VP], CI),
    "lns3d_ctx_set_residual_output": ([VP, VP, CI], CI),
    "lns3d_residual_norm_noalloc": ([VP, VP, ctypes.c_long, VP, VP], CI),
    "lns3d_set_module_state": ([VP], CI),
    "lns3d_drop_lrhs3d": ([VP] * 5, CI),
    "lns3d_drop_lhs1f3d": ([VP] * 8, CI),
    "lns3d_drop_lhs2f3d": ([VP] * 8, CI),
    "lns3d_drop_cgrad": ([CI, CI, CI, VP, VP, VP, CD, CD] + [CI] * 9, CI),
    "lns3d_drop_cgrad2": ([CI, CI, CI, VP, VP, VP, VP, VP, CD, CD] + [CI] * 9, CI),
    "lns3d_read_header": ([ctypes.c_char_p, VP], CI),
    "lns3d_read_restart": ([ctypes.c_char_p, VP, VP, VP, ctypes.POINTER(CI)], CI),
    "lns3d_write_restart": ([ctypes.c_char_p, VP, VP, VP], CI),
    "lns3d_read_mean": ([ctypes.c_char_p, VP, VP], CI),
    "lns3d_write_mean": ([ctypes.c_char_p, VP, VP], CI),
    "lns3d_read_grid": ([ctypes.c_char_p, ctypes.POINTER(CI), ctypes.POINTER(CI), ctypes.POINTER(CI), VP, VP], CI),
    "lns3d_write_grid": ([ctypes.c_char_p, CI, CI, CI, VP, VP], CI),
    "lns3d_read_metric": ([ctypes.c_char_p, CI, CI, VP], CI),
    "lns3d_write_metric": ([ctypes.c_char_p, CI, CI, VP], CI),
}

_lib = None


class LNS3DError(RuntimeError):
    """a non-zero status of the library; .code is the LNS3D_E* value"""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


EINVAL, EHIP, ENOTFACTORED, EUNSUPPORTED, ENONFINITE = 1, 2, 3, 4, 5


def lib():
    """Load the native library (once)."""
    global _lib
    if _lib is None:
        # torch carries its own HIP runtime beside ROCm's libamdhip64.so.7,
        # which this library links; a process that loaded this library first
        # and torch's runtime after it saw the library's hipMalloc fail, while
        # the other order works.  Load torch first where it is installed.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
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
        raise LNS3DError(f"{what}: error {rc}: {msg}", rc)


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


class CParams(ctypes.Structure):
    """lns3d_params of include/lns3d_amd.h"""
    _fields_ = [(n, CI) for n in ("nx", "ny", "optx", "opty", "xper", "yper", "wall", "wallt",
                                  "top", "left", "right", "linear", "mattyp", "loctime")] + \
               [(n, CD) for n in ("Ma", "Re", "Pr", "gamma", "kz", "omega", "cflmax", "dtmax",
                                  "alfa")] + \
               [("periodic_exact", CI), ("precision", CI)] + \
               [(n, CI) for n in ("wall_is", "wall_ie", "idamp", "use_calcd", "update_lhs")]

    @classmethod
    def from_params(cls, p, precision=0):
        s = cls()
        for name, _ in cls._fields_:
            if name == "precision":
                s.precision = precision
            else:
                setattr(s, name, getattr(p, name, 0))
        return s


class Step:
    """Device-resident implicit step context (lns3d_ctx_*)."""

    def __init__(self, p, precision: int = 0, stream=None):
        self.params = p
        self.cp = CParams.from_params(p, precision)
        h = VP()
        check(lib().lns3d_ctx_create(ctypes.byref(self.cp), ctypes.byref(h)), "lns3d_ctx_create")
        self.h = h
        if stream is not None:
            check(lib().lns3d_ctx_set_stream(self.h, stream), "lns3d_ctx_set_stream")

    def setup(self, metric, vm):
        import numpy as np
        met = np.asfortranarray(metric, dtype=np.float64)
        v = np.asfortranarray(vm, dtype=np.float64)
        check(lib().lns3d_setup(self.h, met.ctypes.data, v.ctypes.data), "lns3d_setup")

    def set_time_scheme(self, impl):
        """the deck's impl: 1 backward Euler, 2 two-step backward, 3 alfa 1/2"""
        check(lib().lns3d_set_time_scheme(self.h, impl), "lns3d_set_time_scheme")

    def set_inflow(self, x=None, x0=0.0, eig=None, lalpha=0j):
        """left = 1 / left = 4 / right = 4 boundary values: x (nx, ny), eig
        (10, ny) (rhor, rhoi, ur, ui, vr, vi, wr, wi, tr, ti), lalpha complex"""
        import numpy as np
        xa = None if x is None else np.asfortranarray(x, dtype=np.float64)
        ea = None if eig is None else np.asfortranarray(eig, dtype=np.float64)
        la = np.array([complex(lalpha).real, complex(lalpha).imag])
        hp = lambda a: None if a is None else a.ctypes.data
        check(lib().lns3d_set_inflow(self.h, hp(xa), float(x0), hp(ea), hp(la)), "lns3d_set_inflow")

    def step_forced(self, v_dev, f_dev, vold_dev=None, dv_dev=None):
        """one corrector of the shift-invert driver's pseudo-time steps
        (ebe3D forcing f); vold None: v"""
        check(lib().lns3d_step_forced(self.h, ptr(v_dev), None if vold_dev is None else ptr(vold_dev),
                                      ptr(f_dev), None if dv_dev is None else ptr(dv_dev)),
              "lns3d_step_forced")

    def forced_residual(self, v_dev, vold_dev, f_dev, r_dev):
        """the driver's residual into r_dev; returns sqrt(sum |r|^2 / dtl^2)"""
        out = ctypes.c_double(0.0)
        check(lib().lns3d_forced_residual(self.h, ptr(v_dev), ptr(vold_dev), ptr(f_dev), ptr(r_dev),
                                          ctypes.byref(out)), "lns3d_forced_residual")
        return out.value

    def residual_norm(self, r_dev, norm_dev):
        """sqrt(sum |r|^2 / dtl^2) of r_dev into the device double norm_dev"""
        check(lib().lns3d_residual_norm(self.h, ptr(r_dev), ptr(norm_dev)), "lns3d_residual_norm")

    def set_residual_output(self, norms_dev=None, count=0):
        """every step writes its residual norm to norms_dev[k mod count] (device
        float64 tensor); None stops"""
        check(lib().lns3d_ctx_set_residual_output(self.h, None if norms_dev is None else ptr(norms_dev),
                                                  int(count)), "lns3d_ctx_set_residual_output")
        self._norms_keep = norms_dev

    def set_sutherland(self, S0):
        """Sutherland's constant S0 = 110.4 K / Tref for mattyp = 1; before setup"""
        check(lib().lns3d_set_sutherland(self.h, float(S0)), "lns3d_set_sutherland")

    def set_wall(self, uw):
        """the bump wall values (wall = 3): uw (nx, 5) complex, u1w, u2w, u3w, tw, twp"""
        import numpy as np
        a = np.asfortranarray(uw, dtype=np.complex128)
        check(lib().lns3d_set_wall(self.h, a.ctypes.data), "lns3d_set_wall")

    def set_smoother(self, eps_e, buff=None):
        """smoother eps_e with the damping profile buff (nx, ny); 0: off"""
        import numpy as np
        b = None if buff is None else np.asfortranarray(buff, dtype=np.float64)
        check(lib().lns3d_set_smoother(self.h, float(eps_e), None if b is None else b.ctypes.data),
              "lns3d_set_smoother")

    def set_corrector_iterations(self, niter):
        check(lib().lns3d_set_corrector_iterations(self.h, niter), "lns3d_set_corrector_iterations")

    def set_time_history(self, vold_dev, lstep):
        check(lib().lns3d_set_time_history(self.h, ptr(vold_dev), lstep), "lns3d_set_time_history")

    def set_top_forcing(self, x, wamp=None, x0=0.0):
        """the top = 1 incoming wave (lns3d_set_top_forcing); x (nx,ny), wamp (nx)"""
        import numpy as np
        xa = np.asfortranarray(x, dtype=np.float64)
        wa = None if wamp is None else np.ascontiguousarray(wamp, dtype=np.float64)
        check(lib().lns3d_set_top_forcing(self.h, xa.ctypes.data, None if wa is None else wa.ctypes.data,
                                          float(x0)), "lns3d_set_top_forcing")

    def set_sponge(self, ispg, spg=None, spg2=None, x=None, wamp=None, x0=0.0):
        """sponge profiles (nx,ny), physical x (nx,ny) and wamp (nx); after
        setup, before factor (lns3d_set_sponge)"""
        import numpy as np
        arrs = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (spg, spg2, x, wamp)]
        check(lib().lns3d_set_sponge(self.h, ispg, *[None if a is None else a.ctypes.data for a in arrs],
                                     float(x0)), "lns3d_set_sponge")

    def set_coeffs(self, c, dtl):
        import numpy as np
        arrs = [np.asfortranarray(c[k], dtype=np.float64)
                for k in ("Ah", "Bh", "Dh", "Dhi", "Vh11", "Vh12", "Vh22", "ABhi")]
        d = np.asfortranarray(dtl, dtype=np.float64)
        check(lib().lns3d_set_coeffs(self.h, *[a.ctypes.data for a in arrs], d.ctypes.data),
              "lns3d_set_coeffs")

    def get_coeffs(self):
        import numpy as np
        nx, ny = self.params.nx, self.params.ny
        z = lambda *s: np.zeros(s, order="F")
        c = dict(Ah=z(5, 5, nx, ny), Bh=z(5, 5, nx, ny), Dh=z(5, 5, nx, ny), Dhi=z(5, 5, nx, ny),
                 Vh11=z(6, nx, ny), Vh12=z(6, nx, ny), Vh22=z(6, nx, ny), ABhi=z(6, nx, ny))
        dtl = z(nx, ny)
        check(lib().lns3d_get_coeffs(self.h, *[c[k].ctypes.data for k in
                                               ("Ah", "Bh", "Dh", "Dhi", "Vh11", "Vh12", "Vh22",
                                                "ABhi")], dtl.ctypes.data), "lns3d_get_coeffs")
        return c, dtl

    def factor(self):
        check(lib().lns3d_factor(self.h), "lns3d_factor")

    def assemble(self, direction, mat_dev):
        check(lib().lns3d_assemble(self.h, direction, ptr(mat_dev)), "lns3d_assemble")

    def lrhs3d(self, v_dev, r_dev):
        check(lib().lns3d_lrhs3d(self.h, ptr(v_dev), ptr(r_dev)), "lns3d_lrhs3d")

    def itrbc3d(self, v_dev):
        check(lib().lns3d_itrbc3d(self.h, ptr(v_dev)), "lns3d_itrbc3d")

    def solve(self, r_dev):
        check(lib().lns3d_solve(self.h, ptr(r_dev)), "lns3d_solve")

    def step(self, v_dev, dv_dev=None):
        check(lib().lns3d_step(self.h, ptr(v_dev), ptr(dv_dev) if dv_dev is not None else None),
              "lns3d_step")

    def step_fp32(self, v_dev, dv_dev=None):
        check(lib().lns3d_step_fp32(self.h, ptr(v_dev), ptr(dv_dev) if dv_dev is not None else None),
              "lns3d_step_fp32")

    def step_refined(self, v_dev, dv_dev=None, tol=1e-10, max_iters=8):
        it = CI(0)
        res = CD(0.0)
        check(lib().lns3d_step_refined(self.h, ptr(v_dev),
                                       ptr(dv_dev) if dv_dev is not None else None, tol, max_iters,
                                       ctypes.byref(it), ctypes.byref(res)), "lns3d_step_refined")
        return it.value, res.value

    def step_refined_fixed(self, v_dev, dv_dev=None, corrections=1, check_residual=False):
        check(lib().lns3d_step_refined_fixed(self.h, ptr(v_dev),
                                             ptr(dv_dev) if dv_dev is not None else None,
                                             corrections, 1 if check_residual else 0),
              "lns3d_step_refined_fixed")

    def last_residual(self):
        r = CD(0.0)
        check(lib().lns3d_last_residual(self.h, ctypes.byref(r)), "lns3d_last_residual")
        return r.value

    def set_params(self, p):
        """new physics for this context (same grid size and precision)"""
        cp = CParams.from_params(p, self.cp.precision)
        check(lib().lns3d_ctx_set_params(self.h, ctypes.byref(cp)), "lns3d_ctx_set_params")
        self.cp, self.params = cp, p

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream (or raw hipStream_t as int); None = default"""
        raw = getattr(stream, "cuda_stream", stream)
        check(lib().lns3d_ctx_set_stream(self.h, raw), "lns3d_ctx_set_stream")

    def synchronize(self):
        check(lib().lns3d_ctx_synchronize(self.h), "lns3d_ctx_synchronize")

    def close(self):
        if getattr(self, "h", None):
            lib().lns3d_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ctx_bytes(p, precision=0) -> int:
    """device bytes a factored context of these parameters holds"""
    cp = CParams.from_params(p, precision)
    b = ctypes.c_size_t(0)
    check(lib().lns3d_ctx_bytes(ctypes.byref(cp), ctypes.byref(b)), "lns3d_ctx_bytes")
    return int(b.value)


# line-solve variants the autotuner times (LNS3D_TUNE_VARIANTS)
TUNE_VARIANTS = 7
VARIANT_NAMES = ("1 line/WG", "4 lines/WG", "twisted", "2 lines/WG", "3 lines/WG", "twisted pairs",
                 "twisted half")


def tune_lines(n: int, lines: int, precision: int = 0):
    """time every line-solve variant on synthetic lines of this shape; the
    fastest is used by plans of that shape created afterwards.  Returns
    (variant, [ms per variant, -1 where not applicable])"""
    v = CI(-1)
    t = (ctypes.c_float * TUNE_VARIANTS)()
    check(lib().lns3d_tune_lines(n, lines, precision, ctypes.byref(v), t), "lns3d_tune_lines")
    return v.value, [float(x) for x in t]


def set_line_variant(n: int, lines: int, precision: int, variant: int) -> None:
    """use this solve variant for plans of the shape created afterwards
    (-1: the built-in choice again)"""
    check(lib().lns3d_tune_set(n, lines, precision, variant), "lns3d_tune_set")


MODE_FP64, MODE_FP32, MODE_REFINED = 0, 1, 2
# the refined steps of MODE_REFINED with one launch per stage over all
# contexts of the graph (instead of one branch per context)
MODE_BATCHED = 3


class Graph:
    """nsteps steps of several factored contexts captured once (one branch
    per context) and replayed with one launch (lns3d_graph_*)."""

    def __init__(self, steps, v_devs, dv_devs=None, nsteps=1, mode=MODE_FP64, corrections=1):
        n = len(steps)
        ctxs = (VP * n)(*[s.h.value for s in steps])
        vs = (VP * n)(*[ptr(v) for v in v_devs])
        dvs = (VP * n)(*[ptr(d) for d in dv_devs]) if dv_devs is not None else None
        h = VP()
        check(lib().lns3d_graph_create(ctxs, vs, dvs, n, nsteps, mode, corrections, ctypes.byref(h)),
              "lns3d_graph_create")
        self.h = h
        self._keep = (steps, v_devs, dv_devs)

    def launch(self, stream=None):
        raw = getattr(stream, "cuda_stream", stream)
        check(lib().lns3d_graph_launch(self.h, raw), "lns3d_graph_launch")

    def synchronize(self):
        check(lib().lns3d_graph_synchronize(self.h), "lns3d_graph_synchronize")

    def close(self):
        if getattr(self, "h", None):
            lib().lns3d_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
