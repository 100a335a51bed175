"""ctypes front-end to the oracle: the C restatement (liboracle.so) and, where
it was built, the compiled reference (_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
VP = ctypes.c_void_p


def _ptr(a):
    if a is None:
        return None
    assert a.flags["F_CONTIGUOUS"] or a.ndim <= 1, "arrays must be Fortran-ordered"
    return a.ctypes.data_as(VP)


class _OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("nx", "ny", "optx", "opty", "xper", "yper", "wall", "wallt", "top",
                 "left", "right", "linear", "mattyp", "loctime")] + \
               [(n, ctypes.c_double) for n in
                ("Ma", "Re", "Pr", "gamma", "kz", "omega", "cflmax", "dtmax", "alfa")] + \
               [(n, ctypes.c_int) for n in ("periodic_exact", "wall_is", "wall_ie", "d_in_eta", "idamp")]


class _OrcSponge(ctypes.Structure):
    _fields_ = [("ispg", ctypes.c_int)] + [(n, VP) for n in ("spg", "spg2", "x", "wamp", "vm")] + \
               [("x0", ctypes.c_double)]


class _OrcCoeffs(ctypes.Structure):
    _fields_ = [(n, VP) for n in ("Ah", "Bh", "Dh", "Dhi", "Vh11", "Vh12", "Vh22", "ABhi")]


def coeff_arrays(nx, ny):
    z = lambda *s: np.zeros(s, order="F")
    return dict(Ah=z(5, 5, nx, ny), Bh=z(5, 5, nx, ny), Dh=z(5, 5, nx, ny),
                Dhi=z(5, 5, nx, ny), Vh11=z(6, nx, ny), Vh12=z(6, nx, ny),
                Vh22=z(6, nx, ny), ABhi=z(6, nx, ny))


class Oracle:
    """The C restatement."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.lib = ctypes.CDLL(path)

    def params(self, p):
        s = _OrcParams()
        for name, _ in _OrcParams._fields_:
            setattr(s, name, getattr(p, name, 0))
        return s

    def set_wall_normals(self, bnb=None):
        """wall normals bnb (nx, 2) of the inviscid wall rows; None clears"""
        self._bnb_keep = None if bnb is None else np.asfortranarray(bnb, dtype=np.float64)
        self.lib.orc_set_wall_normals(_ptr(self._bnb_keep))

    def set_sponge(self, ispg, spg=None, spg2=None, x=None, wamp=None, vm=None, x0=0.0):
        """sponge state for later lrhs3d / lhs / step calls (ispg 0: off); the
        arrays are kept referenced here while they are in use"""
        arrs = [None if a is None else np.asfortranarray(a, dtype=np.float64)
                for a in (spg, spg2, x, wamp, vm)]
        s = _OrcSponge(ispg, *(_ptr(a) for a in arrs), x0)
        rc = self.lib.orc_set_sponge(ctypes.byref(s))
        assert rc == 0, "bad sponge arguments"
        self._sponge_keep = arrs

    def set_top(self, x=None, wamp=None, vm=None, x0=0.0):
        """top = 1 forcing data (x, vm None: cleared)"""
        arrs = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (x, wamp, vm)]
        self.lib.orc_set_top.argtypes = [VP, VP, VP, ctypes.c_double]
        self.lib.orc_set_top(*(_ptr(a) for a in arrs), x0)
        self._top_keep = arrs

    def set_sutherland(self, S0=-1.0):
        """Sutherland's constant S0 for mattyp = 1 (negative: cleared)"""
        self.lib.orc_set_sutherland.argtypes = [ctypes.c_double]
        self.lib.orc_set_sutherland(float(S0))

    def set_bump(self, uw=None):
        """bump wall values uw (nx, 5) complex (wall = 3); None: cleared"""
        self._bump_keep = None if uw is None else np.asfortranarray(uw, dtype=np.complex128)
        self.lib.orc_set_bump(_ptr(self._bump_keep))

    def set_forcing(self, f=None):
        """the shift-invert driver's forcing f (5, nx, ny); None: off"""
        self._force_keep = None if f is None else np.asfortranarray(f, dtype=np.complex128)
        self.lib.orc_set_forcing(_ptr(self._force_keep))

    def forced_residual(self, p, c, dtl, v, vold):
        r = np.zeros_like(v, order="F")
        work = np.zeros(v.shape + (5,), complex, order="F")
        nrm = ctypes.c_double(0.0)
        ps, cs = self.params(p), self._coeffs(c)
        assert self.lib.orc_forced_residual(ctypes.byref(ps), ctypes.byref(cs), _ptr(dtl), _ptr(v),
                                            _ptr(np.asfortranarray(vold)), _ptr(r), _ptr(work),
                                            ctypes.byref(nrm)) == 0
        return r, nrm.value

    def set_smoother(self, eps_e=0.0, buff=None):
        """smoother eps_e with the damping profile buff (nx, ny); 0: off"""
        b = None if buff is None else np.asfortranarray(buff, dtype=np.float64)
        self.lib.orc_set_smoother.argtypes = [ctypes.c_double, VP]
        assert self.lib.orc_set_smoother(float(eps_e), _ptr(b)) == 0
        self._smo_keep = b

    def set_inflow(self, x=None, vm=None, x0=0.0, eig=None, lalpha=0j):
        """left = 1 / left = 4 / right = 4 inflow data (None: cleared); eig
        (10, ny): rhor, rhoi, ur, ui, vr, vi, wr, wi, tr, ti"""
        arrs = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (x, vm, eig)]
        la = np.array([complex(lalpha).real, complex(lalpha).imag])
        self.lib.orc_set_inflow.argtypes = [VP, VP, ctypes.c_double, VP, VP]
        self.lib.orc_set_inflow(_ptr(arrs[0]), _ptr(arrs[1]), x0, _ptr(arrs[2]), _ptr(la))
        self._inf_keep = arrs + [la]

    @staticmethod
    def _coeffs(c):
        s = _OrcCoeffs()
        for k in ("Ah", "Bh", "Dh", "Dhi", "Vh11", "Vh12", "Vh22", "ABhi"):
            setattr(s, k, _ptr(c[k]))
        return s

    def cpenta_bc(self, layout, mat, rhs, code):
        nsys, n = (mat.shape[3], mat.shape[4]) if layout == 1 else (mat.shape[4], mat.shape[3])
        rc = self.lib.orc_cpenta_bc(layout, nsys, n, _ptr(mat), _ptr(rhs), code)
        assert rc == 0

    def cpenta_p(self, layout, mat, rhs, per, per2, code, exact=0):
        nsys, np_ = (mat.shape[3], mat.shape[4]) if layout == 1 else (mat.shape[4], mat.shape[3])
        rc = self.lib.orc_cpenta_p(layout, nsys, np_, _ptr(mat), _ptr(rhs), _ptr(per),
                                   _ptr(per2), code, exact)
        assert rc == 0

    def cgrad(self, p, v):
        """first derivatives of a complex (ndof, nx, ny) field (ndof = 5)"""
        g1, g2 = np.zeros_like(v, order="F"), np.zeros_like(v, order="F")
        ps = self.params(p)
        self.lib.orc_cgrad(ctypes.byref(ps), _ptr(v), _ptr(g1), _ptr(g2))
        return g1, g2

    def cgrad2(self, p, v, g1):
        g11, g12, g22 = (np.zeros_like(v, order="F") for _ in range(3))
        ps = self.params(p)
        self.lib.orc_cgrad2(ctypes.byref(ps), _ptr(v), _ptr(g1), _ptr(g11), _ptr(g12), _ptr(g22))
        return g11, g12, g22

    def grad(self, p, v):
        """first derivatives of a real (ndof, nx, ny) field"""
        g1, g2 = np.zeros_like(v, order="F"), np.zeros_like(v, order="F")
        ps = self.params(p)
        self.lib.orc_grad(ctypes.byref(ps), v.shape[0], _ptr(v), _ptr(g1), _ptr(g2))
        return g1, g2

    def grad2(self, p, v, g1):
        g11, g12, g22 = (np.zeros_like(v, order="F") for _ in range(3))
        ps = self.params(p)
        self.lib.orc_grad2(ctypes.byref(ps), v.shape[0], _ptr(v), _ptr(g1), _ptr(g11), _ptr(g12),
                           _ptr(g22))
        return g11, g12, g22

    def genmtrx(self, p, vm, met):
        c = coeff_arrays(p.nx, p.ny)
        cs = self._coeffs(c)
        ps = self.params(p)
        rc = self.lib.orc_genmtrx(ctypes.byref(ps), _ptr(vm), _ptr(met), ctypes.byref(cs))
        assert rc == 0, "unsupported options"
        return c

    def set_time_step(self, p, vm, met):
        dtl = np.zeros((p.nx, p.ny), order="F")
        ps = self.params(p)
        assert self.lib.orc_set_time_step(ctypes.byref(ps), _ptr(vm), _ptr(met), _ptr(dtl)) == 0
        return dtl

    def lrhs3d(self, p, c, v):
        r = np.zeros_like(v, order="F")
        work = np.zeros(v.shape + (5,), complex, order="F")
        ps, cs = self.params(p), self._coeffs(c)
        assert self.lib.orc_lrhs3d(ctypes.byref(ps), ctypes.byref(cs), _ptr(v), _ptr(r),
                                   _ptr(work)) == 0
        return r, work

    def rbe3d(self, p, r, v, vold, dtl):
        ps = self.params(p)
        self.lib.orc_rbe3d(ctypes.byref(ps), _ptr(r), _ptr(v), _ptr(vold), _ptr(dtl))

    def lhs(self, p, direction, c, dtl):
        mat = np.zeros((5, 5, 5, p.nx, p.ny), complex, order="F")
        ps, cs = self.params(p), self._coeffs(c)
        assert self.lib.orc_lhs(ctypes.byref(ps), direction, ctypes.byref(cs), _ptr(dtl),
                                _ptr(mat)) == 0
        return mat

    def rhsbc3d(self, p, r, v):
        ps = self.params(p)
        assert self.lib.orc_rhsbc3d(ctypes.byref(ps), _ptr(r), _ptr(v)) == 0

    def itrbc3d(self, p, v):
        ps = self.params(p)
        assert self.lib.orc_itrbc3d(ctypes.byref(ps), _ptr(v)) == 0

    def factor(self, p, c, dtl):
        nx, ny = p.nx, p.ny
        f = dict(matx=np.zeros((5, 5, 5, ny, nx) if not p.xper else (5, 5, 5, nx, ny),
                               complex, order="F"),
                 maty=np.zeros((5, 5, 5, nx, ny), complex, order="F"),
                 perx=np.zeros((2, 5, 5, nx, ny), complex, order="F"),
                 per2x=np.zeros((2, 5, 5, nx, ny), complex, order="F"),
                 pery=np.zeros((2, 5, 5, nx, ny), complex, order="F"),
                 per2y=np.zeros((2, 5, 5, nx, ny), complex, order="F"))
        ps, cs = self.params(p), self._coeffs(c)
        rc = self.lib.orc_factor(ctypes.byref(ps), ctypes.byref(cs), _ptr(dtl),
                                 *(_ptr(f[k]) for k in ("matx", "maty", "perx", "per2x",
                                                         "pery", "per2y")))
        assert rc == 0
        return f

    def step(self, p, c, dtl, f, v):
        """one step; v updated in place; returns the increment dv"""
        dv = np.zeros_like(v, order="F")
        work = np.zeros(v.shape + (7,), complex, order="F")
        ps, cs = self.params(p), self._coeffs(c)
        rc = self.lib.orc_step(ctypes.byref(ps), ctypes.byref(cs), _ptr(dtl),
                               *(_ptr(f[k]) for k in ("matx", "maty", "perx", "per2x",
                                                       "pery", "per2y")),
                               _ptr(v), _ptr(dv), _ptr(work))
        assert rc == 0
        return dv

    def step_bdf2(self, p, c, dtl, f, v, v2old):
        """an impl = 2 step after the first (v2old: the field at the start of
        the previous step); v updated in place; returns dv"""
        dv = np.zeros_like(v, order="F")
        work = np.zeros(v.shape + (7,), complex, order="F")
        ps, cs = self.params(p), self._coeffs(c)
        rc = self.lib.orc_step_bdf2(ctypes.byref(ps), ctypes.byref(cs), _ptr(dtl),
                                    *(_ptr(f[k]) for k in ("matx", "maty", "perx", "per2x",
                                                            "pery", "per2y")),
                                    _ptr(v), _ptr(np.asfortranarray(v2old)), _ptr(dv), _ptr(work))
        assert rc == 0
        return dv


    def step_iter(self, p, c, dtl, f, v, niter, v2old=None, half=False):
        """a step of niter corrector iterations (impl = 2 with v2old, impl = 3
        with half); v updated in place; returns the last increment"""
        dv = np.zeros_like(v, order="F")
        work = np.zeros(v.shape + (7,), complex, order="F")
        ps, cs = self.params(p), self._coeffs(c)
        self.lib.orc_step_iter.argtypes = [VP] * 10 + [VP, ctypes.c_int, ctypes.c_int, VP, VP]
        rc = self.lib.orc_step_iter(ctypes.byref(ps), ctypes.byref(cs), _ptr(dtl),
                                    *(_ptr(f[k]) for k in ("matx", "maty", "perx", "per2x",
                                                            "pery", "per2y")),
                                    _ptr(v), None if v2old is None else _ptr(np.asfortranarray(v2old)),
                                    niter, 1 if half else 0, _ptr(dv), _ptr(work))
        assert rc == 0
        return dv


class Reference:
    """The reference's own routines, compiled by build_ref.sh (this container
    only; absent where the reference sources are absent)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ctypes.CDLL(path)

    @staticmethod
    def available(path=None):
        return os.path.exists(path or os.path.join(HERE, "_ref", "libref.so"))

    def setup(self, p, x, y, met, vm, usecalcd=0, ispg=0, x0=0.0, impl=1):
        lsym = int(p.left == 7 and not p.xper)           # src/input.f90:235-243
        rsym = int(p.right == 7 and not p.xper)
        # Re = 0 with constant viscosity: Navier = .false., Re = 1e99
        # (src/input.f90:31-36)
        inviscid = p.Re == 0.0 and p.mattyp == 0
        ip = np.array([p.nx, p.ny, p.optx, p.opty, 0, p.xper, p.yper, lsym, rsym, p.wall,
                       p.wallt, p.top, p.left, p.right, p.linear, impl, p.loctime, p.mattyp,
                       ispg, 0 if inviscid else 1, usecalcd or getattr(p, "use_calcd", 0),
                       getattr(p, "idamp", 0)],
                      dtype=np.int32)
        rp = np.array([p.Ma, 1.0e99 if inviscid else p.Re, p.Pr, p.gamma, p.kz, p.omega, p.cflmax,
                       p.dtmax, 0.0, x0])
        self.lib.ref_init(ip.ctypes.data_as(VP), rp.ctypes.data_as(VP))
        wis, wie = getattr(p, "wall_is", 0), getattr(p, "wall_ie", 0)
        self.lib.ref_set_extras(ctypes.c_int(wis or 1), ctypes.c_int(wie or p.nx),
                                ctypes.c_int(getattr(p, "update_lhs", 0)))
        self.lib.ref_set_grid(_ptr(x), _ptr(y), _ptr(met))
        self.lib.ref_set_mean(_ptr(vm))
        self.p = p

    def set_inflow(self, eig, lalpha=0j, lomega=0j):
        """the eigenfunction profiles of src/gendist.f90 (eig (10, ny): rhor,
        rhoi, ur, ui, vr, vi, wr, wi, tr, ti) and the namelist's lalpha, lomega"""
        la = np.array([complex(lalpha).real, complex(lalpha).imag])
        lo = np.array([complex(lomega).real, complex(lomega).imag])
        self.lib.ref_set_inflow(_ptr(np.asfortranarray(eig, dtype=np.float64)), _ptr(la), _ptr(lo))

    def set_smoother(self, eps_e, buff):
        """the deck's eps_e and the damping profile buff (nx, ny)"""
        self.lib.ref_set_smoother.argtypes = [ctypes.POINTER(ctypes.c_double), VP]
        self._buff_keep = np.asfortranarray(buff, dtype=np.float64)
        self.lib.ref_set_smoother(ctypes.byref(ctypes.c_double(eps_e)), _ptr(self._buff_keep))

    def set_forcing(self, f=None):
        """the shift-invert driver's forcing f (5, nx, ny): steps then use ebe3D
        (None: rbe3D again)"""
        if f is None:
            self.lib.ref_set_forcing(None, ctypes.c_int(0))
        else:
            self._f_keep = np.asfortranarray(f, dtype=np.complex128)
            self.lib.ref_set_forcing(_ptr(self._f_keep), ctypes.c_int(1))

    def forced_residual(self, v, vold):
        r = np.zeros_like(v, order="F")
        nrm = ctypes.c_double(0.0)
        self.lib.ref_forced_residual(_ptr(np.asfortranarray(v)), _ptr(np.asfortranarray(vold)), _ptr(r),
                                     ctypes.byref(nrm))
        return r, nrm.value

    def set_sutherland(self, S0):
        """Sutherland's constant S0 (110.4 K / Tref) for mattyp = 1"""
        self.lib.ref_set_sutherland(ctypes.c_double(S0))

    def set_bump(self, uw):
        """the bump wall values uw (nx, 5) complex: u1w, u2w, u3w, tw, twp"""
        self._bump_keep = np.asfortranarray(uw, dtype=np.complex128)
        self.lib.ref_set_bump(_ptr(self._bump_keep))

    def set_niter(self, n):
        """corrector iterations per step (the deck's niter)"""
        self.lib.ref_set_niter(ctypes.c_int(n))

    def set_sponge(self, spg, spg2, wamp):
        """the reference's spg, spg2 (nx,ny) and wamp (nx) module arrays"""
        self.lib.ref_set_sponge(_ptr(np.asfortranarray(spg, dtype=np.float64)),
                                _ptr(np.asfortranarray(spg2, dtype=np.float64)),
                                _ptr(np.ascontiguousarray(wamp, dtype=np.float64)))

    def genmtrx(self):
        self.lib.ref_genmtrx()
        c = coeff_arrays(self.p.nx, self.p.ny)
        self.lib.ref_get_coeffs(*(_ptr(c[k]) for k in
                                  ("Ah", "Bh", "Dh", "Dhi", "Vh11", "Vh12", "Vh22", "ABhi")))
        return c

    def set_coeffs(self, c):
        self.lib.ref_set_coeffs(*(_ptr(c[k]) for k in
                                  ("Ah", "Bh", "Dh", "Dhi", "Vh11", "Vh12", "Vh22", "ABhi")))

    def set_time_step(self):
        dtl = np.zeros((self.p.nx, self.p.ny), order="F")
        self.lib.ref_set_time_step(_ptr(dtl))
        return dtl

    def set_dtl(self, dtl):
        self.lib.ref_set_dtl(_ptr(dtl))

    def lrhs3d(self, v):
        r = np.zeros_like(v, order="F")
        g = np.zeros(v.shape + (5,), complex, order="F")
        self.lib.ref_lrhs3d(_ptr(v), _ptr(r), _ptr(g))
        return r, g

    def rbe3d(self, r, v, vold):
        self.lib.ref_rbe3d(_ptr(r), _ptr(v), _ptr(vold))

    def rhsbc3d(self, r, v):
        self.lib.ref_rhsbc3d(_ptr(r), _ptr(v))

    def itrbc3d(self, v):
        self.lib.ref_itrbc3d(_ptr(v))

    def lhs(self, direction, alfa=1.0):
        mat = np.zeros((5, 5, 5, self.p.nx, self.p.ny), complex, order="F")
        self.lib.ref_lhs(ctypes.c_int(direction), ctypes.c_double(alfa), _ptr(mat))
        return mat

    def step(self, v, first):
        dv = np.zeros_like(v, order="F")
        self.lib.ref_step(_ptr(v), _ptr(dv), ctypes.c_int(1 if first else 0))
        return dv

    def _call(self, name, *args):
        cargs = [ctypes.byref(ctypes.c_int(a)) if isinstance(a, int) else _ptr(a) for a in args]
        getattr(self.lib, name)(*cargs)

    def cpenta1bc(self, mat, rhs, code, shipped=False):
        self._call("cpenta1bc_shipped_" if shipped else "cpenta1bc_",
                   mat.shape[3], mat.shape[4], 5, mat, rhs, code)

    def cpenta2bc(self, mat, rhs, code, shipped=False):
        self._call("cpenta2bc_shipped_" if shipped else "cpenta2bc_",
                   mat.shape[3], mat.shape[4], 5, mat, rhs, code)

    def cpenta1p(self, mat, rhs, per, per2, code):
        self._call("cpenta1p_", mat.shape[3], mat.shape[4], 5, mat, rhs, per, per2, code)

    def cpenta2p(self, mat, rhs, per, per2, code):
        self._call("cpenta2p_", mat.shape[3], mat.shape[4], 5, mat, rhs, per, per2, code)
