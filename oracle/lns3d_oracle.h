/*
 * CPU restatement (the oracle) of the LNS3D implicit approximate-factorization
 * step: block-pentadiagonal line solvers, finite-difference RHS, LHS assembly,
 * boundary conditions, coefficient generation and time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests, the smoke check and the CPU-baseline
 * leg of the benchmark may load it, and only as the checker.  The product
 * library (lns3d_amd) never links or calls it.
 *
 * Arrays use the reference's Fortran (column-major) layouts and its 1-based
 * index meaning, stored 0-based:
 *   v, r      complex (5, nx, ny)            v[d + 5*(i + nx*j)]
 *   Ah,Bh,..  real    (5, 5, nx, ny)         a[row + 5*col + 25*(i + nx*j)]
 *   Vh, ABhi  real    (6, nx, ny)
 *   mat       complex (5, 5, 5, nx, ny)      mat[k + 5*row + 25*col + 125*p]
 * Every function states the reference file:line it follows.
 */
#ifndef LNS3D_ORACLE_H
#define LNS3D_ORACLE_H
#include <complex.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex zdouble;

/* Scheme, physics and boundary selectors: the module state of src/global.f90
 * that the step routines read. */
typedef struct {
  int nx, ny;
  int optx, opty;
  int xper, yper;
  int wall, wallt, top, left, right;
  int linear, mattyp, loctime;
  double Ma, Re, Pr, gamma, kz, omega;
  double cflmax, dtmax;
  double alfa;
  int periodic_exact; /* 1: exact LU for periodic lines; 0: shipped slot order */
  int wall_is, wall_ie; /* wall extent is:ie, 1-based (src/lns3d.f90:42,73); 0: 1, nx */
  int d_in_eta;       /* 1: the D terms in the eta factor and the sponge in the xi
                       * factor (useCalcd's first factors, src/lhs1f3d.f90:27-37) */
  int idamp;          /* 1: the outflow column's damper (src/damper.f90:26-29,
                       * src/genmtrx.f90:87-99,877-881); 0, 2: none on this path */
} orc_params;
/* Re = 0 with mattyp = 0 is the inviscid deck (src/input.f90:31-36:
 * Navier = .false., Re = 1e99): rotated wall rows with the wall normals of
 * orc_set_wall_normals */
int orc_set_wall_normals(const double *bnb);   /* (nx,2) as src/gengrid.f90:133-137; NULL clears */

/* ---- block-pentadiagonal line solvers (src/cpenta1bc.f, cpenta2bc.f,
 *      cpenta1p.f, cpenta2p.f) ---------------------------------------------- */
/* layout 1: mat(5,5,5,nsys,n), rhs(5,nsys,n)   (cpenta1bc / cpenta1p)
 * layout 2: mat(5,5,5,n,nsys), rhs(5,n,nsys)   (cpenta2bc / cpenta2p)
 * code 0 factor, 1 solve, 2 both.  Returns 0, or -1 on bad arguments. */
int orc_cpenta_bc(int layout, int nsys, int n, zdouble *mat, zdouble *rhs, int code);
int orc_cpenta_p(int layout, int nsys, int np, zdouble *mat, zdouble *rhs,
                 zdouble *per, zdouble *per2, int code, int exact);

/* ---- stencils (src/cgrad.f90, src/cgrad2.f90, src/grad.f90, grad2.f90) ---- */
void orc_cgrad(const orc_params *p, const zdouble *v, zdouble *g1, zdouble *g2);
void orc_cgrad2(const orc_params *p, const zdouble *v, const zdouble *g1,
                zdouble *g11, zdouble *g12, zdouble *g22);
void orc_grad(const orc_params *p, int ndof, const double *v, double *g1, double *g2);
void orc_grad2(const orc_params *p, int ndof, const double *v, const double *g1,
               double *g11, double *g12, double *g22);

/* coefficient set produced by src/genmtrx.f90 */
typedef struct {
  double *Ah, *Bh, *Dh, *Dhi;     /* (5,5,nx,ny) */
  double *Vh11, *Vh12, *Vh22;     /* (6,nx,ny) */
  double *ABhi;                   /* (6,nx,ny) */
} orc_coeffs;

/* metrics: m1,m2,n1,n2,m11,m12,m22,n11,n12,n22, each (nx,ny), stacked */
int orc_genmtrx(const orc_params *p, const double *vm, const double *metric,
                orc_coeffs *c);
int orc_set_time_step(const orc_params *p, const double *vm, const double *metric,
                      double *dtl);

/* ---- one implicit step pieces (src/lrhs3d.f90, impdrv3d.f90, lhs*f3d.f90,
 *      lhsbt*f3d.f90, lhsbc*f3d.f90, rhsbc3d.f90, itrbc3d.f90) ------------- */
/* sponge state read by orc_lrhs3d and orc_lhs, as the reference's module
 * state (src/global.f90:132-141, src/sponge.f90): ispg 0 (off), 1, 2, 3 or 5
 * (4, the iterative form of src/sponge.f90:536, is not restated); spg, spg2
 * (nx,ny); x (nx,ny) physical x and vm (5,nx,ny) mean flow for the inflow
 * forcing of ispg 2, 3; wamp (nx) for ispg 3.  The arrays must outlive their
 * use.  Returns -1 on a bad combination. */
typedef struct {
  int ispg;
  const double *spg, *spg2, *x, *wamp, *vm;
  double x0;
} orc_sponge;
int orc_set_sponge(const orc_sponge *s);
/* top = 1 (src/rhsbc3d.f90:129-161, itrbc3d.f90:145-195): physical x
 * (nx,ny), amplitude wamp (nx; NULL for 1), mean flow vm (5,nx,ny), x0 */
int orc_set_top(const double *x, const double *wamp, const double *vm, double x0);
/* left = 1 / left = 4 / right = 4 inflow state: x (nx,ny), vm (5,nx,ny), x0;
 * eig (10,ny) the eigenfunction profiles rhor, rhoi, ur, ui, vr, vi, wr, wi,
 * tr, ti; lalpha (re, im) for right = 4 (NULL: 0) */
int orc_set_inflow(const double *x, const double *vm, double x0, const double *eig,
                   const double *lalpha);
/* the shift-invert driver's forcing f (5,nx,ny) (NULL: off): steps then use
 * ebe3D; orc_forced_residual gives its residual after a step and the norm */
int orc_set_forcing(const zdouble *f);
int orc_forced_residual(const orc_params *p, const orc_coeffs *c, const double *dtl,
                        const zdouble *v, const zdouble *vold, zdouble *r,
                        zdouble *work /* 5 arrays of (5,nx,ny) */, double *norm);
/* bump wall values for wall = 3: uw (nx,5) complex u1w, u2w, u3w, tw, twp */
int orc_set_bump(const zdouble *uw);
/* Sutherland's constant S0 = 110.4 K / Tref for mattyp = 1 */
int orc_set_sutherland(double S0);
/* smoother eps_e with the damping profile buff (nx,ny) (eps_e 0: off) */
int orc_set_smoother(double eps_e, const double *buff);
int orc_lrhs3d(const orc_params *p, const orc_coeffs *c, const zdouble *v,
               zdouble *r, zdouble *work /* 5 arrays of (5,nx,ny) */);
void orc_rbe3d(const orc_params *p, zdouble *r, const zdouble *v,
               const zdouble *vold, const double *dtl);
int orc_lhs(const orc_params *p, int dir, const orc_coeffs *c, const double *dtl,
            zdouble *mat);
int orc_rhsbc3d(const orc_params *p, zdouble *r, const zdouble *v);
int orc_itrbc3d(const orc_params *p, zdouble *v);

/* Factor both directions (dir 1 into the transposed cpenta1bc layout that
 * src/impdrv3d.f90:257-269 uses, dir 2 in place).  matx is (5,5,5,ny,nx),
 * maty is (5,5,5,nx,ny); per* arrays are (2,5,5,nx,ny), only for periodic. */
int orc_factor(const orc_params *p, const orc_coeffs *c, const double *dtl,
               zdouble *matx, zdouble *maty, zdouble *perx, zdouble *per2x,
               zdouble *pery, zdouble *per2y);
/* One step given the factors: v is updated in place, dv receives r.  work
 * holds 7 arrays of (5,nx,ny). */
int orc_step(const orc_params *p, const orc_coeffs *c, const double *dtl,
             zdouble *matx, zdouble *maty, zdouble *perx, zdouble *per2x,
             zdouble *pery, zdouble *per2y, zdouble *v, zdouble *dv,
             zdouble *work);
/* src/impdrv3d.f90:473-501 rtb3D, and the impl = 2 step after the first
 * (two-step backward residual with the field v2old of the previous step's
 * start; factors formed with alfa = 2/3) */
void orc_rtb3d(const orc_params *p, zdouble *r, const zdouble *v, const zdouble *vold,
               const zdouble *v2old, const double *dtl);
int orc_step_bdf2(const orc_params *p, const orc_coeffs *c, const double *dtl,
                  zdouble *matx, zdouble *maty, zdouble *perx, zdouble *per2x,
                  zdouble *pery, zdouble *per2y, zdouble *v, const zdouble *v2old,
                  zdouble *dv, zdouble *work);
/* a step of niter corrector iterations (src/impdrv3d.f90:173-358) with vold
 * the field at its start: v2old NULL for backward Euler, else rtb3D; half:
 * the residual at 0.5 (v + vold) (impl = 3) */
int orc_step_iter(const orc_params *p, const orc_coeffs *c, const double *dtl,
                  zdouble *matx, zdouble *maty, zdouble *perx, zdouble *per2x,
                  zdouble *pery, zdouble *per2y, zdouble *v, const zdouble *v2old, int niter,
                  int half, zdouble *dv, zdouble *work);

#ifdef __cplusplus
}
#endif
#endif
