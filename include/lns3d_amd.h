/*
 * lns3d_amd: MI355X (gfx950) implementation of the LNS3D implicit
 * approximate-factorization step (sscollis/lns3d, ImpDrv3D).  C ABI.
 *
 * Conventions shared by every entry point:
 *  - complex arrays are complex double (16 bytes, real then imaginary) in the
 *    reference's Fortran column-major layouts; sizes are plain ints;
 *  - "_dev" pointers are device (HBM) addresses, other array pointers are host;
 *  - stream arguments are hipStream_t passed as void* (NULL = default stream);
 *  - every function returns 0 on success or one of the LNS3D_E* codes, with a
 *    message available from lns3d_last_error() (per host thread);
 *  - calls on one handle must come from one host thread at a time.
 */
#ifndef LNS3D_AMD_H
#define LNS3D_AMD_H
#ifdef __cplusplus
extern "C" {
#endif

#define LNS3D_OK 0
#define LNS3D_EINVAL 1        /* bad argument or shape                          */
#define LNS3D_EHIP 2          /* HIP runtime error                              */
#define LNS3D_ENOTFACTORED 3  /* solve requested before the factorization       */
#define LNS3D_EUNSUPPORTED 4  /* option the library does not implement          */
#define LNS3D_ENONFINITE 5    /* a result or pivot was not finite               */

const char *lns3d_last_error(void);

/* Scheme, physics and boundary selectors: the module state of src/global.f90
 * that the step routines read (src/input.f90 deck entries of the same names).
 * Supported: optx/opty any; xper/yper; wall in {0,1,2,3} (3: the bump wall,
 * lns3d_set_wall); wallt in {0,1,2};
 * top in {0,1} (1: the incoming wave of lns3d_set_top_forcing); left in
 * {0,1,4,5,7} and right in {0,4,5,7,8,9} (8: the column held at its initial
 * values; 9: extrapolated from the two columns inside, exp(2 log v(nx-1) -
 * log v(nx-2)), src/rhsbc3d.f90:317-320, src/itrbc3d.f90:408-410, which is
 * not finite where both are zero, as at a no-slip wall's corner, in the
 * reference as here; 1: acoustic wave, 5: the undamped wave, right 5 with nx <= ny;
 * 4: eigenfunction inflow; 1, 4 and 5
 * both through lns3d_set_inflow; right 4 with left 4 only; 7: symmetry,
 * src/lhsbs1f3d.f90 and the symmetric closures of src/cgrad.f90,
 * cgrad2.f90, grad.f90, grad2.f90; needs nx >= ny, since
 * the reference's symmetric rows read point i + ny j); linear 1 or 2 (the
 * deck's 2, complex analysis, is 1 once read, src/input.f90:367-371, and the
 * 3D driver runs only with 1: the coefficients take the linear = 1 forms of
 * src/genmtrx.f90:719-820); mattyp in
 * {0,1,2,3} (1: Sutherland's law, lns3d_set_sutherland); loctime in {0,1};
 * Navier-Stokes (viscous) boundary rows; the smoother eps_e through
 * lns3d_set_smoother; sponges through
 * lns3d_set_sponge (ispg 1 to 5).  Other values are rejected with
 * LNS3D_EUNSUPPORTED. */
typedef struct {
  int nx, ny;
  int optx, opty;
  int xper, yper;
  int wall, wallt, top, left, right;
  int linear, mattyp, loctime;
  double Ma, Re, Pr, gamma;
  double kz;        /* spanwise wavenumber (beta) */
  double omega;     /* frequency, enters Dhi as -omega (src/genmtrx.f90:981-989) */
  double cflmax, dtmax;
  double alfa;      /* time-scheme factor: 1 for backward Euler */
  int periodic_exact;
  int precision;    /* 0: complex-double factors; 1: complex-float factors */
  /* &extras namelist (src/lns3d.f90:42) and deck entries; zero selects the
   * reference's default for each */
  int wall_is, wall_ie; /* wall extent is:ie, 1-based (src/lns3d.f90:73
                         * default 1:nx); 0, 0 = the whole wall */
  int idamp;        /* viscous damper (src/input.f90:268, src/damper.f90): 1
                     * zeroes the outflow column's mean xi second derivatives
                     * and Vh11 (src/genmtrx.f90:87-96, 877-881); 2 has no
                     * effect on this path */
  int use_calcd;    /* useCalcd (src/global.f90:34): the D terms alternate
                     * between the factors (src/lhs1f3d.f90:27-37) */
  int update_lhs;   /* updateLHS (src/global.f90:33): factors formed again
                     * every iteration (src/impdrv3d.f90:248,310) */
} lns3d_params;
/* Re = 0 with mattyp = 0 is the reference's inviscid deck (src/input.f90
 * :31-36: Navier = .false., Re = 1e99): the wall row holds the rotated
 * momentum rows of src/rhsbc3d.f90, src/lhsbt1f3d.f90 and src/lhsbt2f3d.f90
 * with the wall normals lns3d_setup forms from the metric (src/gengrid.f90
 * :133-137), itrBC3D's tangency condition, and bcfix after the eta solve
 * (src/impdrv3d.f90:333-336).  update_lhs forms both factors again each
 * iteration, as the build with USE_CPENTA1BC does (src/impdrv3d.f90:248,310);
 * the factors change only with useCalcd's D-term placement, so they are
 * formed again only then.
 * With precision = 1 the refined steps' operator applies the inviscid
 * deck's rotated wall rows as the assembled factors hold them, and with
 * update_lhs the complex-float steps form both precisions' factors again
 * as lns3d_step does.
 * Re < 0, a non-finite Ma, Re, Pr, gamma, kz or omega, Re = 0 with
 * mattyp != 0, idamp outside 0..2 and a wall extent outside
 * 1 <= is <= ie <= nx are rejected with LNS3D_EINVAL. */

/* ======================================================================
 * Block-pentadiagonal line solvers, drop-in form (host arrays).
 * Argument meaning, layouts and code (0 factor, 1 solve, 2 both) as in the
 * reference; factors are kept on the device keyed by the mat address and
 * shape, and mat/per/per2 are not written back (the reference's driver never
 * reads them after factoring).  nblk must be 5.
 * ====================================================================== */
/* replaces src/cpenta1bc.f:2  cpenta1bc(nsys,n,nblk,mat,rhs,code)
 * mat(5,nblk,nblk,nsys,n), rhs(nblk,nsys,n).  The row n-1 update of the
 * forward elimination follows the corrected index (the shipped line
 * src/cpenta1bc.f:675 addresses rhs(iv,r,m)). */
int lns3d_cpenta1bc(int nsys, int n, int nblk, void *mat, void *rhs, int code);
/* replaces src/cpenta2bc.f:2  cpenta2bc(n,nsys,nblk,mat,rhs,code)
 * mat(5,nblk,nblk,n,nsys), rhs(nblk,n,nsys); corrected index as above
 * (shipped line src/cpenta2bc.f:676). */
int lns3d_cpenta2bc(int n, int nsys, int nblk, void *mat, void *rhs, int code);
/* replaces src/cpenta1p.f:2  cpenta1p(nsys,np,nblk,mat,rhs,per,per2,code)
 * mat(5,nblk,nblk,nsys,np), rhs(nblk,nsys,np); np-1 unknowns per line and
 * rhs(:,:,np) = rhs(:,:,1) on return.  per/per2 are accepted and unused. */
int lns3d_cpenta1p(int nsys, int np, int nblk, void *mat, void *rhs, void *per, void *per2,
                   int code);
/* replaces src/cpenta2p.f:2  cpenta2p(np,nsys,nblk,mat,rhs,per,per2,code)
 * mat(5,nblk,nblk,np,nsys), rhs(nblk,np,nsys). */
int lns3d_cpenta2p(int np, int nsys, int nblk, void *mat, void *rhs, void *per, void *per2,
                   int code);
/* Periodic factorization mode for the drop-ins (default 0): 0 reproduces the
 * reference's factor, which adds the row n-5 -> row n-3 update of column n-1
 * into slot 4 (src/cpenta1p.f:284,345); 1 is the exact block LU. */
int lns3d_set_periodic_exact(int exact);
/* free every factor the drop-ins hold */
int lns3d_release_factors(void);

/* ======================================================================
 * Drop-ins for the RHS, the LHS assembly and the derivative routines, with
 * the reference's argument lists (host arrays in its layouts).  The routines
 * that read module state in the reference (lrhs3D, lhs1f3D, lhs2f3D) read it
 * here from the state the caller forwards once through
 * lns3d_set_module_state: the variables of src/global.f90, local2d.f90 and
 * local3d.f90 the step reads.  Arrays are copied over PCIe on every call
 * (the device-resident context below avoids that).
 * ====================================================================== */
typedef struct {
  int nx, ny, ndof;                       /* sizes; ndof must be 5 */
  double alfa, dxi, deta;                 /* scheme (dxi = 1/(nx-1), deta = 1/(ny-1)) */
  int optx, opty, carp;                   /* carp (Carpenter closure, impl = 0 only) must be 0 */
  int xper, yper, lsym, rsym, bsym, tsym; /* lsym/rsym follow left/right = 7; bsym = tsym = 0 */
  int wall, wallt, top, left, right;      /* boundary selectors, as lns3d_params */
  int is, ie, navier;                     /* wall extent (1-based) and Navier */
  int linear, mattyp, loctime, idamp;
  double Ma, Re, Pr, gamma, gamma1, kz, omega, x0, S0, cflmax, dtmax;   /* physics */
  int ispg;                               /* sponge (src/global.f90:132-141) */
  const double *spg, *spg2;               /* (nx,ny); NULL when ispg does not use them */
  double eps_e;                           /* smoother and its damping profile */
  const double *buff;                     /* (nx,ny); NULL when eps_e = 0 */
  int use_calcd, update_lhs;              /* factor control */
  const int *istep, *iter;                /* the caller's step and iteration
                                           * counters (src/global.f90), read at
                                           * each lhs1f3D / lhs2f3D call for
                                           * useCalcd; NULL reads as 1 */
  const double *wamp;                     /* (nx) sponge / top wave amplitude; NULL: 1 */
  const double *vm;                       /* (5,nx,ny) mean flow */
  const double *x;                        /* (nx,ny) physical x */
  const double *bnb;                      /* (nx,2) wall normals (src/gengrid.f90:133-137) */
  const double *eig;                      /* (10,ny) left = 4 inflow profiles; NULL otherwise */
  double lalpha[2];                       /* inflow namelist's lalpha (re, im) */
  const double *bump;                     /* (nx,5) complex wall values for wall = 3; NULL otherwise */
  const double *Ah, *Bh, *Dh, *Dhi;       /* (5,5,nx,ny) src/local2d.f90:34-36 */
  const double *Vh11, *Vh12, *Vh22, *ABhi;/* (6,nx,ny) src/local3d.f90:14 */
  const double *dtl;                      /* (nx,ny) */
} lns3d_module_state;
/* forward the module state; creates (or reuses) the context the drop-ins
 * below compute with.  Options the device path does not implement are
 * refused as lns3d_ctx_create refuses them. */
int lns3d_set_module_state(const lns3d_module_state *s);
/* replaces src/lrhs3d.f90:2 lrhs3D(vl,rl,vml,xl,yl): rl = the linearized
 * operator of vl with the forwarded coefficients, sponge and smoother terms
 * (vml, xl, yl are the module arrays the state already holds; unused) */
int lns3d_drop_lrhs3d(const void *vl, void *rl, const double *vml, const double *xl, const double *yl);
/* replaces src/lhs1f3d.f90:2 lhs1f3D(mat,Ah,Dh,Dhi,Vh11,ABhi,dtl,vl) with
 * lhsbt1f3D/lhsbs1f3D/lhsbc1f3d: mat(5,5,5,nx,ny) assembled from the given
 * arrays and the forwarded state (vl unused, as in the reference) */
int lns3d_drop_lhs1f3d(void *mat, const double *Ah, const double *Dh, const double *Dhi, const double *Vh11,
                       const double *ABhi, const double *dtl, const void *vl);
/* replaces src/lhs2f3d.f90:2 lhs2f3D(mat,Bh,Dh,Dhi,Vh22,ABhi,dtl,vl) with
 * lhsbt2f3d/lhsbc2f3d */
int lns3d_drop_lhs2f3d(void *mat, const double *Bh, const double *Dh, const double *Dhi, const double *Vh22,
                       const double *ABhi, const double *dtl, const void *vl);
/* replaces src/cgrad.f90:2-3 cgrad(ndof,nx,ny,v,g1v,g2v,dx,dy,optx,opty,
 * xper,yper,lsym,rsym,bsym,tsym,carp); logicals as 0/1; ndof = 5, bsym =
 * tsym = carp = 0 */
int lns3d_drop_cgrad(int ndof, int nx, int ny, const void *v, void *g1v, void *g2v, double dx, double dy,
                     int optx, int opty, int xper, int yper, int lsym, int rsym, int bsym, int tsym, int carp);
/* replaces src/cgrad2.f90:2-3 cgrad2(ndof,nx,ny,v,g1v,g11v,g12v,g22v,dx,dy,
 * optx,opty,xper,yper,lsym,rsym,bsym,tsym,carp); g12v = d/deta of g1v */
int lns3d_drop_cgrad2(int ndof, int nx, int ny, const void *v, const void *g1v, void *g11v, void *g12v,
                      void *g22v, double dx, double dy, int optx, int opty, int xper, int yper, int lsym,
                      int rsym, int bsym, int tsym, int carp);

/* ======================================================================
 * Device-resident solver handles: factor once from a device matrix in the
 * reference's layout, then solve device right-hand sides in place.
 * kind 0 = non-periodic (cpenta*bc), 1 = periodic (cpenta*p);
 * layout 1 = system index before row index (cpenta1*), 2 = after (cpenta2*);
 * nrows = n (kind 0) or np (kind 1); precision 0 = complex double factors,
 * 1 = complex float factors (kind 1: records rounded from the complex-
 * double bordered LU, which the handle also holds); exact as above.
 * ====================================================================== */
typedef struct lns3d_penta lns3d_penta;
int lns3d_penta_create(int kind, int layout, int nsys, int nrows, int precision, int exact,
                       lns3d_penta **out);
/* factors, then checks them (synchronises the stream): a line with a zero
 * pivot or a non-finite entry returns LNS3D_ENONFINITE, naming the first
 * such line (the reference divides by the pivot unguarded,
 * src/cpenta1bc.f:80-82).  The drop-ins above report the same status for
 * code 0 and 2 (code 2 still solves, as the reference does). */
int lns3d_penta_factor(lns3d_penta *h, const void *mat_dev, void *stream);
int lns3d_penta_solve(lns3d_penta *h, void *rhs_dev, void *stream);
/* same result, one lane per line in the reference's operation order */
int lns3d_penta_solve_reference_order(lns3d_penta *h, void *rhs_dev, void *stream);
int lns3d_penta_destroy(lns3d_penta *h);

/* ======================================================================
 * Device-resident implicit step (src/impdrv3d.f90:130-406, impl = 1,
 * niter = 1).  One context per problem; every array of the step stays in
 * HBM.  Field arrays v, r, dv are complex (5,nx,ny) in the reference layout.
 * ====================================================================== */
typedef struct lns3d_ctx lns3d_ctx;
int lns3d_ctx_create(const lns3d_params *p, lns3d_ctx **out);
int lns3d_ctx_destroy(lns3d_ctx *c);
int lns3d_ctx_set_stream(lns3d_ctx *c, void *stream);
/* new physics (beta, omega, Ma, Re, ... and boundary selectors) for an
 * existing context of the same grid size, periodicity and precision; form
 * the coefficients and factors again with lns3d_setup / lns3d_factor.  The
 * time scheme restarts (its alfa applies); the sponge and the top wave are
 * cleared (set them again). */
int lns3d_ctx_set_params(lns3d_ctx *c, const lns3d_params *p);
int lns3d_ctx_synchronize(lns3d_ctx *c);
/* metric (nx,ny,10: m1,m2,n1,n2,m11,m12,m22,n11,n12,n22; src/gengrid.f90:76-80)
 * and mean flow vm (5,nx,ny), host arrays; generates on the device the
 * coefficients of src/genmtrx.f90:2 (Ah,Bh,Dh,Dhi,Vh11,Vh12,Vh22,ABhi) and
 * the time step dtl of src/temporal.f90:12 (set_time_step). */
int lns3d_setup(lns3d_ctx *c, const double *metric, const double *vm);
/* Sponge (src/global.f90:132-141 ispg, spg, spg2, wamp, x0): the terms of
 * src/lrhs3d.f90:222-282 in the residual and the diagonal of
 * src/lhs2f3d.f90:207-225 (and its closure rows) in the eta factor, the
 * factor without the D terms.  ispg 0 (none), 1 (spg), 2 and 3 (spg, and
 * spg2 toward the incoming acoustic wave of frequency omega: x physical
 * coordinate, x0, and for 3 the amplitude wamp(nx)) or 5 (spg + spg2).  spg,
 * spg2, x are host (nx,ny) arrays as src/sponge.f90:20 (init_sponge) forms
 * them.  Call after lns3d_setup (the forcing uses the mean flow and the
 * context's omega, Ma, Re, Pr, gamma) and before lns3d_factor; again after
 * lns3d_ctx_set_params.  ispg = 4 (src/sponge.f90:536-606, cspg_it):
 * (spg + spg2)(v - cvic) toward the eigenfunction inflow continued
 * downstream, cvic = profile(j) exp(i lalpha x(i,j)); it needs left = 4 with
 * its profiles and lalpha given first (lns3d_set_inflow), x, and omega != 0. */
int lns3d_set_sponge(lns3d_ctx *c, int ispg, const double *spg, const double *spg2, const double *x,
                     const double *wamp, double x0);
/* top = 1 (src/rhsbc3d.f90:129-161, src/itrbc3d.f90:145-195): the incoming
 * acoustic wave of frequency omega on the top boundary, from the mean flow,
 * the physical x (host (nx,ny); its top row is used), the amplitude wamp (nx;
 * NULL for 1) and x0.  After lns3d_setup; lns3d_factor refuses top = 1
 * without it. */
int lns3d_set_top_forcing(lns3d_ctx *c, const double *x, const double *wamp, double x0);
/* Time scheme, the deck's impl (src/impdrv3d.f90:179-199): 1 backward Euler
 * (default), 2 two-step backward (rtb3D, src/impdrv3d.f90:473-501: the first
 * step backward Euler, then the second-order residual, the factors formed
 * again at the second step with alfa = 2/3), 3 alfa = 1/2.  Starts a new
 * run; lns3d_factor afterwards.  With impl = 2 lns3d_step and lns3d_advance
 * keep the field of the previous step on the device; the complex-float and
 * refined steps and graphs return LNS3D_EUNSUPPORTED. */
int lns3d_set_time_scheme(lns3d_ctx *c, int impl);
/* impl = 2 resumed from a restart holding vold (src/genini3d.f90:55-57):
 * vold_dev is the field at the start of the last step taken (device), lstep
 * the number of steps taken. */
int lns3d_set_time_history(lns3d_ctx *c, const void *vold_dev, long lstep);
/* left = 1, 4, 5 and right = 4, 5 boundary values (src/rhsbc3d.f90:176-307,
 * src/itrbc3d.f90:219-404): left = 1 the acoustic wave on the left column
 * (from the mean flow, x (host (nx,ny)) and x0: lns3d_setup first); left = 5
 * and right = 5 the undamped wave exp(i k x) of that column, x read where the
 * reference reads it (x(j,1), x(j,nx): memory offsets j and j + nx(nx-1));
 * left = 4
 * the eigenfunction profiles eig (host (10,ny): rhor, rhoi, ur, ui, vr, vi,
 * wr, wi, tr, ti, the columns of src/gendist.f90's inflow file; omega != 0);
 * right = 4 the same profiles times exp(i lalpha x(nx,j)), lalpha = (re, im)
 * of the inflow namelist.  lns3d_factor refuses these selectors without it. */
int lns3d_set_inflow(lns3d_ctx *c, const double *x, double x0, const double *eig,
                     const double *lalpha);
/* Sutherland's law, mattyp = 1 (src/getmat.f90:28-40): the constant S0 =
 * 110.4 K / Tref (src/input.f90:41-47).  Before lns3d_setup, which refuses
 * mattyp = 1 without it. */
int lns3d_set_sutherland(lns3d_ctx *c, double S0);
/* the bump wall, wall = 3 (src/rhsbc3d.f90:47-100, src/itrbc3d.f90:48-105,
 * src/lhsbc2f3d.f90:41-48): the wall values src/genbump.f90:74-131 forms
 * from the bump shape, uw host (nx,5) complex: u1w, u2w, u3w, tw, twp.
 * lns3d_factor refuses wall = 3 without them. */
int lns3d_set_wall(lns3d_ctx *c, const double *uw);
/* smoother, the deck's eps_e (src/smoother3d.f90 in the residual, called at
 * src/lrhs3d.f90:293; the implicit damping of src/lhs1f3d.f90:169-201,
 * lhs2f3d.f90:171-203 and the closure and symmetric rows): eps_e and the
 * damping profile buff (host (nx,ny)) that src/buffer.f90 (init_buffer)
 * forms from ibuff.  eps_e = 0 turns it off.  Before lns3d_factor. */
int lns3d_set_smoother(lns3d_ctx *c, double eps_e, const double *buff);
/* the shift-invert driver's pseudo-time steps (src/si_eigdrv3d_new.f90
 * :310-440, the operator an ARPACK shift-invert loop applies): one corrector
 * r = rhsBC3D(vold - v - dtl (lrhs3D(v) - f)) (ebe3D, src/si_eigdrv3d.f90
 * :319-345), the two line solves with the context's factors, v += r,
 * itrBC3D.  f, v, vold are device (5,nx,ny) complex; vold NULL means v (the
 * first corrector of a step); impl = 1. */
int lns3d_step_forced(lns3d_ctx *c, void *v_dev, const void *vold_dev, const void *f_dev, void *dv_dev);
/* the same residual r (into r_dev) and its norm sqrt(sum |r|^2 / dtl^2), the
 * driver's convergence measure (src/si_eigdrv3d_new.f90:359-374) */
int lns3d_forced_residual(lns3d_ctx *c, const void *v_dev, const void *vold_dev, const void *f_dev,
                          void *r_dev, double *norm_out);
/* residual norm sqrt(sum over points and components of |r|^2 / dtl^2) of r
 * in the step's form (dtl times the residual, with the rhsBC3D rows): the
 * measure src/resstat3d.f90:61-89 prints (totres, without its 1/N, which
 * cancels in the ratio to the first step's that it reports).  Written to
 * *norm_dev (device) with no host synchronisation; the sum's order is fixed. */
int lns3d_residual_norm(lns3d_ctx *c, const void *r_dev, double *norm_dev);
/* lns3d_residual_norm without a context and without allocating: for
 * callers that cannot call hipMalloc (code under HIP graph capture, serving
 * loops that forbid allocation).  No counterpart in the reference.  r_dev
 * holds npts x 5 complex (the step's residual), dtl_dev npts doubles, the
 * norm goes to *norm_dev (device); the current device's stream `stream`.
 * The per-block partial sums live in a static device array of 1024 doubles
 * (8 KiB) that each GPU holds until the process ends; every call on one GPU
 * shares it, so calls that overlap on one GPU (other streams, other host
 * threads) race and must be serialised by the caller.  Sizes: 1 <= npts <=
 * LNS3D_NOALLOC_MAX_POINTS, else LNS3D_EINVAL. */
#define LNS3D_NOALLOC_MAX_POINTS (1L << 28)
int lns3d_residual_norm_noalloc(const void *r_dev, const double *dtl_dev, long npts, double *norm_dev,
                                void *stream);
/* record the residual norm of every step: lns3d_step (first corrector),
 * lns3d_step_fp32 and lns3d_step_refined_fixed write the norm of the step's
 * residual (before the line solves, as src/impdrv3d.f90:207 calls
 * resstat3D) into norms_dev[k mod count], k counting the steps since this
 * call; graph-capturable.  NULL stops recording.  No counterpart in the
 * reference (its norm is printed per step on the host). */
int lns3d_ctx_set_residual_output(lns3d_ctx *c, double *norms_dev, int count);
/* corrector iterations per step, the deck's niter (src/impdrv3d.f90:173-358):
 * the residual with vold held at the start of the step, solved with the
 * factors of the first iteration, niter times.  lns3d_step only (impl 1 or
 * 2); dv_dev receives the last iteration's increment. */
int lns3d_set_corrector_iterations(lns3d_ctx *c, int niter);
/* alternatively, take the coefficients the caller already has (host arrays
 * in the module layouts of src/local2d.f90:34-36, src/local3d.f90:14) */
int lns3d_set_coeffs(lns3d_ctx *c, const double *Ah, const double *Bh, const double *Dh,
                     const double *Dhi, const double *Vh11, const double *Vh12,
                     const double *Vh22, const double *ABhi, const double *dtl);
int lns3d_get_coeffs(lns3d_ctx *c, double *Ah, double *Bh, double *Dh, double *Dhi,
                     double *Vh11, double *Vh12, double *Vh22, double *ABhi, double *dtl);
/* the device form of src/lhs1f3d.f90:2 lhs1f3D + src/lhs2f3d.f90:2 lhs2f3D
 * (with lhsbt1f3D/lhsbt2f3d/lhsbc1f3d/lhsbc2f3d) and the factor calls
 * cpenta1bc/cpenta2p(..,0) of src/impdrv3d.f90:252-269,314-319; checks the
 * factors as lns3d_penta_factor does (LNS3D_ENONFINITE: the context stays
 * unfactored) */
int lns3d_factor(lns3d_ctx *c);
/* assembled LHS of direction dir (1 xi, 2 eta) into mat_dev (5,5,5,nx,ny) */
int lns3d_assemble(lns3d_ctx *c, int dir, void *mat_dev);
/* the device form of src/lrhs3d.f90:2 lrhs3D: r = L(v) on device arrays
 * with the context's coefficients (the drop-in with the reference's
 * argument list is lns3d_drop_lrhs3d) */
int lns3d_lrhs3d(lns3d_ctx *c, const void *v_dev, void *r_dev);
/* the device form of src/itrbc3d.f90:2 itrBC3D(vl,vml) */
int lns3d_itrbc3d(lns3d_ctx *c, void *v_dev);
/* xi then eta line solves of r in place (src/impdrv3d.f90:284-327) */
int lns3d_solve(lns3d_ctx *c, void *r_dev);
/* one direction only: dir 1 the xi lines (the cpenta1bc call of
 * src/impdrv3d.f90:295 on the transposed array), dir 2 the eta lines (:327) */
int lns3d_solve_dir(lns3d_ctx *c, int dir, void *r_dev);
/* one step: lrhs3D, rbe3D, rhsBC3D, xi and eta solves, v += r, itrBC3D;
 * v must already satisfy the boundary conditions; dv_dev (may be NULL)
 * receives the increment r */
int lns3d_step(lns3d_ctx *c, void *v_dev, void *dv_dev);
/* the same step with complex-float factors (precision = 1) used as a
 * defect-correction preconditioner of the complex-double AF system until
 * |res|_max <= tol |r|_max or max_iters solves (max_iters <= 0: up to 64).
 * The update is applied either way: *resid_out > tol on return says that
 * the solve count ran out first.  A non-finite residual returns
 * LNS3D_ENONFINITE and leaves v unchanged.  Periodic lines: complex-float
 * records rounded from the complex-double ones; the refinement converges to
 * the cyclic system's solution (with periodic_exact = 0 the shipped factor
 * misses it on short lines), and the increment's duplicate periodic column
 * (xper, which itrBC3D overwrites in v) keeps complex-float accuracy. */
int lns3d_step_refined(lns3d_ctx *c, void *v_dev, void *dv_dev, double tol, int max_iters,
                       int *iters_out, double *resid_out);
/* the step with complex-float factors and no inner refinement */
int lns3d_step_fp32(lns3d_ctx *c, void *v_dev, void *dv_dev);

/* The refined step with a fixed number of corrections and no host
 * synchronisation (graph-capturable); check = 1 adds one operator
 * application whose relative residual lns3d_last_residual returns.  No
 * counterpart in the reference (its step is complex double). */
int lns3d_step_refined_fixed(lns3d_ctx *c, void *v_dev, void *dv_dev, int corrections, int check);
int lns3d_last_residual(lns3d_ctx *c, double *rel);
/* device bytes a factored context of these parameters holds (admission by
 * memory budget in the serving loop).  The optional setters add their own
 * blocks on first use (P = nx ny): lns3d_set_sponge 8 (14 P + nx),
 * lns3d_set_smoother 8 P, lns3d_set_time_scheme(2) and
 * lns3d_set_corrector_iterations(> 1) 80 P each, lns3d_set_top_forcing
 * 96 nx, lns3d_set_inflow 176 ny. */
int lns3d_ctx_bytes(const lns3d_params *p, size_t *bytes);

/* Graphs of steps: nsteps steps of each of count factored contexts,
 * captured once and replayed with one launch (batches of uneven
 * instances).  mode 0 complex double (lns3d_step), 1 complex-float factors
 * (lns3d_step_fp32), 2 complex-float factors with `corrections` defect
 * corrections (lns3d_step_refined_fixed): every context on its own branch.
 * mode 3: the steps of mode 2 with one launch per stage over all contexts
 * (residual, residual norms, line solves grouped by solve kernel,
 * corrections, update), the same arithmetic; with corrections = 0 the steps
 * of mode 1; contexts of the inviscid deck take bcfix between the last
 * correction and the update, as their own steps do, and periodic lines
 * their periodic solve kernels.  Contexts with a residual output
 * (lns3d_ctx_set_residual_output) record each step's norm in their next
 * slot, with the bits of the per-context steps.  A mode-3 graph's tables are
 * copied to the device on the stream of its first launch, ahead of it
 * (creating the graph enqueues no device work).  dv_dev may be NULL.
 * Launch on `stream` (NULL: the graph's own stream).  Every captured step
 * must be alike: updateLHS with useCalcd is refused, impl = 2 and corrector
 * iterations take mode 0 (impl = 2 from its third step, with corrector
 * iterations an even nsteps: the history buffers trade places each step).
 * No counterpart in the reference. */
typedef struct lns3d_graph lns3d_graph;
int lns3d_graph_create(lns3d_ctx **ctxs, void **v_dev, void **dv_dev, int count, int nsteps, int mode,
                       int corrections, lns3d_graph **out);
int lns3d_graph_launch(lns3d_graph *g, void *stream);
int lns3d_graph_synchronize(lns3d_graph *g);
int lns3d_graph_destroy(lns3d_graph *g);
/* autotuner of the line solves: times every solve-kernel variant on
 * synthetic lines of n rows x lines and uses the fastest for plans of that
 * shape created later.  Variants: 0 one line per workgroup, 1 four lines,
 * 2 twisted (two chains per line, n >= 16), 3 two lines, 4 three lines,
 * 5 twisted with chain steps taken in pairs, 6 twisted with each row's
 * columns split over the two halves of a wave.  times_ms (may be NULL)
 * receives LNS3D_TUNE_VARIANTS timings in ms (-1: not applicable).  No
 * counterpart in the reference. */
#define LNS3D_TUNE_VARIANTS 7
int lns3d_tune_lines(int n, long lines, int precision, int *variant, float *times_ms);
/* record a solve variant for the shape (n rows x lines, precision) without
 * timing it, for plans created later (a twisted variant on lines shorter
 * than 16 rows falls back to variant 0); variant -1 drops the record, so
 * that the built-in choice applies again.  No counterpart in the reference. */
int lns3d_tune_set(int n, long lines, int precision, int variant);

/* ======================================================================
 * Reference file formats (Fortran sequential unformatted: 4-byte record
 * length markers, 4-byte integers, 8-byte reals, 16-byte complex), host
 * arrays in the reference layouts.  Readers check every record length.
 * ====================================================================== */
/* first record of restart and mean files (src/impdrv3d.f90:396-397) */
typedef struct {
  int lstep;
  double time;
  int nx, ny, nz, ndof;
  double Re, Ma, Pr, gamma, cv;
} lns3d_field_header;
int lns3d_read_header(const char *path, lns3d_field_header *h);
/* restart file: header, v complex(ndof,nx,ny), and vold when impl = 2
 * (written by src/impdrv3d.f90:393-399, read by src/genini3d.f90:26-57).
 * v must hold ndof*nx*ny*nz complex values (read the header first); vold may
 * be NULL; *got_vold (may be NULL) tells whether the file held vold. */
int lns3d_read_restart(const char *path, lns3d_field_header *h, void *v, void *vold, int *got_vold);
int lns3d_write_restart(const char *path, const lns3d_field_header *h, const void *v, const void *vold);
/* mean-field file: header and vm real(ndof,nx,ny) (src/genmean.f90:28-33) */
int lns3d_read_mean(const char *path, lns3d_field_header *h, double *vm);
int lns3d_write_mean(const char *path, const lns3d_field_header *h, const double *vm);
/* grid file: (nx, ny, nz) then x, y, z of (nx,ny,nz) reals in one record
 * (src/gengrid.f90:17-32).  With x and y NULL only the sizes are read; the
 * writer stores z = 0. */
int lns3d_read_grid(const char *path, int *nx, int *ny, int *nz, double *x, double *y);
int lns3d_write_grid(const char *path, int nx, int ny, int nz, const double *x, const double *y);
/* metric file: m1, m2, n1, n2, m11, m12, m22, n11, n12, n22 of (nx,ny) reals
 * in one record (src/gengrid.f90:76-80), the metric layout of lns3d_setup */
int lns3d_read_metric(const char *path, int nx, int ny, double *metric);
int lns3d_write_metric(const char *path, int nx, int ny, const double *metric);

/* Host-array driving of the time loop of src/impdrv3d.f90:126-406 (for
 * callers that hold v in host memory, such as the Fortran host layer):
 * upload v (5,nx,ny) once, advance nsteps steps on the device (the first
 * advance after an upload applies itrBC3D first, as impdrv3d.f90:126 does),
 * download v when the caller needs it.  dv_host (may be NULL) receives the
 * last increment r, the array impdrv3d.f90:340-348 reads for its norms. */
int lns3d_state_upload(lns3d_ctx *c, const void *v_host);
int lns3d_state_download(lns3d_ctx *c, void *v_host);
int lns3d_advance(lns3d_ctx *c, int nsteps, void *dv_host);

#ifdef __cplusplus
}
#endif
#endif
