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
 * Device-resident solver handles: factor once from a device matrix in the
 * reference's layout, then solve device right-hand sides in place.
 * kind 0 = non-periodic (cpenta*bc), 1 = periodic (cpenta*p);
 * layout 1 = system index before row index (cpenta1*), 2 = after (cpenta2*);
 * nrows = n (kind 0) or np (kind 1); precision 0 = complex double factors,
 * 1 = complex float factors (kind 0 only); exact as above.
 * ====================================================================== */
typedef struct lns3d_penta lns3d_penta;
int lns3d_penta_create(int kind, int layout, int nsys, int nrows, int precision, int exact,
                       lns3d_penta **out);
int lns3d_penta_factor(lns3d_penta *h, const void *mat_dev, void *stream);
int lns3d_penta_solve(lns3d_penta *h, void *rhs_dev, void *stream);
/* same result, one lane per line in the reference's operation order */
int lns3d_penta_solve_reference_order(lns3d_penta *h, void *rhs_dev, void *stream);
int lns3d_penta_destroy(lns3d_penta *h);

#ifdef __cplusplus
}
#endif
#endif
