/*
 * Oracle: finite-difference derivatives, linearized RHS, LHS assembly with
 * boundary closures and boundary conditions, coefficient generation, time
 * step, and the implicit step driver.  TEST INFRASTRUCTURE ONLY (see
 * lns3d_oracle.h).  Supported options are the ones the benchmark and test
 * configurations use (supported() below lists them); anything else returns
 * -1 rather than guessing.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "lns3d_oracle.h"

/* src/stencil.f90 coefficient tables (the values are the scheme) */
static const double ga1 = 8.333333333333333333333E-02, ga2 = -6.666666666666666666667E-01,
                    ga3 = 6.666666666666666666667E-01, ga4 = -8.333333333333333333333E-02;
static const double gc1 = -2.083333333333333333333E+00, gc2 = 4.000000000000000000000E+00,
                    gc3 = -3.000000000000000000000E+00, gc4 = 1.333333333333333333333E+00,
                    gc5 = -2.500000000000000000000E-01;
static const double gb1 = -2.500000000000000000000E-01, gb2 = -8.333333333333333333333E-01,
                    gb3 = 1.500000000000000000000E+00, gb4 = -5.000000000000000000000E-01,
                    gb5 = 8.333333333333333333333E-02;
static const double da1 = -8.333333333333333333333E-02, da2 = 1.333333333333333333333E+00,
                    da3 = -2.500000000000000000000E+00, da4 = 1.333333333333333333333E+00,
                    da5 = -8.333333333333333333333E-02;
static const double db1 = 9.166666666666666666667E-01, db2 = -1.666666666666666666667E+00,
                    db3 = 5.000000000000000000000E-01, db4 = 3.333333333333333333333E-01,
                    db5 = -8.333333333333333333333E-02;
static const double dd1 = 2.916666666666666666667E+00, dd2 = -8.666666666666666666667E+00,
                    dd3 = 9.500000000000000000000E+00, dd4 = -4.666666666666666666667E+00,
                    dd5 = 9.166666666666666666667E-01;
static const double pt66 = 6.6666666666666666666e-1;

/* interior 7-point coefficients, src/cgrad.f90:34-78, src/cgrad2.f90:38-133 */
static void first_coefs(int opt, double dinv, double g[6]) {
  double c;
  if (opt == 0) c = 1.0 / 60.0;
  else if (opt == -1) c = 0.0;
  else {
    double w = 2.0 * 3.1415926535897932385e+0 / 12.0;
    c = (w / 2.0 - 2.0 * sin(w) / 3.0 + sin(2.0 * w) / 12.0) /
        (5.0 * sin(w) - 4.0 * sin(2.0 * w) + sin(3.0 * w));
  }
  double a = (2.0 + 15.0 * c) / 3.0, b = -(1.0 + 48.0 * c) / 12.0;
  g[0] = -c * dinv; g[1] = -b * dinv; g[2] = -a * dinv;
  g[3] = a * dinv;  g[4] = b * dinv;  g[5] = c * dinv;
}
static void second_coefs(int opt, double dsinv, double d[7]) {
  double c;
  if (opt == 0) c = 1.0 / 90.0;
  else if (opt == -1) c = 0.0;
  else {
    double w = 2.0 * 3.1415926535897932385e+0 / 12.0;
    double s1 = sin(w / 2.0), s2 = sin(w), s3 = sin(3.0 * w / 2.0);
    c = -(3.0 * w * w - 16.0 * s1 * s1 + s2 * s2) /
        (12.0 * (-15.0 * s1 * s1 + 6.0 * s2 * s2 - s3 * s3));
  }
  double a = (4.0 + 45.0 * c) / 3.0, b = -(1.0 + 72.0 * c) / 12.0;
  d[0] = c * dsinv; d[1] = b * dsinv; d[2] = a * dsinv;
  d[3] = -2.0 * (a + b + c) * dsinv;
  d[4] = a * dsinv; d[5] = b * dsinv; d[6] = c * dsinv;
}

/* Symmetry boundaries (left/right = 7: src/cgrad.f90:190-262,
 * src/cgrad2.f90:229-310, src/grad.f90:155-290, src/grad2.f90:250-330): the
 * three points next to a symmetric end take the interior stencil, the points
 * beyond the end reflected onto the line (index q -> -q on the left,
 * 2(N-1) - q on the right), component 3 (index 2) odd.  sym: bit 0 left,
 * bit 1 right. */
static int sym_row(int sym, int p, int N) {
  return ((sym & 1) && p <= 2) || ((sym & 2) && p >= N - 3);
}
static double reflect(int *q, int N, int e) {
  double s = e == 2 ? -1.0 : 1.0;
  if (*q < 0) { *q = -*q; return s; }
  if (*q > N - 1) { *q = 2 * (N - 1) - *q; return s; }
  return 1.0;
}
static int xsym(const orc_params *p) {
  return p->xper ? 0 : ((p->left == 7 ? 1 : 0) | (p->right == 7 ? 2 : 0));
}

/* One line of N points: values at x[p*xs + e] for e < ne.  Generic over
 * real/complex via macros instantiated twice below. */
#define DEFINE_LINE_OPS(T, SUF)                                                       \
static void d1_line_##SUF(int N, int per, int sym, const double g[6], double dinv,   \
                          int ne, const T *x, long xs, T *y, long ys) {               \
  for (int p = 0; p < N; p++)                                                         \
    for (int e = 0; e < ne; e++) {                                                    \
      T r;                                                                            \
      if (per) {                                                                      \
        if (p == N - 1) continue;                                                     \
        int m = N - 1; /* distinct points */                                          \
        int q[6] = {p - 3, p - 2, p - 1, p + 1, p + 2, p + 3};                        \
        for (int k = 0; k < 6; k++) q[k] = ((q[k] % m) + m) % m;                       \
        r = g[0] * x[q[0] * xs + e] + g[1] * x[q[1] * xs + e] + g[2] * x[q[2] * xs + e] \
          + g[3] * x[q[3] * xs + e] + g[4] * x[q[4] * xs + e] + g[5] * x[q[5] * xs + e]; \
      } else if (sym_row(sym, p, N)) {                                                \
        int q[6] = {p - 3, p - 2, p - 1, p + 1, p + 2, p + 3};                        \
        double sg[6];                                                                 \
        for (int k = 0; k < 6; k++) sg[k] = reflect(&q[k], N, e);                      \
        r = g[0] * (sg[0] * x[q[0] * xs + e]) + g[1] * (sg[1] * x[q[1] * xs + e])     \
          + g[2] * (sg[2] * x[q[2] * xs + e]) + g[3] * (sg[3] * x[q[3] * xs + e])     \
          + g[4] * (sg[4] * x[q[4] * xs + e]) + g[5] * (sg[5] * x[q[5] * xs + e]);    \
      } else if (p == 0) {                                                            \
        r = (gc1 * x[0 * xs + e] + gc2 * x[1 * xs + e] + gc3 * x[2 * xs + e] +        \
             gc4 * x[3 * xs + e] + gc5 * x[4 * xs + e]) * dinv;                       \
      } else if (p == 1) {                                                            \
        r = (gb1 * x[0 * xs + e] + gb2 * x[1 * xs + e] + gb3 * x[2 * xs + e] +        \
             gb4 * x[3 * xs + e] + gb5 * x[4 * xs + e]) * dinv;                       \
      } else if (p == 2) {                                                            \
        r = (ga1 * x[0 * xs + e] + ga2 * x[1 * xs + e] + ga3 * x[3 * xs + e] +        \
             ga4 * x[4 * xs + e]) * dinv;                                             \
      } else if (p == N - 3) {                                                        \
        r = (ga1 * x[(N - 5) * xs + e] + ga2 * x[(N - 4) * xs + e] +                  \
             ga3 * x[(N - 2) * xs + e] + ga4 * x[(N - 1) * xs + e]) * dinv;           \
      } else if (p == N - 2) {                                                        \
        r = -(gb1 * x[(N - 1) * xs + e] + gb2 * x[(N - 2) * xs + e] +                 \
              gb3 * x[(N - 3) * xs + e] + gb4 * x[(N - 4) * xs + e] +                 \
              gb5 * x[(N - 5) * xs + e]) * dinv;                                      \
      } else if (p == N - 1) {                                                        \
        r = -(gc1 * x[(N - 1) * xs + e] + gc2 * x[(N - 2) * xs + e] +                 \
              gc3 * x[(N - 3) * xs + e] + gc4 * x[(N - 4) * xs + e] +                 \
              gc5 * x[(N - 5) * xs + e]) * dinv;                                      \
      } else {                                                                        \
        r = g[0] * x[(p - 3) * xs + e] + g[1] * x[(p - 2) * xs + e] +                 \
            g[2] * x[(p - 1) * xs + e] + g[3] * x[(p + 1) * xs + e] +                 \
            g[4] * x[(p + 2) * xs + e] + g[5] * x[(p + 3) * xs + e];                  \
      }                                                                               \
      y[p * ys + e] = r;                                                              \
    }                                                                                 \
  if (per)                                                                            \
    for (int e = 0; e < ne; e++) y[(N - 1) * ys + e] = y[e];                          \
}                                                                                     \
static void d2_line_##SUF(int N, int per, int sym, const double d[7], double dsinv,   \
                          int ne, const T *x, long xs, T *y, long ys) {               \
  for (int p = 0; p < N; p++)                                                         \
    for (int e = 0; e < ne; e++) {                                                    \
      T r;                                                                            \
      if (per) {                                                                      \
        if (p == N - 1) continue;                                                     \
        int m = N - 1;                                                                \
        int q[7];                                                                     \
        for (int k = 0; k < 7; k++) q[k] = (((p - 3 + k) % m) + m) % m;                \
        r = d[0] * x[q[0] * xs + e] + d[1] * x[q[1] * xs + e] + d[2] * x[q[2] * xs + e] \
          + d[3] * x[q[3] * xs + e] + d[4] * x[q[4] * xs + e] + d[5] * x[q[5] * xs + e] \
          + d[6] * x[q[6] * xs + e];                                                  \
      } else if (sym_row(sym, p, N)) {                                                \
        int q[7];                                                                     \
        double sg[7];                                                                 \
        for (int k = 0; k < 7; k++) { q[k] = p - 3 + k; sg[k] = reflect(&q[k], N, e); } \
        r = d[0] * (sg[0] * x[q[0] * xs + e]) + d[1] * (sg[1] * x[q[1] * xs + e])     \
          + d[2] * (sg[2] * x[q[2] * xs + e]) + d[3] * (sg[3] * x[q[3] * xs + e])     \
          + d[4] * (sg[4] * x[q[4] * xs + e]) + d[5] * (sg[5] * x[q[5] * xs + e])     \
          + d[6] * (sg[6] * x[q[6] * xs + e]);                                        \
      } else if (p == 0) {                                                            \
        r = (dd1 * x[0 * xs + e] + dd2 * x[1 * xs + e] + dd3 * x[2 * xs + e] +        \
             dd4 * x[3 * xs + e] + dd5 * x[4 * xs + e]) * dsinv;                      \
      } else if (p == 1) {                                                            \
        r = (db1 * x[0 * xs + e] + db2 * x[1 * xs + e] + db3 * x[2 * xs + e] +        \
             db4 * x[3 * xs + e] + db5 * x[4 * xs + e]) * dsinv;                      \
      } else if (p == 2) {                                                            \
        r = (da1 * x[0 * xs + e] + da2 * x[1 * xs + e] + da3 * x[2 * xs + e] +        \
             da4 * x[3 * xs + e] + da5 * x[4 * xs + e]) * dsinv;                      \
      } else if (p == N - 3) {                                                        \
        r = (da1 * x[(N - 5) * xs + e] + da2 * x[(N - 4) * xs + e] +                  \
             da3 * x[(N - 3) * xs + e] + da4 * x[(N - 2) * xs + e] +                  \
             da5 * x[(N - 1) * xs + e]) * dsinv;                                      \
      } else if (p == N - 2) {                                                        \
        r = (db1 * x[(N - 1) * xs + e] + db2 * x[(N - 2) * xs + e] +                  \
             db3 * x[(N - 3) * xs + e] + db4 * x[(N - 4) * xs + e] +                  \
             db5 * x[(N - 5) * xs + e]) * dsinv;                                      \
      } else if (p == N - 1) {                                                        \
        r = (dd1 * x[(N - 1) * xs + e] + dd2 * x[(N - 2) * xs + e] +                  \
             dd3 * x[(N - 3) * xs + e] + dd4 * x[(N - 4) * xs + e] +                  \
             dd5 * x[(N - 5) * xs + e]) * dsinv;                                      \
      } else {                                                                        \
        r = d[0] * x[(p - 3) * xs + e] + d[1] * x[(p - 2) * xs + e] +                 \
            d[2] * x[(p - 1) * xs + e] + d[3] * x[p * xs + e] +                       \
            d[4] * x[(p + 1) * xs + e] + d[5] * x[(p + 2) * xs + e] +                 \
            d[6] * x[(p + 3) * xs + e];                                               \
      }                                                                               \
      y[p * ys + e] = r;                                                              \
    }                                                                                 \
  if (per)                                                                            \
    for (int e = 0; e < ne; e++) y[(N - 1) * ys + e] = y[e];                          \
}

DEFINE_LINE_OPS(zdouble, z)
DEFINE_LINE_OPS(double, d)

/* first and second derivatives of an ndof-component field (nx,ny):
 * src/cgrad.f90:31-461 and src/cgrad2.f90:36-620 (non-Carpenter branches) */
#define DEFINE_FIELD_OPS(T, SUF)                                                      \
static void grad_##SUF(const orc_params *p, int nd, const T *v, T *g1, T *g2) {       \
  int nx = p->nx, ny = p->ny;                                                         \
  double dxi = 1.0 / (double)(nx - 1), deta = 1.0 / (double)(ny - 1);                 \
  double dxinv = 1.0 / dxi, dyinv = 1.0 / deta, gx[6], gy[6];                         \
  first_coefs(p->optx, dxinv, gx);                                                    \
  first_coefs(p->opty, dyinv, gy);                                                    \
  for (int j = 0; j < ny; j++)                                                        \
    d1_line_##SUF(nx, p->xper, xsym(p), gx, dxinv, nd, v + (long)nd * nx * j, nd, g1 + (long)nd * nx * j, nd); \
  for (int i = 0; i < nx; i++)                                                        \
    d1_line_##SUF(ny, p->yper, 0, gy, dyinv, nd, v + (long)nd * i, (long)nd * nx, g2 + (long)nd * i, (long)nd * nx); \
}                                                                                     \
static void grad2_##SUF(const orc_params *p, int nd, const T *v, const T *g1,         \
                        T *g11, T *g12, T *g22) {                                     \
  int nx = p->nx, ny = p->ny;                                                         \
  double dxi = 1.0 / (double)(nx - 1), deta = 1.0 / (double)(ny - 1);                 \
  double dxinv = 1.0 / dxi, dyinv = 1.0 / deta;                                       \
  double dxsinv = 1.0 / (dxi * dxi), dysinv = 1.0 / (deta * deta);                    \
  double gy[6], dx[7], dy[7];                                                         \
  first_coefs(p->opty, dyinv, gy);                                                    \
  second_coefs(p->optx, dxsinv, dx);                                                  \
  second_coefs(p->opty, dysinv, dy);                                                  \
  for (int j = 0; j < ny; j++)                                                        \
    d2_line_##SUF(nx, p->xper, xsym(p), dx, dxsinv, nd, v + (long)nd * nx * j, nd, g11 + (long)nd * nx * j, nd); \
  for (int i = 0; i < nx; i++) {                                                      \
    d2_line_##SUF(ny, p->yper, 0, dy, dysinv, nd, v + (long)nd * i, (long)nd * nx, g22 + (long)nd * i, (long)nd * nx); \
    d1_line_##SUF(ny, p->yper, 0, gy, dyinv, nd, g1 + (long)nd * i, (long)nd * nx, g12 + (long)nd * i, (long)nd * nx); \
  }                                                                                   \
}

DEFINE_FIELD_OPS(zdouble, z)
DEFINE_FIELD_OPS(double, d)

void orc_cgrad(const orc_params *p, const zdouble *v, zdouble *g1, zdouble *g2) {
  grad_z(p, 5, v, g1, g2);
}
void orc_cgrad2(const orc_params *p, const zdouble *v, const zdouble *g1,
                zdouble *g11, zdouble *g12, zdouble *g22) {
  grad2_z(p, 5, v, g1, g11, g12, g22);
}
void orc_grad(const orc_params *p, int ndof, const double *v, double *g1, double *g2) {
  grad_d(p, ndof, v, g1, g2);
}
void orc_grad2(const orc_params *p, int ndof, const double *v, const double *g1,
               double *g11, double *g12, double *g22) {
  grad2_d(p, ndof, v, g1, g11, g12, g22);
}

static int supported(const orc_params *p) {
  if (p->nx < 7 || p->ny < 7) return 0;
  if (p->wall < 0 || p->wall > 3 || p->wallt < 0 || p->wallt > 2) return 0;
  if ((p->top != 0 && p->top != 1) ||
      (p->left != 0 && p->left != 1 && p->left != 4 && p->left != 5 && p->left != 7) ||
      (p->right != 0 && p->right != 4 && p->right != 5 && p->right != 7 && p->right != 8 && p->right != 9))
    return 0;
  if (!p->xper && p->right == 5 && p->ny - 1 + (long)p->nx * (p->nx - 1) >= (long)p->nx * p->ny)
    return 0;                                    /* x(j,nx) off the grid (nx > ny) */
  if (p->right == 4 && p->left != 4) return 0;   /* the profiles exist for left = 4 only */
  /* symmetric rows read the spanwise cross entries through a transposed
   * shape (src/lhsbs1f3d.f90:17, Ahi(6,ny,nx)): point i + ny j, which must
   * lie on the grid (it does for nx >= ny) */
  if (!p->xper && p->left == 7 && 1 + (long)p->ny * (p->ny - 1) >= (long)p->nx * p->ny) return 0;
  if (!p->xper && p->right == 7 && p->nx - 1 + (long)p->ny * (p->ny - 1) >= (long)p->nx * p->ny)
    return 0;
  /* the deck's linear = 2 (complex) becomes linear = 1 after input
   * (src/input.f90:367-371), and the 3D driver runs only with linear = 1
   * (src/lns3d.f90:153-167): both decks take the linear = 1 forms */
  if (p->linear != 1 && p->linear != 2) return 0;
  if (p->idamp < 0 || p->idamp > 2) return 0;
  if (p->mattyp < 0 || p->mattyp > 3) return 0;
  return 1;
}

/* Navier (src/input.f90:31-36): Re = 0 with constant viscosity is the
 * inviscid deck, which holds Re = 1e99 */
static int navier(const orc_params *p) { return !(p->Re == 0.0 && p->mattyp == 0); }
static double re_eff(const orc_params *p) { return navier(p) ? p->Re : 1.0e99; }
/* the wall extent is:ie (0-based, inclusive) */
static int wall_lo(const orc_params *p) { return p->wall_is ? p->wall_is - 1 : 0; }
static int wall_hi(const orc_params *p) { return p->wall_ie ? p->wall_ie - 1 : p->nx - 1; }
static int on_wall(const orc_params *p, int i) { return i >= wall_lo(p) && i <= wall_hi(p); }

/* wall normals bnb(nx,2) (src/gengrid.f90:133-137) for the inviscid rows */
static const double *g_bnb;
int orc_set_wall_normals(const double *bnb) {
  g_bnb = bnb;
  return 0;
}

/* ---- sponge: the module state src/lrhs3d.f90:222-289 and the LHS routines
 *      read (spg, spg2 of src/sponge.f90:20, wamp, x0, omega) ----------------- */
static orc_sponge g_spg;          /* ispg = 0: no sponge */
/* the bump wall (wall = 3): the wall values of src/genbump.f90:74-131, uw
 * (nx,5) complex: u1w, u2w, u3w, tw, twp */
static const zdouble *g_bump;

int orc_set_bump(const zdouble *uw) {
  g_bump = uw;
  return 0;
}
static struct { const double *x, *vm, *eig; double x0; zdouble lalpha; } g_inf;   /* see orc_set_inflow */

int orc_set_sponge(const orc_sponge *s) {
  if (!s || s->ispg == 0) { memset(&g_spg, 0, sizeof g_spg); return 0; }
  if (s->ispg < 0 || s->ispg > 5 || !s->spg) return -1;
  if (s->ispg >= 2 && !s->spg2) return -1;
  /* ispg = 4 sponges towards the eigenfunction inflow continued downstream:
   * it needs the profiles and lalpha (orc_set_inflow first) and x */
  if (s->ispg == 4 && (!s->x || !g_inf.eig)) return -1;
  if ((s->ispg == 2 || s->ispg == 3) && (!s->x || !s->vm)) return -1;
  if (s->ispg == 3 && !s->wamp) return -1;
  g_spg = *s;
  return 0;
}

/* LHS diagonal sponge of the factor without the D terms (src/lhs1f3d.f90:207-225,
 * lhsbt1f3d.f90:156-174): spg (ispg = 1) or spg + spg2 (ispg >= 2) */
static double lhs_sponge(long q) {
  if (g_spg.ispg == 0) return 0.0;
  return g_spg.ispg == 1 ? g_spg.spg[q] : g_spg.spg[q] + g_spg.spg2[q];
}

/* the incoming acoustic wave of src/lrhs3d.f90:248-282 and
 * src/rhsbc3d.f90:131-159: amplitude amp, mean state vml at physical x xl;
 * f[m] is the wave's disturbance in component m */
static void wave_terms(const orc_params *p, const double *vml, zdouble c3, zdouble f[5]) {
  const double cm = sqrt(vml[4]) / p->Ma;
  const double rho = vml[0], T = vml[4];
  const double gM2 = p->gamma * p->Ma * p->Ma;
  f[0] = 0.5 * c3 / (cm * cm);
  f[1] = c3 * 0.5 / (rho * cm);
  f[2] = 0;
  f[3] = 0;
  f[4] = (gM2 * c3 * 0.5 - T * 0.5 * c3 / (cm * cm)) / rho;
}

static void acoustic_wave(const orc_params *p, const double *vml, double xl, double x0, double amp,
                          zdouble f[5]) {
  const double cm = sqrt(vml[4]) / p->Ma, um = vml[1];
  const double d = 0.5 * (1.3333333333333333333 + (p->gamma - 1.0) / p->Pr) / re_eff(p);
  const double kk = p->omega / (cm + um);
  const double a = p->omega * p->omega * d / ((cm + um) * (cm + um) * (cm + um));
  const zdouble c3 = amp * (exp(-a * (xl - x0)) * cexp(_Complex_I * (kk * xl)));
  wave_terms(p, vml, c3, f);
}

/* left = 5 / right = 5 (src/rhsbc3d.f90:248-270, 286-307, itrbc3d.f90:320-340,
 * 382-404): the undamped wave exp(i k x) of amplitude 1 */
static void plain_wave(const orc_params *p, const double *vml, double xl, zdouble f[5]) {
  const double cm = sqrt(vml[4]) / p->Ma, um = vml[1];
  const double kk = p->omega / (cm + um);
  wave_terms(p, vml, cexp(_Complex_I * (kk * xl)), f);
}

/* left/right inflow state: left = 1 (the acoustic wave on x(1,j), from the
 * mean flow, x and x0: src/rhsbc3d.f90:176-205, src/itrbc3d.f90:219-245),
 * left = 4 (the eigenfunction profiles rhor, rhoi, ..., ti of
 * src/gendist.f90, eig[10 j + 2 d] + i eig[10 j + 2 d + 1]: src/rhsbc3d.f90
 * :207-246, src/itrbc3d.f90:247-318, omega != 0 so no transient factor) and
 * right = 4 (the same profiles times exp(i lalpha x(nx,j)):
 * src/itrbc3d.f90:352-376) */
int orc_set_inflow(const double *x, const double *vm, double x0, const double *eig,
                   const double *lalpha) {
  g_inf.x = x; g_inf.vm = vm; g_inf.x0 = x0; g_inf.eig = eig;
  g_inf.lalpha = lalpha ? lalpha[0] + _Complex_I * lalpha[1] : 0;
  return 0;
}

/* the values the left column (side 0) or the right column (side 1) is held
 * at on line j; -1 when the state they need is missing */
static int side_values(const orc_params *p, int side, int j, zdouble f[5]) {
  const long q = (side ? p->nx - 1 : 0) + (long)p->nx * j;
  const int sel = side ? p->right : p->left;
  if (sel == 1) {
    if (!g_inf.x || !g_inf.vm) return -1;
    acoustic_wave(p, g_inf.vm + 5 * q, g_inf.x[q], g_inf.x0, 1.0, f);
    return 0;
  }
  if (sel == 5) {
    /* the reference reads x(j,1) and x(j,nx): memory offsets j - 1 and
     * j - 1 + nx (nx - 1) of x(nx,ny), transposed; on the grid for left = 5,
     * and for right = 5 when nx <= ny (supported() checks) */
    if (!g_inf.x || !g_inf.vm) return -1;
    const long xo = side ? j + (long)p->nx * (p->nx - 1) : j;
    plain_wave(p, g_inf.vm + 5 * q, g_inf.x[xo], f);
    return 0;
  }
  if (sel == 4) {
    if (!g_inf.eig || p->omega == 0.0) return -1;
    for (int d = 0; d < 5; d++) f[d] = g_inf.eig[10 * j + 2 * d] + _Complex_I * g_inf.eig[10 * j + 2 * d + 1];
    if (side) {
      if (!g_inf.x) return -1;
      const zdouble e = cexp(_Complex_I * g_inf.lalpha * g_inf.x[q]);
      for (int d = 0; d < 5; d++) f[d] = f[d] * e;
    }
    return 0;
  }
  return -1;
}

/* top = 1 forcing state (src/global.f90 x, wamp, x0): wamp NULL means 1 */
static struct { const double *x, *wamp, *vm; double x0; } g_top;

int orc_set_top(const double *x, const double *wamp, const double *vm, double x0) {
  g_top.x = x; g_top.wamp = wamp; g_top.vm = vm; g_top.x0 = x0;
  return 0;
}

/* the wave on the top boundary at column i (src/rhsbc3d.f90:131-150) */
static int top_wave(const orc_params *p, int i, zdouble f[5]) {
  if (!g_top.x || !g_top.vm) return -1;
  const long q = i + (long)p->nx * (p->ny - 1);
  acoustic_wave(p, g_top.vm + 5 * q, g_top.x[q], g_top.x0, g_top.wamp ? g_top.wamp[i] : 1.0, f);
  return 0;
}

/* src/lrhs3d.f90:222-282: the sponge terms of point q (i = q mod nx) */
static void rhs_sponge(const orc_params *p, long q, int i, const zdouble *vq, zdouble rr[5]) {
  const int ispg = g_spg.ispg;
  if (ispg == 1) {
    for (int m = 0; m < 5; m++) rr[m] = rr[m] + g_spg.spg[q] * vq[m];
  } else if (ispg > 4) {
    for (int m = 0; m < 5; m++) rr[m] = rr[m] + (g_spg.spg[q] + g_spg.spg2[q]) * vq[m];
  } else if (ispg == 4) {
    /* src/sponge.f90:536-606 (cspg_it; omega != 0, so no transient factor):
     * (spg + spg2) (v - cvic), cvic = profile(j) exp(i lalpha x(i,j)) */
    const int j = (int)(q / p->nx);
    const zdouble e = cexp(_Complex_I * g_inf.lalpha * g_spg.x[q]);
    const double s = g_spg.spg[q] + g_spg.spg2[q];
    for (int m = 0; m < 5; m++) {
      const zdouble cv = (g_inf.eig[10 * j + 2 * m] + _Complex_I * g_inf.eig[10 * j + 2 * m + 1]) * e;
      rr[m] = rr[m] + s * (vq[m] - cv);
    }
  } else if (ispg == 2 || ispg == 3) {
    for (int m = 0; m < 5; m++) rr[m] = rr[m] + g_spg.spg[q] * vq[m];       /* outflow */
    zdouble f[5];                                                           /* inflow */
    acoustic_wave(p, g_spg.vm + 5 * q, g_spg.x[q], g_spg.x0, ispg == 3 ? g_spg.wamp[i] : 1.0, f);
    const double s2 = g_spg.spg2[q];
    for (int m = 0; m < 5; m++) rr[m] = rr[m] + s2 * (vq[m] - f[m]);
  }
}

/* ---- smoother (eps_e != 0): src/smoother3d.f90 in the residual and the
 *      implicit damping of src/lhs1f3d.f90:169-201, lhs2f3d.f90:171-203,
 *      lhsbt1f3d.f90:284-304, 451-471, lhsbt2f3d.f90:277-297, 431-451 and
 *      lhsbs1f3d.f90; buff is the damping profile init_buffer forms ------ */
static struct { double eps; const double *buff; } g_smo;

int orc_set_smoother(double eps_e, const double *buff) {
  if (eps_e != 0.0 && !buff) return -1;
  g_smo.eps = eps_e;
  g_smo.buff = eps_e != 0.0 ? buff : NULL;
  return 0;
}

static const double fa[7] = {-1.666666666666667E-01, 2.0, -6.5, 9.333333333333333E+00, -6.5, 2.0,
                             -1.666666666666667E-01};
static const double fb[5] = {1.0, -4.0, 6.0, -4.0, 1.0};

/* the dissipation of src/smoother3d.f90 along one line at position p of N
 * (value x(k) of point k of the line, component e), without eps buff:
 * order 2 at the second point from an end, 4 at the third, the 7-point
 * sixth difference inside; next to a symmetric end (sym) the 7-point form
 * with reflected points; 0 on the end points */
#define SMOOTH_LINE(X)                                                                       \
  do {                                                                                       \
    if (sym_row(sym, p, N)) {                                                                \
      zdouble acc = 0; int first = 1;                                                        \
      for (int k = 0; k < 7; k++) {                                                          \
        int q = p - 3 + k; double sg = reflect(&q, N, e);                                    \
        zdouble t = (sg * fa[k]) * X(q);                                                     \
        acc = first ? t : acc + t; first = 0;                                                \
      }                                                                                      \
      out = acc;                                                                             \
    } else if (p == 1 || p == N - 2) {                                                       \
      out = -1.0 * X(p - 1) + 2.0 * X(p) - 1.0 * X(p + 1);                                   \
    } else if (p == 2 || p == N - 3) {                                                       \
      out = X(p - 2) - 4.0 * X(p - 1) + 6.0 * X(p) - 4.0 * X(p + 1) + 1.0 * X(p + 2);         \
    } else if (p >= 3 && p <= N - 4) {                                                       \
      out = fa[0] * X(p - 3) + fa[1] * X(p - 2) + fa[2] * X(p - 1) + fa[3] * X(p) +          \
            fa[4] * X(p + 1) + fa[5] * X(p + 2) + fa[6] * X(p + 3);                          \
    } else {                                                                                 \
      out = 0; has = 0;                                                                      \
    }                                                                                        \
  } while (0)

static void rhs_smoother(const orc_params *p_, long q, const zdouble *v, zdouble rr[5]) {
  if (g_smo.eps == 0.0) return;
  const int nx = p_->nx, ny = p_->ny, i = (int)(q % nx), j = (int)(q / nx);
  const double eb = g_smo.eps * g_smo.buff[q];
  for (int e = 0; e < 5; e++) {
    zdouble out; int has = 1;
    {                                            /* eta (no symmetry there) */
      int p = j, N = ny, sym = 0;
#define XE(k) v[5 * (i + (long)nx * (k)) + e]
      SMOOTH_LINE(XE);
#undef XE
      if (has) rr[e] = rr[e] + eb * out;
    }
    has = 1;
    {                                            /* xi */
      int p = i, N = nx, sym = xsym(p_);
#define XX(k) v[5 * ((k) + (long)nx * j) + e]
      SMOOTH_LINE(XX);
#undef XX
      if (has) rr[e] = rr[e] + eb * out;
    }
  }
}

/* implicit damping weight of slot k (point pos + k - 2 in the interior
 * layout) on the diagonal of component cc, for a row of a line of N points:
 * fb in the interior, the second difference on the second and second-last
 * rows, nothing on the end rows; symmetric rows (sym) reflect fb; fb on
 * every row of a periodic line (src/lhsbt1f3d.f90:30 returns for xper) */
static double damp_weight(int pos, int N, int per, int sym, int k, int cc) {
  if (per) return fb[k];                         /* every row of a periodic line */
  if (sym_row(sym, pos, N) && (pos <= 1 || pos >= N - 2)) {
    double w = 0.0;
    for (int o = 0; o < 5; o++) {
      int q = pos + o - 2;
      double s = reflect(&q, N, cc);
      if (q - pos + 2 == k) w += s * fb[o];
    }
    return w;
  }
  if (pos == 0 || pos == N - 1) return 0.0;
  if (pos == 1 || pos == N - 2) return k == 1 ? -1.0 : k == 2 ? 2.0 : k == 3 ? -1.0 : 0.0;
  return fb[k];
}

/* ---- src/lrhs3d.f90:45-293 (point operator) ------------------------------ */
int orc_lrhs3d(const orc_params *p, const orc_coeffs *c, const zdouble *v,
               zdouble *r, zdouble *work) {
  if (!supported(p)) return -1;
  int nx = p->nx, ny = p->ny;
  long P = (long)nx * ny;
  zdouble *c1 = work, *c2 = work + 5 * P, *c11 = work + 10 * P, *c12 = work + 15 * P,
          *c22 = work + 20 * P;
  orc_cgrad(p, v, c1, c2);
  orc_cgrad2(p, v, c1, c11, c12, c22);
  if (navier(p) && (p->wallt == 1 || p->wallt == 2))   /* src/cgradbc.f90:15, src/lrhs3d.f90:43 */
    for (int i = 0; i < nx; i++) c2[4 + 5 * i] = 0;
  const zdouble ci = _Complex_I;
  for (long q = 0; q < P; q++) {
    const double *A = c->Ah + 25 * q, *B = c->Bh + 25 * q, *D = c->Dh + 25 * q,
                 *Di = c->Dhi + 25 * q;
    const double *V11 = c->Vh11 + 6 * q, *V12 = c->Vh12 + 6 * q, *V22 = c->Vh22 + 6 * q,
                 *ABi = c->ABhi + 6 * q;
    const zdouble *vq = v + 5 * q, *g1 = c1 + 5 * q, *g2 = c2 + 5 * q, *h11 = c11 + 5 * q,
                  *h12 = c12 + 5 * q, *h22 = c22 + 5 * q;
    zdouble rr[5];
    for (int m = 0; m < 5; m++)
      rr[m] = A[m] * g1[0] + A[m + 5] * g1[1] + A[m + 10] * g1[2] + A[m + 15] * g1[3] +
              A[m + 20] * g1[4];
    for (int m = 0; m < 5; m++)
      rr[m] = rr[m] + B[m] * g2[0] + B[m + 5] * g2[1] + B[m + 10] * g2[2] +
              B[m + 15] * g2[3] + B[m + 20] * g2[4];
    for (int m = 0; m < 5; m++)
      rr[m] = rr[m] + D[m] * vq[0] + D[m + 5] * vq[1] + D[m + 10] * vq[2] +
              D[m + 15] * vq[3] + D[m + 20] * vq[4];
    const double *Vs[3] = {V11, V12, V22};
    const zdouble *hs[3] = {h11, h12, h22};
    for (int t = 0; t < 3; t++) {
      const double *W = Vs[t];
      const zdouble *h = hs[t];
      rr[1] = rr[1] - W[0] * h[1] - W[4] * h[2];
      rr[2] = rr[2] - W[5] * h[1] - W[1] * h[2];
      rr[3] = rr[3] - W[2] * h[3];
      rr[4] = rr[4] - W[3] * h[4];
    }
    rr[1] = rr[1] - ci * (ABi[0] * g1[3] + ABi[2] * g2[3]);
    rr[2] = rr[2] - ci * (ABi[1] * g1[3] + ABi[3] * g2[3]);
    rr[3] = rr[3] - ci * (ABi[0] * g1[1] + ABi[1] * g1[2] + ABi[2] * g2[1] + ABi[3] * g2[2]);
    for (int m = 0; m < 5; m++)
      rr[m] = rr[m] + ci * (Di[m] * vq[0] + Di[m + 5] * vq[1] + Di[m + 10] * vq[2] +
                           Di[m + 15] * vq[3] + Di[m + 20] * vq[4]);
    rhs_sponge(p, q, (int)(q % nx), vq, rr);
    rhs_smoother(p, q, v, rr);                   /* src/lrhs3d.f90:293 */
    for (int m = 0; m < 5; m++) r[5 * q + m] = rr[m];
  }
  return 0;
}

/* src/impdrv3d.f90:444-470 */
void orc_rbe3d(const orc_params *p, zdouble *r, const zdouble *v, const zdouble *vold,
               const double *dtl) {
  long P = (long)p->nx * p->ny;
  for (long q = 0; q < P; q++)
    for (int m = 0; m < 5; m++)
      r[5 * q + m] = vold[5 * q + m] - v[5 * q + m] - dtl[q] * r[5 * q + m];
}

/* ---- LHS assembly: src/lhs1f3d.f90 / lhs2f3d.f90 (interior), lhsbt1f3d.f90 /
 *      lhsbt2f3d.f90 (rows 1,2,n-1,n), lhsbc1f3d.f90 / lhsbc2f3d.f90 (BCs) --- */
#define MAT(k, r, cc) mat[(k) + 5 * (r) + 25 * (cc) + 125 * pt]

/* row of a boundary-closure line: first-derivative coefficient F[q] and
 * second-derivative coefficient S[q] act on the column held in slot sl[q];
 * q = 0 is the diagonal */
static void closure_row(zdouble *mat, long pt, const double F[5], const double S[5],
                        const int sl[5], double alfa, double dinv, double dsinv,
                        double dtl, const double *X, const double *Xi2 /* 2 */,
                        const double *D, const double *Di, const double *Vh, double rcalcd,
                        double spgd /* alfa dtl (spg [+ spg2]) or 0 */) {
  double a[5], cim[5];
  for (int q = 0; q < 5; q++) a[q] = alfa * F[q] * dinv;
  for (int r = 0; r < 5; r++)
    for (int cc = 0; cc < 5; cc++) {
      for (int q = 1; q < 5; q++) MAT(sl[q], r, cc) = a[q] * dtl * X[r + 5 * cc];
      double re = dtl * (a[0] * X[r + 5 * cc] + rcalcd * alfa * D[r + 5 * cc]);
      double im = dtl * (rcalcd * alfa * Di[r + 5 * cc]);
      MAT(sl[0], r, cc) = re + im * _Complex_I;
    }
  for (int q = 0; q < 5; q++) cim[q] = alfa * F[q] * dinv;   /* imaginary coefficient */
  for (int q = 0; q < 5; q++) {
    int k = sl[q];
    MAT(k, 1, 3) -= _Complex_I * (cim[q] * dtl * Xi2[0]);
    MAT(k, 2, 3) -= _Complex_I * (cim[q] * dtl * Xi2[1]);
    MAT(k, 3, 1) -= _Complex_I * (cim[q] * dtl * Xi2[0]);
    MAT(k, 3, 2) -= _Complex_I * (cim[q] * dtl * Xi2[1]);
  }
  for (int q = 0; q < 5; q++) {
    int k = sl[q];
    double b = alfa * S[q] * dsinv;
    MAT(k, 1, 1) -= b * dtl * Vh[0];
    MAT(k, 1, 2) -= b * dtl * Vh[4];
    MAT(k, 2, 1) -= b * dtl * Vh[5];
    MAT(k, 2, 2) -= b * dtl * Vh[1];
    MAT(k, 3, 3) -= b * dtl * Vh[2];
    MAT(k, 4, 4) -= b * dtl * Vh[3];
  }
  for (int r = 0; r < 5; r++) MAT(sl[0], r, r) += 1.0;
  for (int r = 0; r < 5; r++) MAT(sl[0], r, r) += spgd;
}

/* a row of the xi factor next to a symmetric end (src/lhsbs1f3d.f90:36-329
 * left, :331-624 right: the D terms times rcalcd, the sponge spgd on the
 * diagonal without them (:154-170); Xi2 read where that routine reads it,
 * see orc_lhs): the interior row of
 * src/lhs1f3d.f90 with the weight of each point beyond the end added to the
 * point it reflects onto, negated for column component 3 (index 2); slots
 * keep their interior meaning (slot k: point i + k - 2), the slots that then
 * hold nothing are zero.  a[], b[]: interior first/second-derivative weights
 * times alfa. */
static void sym_lhs_row(zdouble *mat, long pt, int i, int N, const double a[5], const double b[5],
                        double dt, const double *X, const double *Xi2, const double *D,
                        const double *Di, const double *Vh, double alfa, double rc, double spgd) {
  double aw[5][5], bw[5][5];                   /* [slot][column component] */
  for (int k = 0; k < 5; k++)
    for (int cc = 0; cc < 5; cc++) aw[k][cc] = bw[k][cc] = 0.0;
  for (int o = 0; o < 5; o++) {
    int col = i + o - 2;
    for (int cc = 0; cc < 5; cc++) {
      int q = col;
      double s = reflect(&q, N, cc);
      int t = q - i + 2;
      aw[t][cc] += s * a[o];
      bw[t][cc] += s * b[o];
    }
  }
  for (int e = 0; e < 125; e++) mat[125 * pt + e] = 0;
  for (int k = 0; k < 5; k++)
    for (int r = 0; r < 5; r++)
      for (int cc = 0; cc < 5; cc++) MAT(k, r, cc) = dt * aw[k][cc] * X[r + 5 * cc];
  for (int r = 0; r < 5; r++)
    for (int cc = 0; cc < 5; cc++)
      MAT(2, r, cc) += rc * alfa * dt * D[r + 5 * cc] + _Complex_I * (rc * alfa * dt * Di[r + 5 * cc]);
  for (int k = 0; k < 5; k++) {
    MAT(k, 1, 3) -= _Complex_I * (aw[k][3] * dt * Xi2[0]);
    MAT(k, 2, 3) -= _Complex_I * (aw[k][3] * dt * Xi2[1]);
    MAT(k, 3, 1) -= _Complex_I * (aw[k][1] * dt * Xi2[0]);
    MAT(k, 3, 2) -= _Complex_I * (aw[k][2] * dt * Xi2[1]);
    MAT(k, 1, 1) -= dt * bw[k][1] * Vh[0];
    MAT(k, 1, 2) -= dt * bw[k][2] * Vh[4];
    MAT(k, 2, 1) -= dt * bw[k][1] * Vh[5];
    MAT(k, 2, 2) -= dt * bw[k][2] * Vh[1];
    MAT(k, 3, 3) -= dt * bw[k][3] * Vh[2];
    MAT(k, 4, 4) -= dt * bw[k][4] * Vh[3];
  }
  for (int r = 0; r < 5; r++) MAT(2, r, r) += 1.0;
  for (int r = 0; r < 5; r++) MAT(2, r, r) += spgd;
}

int orc_lhs(const orc_params *p, int dir, const orc_coeffs *c, const double *dtl,
            zdouble *mat) {
  if (!supported(p) || (dir != 1 && dir != 2)) return -1;
  int nx = p->nx, ny = p->ny;
  double alfa = p->alfa;
  double d = dir == 1 ? 1.0 / (double)(nx - 1) : 1.0 / (double)(ny - 1);
  const double *X = dir == 1 ? c->Ah : c->Bh;
  const double *Vh = dir == 1 ? c->Vh11 : c->Vh22;
  int off = dir == 1 ? 0 : 2;                  /* ABhi(1:2) or ABhi(3:4) */
  /* the D terms ride in the xi factor, or in the eta factor for useCalcd's
   * first factors (src/lhs1f3d.f90:27-37, lhs2f3d.f90:27-37) */
  int calcd = (dir == 1) != (p->d_in_eta != 0);
  double a[5] = {alfa * ga1 / d, alfa * ga2 / d, alfa * 0.0 / d, alfa * ga3 / d, alfa * ga4 / d};
  double b[5] = {alfa * da1 / (d * d), alfa * da2 / (d * d), alfa * da3 / (d * d),
                 alfa * da4 / (d * d), alfa * da5 / (d * d)};
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      long pt = i + (long)nx * j;
      const double *Xp = X + 25 * pt, *Dp = c->Dh + 25 * pt, *Dip = c->Dhi + 25 * pt;
      const double *Vp = Vh + 6 * pt, *ABp = c->ABhi + 6 * pt + off;
      double dt = dtl[pt];
      for (int k = 0; k < 5; k++)
        for (int r = 0; r < 5; r++)
          for (int cc = 0; cc < 5; cc++) MAT(k, r, cc) = a[k] * dt * Xp[r + 5 * cc];
      for (int k = 0; k < 5; k++) {
        MAT(k, 1, 3) -= _Complex_I * (a[k] * dt * ABp[0]);
        MAT(k, 2, 3) -= _Complex_I * (a[k] * dt * ABp[1]);
        MAT(k, 3, 1) -= _Complex_I * (a[k] * dt * ABp[0]);
        MAT(k, 3, 2) -= _Complex_I * (a[k] * dt * ABp[1]);
      }
      if (calcd) {
        for (int r = 0; r < 5; r++)
          for (int cc = 0; cc < 5; cc++) MAT(2, r, cc) += alfa * dt * Dp[r + 5 * cc];
        for (int r = 0; r < 5; r++)
          for (int cc = 0; cc < 5; cc++)
            MAT(2, r, cc) += _Complex_I * (alfa * dt * Dip[r + 5 * cc]);
      }
      for (int k = 0; k < 5; k++) {
        MAT(k, 1, 1) -= b[k] * dt * Vp[0];
        MAT(k, 1, 2) -= b[k] * dt * Vp[4];
        MAT(k, 2, 1) -= b[k] * dt * Vp[5];
        MAT(k, 2, 2) -= b[k] * dt * Vp[1];
        MAT(k, 3, 3) -= b[k] * dt * Vp[2];
        MAT(k, 4, 4) -= b[k] * dt * Vp[3];
      }
      for (int r = 0; r < 5; r++) MAT(2, r, r) += 1.0;
      if (!calcd && g_spg.ispg)                  /* src/lhs2f3d.f90:207-225 */
        for (int r = 0; r < 5; r++) MAT(2, r, r) += alfa * dt * lhs_sponge(pt);
    }

  int per = dir == 1 ? p->xper : p->yper;
  if (!per) {
    int N = dir == 1 ? nx : ny, L = dir == 1 ? ny : nx;
    double dinv = 1.0 / d, dsinv = 1.0 / (d * d), rc = calcd ? 1.0 : 0.0;
    const double Fc[5] = {gc1, gc2, gc3, gc4, gc5}, Fb[5] = {gb1, gb2, gb3, gb4, gb5};
    const double Sd[5] = {dd1, dd2, dd3, dd4, dd5}, Sb[5] = {db1, db2, db3, db4, db5};
    /* rows 1, 2, n-1, n: diagonal first in each list (Appendix A slots) */
    const double Fr0[5] = {Fc[0], Fc[1], Fc[2], Fc[3], Fc[4]};
    const int sl0[5] = {2, 3, 4, 0, 1};
    const double Fr1[5] = {Fb[1], Fb[0], Fb[2], Fb[3], Fb[4]};
    const double Sr1[5] = {Sb[1], Sb[0], Sb[2], Sb[3], Sb[4]};
    const int sl1[5] = {2, 1, 3, 4, 0};
    const double Frm[5] = {-Fb[1], -Fb[0], -Fb[2], -Fb[3], -Fb[4]};
    const int slm[5] = {2, 3, 1, 0, 4};
    const double Frn[5] = {-Fc[0], -Fc[1], -Fc[2], -Fc[3], -Fc[4]};
    const int sln[5] = {2, 1, 0, 4, 3};
    for (int l = 0; l < L; l++) {
      int rows[4] = {0, 1, N - 2, N - 1};
      const double *Fs[4] = {Fr0, Fr1, Frm, Frn};
      const double *Ss[4] = {Sd, Sr1, Sr1, Sd};
      const int *sls[4] = {sl0, sl1, slm, sln};
      for (int w = 0; w < 4; w++) {
        int i = dir == 1 ? rows[w] : l, j = dir == 1 ? l : rows[w];
        long pt = i + (long)nx * j;
        closure_row(mat, pt, Fs[w], Ss[w], sls[w], alfa, dinv, dsinv, dtl[pt],
                    X + 25 * pt, c->ABhi + 6 * pt + off, c->Dh + 25 * pt, c->Dhi + 25 * pt,
                    Vh + 6 * pt, rc, (!calcd && g_spg.ispg) ? alfa * dtl[pt] * lhs_sponge(pt) : 0.0);
      }
    }
  }
  if (dir == 1 && xsym(p)) {                   /* src/lhsbs1f3d.f90 (xi factor only) */
    const int sym = xsym(p);
    for (int j = 0; j < ny; j++) {
      const int rows[4] = {0, 1, nx - 2, nx - 1};
      for (int w = 0; w < 4; w++) {
        if (!((w < 2 && (sym & 1)) || (w >= 2 && (sym & 2)))) continue;
        long pt = rows[w] + (long)nx * j;
        const long pa = rows[w] + (long)ny * j;     /* Ahi(6,ny,nx) of src/lhsbs1f3d.f90:17 */
        sym_lhs_row(mat, pt, rows[w], nx, a, b, dtl[pt], X + 25 * pt, c->ABhi + 6 * pa,
                    c->Dh + 25 * pt, c->Dhi + 25 * pt, Vh + 6 * pt, alfa, calcd ? 1.0 : 0.0,
                    (!calcd && g_spg.ispg) ? alfa * dtl[pt] * lhs_sponge(pt) : 0.0);
      }
    }
  }

  if (g_smo.eps != 0.0) {                      /* implicit damping, all rows */
    const int N = dir == 1 ? nx : ny, sym = dir == 1 ? xsym(p) : 0;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long pt = i + (long)nx * j;
        const int pos = dir == 1 ? i : j;
        const double f = alfa * dtl[pt] * g_smo.eps * g_smo.buff[pt];
        for (int k = 0; k < 5; k++)
          for (int r = 0; r < 5; r++)
            MAT(k, r, r) += f * damp_weight(pos, N, dir == 1 ? p->xper : p->yper, sym, k, r);
      }
  }

  /* boundary conditions */
  if (!p->yper && !navier(p)) {
    /* inviscid wall rows (src/lhsbc1f3d.f90:55-393, src/lhsbc2f3d.f90:75-120),
     * rows i of is:ie: the blocks of the wall's unknowns rotated to the body
     * normal coordinates (u, v components), then the u and v rows rotated and
     * the normal velocity held at zero */
    if (!g_bnb) return -1;
    const int cols0[5] = {3, 4, 0, 1, 2}, cols1[5] = {4, 0, 1, 2, 3};
    for (int i = wall_lo(p); i <= wall_hi(p); i++) {
      for (int k = 0; k < 5; k++) {
        long pt;
        int col;
        if (dir == 1) {                 /* every block of the wall line: its column's normal */
          pt = i;
          col = i == 0 ? cols0[k] : i == 1 ? cols1[k] :
                i == nx - 2 ? (int[5]){nx - 4, nx - 3, nx - 2, nx - 1, nx - 5}[k] :
                i == nx - 1 ? (int[5]){nx - 3, nx - 2, nx - 1, nx - 5, nx - 4}[k] : i - 2 + k;
        } else {                        /* the blocks of column j = 0: rows 0, 1, 2 */
          if (k > 2) continue;
          pt = i + (long)nx * (2 - k);
          col = i;
        }
        const double b1 = g_bnb[col], b2 = g_bnb[col + nx];
        for (int r = 0; r < 5; r++) {
          const zdouble u = MAT(k, r, 1), w = MAT(k, r, 2);
          MAT(k, r, 1) = u * b2 - w * b1;
          MAT(k, r, 2) = u * b1 + w * b2;
        }
      }
      long pt = i;
      const double b1 = g_bnb[i], b2 = g_bnb[i + nx];
      for (int k = 0; k < 5; k++)
        for (int cc = 0; cc < 5; cc++) {
          MAT(k, 1, cc) = b2 * MAT(k, 1, cc) - b1 * MAT(k, 2, cc);
          MAT(k, 2, cc) = 0;
        }
      MAT(2, 2, 2) = 1.0;
    }
  }
  if (!p->yper) {
    for (int i = 0; i < nx; i++) {
      long pt = i;
      if (!navier(p) || !on_wall(p, i)) continue;    /* the viscous rows of is:ie */
      if (dir == 1 && (p->wall == 1 || p->wall == 2 || p->wall == 3)) {
        for (int k = 0; k < 5; k++)
          for (int cc = 0; cc < 5; cc++) MAT(k, 0, cc) = 0;
        MAT(2, 0, 0) = 1.0;
      }
      if (dir == 2 && p->wall == 1) {
        for (int k = 0; k < 5; k++)
          for (int cc = 0; cc < 5; cc++) MAT(k, 0, cc) = 0;
        MAT(2, 0, 0) = -gc1; MAT(3, 0, 0) = -gc2; MAT(4, 0, 0) = -gc3;
        MAT(0, 0, 0) = -gc4; MAT(1, 0, 0) = -gc5;
      }
      if (dir == 2 && p->wall == 3) {            /* src/lhsbc2f3d.f90:41-48 */
        for (int k = 0; k < 5; k++)
          for (int cc = 0; cc < 5; cc++) MAT(k, 0, cc) = 0;
        MAT(2, 0, 0) = -1.0; MAT(3, 0, 0) = 4.0; MAT(4, 0, 0) = -6.0;
        MAT(0, 0, 0) = 4.0; MAT(1, 0, 0) = -1.0;
      }
      for (int r = 1; r <= 3; r++)
        for (int k = 0; k < 5; k++)
          for (int cc = 0; cc < 5; cc++) MAT(k, r, cc) = 0;
      MAT(2, 1, 1) = 1.0; MAT(2, 2, 2) = 1.0; MAT(2, 3, 3) = 1.0;
      if (p->wallt == 0 || (dir == 1 && p->wallt == 1)) {
        for (int k = 0; k < 5; k++)
          for (int cc = 0; cc < 5; cc++) MAT(k, 4, cc) = 0;
        MAT(2, 4, 4) = 1.0;
      }
      if (dir == 2 && p->wallt == 1) {
        for (int k = 0; k < 5; k++)
          for (int cc = 0; cc < 5; cc++) MAT(k, 4, cc) = 0;
        MAT(2, 4, 4) = -gc1; MAT(3, 4, 4) = -gc2; MAT(4, 4, 4) = -gc3;
        MAT(0, 4, 4) = -gc4; MAT(1, 4, 4) = -gc5;
      }
    }
    for (int i = 0; i < nx; i++) {                        /* top = 0 */
      long pt = i + (long)nx * (ny - 1);
      for (int e = 0; e < 125; e++) mat[125 * pt + e] = 0;
      for (int r = 0; r < 5; r++) MAT(2, r, r) = 1.0;
    }
  }
  if (!p->xper) {
    for (int j = 0; j < ny; j++) {                        /* left, right = 0 */
      long pts[2] = {(long)nx * j, nx - 1 + (long)nx * j};
      for (int s = 0; s < 2; s++) {
        const int sel = s == 0 ? p->left : p->right;   /* src/lhsbc1f3d.f90:414-446 */
        if (s == 1 && sel == 8 && dir == 1) {      /* hold the initial condition: -I */
          long pt = pts[s];
          for (int e = 0; e < 125; e++) mat[125 * pt + e] = 0;
          for (int r = 0; r < 5; r++) MAT(2, r, r) = -1.0;
          continue;
        }
        if (sel != 0 && sel != 4 && sel != 5) continue;
        long pt = pts[s];
        for (int e = 0; e < 125; e++) mat[125 * pt + e] = 0;
        for (int r = 0; r < 5; r++) MAT(2, r, r) = 1.0;
      }
    }
  }
  return 0;
}
#undef MAT

/* src/rhsbc3d.f90:34-215, 271, 281-284 (wall; top 0, 1; left 0, 1, 4, 7;
 * right 0, 4, 7) */
int orc_rhsbc3d(const orc_params *p, zdouble *r, const zdouble *v) {
  if (!supported(p)) return -1;
  int nx = p->nx, ny = p->ny;
#define R(d, i, j) r[(d) + 5 * ((i) + (long)nx * (j))]
#define V(d, i, j) v[(d) + 5 * ((i) + (long)nx * (j))]
  if (p->yper) {
    for (int i = 0; i < nx; i++)
      for (int d = 0; d < 5; d++) R(d, i, ny - 1) = 0;
  } else {
    if (p->wall == 3 && !g_bump) return -1;
    if (!navier(p)) {                            /* src/rhsbc3d.f90:108-118 */
      if (!g_bnb) return -1;
      for (int i = wall_lo(p); i <= wall_hi(p); i++) {
        const double b1 = g_bnb[i], b2 = g_bnb[i + nx];
        R(1, i, 0) = b2 * R(1, i, 0) - b1 * R(2, i, 0);
        R(2, i, 0) = -(b1 * V(1, i, 0) + b2 * V(2, i, 0));
      }
    }
    for (int i = 0; i < nx; i++) {
      if (!navier(p) || !on_wall(p, i)) continue;   /* the viscous rows of is:ie */
      /* the bump's wall values enter these rows as whole arrays against the
       * is:ie section (vl(2,is:ie,1) - u1w, src/rhsbc3d.f90:47-50): position
       * i of the section takes element i - is of the array */
      const long ib = i - wall_lo(p);
      for (int d = 1; d <= 3; d++)              /* wall = 3: the bump's velocities */
        R(d, i, 0) = -(V(d, i, 0) - (p->wall == 3 ? g_bump[ib + (long)nx * (d - 1)] : 0.0));
      if (p->wall == 1)
        R(0, i, 0) = gc1 * V(0, i, 0) + gc2 * V(0, i, 1) + gc3 * V(0, i, 2) +
                     gc4 * V(0, i, 3) + gc5 * V(0, i, 4);
      if (p->wall == 2) R(0, i, 0) = V(0, i, 0) - 2.0 * V(0, i, 1) + V(0, i, 2);
      if (p->wall == 3)                          /* src/rhsbc3d.f90:74-79 */
        R(0, i, 0) = V(0, i, 0) - (4.0 * V(0, i, 1) - 6.0 * V(0, i, 2) + 4.0 * V(0, i, 3) - V(0, i, 4));
      if (p->wallt == 0) R(4, i, 0) = p->wall == 3 ? -(V(4, i, 0) - g_bump[ib + 3L * nx]) : 0;
      if (p->wallt == 1) {
        R(4, i, 0) = gc1 * V(4, i, 0) + gc2 * V(4, i, 1) + gc3 * V(4, i, 2) +
                     gc4 * V(4, i, 3) + gc5 * V(4, i, 4);
        if (p->wall == 3) R(4, i, 0) = R(4, i, 0) - g_bump[ib + 4L * nx];
      }
    }
    for (int i = 0; i < nx; i++) {
      if (p->top == 1) {                           /* src/rhsbc3d.f90:129-161 */
        zdouble f[5];
        if (top_wave(p, i, f)) return -1;
        for (int d = 0; d < 5; d++) R(d, i, ny - 1) = -(V(d, i, ny - 1) - f[d]);
      } else {
        for (int d = 0; d < 5; d++) R(d, i, ny - 1) = 0;
      }
    }
  }
  if (p->xper) {
    for (int j = 0; j < ny; j++)
      for (int d = 0; d < 5; d++) R(d, nx - 1, j) = 0;
  } else {
    for (int j = 0; j < ny; j++) {            /* left/right = 7: the operator's row */
      zdouble f[5];
      if (p->left == 1 || p->left == 4 || p->left == 5) {
        if (side_values(p, 0, j, f)) return -1;
        for (int d = 0; d < 5; d++) R(d, 0, j) = -(V(d, 0, j) - f[d]);
      }
      if (p->right == 5) {
        if (side_values(p, 1, j, f)) return -1;
        for (int d = 0; d < 5; d++) R(d, nx - 1, j) = -(V(d, nx - 1, j) - f[d]);
      }
      for (int d = 0; d < 5; d++) {
        if (p->left == 0) R(d, 0, j) = 0;
        if (p->right == 0 || p->right == 4 || p->right == 8) R(d, nx - 1, j) = 0;
        if (p->right == 9)                       /* src/rhsbc3d.f90:317-320 */
          R(d, nx - 1, j) = -(V(d, nx - 1, j) - cexp(2.0 * clog(V(d, nx - 2, j)) - clog(V(d, nx - 3, j))));
      }
    }
  }
  return 0;
}

/* src/itrbc3d.f90:34-376 (same selectors; right = 8 holds the column) */
int orc_itrbc3d(const orc_params *p, zdouble *v) {
  if (!supported(p)) return -1;
  int nx = p->nx, ny = p->ny;
  if (p->yper) {
    for (int i = 0; i < nx; i++)
      for (int d = 0; d < 5; d++) V(d, i, ny - 1) = V(d, i, 0);
  } else {
    if (p->wall == 3 && !g_bump) return -1;
    if (!navier(p)) {                            /* src/itrbc3d.f90:120-132 */
      if (!g_bnb) return -1;
      for (int i = wall_lo(p); i <= wall_hi(p); i++) {
        const double b1 = g_bnb[i], b2 = g_bnb[i + nx];
        const zdouble ub = b2 * V(1, i, 0) - b1 * V(2, i, 0);
        V(1, i, 0) = b2 * ub;
        V(2, i, 0) = -b1 * ub;
      }
    }
    for (int i = 0; i < nx; i++) {
      if (!navier(p)) break;
      /* wall = 3 holds the velocities on the whole row, the other wall rows
       * on is:ie only (src/itrbc3d.f90:48-58) */
      if (p->wall == 3)
        for (int d = 1; d <= 3; d++) V(d, i, 0) = g_bump[i + (long)nx * (d - 1)];
      if (!on_wall(p, i)) continue;
      if (p->wall != 3)
        for (int d = 1; d <= 3; d++) V(d, i, 0) = 0;
      /* the is:ie section against the whole arrays tw, twp
       * (src/itrbc3d.f90:86-104): element i - is of the array */
      const long ib = i - wall_lo(p);
      if (p->wall == 1)
        V(0, i, 0) = -(gc2 * V(0, i, 1) + gc3 * V(0, i, 2) + gc4 * V(0, i, 3) +
                       gc5 * V(0, i, 4)) / gc1;
      if (p->wall == 2) V(0, i, 0) = 2.0 * V(0, i, 1) - V(0, i, 2);
      if (p->wall == 3)                          /* src/itrbc3d.f90:76-81 */
        V(0, i, 0) = 4.0 * V(0, i, 1) - 6.0 * V(0, i, 2) + 4.0 * V(0, i, 3) - V(0, i, 4);
      if (p->wallt == 0) V(4, i, 0) = p->wall == 3 ? g_bump[ib + 3L * nx] : 0;
      if (p->wallt == 1) {
        if (p->wall == 3)                        /* src/itrbc3d.f90:93-104 */
          V(4, i, 0) = (g_bump[ib + 4L * nx] - (gc2 * V(4, i, 1) + gc3 * V(4, i, 2) + gc4 * V(4, i, 3) +
                                                gc5 * V(4, i, 4))) / gc1;
        else
          V(4, i, 0) = -(gc2 * V(4, i, 1) + gc3 * V(4, i, 2) + gc4 * V(4, i, 3) +
                         gc5 * V(4, i, 4)) / gc1;
      }
    }
    for (int i = 0; i < nx; i++) {
      if (p->top == 1) {                           /* src/itrbc3d.f90:145-195 */
        zdouble f[5];
        if (top_wave(p, i, f)) return -1;
        for (int d = 0; d < 5; d++) V(d, i, ny - 1) = f[d];
      } else {
        for (int d = 0; d < 5; d++) V(d, i, ny - 1) = 0;
      }
    }
  }
  if (p->xper) {
    for (int j = 0; j < ny; j++)
      for (int d = 0; d < 5; d++) V(d, nx - 1, j) = V(d, 0, j);
  } else {
    for (int j = 0; j < ny; j++) {            /* left/right = 7: left as they are */
      zdouble f[5];
      if (p->left == 1 || p->left == 4 || p->left == 5) {
        if (side_values(p, 0, j, f)) return -1;
        for (int d = 0; d < 5; d++) V(d, 0, j) = f[d];
      }
      if (p->right == 4 || p->right == 5) {
        if (side_values(p, 1, j, f)) return -1;
        for (int d = 0; d < 5; d++) V(d, nx - 1, j) = f[d];
      }

      for (int d = 0; d < 5; d++) {
        if (p->left == 0) V(d, 0, j) = 0;
        if (p->right == 0) V(d, nx - 1, j) = 0;
        if (p->right == 9)                       /* src/itrbc3d.f90:408-410 */
          V(d, nx - 1, j) = cexp(2.0 * clog(V(d, nx - 2, j)) - clog(V(d, nx - 3, j)));
      }
    }
  }
  return 0;
#undef R
#undef V
}

/* ---- src/getmat.f90 (scalar form) --------------------------------------- */
/* Sutherland's constant S0 (the deck's 110.4 K over Tref, src/input.f90:41-47)
 * for mattyp = 1; negative: not given */
static double g_S0 = -1.0;

int orc_set_sutherland(double S0) {
  g_S0 = S0;
  return 0;
}

static void getmat(int mattyp, double t, double *mu, double *lm, double *con, double *dmu,
                   double *d2mu, double *dlm, double *d2lm, double *dcon, double *d2con) {
  if (mattyp == 0) {
    *mu = 1.0; *dmu = 0.0; *d2mu = 0.0;
  } else if (mattyp == 1) {                      /* Sutherland's law, src/getmat.f90:28-40 */
    const double S0 = g_S0, ts = S0 + t;
    *mu = pow(t, 1.5) * (1.0 + S0) / ts;
    *dmu = sqrt(t) * (1.0 + S0) * (3.0 * S0 + t) * 0.5 / (ts * ts);
    *d2mu = (1.0 + S0) * (3.0 * (S0 * S0) - 6.0 * S0 * t - t * t) * 0.25 / sqrt(t) / (ts * ts * ts);
  } else {  /* 2, 3: linear law */
    *mu = t; *dmu = 1.0; *d2mu = 0.0;
  }
  *con = *mu; *dcon = *dmu; *d2con = *d2mu;
  if (mattyp == 3) { *lm = 0.0; *dlm = 0.0; *d2lm = 0.0; }
  else { *lm = -pt66 * *mu; *dlm = -pt66 * *dmu; *d2lm = -pt66 * *d2mu; }
}

/* ---- src/genmtrx.f90:72-1000 (general metrics, linear = 2, idamp = 0) ----- */
int orc_genmtrx(const orc_params *p, const double *vm, const double *metric, orc_coeffs *c) {
  if (!supported(p) || (p->mattyp == 1 && g_S0 < 0.0)) return -1;
  int nx = p->nx, ny = p->ny;
  long P = (long)nx * ny;
  double *g = calloc(25 * P, sizeof(double));
  if (!g) return -1;
  double *g1 = g, *g2 = g + 5 * P, *g11 = g + 10 * P, *g12 = g + 15 * P, *g22 = g + 20 * P;
  orc_grad(p, 5, vm, g1, g2);
  orc_grad2(p, 5, vm, g1, g11, g12, g22);
  if (navier(p) && (p->wallt == 1 || p->wallt == 2))     /* src/gradbc.f90:15, src/genmtrx.f90:82 */
    for (int i = 0; i < nx; i++) g2[4 + 5 * i] = 0;
  /* idamp = 1: damp = 0 on the last column (src/damper.f90:26-29) times the
   * xi second derivatives of the mean (src/genmtrx.f90:87-96); calcp's gpr
   * enters only the nonlinear forms (:273, 646) */
  if (p->idamp == 1)
    for (int j = 0; j < ny; j++)
      for (int d = 0; d < 5; d++) g11[5 * (nx - 1 + (long)nx * j) + d] = 0.0 * g11[5 * (nx - 1 + (long)nx * j) + d];
  const double *m1 = metric, *m2 = metric + P, *n1 = metric + 2 * P, *n2 = metric + 3 * P,
               *m11 = metric + 4 * P, *m12 = metric + 5 * P, *m22 = metric + 6 * P,
               *n11 = metric + 7 * P, *n12 = metric + 8 * P, *n22 = metric + 9 * P;
  double gamma = p->gamma, gamma1 = gamma - 1.0, Ma = p->Ma, Re = re_eff(p), Pr = p->Pr;
  double kz = p->kz;
  for (long q = 0; q < P; q++) {
    double M1 = m1[q], M2 = m2[q], N1 = n1[q], N2 = n2[q];
    double m1m1 = M1 * M1, m1m2 = M1 * M2, m2m2 = M2 * M2;
    double n1n1 = N1 * N1, n1n2 = N1 * N2, n2n2 = N2 * N2;
    double m1n1 = M1 * N1, m1n2 = M1 * N2, m2n1 = M2 * N1, m2n2 = M2 * N2;
    double g11l[5], g12l[5], g22l[5], g1l[5], g2l[5];
    for (int d = 0; d < 5; d++) {
      double a11 = g11[5 * q + d], a12 = g12[5 * q + d], a22 = g22[5 * q + d];
      double a1 = g1[5 * q + d], a2 = g2[5 * q + d];
      g11l[d] = a11 * m1m1 + 2.0 * a12 * m1n1 + a22 * n1n1 + a1 * m11[q] + a2 * n11[q];
      g12l[d] = a11 * m1m2 + a12 * (m1n2 + m2n1) + a22 * n1n2 + a1 * m12[q] + a2 * n12[q];
      g22l[d] = a11 * m2m2 + 2.0 * a12 * m2n2 + a22 * n2n2 + a1 * m22[q] + a2 * n22[q];
      g1l[d] = a1 * M1 + a2 * N1;
      g2l[d] = a1 * M2 + a2 * N2;
    }
    double fact = 1.0 / (gamma * Ma * Ma);
    double rho = vm[5 * q + 0], u1 = vm[5 * q + 1], u2 = vm[5 * q + 2], u3 = vm[5 * q + 3],
           t = vm[5 * q + 4];
    double rhoinv = 1.0 / rho;
    double grho[3] = {g1l[0], g2l[0], 0.0};
    double gu[3][3] = {{g1l[1], g2l[1], 0.0}, {g1l[2], g2l[2], 0.0}, {g1l[3], g2l[3], 0.0}};
    double gt[3] = {g1l[4], g2l[4], 0.0};
    double g1divu = g11l[1] + g12l[2], g2divu = g12l[1] + g22l[2], g3divu = 0.0;
    double S1jj = 0.5 * (g11l[1] + g11l[1] + g22l[1] + g12l[2]);
    double S2jj = 0.5 * (g11l[2] + g12l[1] + g22l[2] + g22l[2]);
    double S3jj = 0.5 * (g11l[3] + g22l[3]);
    double Lapt = g11l[4] + g22l[4];
    double divu = gu[0][0] + gu[1][1] + gu[2][2];
    double S[3][3] = {{0}};
    S[0][0] = gu[0][0];
    S[1][0] = 0.5 * (gu[1][0] + gu[0][1]);
    S[2][0] = 0.5 * gu[2][0];
    S[0][1] = 0.5 * (gu[0][1] + gu[1][0]);
    S[1][1] = gu[1][1];
    S[2][1] = 0.5 * gu[2][1];
    S[0][2] = 0.5 * gu[2][0];
    S[1][2] = 0.5 * gu[2][1];
    double mu, lm, con, dmu, d2mu, dlm, d2lm, dcon, d2con;
    getmat(p->mattyp, t, &mu, &lm, &con, &dmu, &d2mu, &dlm, &d2lm, &dcon, &d2con);
    double gmu[3], gdmu[3], gcon[3], gdcon[3], glm[3], gdlm[3];
    for (int k = 0; k < 3; k++) {
      gmu[k] = dmu * gt[k]; gdmu[k] = d2mu * gt[k];
      gcon[k] = dcon * gt[k]; gdcon[k] = d2con * gt[k];
      glm[k] = dlm * gt[k]; gdlm[k] = d2lm * gt[k];
    }
    double A[5][5], B[5][5], C[5][5], D[5][5], Vij[4];
    double f1, f2, gmsinv;
    /* A, genmtrx.f90:430-480 */
    A[0][0] = u1; A[0][1] = rho; A[0][2] = 0; A[0][3] = 0; A[0][4] = 0;
    f1 = rhoinv / Re; gmsinv = 1.0 / (gamma * Ma * Ma);
    A[1][0] = rhoinv * t * gmsinv;
    A[1][1] = u1 - f1 * glm[0] - f1 * 2.0 * gmu[0];
    A[1][2] = -f1 * gmu[1];
    A[1][3] = -f1 * gmu[2];
    A[1][4] = gmsinv - f1 * dlm * divu - f1 * dmu * 2.0 * S[0][0];
    A[2][0] = 0; A[2][1] = -f1 * glm[1]; A[2][2] = u1 - f1 * gmu[0]; A[2][3] = 0;
    A[2][4] = -f1 * dmu * 2.0 * S[1][0];
    A[3][0] = 0; A[3][1] = -f1 * glm[2]; A[3][2] = 0; A[3][3] = u1 - f1 * gmu[0];
    A[3][4] = -f1 * dmu * 2.0 * S[2][0];
    f1 = gamma * rhoinv / (Pr * Re);
    f2 = gamma * gamma1 * Ma * Ma * rhoinv / Re;
    A[4][0] = 0;
    A[4][1] = gamma1 * t - f2 * 2.0 * lm * divu - 4.0 * f2 * mu * S[0][0];
    A[4][2] = -4.0 * f2 * mu * S[1][0];
    A[4][3] = -4.0 * f2 * mu * S[2][0];
    A[4][4] = u1 - f1 * (gcon[0] + dcon * gt[0]);
    /* B, :525-570 */
    B[0][0] = u2; B[0][1] = 0; B[0][2] = rho; B[0][3] = 0; B[0][4] = 0;
    f1 = rhoinv / Re;
    B[1][0] = 0; B[1][1] = u2 - f1 * gmu[1]; B[1][2] = -f1 * glm[0]; B[1][3] = 0;
    B[1][4] = -f1 * dmu * 2.0 * S[0][1];
    B[2][0] = rhoinv * t * gmsinv; B[2][1] = -f1 * gmu[0];
    B[2][2] = u2 - f1 * glm[1] - f1 * 2.0 * gmu[1];
    B[2][3] = -f1 * gmu[2];
    B[2][4] = gmsinv - f1 * dlm * divu - f1 * dmu * 2.0 * S[1][1];
    B[3][0] = 0; B[3][1] = 0; B[3][2] = -f1 * glm[2]; B[3][3] = u2 - f1 * gmu[1];
    B[3][4] = -f1 * dmu * 2.0 * S[2][1];
    f1 = gamma * rhoinv / (Pr * Re);
    f2 = gamma * gamma1 * Ma * Ma * rhoinv / Re;
    B[4][0] = 0;
    B[4][1] = -4.0 * f2 * mu * S[0][1];
    B[4][2] = gamma1 * t - f2 * 2.0 * lm * divu - 4.0 * f2 * mu * S[1][1];
    B[4][3] = -4.0 * f2 * mu * S[2][1];
    B[4][4] = u2 - f1 * (gcon[1] + dcon * gt[1]);
    /* Vij, :580-585 */
    f1 = 1.0 / Re; f2 = gamma / (Pr * Re);
    Vij[0] = f1 * rhoinv * (lm + 2.0 * mu);
    Vij[1] = f1 * rhoinv * mu;
    Vij[2] = f2 * rhoinv * con;
    Vij[3] = f1 * rhoinv * (lm + mu);
    /* Ah, Bh, :590-690 */
    double *Ah = c->Ah + 25 * q, *Bh = c->Bh + 25 * q;
#define AH(r, cc) Ah[(r) + 5 * (cc)]
#define BH(r, cc) Bh[(r) + 5 * (cc)]
    const double *mm[2][5] = {{m1 + q, m2 + q, m11 + q, m12 + q, m22 + q},
                              {n1 + q, n2 + q, n11 + q, n12 + q, n22 + q}};
    for (int w = 0; w < 2; w++) {
      double e1 = *mm[w][0], e2 = *mm[w][1], e11 = *mm[w][2], e12 = *mm[w][3], e22 = *mm[w][4];
      double *H = w == 0 ? Ah : Bh;
#define HH(r, cc) H[(r) + 5 * (cc)]
      HH(0, 0) = A[0][0] * e1 + B[0][0] * e2;
      HH(0, 1) = A[0][1] * e1;
      HH(0, 2) = B[0][2] * e2;
      HH(0, 3) = 0; HH(0, 4) = 0;
      HH(1, 0) = A[1][0] * e1;
      HH(1, 1) = A[1][1] * e1 + B[1][1] * e2 - Vij[0] * e11 - Vij[1] * e22;
      HH(1, 2) = A[1][2] * e1 + B[1][2] * e2 - Vij[3] * e12;
      HH(1, 3) = A[1][3] * e1;
      HH(1, 4) = A[1][4] * e1 + B[1][4] * e2;
      HH(2, 0) = B[2][0] * e2;
      HH(2, 1) = A[2][1] * e1 + B[2][1] * e2 - Vij[3] * e12;
      HH(2, 2) = A[2][2] * e1 + B[2][2] * e2 - Vij[1] * e11 - Vij[0] * e22;
      HH(2, 3) = B[2][3] * e2;
      HH(2, 4) = A[2][4] * e1 + B[2][4] * e2;
      HH(3, 0) = 0;
      HH(3, 1) = A[3][1] * e1;
      HH(3, 2) = B[3][2] * e2;
      HH(3, 3) = A[3][3] * e1 + B[3][3] * e2 - Vij[1] * e11 - Vij[1] * e22;
      HH(3, 4) = A[3][4] * e1 + B[3][4] * e2;
      for (int cc = 0; cc < 4; cc++) HH(4, cc) = A[4][cc] * e1 + B[4][cc] * e2;
      HH(4, 4) = A[4][4] * e1 + B[4][4] * e2 - Vij[2] * e11 - Vij[2] * e22;
#undef HH
    }
    if (!p->yper && p->wallt == 2 && q < nx)
      for (int r = 0; r < 5; r++) BH(r, 4) = 0;
#undef AH
#undef BH
    /* D, :700-820, the linear = 1 forms of D(2:4,1), D(5,1) and D(5,5)
     * (:719-723, 743-747, 767-771, 794-797, 817-820) */
    f1 = rhoinv / Re;
    D[0][0] = divu; D[0][1] = grho[0]; D[0][2] = grho[1]; D[0][3] = grho[2]; D[0][4] = 0;
    double Sjj[3] = {S1jj, S2jj, S3jj}, gdivu[3] = {g1divu, g2divu, g3divu};
    for (int k = 0; k < 3; k++) {
      D[1 + k][0] = rhoinv * (u1 * gu[k][0] + u2 * gu[k][1] + u3 * gu[k][2] + gmsinv * gt[k]);
      D[1 + k][1] = gu[k][0]; D[1 + k][2] = gu[k][1]; D[1 + k][3] = gu[k][2];
      D[1 + k][4] = rhoinv * gmsinv * grho[k] - f1 * (gdlm[k] * divu + dlm * gdivu[k]) -
                    f1 * 2.0 * (gdmu[0] * S[k][0] + gdmu[1] * S[k][1] + gdmu[2] * S[k][2] +
                                dmu * Sjj[k]);
    }
    f1 = gamma * rhoinv / (Pr * Re);
    f2 = gamma * gamma1 * Ma * Ma * rhoinv / Re;
    double SS = S[0][0] * S[0][0] + S[0][1] * S[0][1] + S[0][2] * S[0][2] +
                S[1][0] * S[1][0] + S[1][1] * S[1][1] + S[1][2] * S[1][2] +
                S[2][0] * S[2][0] + S[2][1] * S[2][1] + S[2][2] * S[2][2];
    D[4][0] = rhoinv * (u1 * gt[0] + u2 * gt[1] + u3 * gt[2] + gamma1 * t * divu);
    D[4][1] = gt[0]; D[4][2] = gt[1]; D[4][3] = gt[2];
    D[4][4] = -gamma1 * rhoinv * (u1 * grho[0] + u2 * grho[1] + u3 * grho[2]);
    D[4][4] = D[4][4] - f1 * (gdcon[0] * gt[0] + gdcon[1] * gt[1] + gdcon[2] * gt[2] + dcon * Lapt) -
              f2 * dlm * (divu * divu) - 2.0 * f2 * dmu * SS;
    double *Dh = c->Dh + 25 * q;
    for (int r = 0; r < 5; r++)
      for (int cc = 0; cc < 5; cc++) Dh[r + 5 * cc] = D[r][cc];
    if (kz != 0.0) {
      Dh[1 + 5 * 1] += kz * kz * Vij[1];
      Dh[2 + 5 * 2] += kz * kz * Vij[1];
      Dh[3 + 5 * 3] += kz * kz * Vij[0];
      Dh[4 + 5 * 4] += kz * kz * Vij[2];
    }
    /* compact viscous matrices, :850-913 */
    double *V11 = c->Vh11 + 6 * q, *V12 = c->Vh12 + 6 * q, *V22 = c->Vh22 + 6 * q;
    V11[0] = Vij[0] * m1m1 + Vij[1] * m2m2;
    V11[1] = Vij[1] * m1m1 + Vij[0] * m2m2;
    V11[2] = Vij[1] * m1m1 + Vij[1] * m2m2;
    V11[3] = Vij[2] * m1m1 + Vij[2] * m2m2;
    V11[4] = Vij[3] * m1m2;
    V11[5] = Vij[3] * m1m2;
    if (p->idamp == 1 && q % nx == nx - 1)      /* src/genmtrx.f90:877-881 */
      for (int k = 0; k < 6; k++) V11[k] = 0.0 * V11[k];
    V12[0] = 2.0 * (Vij[0] * m1n1 + Vij[1] * m2n2);
    V12[1] = 2.0 * (Vij[1] * m1n1 + Vij[0] * m2n2);
    V12[2] = 2.0 * (Vij[1] * m1n1 + Vij[1] * m2n2);
    V12[3] = 2.0 * (Vij[2] * m1n1 + Vij[2] * m2n2);
    V12[4] = Vij[3] * (m1n2 + m2n1);
    V12[5] = Vij[3] * (m1n2 + m2n1);
    V22[0] = Vij[0] * n1n1 + Vij[1] * n2n2;
    V22[1] = Vij[1] * n1n1 + Vij[0] * n2n2;
    V22[2] = Vij[1] * n1n1 + Vij[1] * n2n2;
    V22[3] = Vij[2] * n1n1 + Vij[2] * n2n2;
    V22[4] = Vij[3] * n1n2;
    V22[5] = Vij[3] * n1n2;
    /* 3-d terms, :918-990 */
    double *ABi = c->ABhi + 6 * q;
    ABi[0] = kz * Vij[3] * M1;
    ABi[1] = kz * Vij[3] * M2;
    ABi[2] = kz * Vij[3] * N1;
    ABi[3] = kz * Vij[3] * N2;
    ABi[4] = 0; ABi[5] = 0;
    f1 = rhoinv / Re;
    C[0][0] = u3; C[0][1] = 0; C[0][2] = 0; C[0][3] = rho; C[0][4] = 0;
    C[1][0] = 0; C[1][1] = u3 - f1 * gmu[2]; C[1][2] = 0; C[1][3] = -f1 * glm[0];
    C[1][4] = -f1 * dmu * 2.0 * S[0][2];
    C[2][0] = 0; C[2][1] = 0; C[2][2] = u3 - f1 * gmu[2]; C[2][3] = -f1 * glm[1];
    C[2][4] = -f1 * dmu * 2.0 * S[1][2];
    C[3][0] = rhoinv * t * gmsinv; C[3][1] = -f1 * gmu[0]; C[3][2] = -f1 * gmu[1];
    C[3][3] = u3 - f1 * glm[2] - f1 * 2.0 * gmu[2];
    C[3][4] = gmsinv - f1 * dlm * divu - f1 * dmu * 2.0 * S[2][2];
    f1 = gamma * rhoinv / (Pr * Re);
    f2 = gamma * gamma1 * Ma * Ma * rhoinv / Re;
    C[4][0] = 0;
    C[4][1] = -4.0 * f2 * mu * S[0][2];
    C[4][2] = -4.0 * f2 * mu * S[1][2];
    C[4][3] = gamma1 * t - f2 * 2.0 * lm * divu - 4.0 * f2 * mu * S[2][2];
    C[4][4] = u3 - f1 * (gcon[2] + dcon * gt[2]);
    double *Dhi = c->Dhi + 25 * q;
    for (int r = 0; r < 5; r++)
      for (int cc = 0; cc < 5; cc++) Dhi[r + 5 * cc] = kz * C[r][cc];
    for (int r = 0; r < 5; r++) Dhi[r + 5 * r] = Dhi[r + 5 * r] - p->omega;
  }
  free(g);
  return 0;
}

/* src/temporal.f90:40-160 (acoustic CFL, global or local time step) */
int orc_set_time_step(const orc_params *p, const double *vm, const double *metric,
                      double *dtl) {
  if (p->mattyp == 1 && g_S0 < 0.0) return -1;
  int nx = p->nx, ny = p->ny;
  long P = (long)nx * ny;
  double dxi = 1.0 / (double)(nx - 1), deta = 1.0 / (double)(ny - 1);
  const double *m1 = metric, *m2 = metric + P, *n1 = metric + 2 * P, *n2 = metric + 3 * P;
  double cfla = -1.0;
  for (long q = 0; q < P; q++) {
    double vinv = 1.0 / (m1[q] * n2[q] - m2[q] * n1[q]);
    double sxj = n2[q] * dxi * vinv, syj = -m2[q] * dxi * vinv;
    double sxi = -n1[q] * deta * vinv, syi = m1[q] * deta * vinv;
    double volinv = 1.0 / (dxi * deta * vinv);
    double acoustic = sqrt(sxi * sxi + 2.0 * sxi * syi + syi * syi + sxj * sxj +
                           2.0 * sxj * syj + syj * syj) / p->Ma;
    double u1 = vm[5 * q + 1], u2 = vm[5 * q + 2], t = vm[5 * q + 4];
    double ct = sqrt(t);
    dtl[q] = (fabs(u1 * sxi + u2 * syi) + fabs(u1 * sxj + u2 * syj) + ct * acoustic) * volinv;
    if (dtl[q] > cfla) cfla = dtl[q];
  }
  double cfl = cfla, delt;
  if (cfl * p->dtmax <= p->cflmax) {
    for (long q = 0; q < P; q++) dtl[q] = p->dtmax;
  } else {
    delt = p->cflmax / cfl;
    if (p->loctime == 0)
      for (long q = 0; q < P; q++) dtl[q] = delt;
    else
      for (long q = 0; q < P; q++) dtl[q] = p->cflmax / dtl[q];
  }
  return 0;
}

/* ---- step driver: src/impdrv3d.f90:146-356 (impl = 1, niter = 1) --------- */
int orc_factor(const orc_params *p, const orc_coeffs *c, const double *dtl,
               zdouble *matx, zdouble *maty, zdouble *perx, zdouble *per2x,
               zdouble *pery, zdouble *per2y) {
  int nx = p->nx, ny = p->ny;
  long P = (long)nx * ny;
  if (orc_lhs(p, 1, c, dtl, maty) != 0) return -1;     /* maty as scratch */
  if (p->xper) {
    memcpy(matx, maty, sizeof(zdouble) * 125 * P);
    if (orc_cpenta_p(2, ny, nx, matx, NULL, perx, per2x, 0, p->periodic_exact)) return -1;
  } else {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        memcpy(matx + 125 * (j + (long)ny * i), maty + 125 * (i + (long)nx * j),
               sizeof(zdouble) * 125);
    if (orc_cpenta_bc(1, ny, nx, matx, NULL, 0)) return -1;
  }
  if (orc_lhs(p, 2, c, dtl, maty) != 0) return -1;
  if (p->yper) return orc_cpenta_p(1, nx, ny, maty, NULL, pery, per2y, 0, p->periodic_exact);
  return orc_cpenta_bc(1, nx, ny, maty, NULL, 0);
}

/* src/impdrv3d.f90:473-501 (rtb3D): r = 4/3 vold - 1/3 v2old - v - 2/3 dtl r */
void orc_rtb3d(const orc_params *p, zdouble *r, const zdouble *v, const zdouble *vold,
               const zdouble *v2old, const double *dtl) {
  const double onept33 = 1.3333333333333333333, pt33 = 3.3333333333333333333e-1,
               pt66 = 6.6666666666666666666e-1;
  long P = (long)p->nx * p->ny;
  for (long q = 0; q < P; q++)
    for (int m = 0; m < 5; m++)
      r[5 * q + m] = onept33 * vold[5 * q + m] - pt33 * v2old[5 * q + m] - v[5 * q + m] -
                     pt66 * dtl[q] * r[5 * q + m];
}

/* rhsBC3D rows, the xi and eta line solves of r, v += r, itrBC3D */
static int solve_update(const orc_params *p, zdouble *matx, zdouble *maty, zdouble *perx,
                        zdouble *per2x, zdouble *pery, zdouble *per2y, zdouble *v, zdouble *r,
                        zdouble *dv) {
  int nx = p->nx, ny = p->ny;
  long P = (long)nx * ny;
  if (orc_rhsbc3d(p, r, v)) return -1;
  if (p->xper) {
    if (orc_cpenta_p(2, ny, nx, matx, r, perx, per2x, 1, p->periodic_exact)) return -1;
  } else {
    zdouble *t = dv;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int d = 0; d < 5; d++) t[d + 5 * (j + (long)ny * i)] = r[d + 5 * (i + (long)nx * j)];
    if (orc_cpenta_bc(1, ny, nx, matx, t, 1)) return -1;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int d = 0; d < 5; d++) r[d + 5 * (i + (long)nx * j)] = t[d + 5 * (j + (long)ny * i)];
  }
  if (p->yper) {
    if (orc_cpenta_p(1, nx, ny, maty, r, pery, per2y, 1, p->periodic_exact)) return -1;
  } else if (orc_cpenta_bc(1, nx, ny, maty, r, 1)) return -1;
  if (!navier(p) && !p->yper) {                /* bcfix, src/impdrv3d.f90:333-336, src/bcfix.f90 */
    if (!g_bnb) return -1;
    for (int i = 0; i < nx; i++) {
      const double b1 = g_bnb[i], b2 = g_bnb[i + nx];
      const zdouble t = b2 * r[1 + 5 * i] + b1 * r[2 + 5 * i];
      r[2 + 5 * i] = -b1 * r[1 + 5 * i] + b2 * r[2 + 5 * i];
      r[1 + 5 * i] = t;
    }
  }
  for (long e = 0; e < 5 * P; e++) { v[e] += r[e]; dv[e] = r[e]; }
  return orc_itrbc3d(p, v);
}

/* one step of niter corrector iterations (src/impdrv3d.f90:173-358; vold is
 * the field at the start of the step); v2old NULL: backward Euler (rbe3D),
 * else the two-step backward residual of impl = 2 (rtb3D,
 * src/impdrv3d.f90:187-190); half: the residual at 0.5 (v + vold) (impl = 3) */
/* the shift-invert driver's forcing (src/si_eigdrv3d_new.f90:325-374): when
 * set, the steps use ebe3D, r = vold - v - dtl (r - f) (src/si_eigdrv3d.f90
 * :333-342), in place of rbe3D */
static const zdouble *g_force;

int orc_set_forcing(const zdouble *f) {
  g_force = f;
  return 0;
}

static void ebe3d(const orc_params *p, zdouble *r, const zdouble *v, const zdouble *vold,
                  const double *dtl) {
  const long P = (long)p->nx * p->ny;
  for (long q = 0; q < P; q++)
    for (int m = 0; m < 5; m++)
      r[5 * q + m] = vold[5 * q + m] - v[5 * q + m] - dtl[q] * (r[5 * q + m] - g_force[5 * q + m]);
}

/* the driver's residual after a step and its norm sqrt(sum |r|^2 / dtl^2)
 * (src/si_eigdrv3d_new.f90:359-374) */
int orc_forced_residual(const orc_params *p, const orc_coeffs *c, const double *dtl,
                        const zdouble *v, const zdouble *vold, zdouble *r, zdouble *work,
                        double *norm) {
  if (!g_force) return -1;
  const long P = (long)p->nx * p->ny;
  int rc = orc_lrhs3d(p, c, v, r, work);
  if (rc) return rc;
  ebe3d(p, r, v, vold, dtl);
  rc = orc_rhsbc3d(p, r, v);
  if (rc) return rc;
  double s = 0.0;
  for (long q = 0; q < P; q++)
    for (int m = 0; m < 5; m++) {
      const zdouble e = r[5 * q + m];
      s += (creal(e) * creal(e) + cimag(e) * cimag(e)) / (dtl[q] * dtl[q]);
    }
  *norm = sqrt(s);
  return 0;
}

static int step_core(const orc_params *p, const orc_coeffs *c, const double *dtl,
                     zdouble *matx, zdouble *maty, zdouble *perx, zdouble *per2x,
                     zdouble *pery, zdouble *per2y, zdouble *v, const zdouble *v2old,
                     int niter, int half, zdouble *dv, zdouble *work) {
  int nx = p->nx, ny = p->ny;
  long P = (long)nx * ny;
  zdouble *vold = work, *r = work + 5 * P, *g = work + 10 * P;
  zdouble *w = half ? (zdouble *)malloc(sizeof(zdouble) * 5 * P) : NULL;
  if (half && !w) return -1;
  memcpy(vold, v, sizeof(zdouble) * 5 * P);
  int rc = orc_itrbc3d(p, v);
  for (int it = 0; it < niter && !rc; it++) {
    if (half) {
      for (long e = 0; e < 5 * P; e++) w[e] = 0.5 * (v[e] + vold[e]);
      rc = orc_lrhs3d(p, c, w, r, g);
    } else {
      rc = orc_lrhs3d(p, c, v, r, g);
    }
    if (rc) break;
    if (v2old) orc_rtb3d(p, r, v, vold, v2old, dtl);
    else if (g_force) ebe3d(p, r, v, vold, dtl);
    else orc_rbe3d(p, r, v, vold, dtl);
    rc = solve_update(p, matx, maty, perx, per2x, pery, per2y, v, r, dv);
  }
  free(w);
  return rc;
}

int orc_step(const orc_params *p, const orc_coeffs *c, const double *dtl,
             zdouble *matx, zdouble *maty, zdouble *perx, zdouble *per2x,
             zdouble *pery, zdouble *per2y, zdouble *v, zdouble *dv, zdouble *work) {
  return step_core(p, c, dtl, matx, maty, perx, per2x, pery, per2y, v, NULL, 1, 0, dv, work);
}

int orc_step_bdf2(const orc_params *p, const orc_coeffs *c, const double *dtl,
                  zdouble *matx, zdouble *maty, zdouble *perx, zdouble *per2x,
                  zdouble *pery, zdouble *per2y, zdouble *v, const zdouble *v2old,
                  zdouble *dv, zdouble *work) {
  return step_core(p, c, dtl, matx, maty, perx, per2x, pery, per2y, v, v2old, 1, 0, dv, work);
}

int orc_step_iter(const orc_params *p, const orc_coeffs *c, const double *dtl,
                  zdouble *matx, zdouble *maty, zdouble *perx, zdouble *per2x,
                  zdouble *pery, zdouble *per2y, zdouble *v, const zdouble *v2old, int niter,
                  int half, zdouble *dv, zdouble *work) {
  if (niter < 1) return -1;
  return step_core(p, c, dtl, matx, maty, perx, per2x, pery, per2y, v, v2old, niter, half, dv, work);
}
