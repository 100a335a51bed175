// Kernels of one implicit approximate-factorization step on gfx950, other
// than the line solves (penta.hip):
//   coefficient generation   src/genmtrx.f90 (with grad/grad2/gradbc/getmat)
//   local time step          src/temporal.f90:40-160
//   LHS assembly             src/lhs1f3d.f90, lhs2f3d.f90 (interior),
//                            lhsbt1f3d.f90, lhsbt2f3d.f90 (rows 1,2,n-1,n),
//                            lhsbc1f3d.f90, lhsbc2f3d.f90 (boundary rows)
//   RHS                      src/lrhs3d.f90 with cgrad/cgrad2/cgradbc, then
//                            rbe3D (impdrv3d.f90:444-470) and rhsBC3D
//   update + BCs             impdrv3d.f90:340-356, src/itrbc3d.f90
// One thread per grid point.  Coefficients are structure-of-arrays ([e][P])
// so a wave's loads of one coefficient are contiguous.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "step.hpp"

namespace lns3d {

using z = cx<double>;

// --------------------------------------------------------------------------
// scalar/complex helpers for the stencil templates
__device__ __forceinline__ double sc(double a, double x) { return a * x; }
__device__ __forceinline__ z sc(double a, z x) { return {a * x.re, a * x.im}; }
__device__ __forceinline__ double ld(const double *p) { return *p; }
__device__ __forceinline__ z ld(const double2 *p) { double2 v = *p; return {v.x, v.y}; }
// right = 9 (src/rhsbc3d.f90:317-320, src/itrbc3d.f90:408-410): the value
// exp(2 log a - log b) of the column from the two columns inside it, formed
// as a^2 / b (the same value, zero for a = 0, not finite for b = 0)
__device__ __forceinline__ z extrap_log(z a, z b) {
  const z a2{a.re * a.re - a.im * a.im, 2.0 * a.re * a.im};
  const double d = b.re * b.re + b.im * b.im;
  return {(a2.re * b.re + a2.im * b.im) / d, (a2.im * b.re - a2.re * b.im) / d};
}

// first derivative along a line at position p (N points), value stride es
// between consecutive points; component e selected by base offset.
// Formulas of src/cgrad.f90:84-283 (periodic and non-periodic branches).
// Symmetric ends (left/right = 7; sym bit 0 left, bit 1 right): the three
// points next to the end take the interior stencil with the points beyond it
// reflected onto the line, negated by sgn (-1 for component 3), as
// src/cgrad.f90:190-262, cgrad2.f90:229-310, grad.f90:155-290, grad2.f90:250-330.
__device__ __forceinline__ bool sym_near(int sym, int p, int N) {
  return ((sym & 1) && p <= 2) || ((sym & 2) && p >= N - 3);
}

// dissipation of src/smoother3d.f90 along a line at position p of N (without
// eps_e buff): second difference at the second point from an end, fourth at
// the third, the 7-point sixth difference inside, the 7-point form with
// reflected points next to a symmetric end; false on the end points
__device__ __forceinline__ bool smooth_at(int N, int p, int sym, double sgn, const double2 *base, long es,
                                          z &out) {
  constexpr double f[7] = {-1.666666666666667E-01, 2.0, -6.5, 9.333333333333333E+00, -6.5, 2.0,
                           -1.666666666666667E-01};
  auto x = [&](int q) { return ld(base + (long)q * es); };
  if (sym && sym_near(sym, p, N)) {
    auto w = [&](int q) { return q < 0 ? sc(sgn, x(-q)) : q > N - 1 ? sc(sgn, x(2 * (N - 1) - q)) : x(q); };
    out = sc(f[0], w(p - 3)) + sc(f[1], w(p - 2)) + sc(f[2], w(p - 1)) + sc(f[3], w(p)) + sc(f[4], w(p + 1)) +
          sc(f[5], w(p + 2)) + sc(f[6], w(p + 3));
    return true;
  }
  if (p == 1 || p == N - 2) {
    out = sc(-1.0, x(p - 1)) + sc(2.0, x(p)) - sc(1.0, x(p + 1));
    return true;
  }
  if (p == 2 || p == N - 3) {
    out = x(p - 2) - sc(4.0, x(p - 1)) + sc(6.0, x(p)) - sc(4.0, x(p + 1)) + sc(1.0, x(p + 2));
    return true;
  }
  if (p >= 3 && p <= N - 4) {
    out = sc(f[0], x(p - 3)) + sc(f[1], x(p - 2)) + sc(f[2], x(p - 1)) + sc(f[3], x(p)) + sc(f[4], x(p + 1)) +
          sc(f[5], x(p + 2)) + sc(f[6], x(p + 3));
    return true;
  }
  return false;
}

template <typename S, typename V>
__device__ __forceinline__ V d1(int N, int per, int p, const double *g, double dinv,
                                const S *base, long es, int sym = 0, double sgn = 1.0) {
  using namespace st;
  auto x = [&](int q) { return ld(base + (long)q * es); };
  if (per) {
    int m = N - 1;
    if (p == N - 1) p = 0;
    auto w = [&](int q) { return x(((q % m) + m) % m); };
    return sc(g[0], w(p - 3)) + sc(g[1], w(p - 2)) + sc(g[2], w(p - 1)) + sc(g[3], w(p + 1)) +
           sc(g[4], w(p + 2)) + sc(g[5], w(p + 3));
  }
  if (sym && sym_near(sym, p, N)) {
    auto w = [&](int q) { return q < 0 ? sc(sgn, x(-q)) : q > N - 1 ? sc(sgn, x(2 * (N - 1) - q)) : x(q); };
    return sc(g[0], w(p - 3)) + sc(g[1], w(p - 2)) + sc(g[2], w(p - 1)) + sc(g[3], w(p + 1)) +
           sc(g[4], w(p + 2)) + sc(g[5], w(p + 3));
  }
  if (p == 0) return sc(dinv, sc(gc1, x(0)) + sc(gc2, x(1)) + sc(gc3, x(2)) + sc(gc4, x(3)) + sc(gc5, x(4)));
  if (p == 1) return sc(dinv, sc(gb1, x(0)) + sc(gb2, x(1)) + sc(gb3, x(2)) + sc(gb4, x(3)) + sc(gb5, x(4)));
  if (p == 2) return sc(dinv, sc(ga1, x(0)) + sc(ga2, x(1)) + sc(ga3, x(3)) + sc(ga4, x(4)));
  if (p == N - 3)
    return sc(dinv, sc(ga1, x(N - 5)) + sc(ga2, x(N - 4)) + sc(ga3, x(N - 2)) + sc(ga4, x(N - 1)));
  if (p == N - 2)
    return sc(-dinv, sc(gb1, x(N - 1)) + sc(gb2, x(N - 2)) + sc(gb3, x(N - 3)) + sc(gb4, x(N - 4)) +
                         sc(gb5, x(N - 5)));
  if (p == N - 1)
    return sc(-dinv, sc(gc1, x(N - 1)) + sc(gc2, x(N - 2)) + sc(gc3, x(N - 3)) + sc(gc4, x(N - 4)) +
                         sc(gc5, x(N - 5)));
  return sc(g[0], x(p - 3)) + sc(g[1], x(p - 2)) + sc(g[2], x(p - 1)) + sc(g[3], x(p + 1)) +
         sc(g[4], x(p + 2)) + sc(g[5], x(p + 3));
}

// second derivative, src/cgrad2.f90:134-326
template <typename S, typename V>
__device__ __forceinline__ V d2(int N, int per, int p, const double *d, double dsinv,
                                const S *base, long es, int sym = 0, double sgn = 1.0) {
  using namespace st;
  auto x = [&](int q) { return ld(base + (long)q * es); };
  if (per) {
    int m = N - 1;
    if (p == N - 1) p = 0;
    auto w = [&](int q) { return x(((q % m) + m) % m); };
    return sc(d[0], w(p - 3)) + sc(d[1], w(p - 2)) + sc(d[2], w(p - 1)) + sc(d[3], w(p)) +
           sc(d[4], w(p + 1)) + sc(d[5], w(p + 2)) + sc(d[6], w(p + 3));
  }
  if (sym && sym_near(sym, p, N)) {
    auto w = [&](int q) { return q < 0 ? sc(sgn, x(-q)) : q > N - 1 ? sc(sgn, x(2 * (N - 1) - q)) : x(q); };
    return sc(d[0], w(p - 3)) + sc(d[1], w(p - 2)) + sc(d[2], w(p - 1)) + sc(d[3], w(p)) +
           sc(d[4], w(p + 1)) + sc(d[5], w(p + 2)) + sc(d[6], w(p + 3));
  }
  if (p == 0) return sc(dsinv, sc(dd1, x(0)) + sc(dd2, x(1)) + sc(dd3, x(2)) + sc(dd4, x(3)) + sc(dd5, x(4)));
  if (p == 1) return sc(dsinv, sc(db1, x(0)) + sc(db2, x(1)) + sc(db3, x(2)) + sc(db4, x(3)) + sc(db5, x(4)));
  if (p == 2) return sc(dsinv, sc(da1, x(0)) + sc(da2, x(1)) + sc(da3, x(2)) + sc(da4, x(3)) + sc(da5, x(4)));
  if (p == N - 3)
    return sc(dsinv, sc(da1, x(N - 5)) + sc(da2, x(N - 4)) + sc(da3, x(N - 3)) + sc(da4, x(N - 2)) +
                         sc(da5, x(N - 1)));
  if (p == N - 2)
    return sc(dsinv, sc(db1, x(N - 1)) + sc(db2, x(N - 2)) + sc(db3, x(N - 3)) + sc(db4, x(N - 4)) +
                         sc(db5, x(N - 5)));
  if (p == N - 1)
    return sc(dsinv, sc(dd1, x(N - 1)) + sc(dd2, x(N - 2)) + sc(dd3, x(N - 3)) + sc(dd4, x(N - 4)) +
                         sc(dd5, x(N - 5)));
  return sc(d[0], x(p - 3)) + sc(d[1], x(p - 2)) + sc(d[2], x(p - 1)) + sc(d[3], x(p)) +
         sc(d[4], x(p + 1)) + sc(d[5], x(p + 2)) + sc(d[6], x(p + 3));
}

Stencil make_stencil(const lns3d_params &p) {
  // src/gengrid.f90:36,44 (dxi = 1/(nx-1), deta = 1/(ny-1)); symmetric ends
  // from left/right = 7 (src/input.f90:235-243, off for periodic x)
  return make_stencil_spacing(p.optx, p.opty, 1.0 / (double)(p.nx - 1), 1.0 / (double)(p.ny - 1),
                              p.xper ? 0 : ((p.left == 7 ? 1 : 0) | (p.right == 7 ? 2 : 0)));
}

Stencil make_stencil_spacing(int optx, int opty, double dxi, double deta, int sym) {
  Stencil s;
  s.dxi = dxi;
  s.deta = deta;
  s.dxinv = 1.0 / s.dxi;
  s.dyinv = 1.0 / s.deta;
  s.dxsinv = 1.0 / (s.dxi * s.dxi);
  s.dysinv = 1.0 / (s.deta * s.deta);
  s.sym = sym;
  auto first = [](int opt, double dinv, double *g) {
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
    g[3] = a * dinv; g[4] = b * dinv; g[5] = c * dinv;
  };
  auto second = [](int opt, double dsinv, double *d) {
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
  };
  first(optx, s.dxinv, s.gx);
  first(opty, s.dyinv, s.gy);
  second(optx, s.dxsinv, s.dx);
  second(opty, s.dysinv, s.dy);
  return s;
}

static inline int nblocks(long n, int b) { return (int)((n + b - 1) / b); }

// ==========================================================================
// coefficient generation (src/genmtrx.f90:72-1000, general metrics,
// linear != 1, idamp = 0)
// first derivative in x of the mean flow into g1 (5,nx,ny)
__global__ void k_mean_d1x(int nx, int ny, int xper, Stencil st, const double *vm, double *g1) {
#pragma clang fp contract(off)
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)nx * ny) return;
  int i = q % nx, j = q / nx;
  const double *base = vm + 5L * nx * j;
  for (int e = 0; e < 5; e++)
    g1[5 * q + e] = d1<double, double>(nx, xper, i, st.gx, st.dxinv, base + e, 5, st.sym, e == 2 ? -1.0 : 1.0);
}

// material properties (src/getmat.f90:14-70): constant (0), Sutherland's law
// with the nondimensional constant S0 (1), linear law (2, 3)
__device__ __forceinline__ void getmat(int mattyp, double S0, double t, double &mu, double &lm, double &con,
                                       double &dmu, double &d2mu, double &dlm, double &d2lm,
                                       double &dcon, double &d2con) {
  if (mattyp == 0) { mu = 1.0; dmu = 0.0; d2mu = 0.0; }
  else if (mattyp == 1) {
    const double ts = S0 + t;
    mu = pow(t, 1.5) * (1.0 + S0) / ts;
    dmu = sqrt(t) * (1.0 + S0) * (3.0 * S0 + t) * 0.5 / (ts * ts);
    d2mu = (1.0 + S0) * (3.0 * (S0 * S0) - 6.0 * S0 * t - t * t) * 0.25 / sqrt(t) / (ts * ts * ts);
  }
  else { mu = t; dmu = 1.0; d2mu = 0.0; }
  con = mu; dcon = dmu; d2con = d2mu;
  if (mattyp == 3) { lm = 0.0; dlm = 0.0; d2lm = 0.0; }
  else { lm = -st::pt66 * mu; dlm = -st::pt66 * dmu; d2lm = -st::pt66 * d2mu; }
}

__global__ __launch_bounds__(128) void k_genmtrx(lns3d_params p, Stencil st, const double *vm, const double *met,
                          const double *g1a, Coeffs c) {
  // set-up kernel: rounded operation by operation (no fused multiply-add) so
  // the coefficients follow the reference's arithmetic; several of them are
  // differences of nearly equal stencil sums
#pragma clang fp contract(off)
  const int nx = p.nx, ny = p.ny;
  const long P = (long)nx * ny;
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P) return;
  int i = q % nx, j = q / nx;
  double g1[5], g2[5], g11[5], g12[5], g22[5];
  const double *colv = vm + 5L * i, *colg = g1a + 5L * i, *rowv = vm + 5L * nx * j;
  for (int e = 0; e < 5; e++) {
    g1[e] = g1a[5 * q + e];
    g2[e] = d1<double, double>(ny, p.yper, j, st.gy, st.dyinv, colv + e, 5L * nx);
    g11[e] = d2<double, double>(nx, p.xper, i, st.dx, st.dxsinv, rowv + e, 5, st.sym, e == 2 ? -1.0 : 1.0);
    g22[e] = d2<double, double>(ny, p.yper, j, st.dy, st.dysinv, colv + e, 5L * nx);
    g12[e] = d1<double, double>(ny, p.yper, j, st.gy, st.dyinv, colg + e, 5L * nx);
  }
  if (!c.bnb && (p.wallt == 1 || p.wallt == 2) && j == 0) g2[4] = 0.0;   // src/gradbc.f90:15 (Navier)
  // idamp = 1: the damper's zero on the last column times the xi second
  // derivatives of the mean (src/damper.f90:26-29, src/genmtrx.f90:87-96)
  if (p.idamp == 1 && i == nx - 1)
    for (int e = 0; e < 5; e++) g11[e] = 0.0 * g11[e];
  const double M1 = met[q], M2 = met[P + q], N1 = met[2 * P + q], N2 = met[3 * P + q];
  const double m11 = met[4 * P + q], m12 = met[5 * P + q], m22 = met[6 * P + q];
  const double n11 = met[7 * P + q], n12 = met[8 * P + q], n22 = met[9 * P + q];
  const double m1m1 = M1 * M1, m1m2 = M1 * M2, m2m2 = M2 * M2;
  const double n1n1 = N1 * N1, n1n2 = N1 * N2, n2n2 = N2 * N2;
  const double m1n1 = M1 * N1, m1n2 = M1 * N2, m2n1 = M2 * N1, m2n2 = M2 * N2;
  double g11l[5], g12l[5], g22l[5], g1l[5], g2l[5];
  for (int d = 0; d < 5; d++) {
    g11l[d] = g11[d] * m1m1 + 2.0 * g12[d] * m1n1 + g22[d] * n1n1 + g1[d] * m11 + g2[d] * n11;
    g12l[d] = g11[d] * m1m2 + g12[d] * (m1n2 + m2n1) + g22[d] * n1n2 + g1[d] * m12 + g2[d] * n12;
    g22l[d] = g11[d] * m2m2 + 2.0 * g12[d] * m2n2 + g22[d] * n2n2 + g1[d] * m22 + g2[d] * n22;
    g1l[d] = g1[d] * M1 + g2[d] * N1;
    g2l[d] = g1[d] * M2 + g2[d] * N2;
  }
  const double gamma = p.gamma, gamma1 = gamma - 1.0, Ma = p.Ma, Re = p.Re, Pr = p.Pr, kz = p.kz;
  const double fact = 1.0 / (gamma * Ma * Ma);
  const double rho = vm[5 * q], u1 = vm[5 * q + 1], u2 = vm[5 * q + 2], u3 = vm[5 * q + 3],
               t = vm[5 * q + 4];
  const double rhoinv = 1.0 / rho;
  const double grho[3] = {g1l[0], g2l[0], 0.0};
  const double gu[3][3] = {{g1l[1], g2l[1], 0.0}, {g1l[2], g2l[2], 0.0}, {g1l[3], g2l[3], 0.0}};
  const double gt[3] = {g1l[4], g2l[4], 0.0};
  const double gdivu[3] = {g11l[1] + g12l[2], g12l[1] + g22l[2], 0.0};
  const double Sjj[3] = {0.5 * (g11l[1] + g11l[1] + g22l[1] + g12l[2]),
                         0.5 * (g11l[2] + g12l[1] + g22l[2] + g22l[2]), 0.5 * (g11l[3] + g22l[3])};
  const double Lapt = g11l[4] + g22l[4];
  const double divu = gu[0][0] + gu[1][1] + gu[2][2];
  double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  S[0][0] = gu[0][0];
  S[1][0] = 0.5 * (gu[1][0] + gu[0][1]);
  S[2][0] = 0.5 * gu[2][0];
  S[0][1] = 0.5 * (gu[0][1] + gu[1][0]);
  S[1][1] = gu[1][1];
  S[2][1] = 0.5 * gu[2][1];
  S[0][2] = 0.5 * gu[2][0];
  S[1][2] = 0.5 * gu[2][1];
  double mu, lm, con, dmu, d2mu, dlm, d2lm, dcon, d2con;
  getmat(p.mattyp, c.S0, t, mu, lm, con, dmu, d2mu, dlm, d2lm, dcon, d2con);
  double gmu[3], gdmu[3], gcon[3], gdcon[3], glm[3], gdlm[3];
  for (int k = 0; k < 3; k++) {
    gmu[k] = dmu * gt[k]; gdmu[k] = d2mu * gt[k];
    gcon[k] = dcon * gt[k]; gdcon[k] = d2con * gt[k];
    glm[k] = dlm * gt[k]; gdlm[k] = d2lm * gt[k];
  }
  double A[5][5], B[5][5], D[5][5], Vij[4], f1, f2;
  const double gmsinv = 1.0 / (gamma * Ma * Ma);
  // A and B (src/genmtrx.f90:430-570)
  f1 = rhoinv / Re;
  A[0][0] = u1; A[0][1] = rho; A[0][2] = 0; A[0][3] = 0; A[0][4] = 0;
  A[1][0] = rhoinv * t * gmsinv;
  A[1][1] = u1 - f1 * glm[0] - f1 * 2.0 * gmu[0];
  A[1][2] = -f1 * gmu[1];
  A[1][3] = -f1 * gmu[2];
  A[1][4] = gmsinv - f1 * dlm * divu - f1 * dmu * 2.0 * S[0][0];
  A[2][0] = 0; A[2][1] = -f1 * glm[1]; A[2][2] = u1 - f1 * gmu[0]; A[2][3] = 0;
  A[2][4] = -f1 * dmu * 2.0 * S[1][0];
  A[3][0] = 0; A[3][1] = -f1 * glm[2]; A[3][2] = 0; A[3][3] = u1 - f1 * gmu[0];
  A[3][4] = -f1 * dmu * 2.0 * S[2][0];
  B[0][0] = u2; B[0][1] = 0; B[0][2] = rho; B[0][3] = 0; B[0][4] = 0;
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
  A[4][0] = 0;
  A[4][1] = gamma1 * t - f2 * 2.0 * lm * divu - 4.0 * f2 * mu * S[0][0];
  A[4][2] = -4.0 * f2 * mu * S[1][0];
  A[4][3] = -4.0 * f2 * mu * S[2][0];
  A[4][4] = u1 - f1 * (gcon[0] + dcon * gt[0]);
  B[4][0] = 0;
  B[4][1] = -4.0 * f2 * mu * S[0][1];
  B[4][2] = gamma1 * t - f2 * 2.0 * lm * divu - 4.0 * f2 * mu * S[1][1];
  B[4][3] = -4.0 * f2 * mu * S[2][1];
  B[4][4] = u2 - f1 * (gcon[1] + dcon * gt[1]);
  f1 = 1.0 / Re; f2 = gamma / (Pr * Re);
  Vij[0] = f1 * rhoinv * (lm + 2.0 * mu);
  Vij[1] = f1 * rhoinv * mu;
  Vij[2] = f2 * rhoinv * con;
  Vij[3] = f1 * rhoinv * (lm + mu);
  // Ah, Bh (src/genmtrx.f90:590-690)
  for (int w = 0; w < 2; w++) {
    double e1 = w ? N1 : M1, e2 = w ? N2 : M2, e11 = w ? n11 : m11, e12 = w ? n12 : m12,
           e22 = w ? n22 : m22;
    double H[5][5];
    H[0][0] = A[0][0] * e1 + B[0][0] * e2;
    H[0][1] = A[0][1] * e1;
    H[0][2] = B[0][2] * e2;
    H[0][3] = 0; H[0][4] = 0;
    H[1][0] = A[1][0] * e1;
    H[1][1] = A[1][1] * e1 + B[1][1] * e2 - Vij[0] * e11 - Vij[1] * e22;
    H[1][2] = A[1][2] * e1 + B[1][2] * e2 - Vij[3] * e12;
    H[1][3] = A[1][3] * e1;
    H[1][4] = A[1][4] * e1 + B[1][4] * e2;
    H[2][0] = B[2][0] * e2;
    H[2][1] = A[2][1] * e1 + B[2][1] * e2 - Vij[3] * e12;
    H[2][2] = A[2][2] * e1 + B[2][2] * e2 - Vij[1] * e11 - Vij[0] * e22;
    H[2][3] = B[2][3] * e2;
    H[2][4] = A[2][4] * e1 + B[2][4] * e2;
    H[3][0] = 0;
    H[3][1] = A[3][1] * e1;
    H[3][2] = B[3][2] * e2;
    H[3][3] = A[3][3] * e1 + B[3][3] * e2 - Vij[1] * e11 - Vij[1] * e22;
    H[3][4] = A[3][4] * e1 + B[3][4] * e2;
    for (int cc = 0; cc < 4; cc++) H[4][cc] = A[4][cc] * e1 + B[4][cc] * e2;
    H[4][4] = A[4][4] * e1 + B[4][4] * e2 - Vij[2] * e11 - Vij[2] * e22;
    if (w == 1 && !p.yper && p.wallt == 2 && j == 0)
      for (int r = 0; r < 5; r++) H[r][4] = 0.0;
    double *dst = w ? c.Bh : c.Ah;
    for (int r = 0; r < 5; r++)
      for (int cc = 0; cc < 5; cc++) dst[(long)(r + 5 * cc) * P + q] = H[r][cc];
  }
  // D (src/genmtrx.f90:700-820) in its linear = 1 forms (:719-723, 743-747,
  // 767-771, 794-797, 817-820): the deck's linear = 2 is linear = 1 once read
  // (src/input.f90:367-371), the only value the 3D driver runs with
  f1 = rhoinv / Re;
  D[0][0] = divu; D[0][1] = grho[0]; D[0][2] = grho[1]; D[0][3] = grho[2]; D[0][4] = 0;
  for (int k = 0; k < 3; k++) {
    D[1 + k][0] = rhoinv * (u1 * gu[k][0] + u2 * gu[k][1] + u3 * gu[k][2] + gmsinv * gt[k]);
    D[1 + k][1] = gu[k][0]; D[1 + k][2] = gu[k][1]; D[1 + k][3] = gu[k][2];
    D[1 + k][4] = rhoinv * gmsinv * grho[k] - f1 * (gdlm[k] * divu + dlm * gdivu[k]) -
                  f1 * 2.0 * (gdmu[0] * S[k][0] + gdmu[1] * S[k][1] + gdmu[2] * S[k][2] + dmu * Sjj[k]);
  }
  f1 = gamma * rhoinv / (Pr * Re);
  f2 = gamma * gamma1 * Ma * Ma * rhoinv / Re;
  const double SS = S[0][0] * S[0][0] + S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][0] * S[1][0] +
                    S[1][1] * S[1][1] + S[1][2] * S[1][2] + S[2][0] * S[2][0] + S[2][1] * S[2][1] +
                    S[2][2] * S[2][2];
  D[4][0] = rhoinv * (u1 * gt[0] + u2 * gt[1] + u3 * gt[2] + gamma1 * t * divu);
  D[4][1] = gt[0]; D[4][2] = gt[1]; D[4][3] = gt[2];
  D[4][4] = -gamma1 * rhoinv * (u1 * grho[0] + u2 * grho[1] + u3 * grho[2]);
  D[4][4] = D[4][4] - f1 * (gdcon[0] * gt[0] + gdcon[1] * gt[1] + gdcon[2] * gt[2] + dcon * Lapt) -
            f2 * dlm * (divu * divu) - 2.0 * f2 * dmu * SS;
  if (kz != 0.0) {
    D[1][1] += kz * kz * Vij[1];
    D[2][2] += kz * kz * Vij[1];
    D[3][3] += kz * kz * Vij[0];
    D[4][4] += kz * kz * Vij[2];
  }
  for (int r = 0; r < 5; r++)
    for (int cc = 0; cc < 5; cc++) c.Dh[(long)(r + 5 * cc) * P + q] = D[r][cc];
  // compact viscous matrices (src/genmtrx.f90:850-913)
  double V11[6] = {Vij[0] * m1m1 + Vij[1] * m2m2, Vij[1] * m1m1 + Vij[0] * m2m2,
                   Vij[1] * m1m1 + Vij[1] * m2m2, Vij[2] * m1m1 + Vij[2] * m2m2,
                   Vij[3] * m1m2, Vij[3] * m1m2};
  double V12[6] = {2.0 * (Vij[0] * m1n1 + Vij[1] * m2n2), 2.0 * (Vij[1] * m1n1 + Vij[0] * m2n2),
                   2.0 * (Vij[1] * m1n1 + Vij[1] * m2n2), 2.0 * (Vij[2] * m1n1 + Vij[2] * m2n2),
                   Vij[3] * (m1n2 + m2n1), Vij[3] * (m1n2 + m2n1)};
  double V22[6] = {Vij[0] * n1n1 + Vij[1] * n2n2, Vij[1] * n1n1 + Vij[0] * n2n2,
                   Vij[1] * n1n1 + Vij[1] * n2n2, Vij[2] * n1n1 + Vij[2] * n2n2,
                   Vij[3] * n1n2, Vij[3] * n1n2};
  if (p.idamp == 1 && i == nx - 1)        // src/genmtrx.f90:877-881
    for (int e = 0; e < 6; e++) V11[e] = 0.0 * V11[e];
  for (int e = 0; e < 6; e++) {
    c.V11[e * P + q] = V11[e];
    c.V12[e * P + q] = V12[e];
    c.V22[e * P + q] = V22[e];
  }
  // 3-d terms (src/genmtrx.f90:918-990)
  c.ABhi[q] = kz * Vij[3] * M1;
  c.ABhi[P + q] = kz * Vij[3] * M2;
  c.ABhi[2 * P + q] = kz * Vij[3] * N1;
  c.ABhi[3 * P + q] = kz * Vij[3] * N2;
  double C[5][5];
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
  for (int r = 0; r < 5; r++)
    for (int cc = 0; cc < 5; cc++)
      c.Dhi[(long)(r + 5 * cc) * P + q] = kz * C[r][cc] - (r == cc ? p.omega : 0.0);
}

int launch_genmtrx(const lns3d_params &p, const Stencil &st, Coeffs &c, double *work, hipStream_t s) {
  long P = (long)p.nx * p.ny;
  hipLaunchKernelGGL(k_mean_d1x, dim3(nblocks(P, 256)), dim3(256), 0, s, p.nx, p.ny, p.xper, st,
                     (const double *)c.vm, work);
  LNS3D_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_genmtrx, dim3(nblocks(P, 128)), dim3(128), 0, s, p, st, (const double *)c.vm,
                     (const double *)c.metric, (const double *)work, c);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// ==========================================================================
// time step (src/temporal.f90:40-160): acoustic CFL per point, global max
__global__ void k_dt_raw(lns3d_params p, const double *vm, const double *met, double *dtl,
                         unsigned long long *maxbits) {
#pragma clang fp contract(off)
  const long P = (long)p.nx * p.ny;
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (q < P) {
    double dxi = 1.0 / (double)(p.nx - 1), deta = 1.0 / (double)(p.ny - 1);
    double m1 = met[q], m2 = met[P + q], n1 = met[2 * P + q], n2 = met[3 * P + q];
    double vinv = 1.0 / (m1 * n2 - m2 * n1);
    double sxj = n2 * dxi * vinv, syj = -m2 * dxi * vinv;
    double sxi = -n1 * deta * vinv, syi = m1 * deta * vinv;
    double volinv = 1.0 / (dxi * deta * vinv);
    double acoustic = sqrt(sxi * sxi + 2.0 * sxi * syi + syi * syi + sxj * sxj + 2.0 * sxj * syj +
                           syj * syj) / p.Ma;
    double u1 = vm[5 * q + 1], u2 = vm[5 * q + 2], ct = sqrt(vm[5 * q + 4]);
    v = (fabs(u1 * sxi + u2 * syi) + fabs(u1 * sxj + u2 * syj) + ct * acoustic) * volinv;
    dtl[q] = v;
  }
  // block max of non-negative values, then one atomic per block on the bits
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_dt_final(lns3d_params p, double *dtl, const unsigned long long *maxbits) {
  const long P = (long)p.nx * p.ny;
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P) return;
  double cfl = __longlong_as_double((long long)*maxbits);
  if (cfl * p.dtmax <= p.cflmax) dtl[q] = p.dtmax;
  else if (p.loctime == 0) dtl[q] = p.cflmax / cfl;
  else dtl[q] = p.cflmax / dtl[q];
}

__global__ void k_zero_u64(unsigned long long *x) { *x = 0ull; }

int launch_timestep(const lns3d_params &p, Coeffs &c, double *work, hipStream_t s) {
  long P = (long)p.nx * p.ny;
  auto *mx = (unsigned long long *)work;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, mx);
  hipLaunchKernelGGL(k_dt_raw, dim3(nblocks(P, 256)), dim3(256), 0, s, p, (const double *)c.vm,
                     (const double *)c.metric, c.dtl, mx);
  hipLaunchKernelGGL(k_dt_final, dim3(nblocks(P, 256)), dim3(256), 0, s, p, c.dtl,
                     (const unsigned long long *)mx);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// ==========================================================================
// inflow forcing of the sponge, src/lrhs3d.f90:248-282: per point the
// complex amplitude c3 of the incoming acoustic wave and the five terms
// subtracted from v under spg2 (components 3 and 4 have none)
__device__ void acoustic_wave(const lns3d_params &p, const double *vml, double xl, double x0, double amp,
                              z (&f)[5]) {
  const double cm = sqrt(vml[4]) / p.Ma, um = vml[1];
  const double d = 0.5 * (1.3333333333333333333 + (p.gamma - 1.0) / p.Pr) / p.Re;
  const double kk = p.omega / (cm + um);
  const double a = p.omega * p.omega * d / ((cm + um) * (cm + um) * (cm + um));
  double sn, cs;
  sincos(kk * xl, &sn, &cs);
  const double e = exp(-a * (xl - x0));
  const z c3 = sc(amp, z{e * cs, e * sn});
  const double rho = vml[0], T = vml[4], cm2 = cm * cm, gM2 = p.gamma * p.Ma * p.Ma;
  f[0] = sc(0.5 / cm2, c3);
  f[1] = sc(0.5 / (rho * cm), c3);
  f[2] = z{0, 0};
  f[3] = z{0, 0};
  f[4] = sc(1.0 / rho, sc(gM2 * 0.5, c3) - sc(T * 0.5 / cm2, c3));
}

__global__ void k_sponge_forcing(lns3d_params p, int ispg, const double *vm, const double *x,
                                 const double *wamp, double x0, double *spf) {
  const long P = (long)p.nx * p.ny;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P) return;
  const int i = (int)(q % p.nx);
  z f[5];
  acoustic_wave(p, vm + 5 * q, x[q], x0, ispg == 3 ? wamp[i] : 1.0, f);
  for (int m = 0; m < 5; m++) {
    spf[2 * (m * P + q)] = f[m].re;
    spf[2 * (m * P + q) + 1] = f[m].im;
  }
}

// the top = 1 wave at the top row, src/rhsbc3d.f90:131-150
__global__ void k_top_forcing(lns3d_params p, const double *vm, const double *xtop, const double *wamp,
                              double x0, double *ftop) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.nx) return;
  const long q = i + (long)p.nx * (p.ny - 1);
  z f[5];
  acoustic_wave(p, vm + 5 * q, xtop[i], x0, wamp ? wamp[i] : 1.0, f);
  for (int m = 0; m < 5; m++) {
    ftop[2 * (m * p.nx + i)] = f[m].re;
    ftop[2 * (m * p.nx + i) + 1] = f[m].im;
  }
}

// left = 5 / right = 5 (src/rhsbc3d.f90:248-270, 286-307): the same wave
// undamped, exp(i k x), from the mean flow of that column
__device__ void plain_wave(const lns3d_params &p, const double *vml, double xl, z (&f)[5]) {
  const double cm = sqrt(vml[4]) / p.Ma, um = vml[1];
  const double kk = p.omega / (cm + um);
  double sn, cs;
  sincos(kk * xl, &sn, &cs);
  const z c3{cs, sn};
  const double rho = vml[0], T = vml[4], cm2 = cm * cm, gM2 = p.gamma * p.Ma * p.Ma;
  f[0] = sc(0.5 / cm2, c3);
  f[1] = sc(0.5 / (rho * cm), c3);
  f[2] = z{0, 0};
  f[3] = z{0, 0};
  f[4] = sc(1.0 / rho, sc(gM2 * 0.5, c3) - sc(T * 0.5 / cm2, c3));
}

// the wave on the left (side 0) or right (side 1) column into fside: damped
// (left = 1, src/rhsbc3d.f90:176-205, src/itrbc3d.f90:219-245: the top wave's
// form with amplitude 1) or not (left/right = 5); xcol the x value each line
// j reads
__global__ void k_side_wave(lns3d_params p, const double *vm, const double *xcol, double x0, int side,
                            int damped, double *fside) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p.ny) return;
  const long q = (side ? p.nx - 1 : 0) + (long)p.nx * j;
  z f[5];
  if (damped) acoustic_wave(p, vm + 5 * q, xcol[j], x0, 1.0, f);
  else plain_wave(p, vm + 5 * q, xcol[j], f);
  for (int m = 0; m < 5; m++) {
    fside[2 * ((5 * side + m) * p.ny + j)] = f[m].re;
    fside[2 * ((5 * side + m) * p.ny + j) + 1] = f[m].im;
  }
}

int launch_side_wave(const lns3d_params &p, const Coeffs &c, const double *xcol, double x0, int side,
                     int damped, hipStream_t s) {
  hipLaunchKernelGGL(k_side_wave, dim3(nblocks(p.ny, 128)), dim3(128), 0, s, p, (const double *)c.vm,
                     xcol, x0, side, damped, c.fside);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int launch_top_forcing(const lns3d_params &p, const Coeffs &c, const double *xtop, const double *wamp,
                       double x0, hipStream_t s) {
  hipLaunchKernelGGL(k_top_forcing, dim3(nblocks(p.nx, 128)), dim3(128), 0, s, p, (const double *)c.vm,
                     xtop, wamp, x0, c.ftop);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int launch_sponge_forcing(const lns3d_params &p, int ispg, const Coeffs &c, const double *x,
                          const double *wamp, double x0, hipStream_t s) {
  const long P = (long)p.nx * p.ny;
  hipLaunchKernelGGL(k_sponge_forcing, dim3(nblocks(P, 256)), dim3(256), 0, s, p, ispg,
                     (const double *)c.vm, x, wamp, x0, c.spf);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// ==========================================================================
// LHS assembly: value of entry (slot k, row r, col cc) at point (i, j)
struct AsmCtx {
  lns3d_params p;
  double d, dinv, dsinv;        // spacing along the line
  double a[5], b[5];            // interior first/second derivative weights x alfa
  int sym;                      // symmetric ends of the xi lines (bit 0 left, bit 1 right)
  int calcd;                    // this factor carries the D terms (else the sponge)
};

// weights of slot k (point pos + k - 2) for column component cc in a row next
// to a symmetric end: interior weights, those of points beyond the end moved
// onto their reflections (-q on the left, 2(N-1) - q on the right), negated
// for component 3 (src/lhsbs1f3d.f90:44-120, 178-270)
__device__ __forceinline__ void sym_weights(const AsmCtx &A, int pos, int N, int k, int cc,
                                            double &aw, double &bw) {
  aw = 0.0;
  bw = 0.0;
  const double s = cc == 2 ? -1.0 : 1.0;
  for (int o = 0; o < 5; o++) {
    int col = pos + o - 2;
    double sg = 1.0;
    if (col < 0) { col = -col; sg = s; }
    else if (col > N - 1) { col = 2 * (N - 1) - col; sg = s; }
    if (col - pos + 2 == k) { aw += sg * A.a[o]; bw += sg * A.b[o]; }
  }
}

// implicit damping weight of slot k on the diagonal of component cc
// (src/lhs1f3d.f90:169-201, lhs2f3d.f90:171-203; closure rows
// lhsbt1f3d.f90:284-304, 451-471, lhsbt2f3d.f90:277-297, 431-451;
// symmetric rows lhsbs1f3d.f90): the fourth difference fb inside and on every
// row of a periodic line, the second difference on the second and
// second-last rows, none on the end rows, fb reflected next to a symmetric end
__device__ __forceinline__ double damp_w(int pos, int N, bool per, int sym, int k, int cc) {
  constexpr double fb[5] = {1.0, -4.0, 6.0, -4.0, 1.0};
  if (per) return fb[k];
  if (sym && (((sym & 1) && pos <= 1) || ((sym & 2) && pos >= N - 2))) {
    const double s = cc == 2 ? -1.0 : 1.0;
    double w = 0.0;
    for (int o = 0; o < 5; o++) {
      int col = pos + o - 2;
      double sg = 1.0;
      if (col < 0) { col = -col; sg = s; }
      else if (col > N - 1) { col = 2 * (N - 1) - col; sg = s; }
      if (col - pos + 2 == k) w += sg * fb[o];
    }
    return w;
  }
  if (pos == 0 || pos == N - 1) return 0.0;
  if (pos == 1 || pos == N - 2) return k == 1 ? -1.0 : k == 2 ? 2.0 : k == 3 ? -1.0 : 0.0;
  return fb[k];
}

__device__ __forceinline__ z lhs_entry(const AsmCtx &A, const Coeffs &c, int dir, int i, int j, int k,
                                       int r, int cc) {
  using namespace st;
  const lns3d_params &p = A.p;
  const long P = (long)p.nx * p.ny;
  const long q = i + (long)p.nx * j;
  const int N = dir == 1 ? p.nx : p.ny;
  const int pos = dir == 1 ? i : j;
  const bool per = dir == 1 ? p.xper : p.yper;
  const double alfa = p.alfa;
  // boundary-condition rows (src/lhsbc1f3d.f90, src/lhsbc2f3d.f90); the
  // left/right rows are applied last, then top, then the wall
  const z one{(k == 2 && r == cc) ? 1.0 : 0.0, 0.0};
  // left/right 0 and 4: identity rows (src/lhsbc1f3d.f90:414-436); left 1
  // keeps the closure row, 7 the symmetric row
  if (!p.xper && ((i == 0 && (p.left == 0 || p.left == 4 || p.left == 5)) ||
                  (i == p.nx - 1 && (p.right == 0 || p.right == 4 || p.right == 5))))
    return one;
  // right = 8 holds the initial condition: -I in the xi factor only
  // (src/lhsbc1f3d.f90:439-446; the eta factor keeps its rows)
  if (!p.xper && dir == 1 && i == p.nx - 1 && p.right == 8) return z{(k == 2 && r == cc) ? -1.0 : 0.0, 0.0};
  if (!p.yper) {
    if (j == p.ny - 1) return one;
    if (j == 0 && !c.bnb && i >= c.wlo && i <= c.whi) {   // the viscous wall rows of is:ie
      if (r >= 1 && r <= 3) return one;
      if (r == 0) {
        if (dir == 1 && (p.wall == 1 || p.wall == 2 || p.wall == 3)) return one;
        if (dir == 2 && p.wall == 1) {
          const double w[5] = {-gc4, -gc5, -gc1, -gc2, -gc3};   // slots 1..5
          return z{cc == 0 ? w[k] : 0.0, 0.0};
        }
        if (dir == 2 && p.wall == 3) {   // third-order extrapolation, src/lhsbc2f3d.f90:41-48
          const double w[5] = {4.0, -1.0, -1.0, 4.0, -6.0};
          return z{cc == 0 ? w[k] : 0.0, 0.0};
        }
      }
      if (r == 4) {
        if (p.wallt == 0 || (dir == 1 && p.wallt == 1)) return one;
        if (dir == 2 && p.wallt == 1) {
          const double w[5] = {-gc4, -gc5, -gc1, -gc2, -gc3};
          return z{cc == 4 ? w[k] : 0.0, 0.0};
        }
      }
    }
  }
  const double dt = c.dtl[q];
  const double X = (dir == 1 ? c.Ah : c.Bh)[(long)(r + 5 * cc) * P + q];
  const double *Vh = dir == 1 ? c.V11 : c.V22;
  const int ab = dir == 1 ? 0 : 2;
  const bool calcd = A.calcd != 0;
  // which viscous / spanwise entry (r, cc) is
  int vi = -1;
  if (r == 1 && cc == 1) vi = 0;
  else if (r == 1 && cc == 2) vi = 4;
  else if (r == 2 && cc == 1) vi = 5;
  else if (r == 2 && cc == 2) vi = 1;
  else if (r == 3 && cc == 3) vi = 2;
  else if (r == 4 && cc == 4) vi = 3;
  int ai = -1;
  if ((r == 1 && cc == 3) || (r == 3 && cc == 1)) ai = 0;
  else if ((r == 2 && cc == 3) || (r == 3 && cc == 2)) ai = 1;
  const double Vv = vi >= 0 ? Vh[(long)vi * P + q] : 0.0;
  const double ABv = ai >= 0 ? c.ABhi[(long)(ab + ai) * P + q] : 0.0;

  const int sym = dir == 1 ? A.sym : 0;
  if (sym && (((sym & 1) && pos <= 1) || ((sym & 2) && pos >= N - 2))) {
    // rows next to a symmetric end (src/lhsbs1f3d.f90, xi factor): the
    // interior row with the weight of each point beyond the end added to
    // the point it reflects onto, negated for column component 3; the D
    // terms times rcalcd, the sponge without them (:154-170); the spanwise
    // cross entries are read through that routine's Ahi(6,ny,nx) shape
    // (point i + ny j; check_params keeps it on the grid)
    double aw, bw;
    sym_weights(A, pos, N, k, cc, aw, bw);
    const double ABs = ai >= 0 ? c.ABhi[(long)(ab + ai) * P + pos + (long)p.ny * j] : 0.0;
    double re = dt * aw * X, im = 0.0;
    if (k == 2) {
      const double rc = calcd ? 1.0 : 0.0;
      re = re + rc * alfa * dt * c.Dh[(long)(r + 5 * cc) * P + q];
      im = rc * alfa * dt * c.Dhi[(long)(r + 5 * cc) * P + q];
    }
    if (ai >= 0) im = im - aw * dt * ABs;
    if (vi >= 0) re = re - dt * bw * Vv;
    if (k == 2 && r == cc) re = re + 1.0;
    if (!calcd && k == 2 && r == cc && c.spl) re = re + alfa * dt * c.spl[q];
    if (c.sbuf && r == cc) re = re + alfa * dt * c.sbuf[q] * damp_w(pos, N, per, sym, k, cc);
    return z{re, im};
  }
  if (!per && (pos <= 1 || pos >= N - 2)) {
    // boundary closure rows (src/lhsbt1f3d.f90, lhsbt2f3d.f90): first- and
    // second-derivative weights of the column held in slot k
    double F, S;
    bool diag = k == 2;
    if (pos == 0) {
      const double Fs[5] = {gc4, gc5, gc1, gc2, gc3}, Ss[5] = {dd4, dd5, dd1, dd2, dd3};
      F = Fs[k]; S = Ss[k];
    } else if (pos == 1) {
      const double Fs[5] = {gb5, gb1, gb2, gb3, gb4}, Ss[5] = {db5, db1, db2, db3, db4};
      F = Fs[k]; S = Ss[k];
    } else if (pos == N - 2) {
      const double Fs[5] = {-gb4, -gb3, -gb2, -gb1, -gb5}, Ss[5] = {db4, db3, db2, db1, db5};
      F = Fs[k]; S = Ss[k];
    } else {
      const double Fs[5] = {-gc3, -gc2, -gc1, -gc5, -gc4}, Ss[5] = {dd3, dd2, dd1, dd5, dd4};
      F = Fs[k]; S = Ss[k];
    }
    double af = alfa * F * A.dinv;
    double re, im = 0.0;
    if (diag) {
      double rc = calcd ? 1.0 : 0.0;
      double D = c.Dh[(long)(r + 5 * cc) * P + q], Di = c.Dhi[(long)(r + 5 * cc) * P + q];
      re = dt * (af * X + rc * alfa * D);
      im = dt * (rc * alfa * Di);
    } else {
      re = af * dt * X;
    }
    if (ai >= 0) im = im - af * dt * ABv;
    if (vi >= 0) re = re - alfa * S * A.dsinv * dt * Vv;
    if (diag && r == cc) re = re + 1.0;
    if (!calcd && diag && r == cc && c.spl) re = re + alfa * dt * c.spl[q];   // lhsbt2f3d.f90:156-174
    if (c.sbuf && r == cc) re = re + alfa * dt * c.sbuf[q] * damp_w(pos, N, per, 0, k, cc);
    return z{re, im};
  }
  // interior (src/lhs1f3d.f90:59-230)
  double re = A.a[k] * dt * X, im = 0.0;
  if (ai >= 0) im = im - A.a[k] * dt * ABv;
  if (calcd && k == 2) {
    re = re + alfa * dt * c.Dh[(long)(r + 5 * cc) * P + q];
    im = im + alfa * dt * c.Dhi[(long)(r + 5 * cc) * P + q];
  }
  if (vi >= 0) re = re - A.b[k] * dt * Vv;
  if (k == 2 && r == cc) re = re + 1.0;
  if (!calcd && k == 2 && r == cc && c.spl) re = re + alfa * dt * c.spl[q];  // lhs2f3d.f90:207-225
  if (c.sbuf && r == cc) re = re + alfa * dt * c.sbuf[q] * damp_w(pos, N, per, 0, k, cc);
  return z{re, im};
}

// The inviscid wall rows (Navier = .false.; src/lhsbc1f3d.f90:55-393,
// src/lhsbc2f3d.f90:75-120) on rows i of is:ie: the blocks of the wall
// row's unknowns rotated to the body normal coordinates -- in the xi factor
// every block of the wall line, with the normal of the block's column (the
// non-periodic slot map, as the reference indexes it); in the eta factor the
// blocks of column j = 0 in rows 0, 1, 2 -- then, on the wall row, the u row
// rotated and the v row (normal velocity) held at zero.  Elsewhere lhs_entry.
__device__ __forceinline__ z lhs_entry_bc(const AsmCtx &A, const Coeffs &c, int dir, int i, int j, int k,
                                          int r, int cc) {
  const lns3d_params &p = A.p;
  const int nx = p.nx;
  const bool lr = !p.xper && ((i == 0 && (p.left == 0 || p.left == 4 || p.left == 5)) ||
                              (i == nx - 1 && (p.right == 0 || p.right == 4 || p.right == 5 ||
                                               (p.right == 8 && dir == 1))));
  const bool zone = c.bnb && !p.yper && !lr && i >= c.wlo && i <= c.whi &&
                    ((dir == 1 && j == 0) || (dir == 2 && j <= 2));
  if (!zone) return lhs_entry(A, c, dir, i, j, k, r, cc);
  if (j == 0 && r == 2) return z{(k == 2 && cc == 2) ? 1.0 : 0.0, 0.0};
  int col = -1;                                   // the wall point whose normal rotates block k
  if (dir == 1) {
    const int t0[5] = {3, 4, 0, 1, 2}, t1[5] = {4, 0, 1, 2, 3};
    const int tm[5] = {nx - 4, nx - 3, nx - 2, nx - 1, nx - 5}, tn[5] = {nx - 3, nx - 2, nx - 1, nx - 5, nx - 4};
    col = i == 0 ? t0[k] : i == 1 ? t1[k] : i == nx - 2 ? tm[k] : i == nx - 1 ? tn[k] : i - 2 + k;
  } else if (k == 2 - j) {
    col = i;
  }
  auto cr = [&](int rr) -> z {                    // the column-rotated entry of row rr
    if (col < 0 || (cc != 1 && cc != 2)) return lhs_entry(A, c, dir, i, j, k, rr, cc);
    const double b1 = c.bnb[col], b2 = c.bnb[nx + col];
    const z u = lhs_entry(A, c, dir, i, j, k, rr, 1), w = lhs_entry(A, c, dir, i, j, k, rr, 2);
    return cc == 1 ? sc(b2, u) - sc(b1, w) : sc(b1, u) + sc(b2, w);
  };
  if (j == 0 && r == 1) {
    const double b1 = c.bnb[i], b2 = c.bnb[nx + i];
    return sc(b2, cr(1)) - sc(b1, cr(2));
  }
  return cr(r);
}

static AsmCtx make_asm(const lns3d_params &p, const Coeffs &c, int dir) {
  using namespace st;
  AsmCtx A;
  A.p = p;
  A.calcd = dir == 1 ? c.calcd_xi : !c.calcd_xi;
  A.d = dir == 1 ? 1.0 / (double)(p.nx - 1) : 1.0 / (double)(p.ny - 1);
  A.dinv = 1.0 / A.d;
  A.dsinv = 1.0 / (A.d * A.d);
  A.sym = (dir == 1 && !p.xper) ? ((p.left == 7 ? 1 : 0) | (p.right == 7 ? 2 : 0)) : 0;
  double alfa = p.alfa, d = A.d;
  double ga[5] = {ga1, ga2, 0.0, ga3, ga4}, da[5] = {da1, da2, da3, da4, da5};
  for (int k = 0; k < 5; k++) {
    A.a[k] = alfa * ga[k] / d;
    A.b[k] = alfa * da[k] / (d * d);
  }
  return A;
}

// into the working copy of the line solver: line-interleaved, element
// ((row*5 + k)*25 + r*5 + cc)*L + line (periodic lines), or line-major,
// element ((line*n + row)*5 + k)*25 + r*5 + cc
__global__ __launch_bounds__(128) void k_assemble_lines(AsmCtx A, Coeffs c, int dir, int line_major, z *lu) {
  const int nx = A.p.nx, ny = A.p.ny;
  const long L = dir == 1 ? ny : nx, n = dir == 1 ? nx : ny;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L * n) return;
  long line = t % L, row = t / L;
  int i = dir == 1 ? (int)row : (int)line, j = dir == 1 ? (int)line : (int)row;
  for (int k = 0; k < 5; k++)
    for (int r = 0; r < 5; r++)
      for (int cc = 0; cc < 5; cc++) {
        const long e = line_major ? ((line * n + row) * 5 + k) * 25 + r * 5 + cc
                                  : ((row * 5 + k) * 25 + r * 5 + cc) * L + line;
        lu[e] = lhs_entry_bc(A, c, dir, i, j, k, r, cc);
      }
}

__global__ __launch_bounds__(128) void k_assemble_ref(AsmCtx A, Coeffs c, int dir, double2 *mat) {
  const int nx = A.p.nx, ny = A.p.ny;
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)nx * ny) return;
  int i = q % nx, j = q / nx;
  for (int k = 0; k < 5; k++)
    for (int r = 0; r < 5; r++)
      for (int cc = 0; cc < 5; cc++) {
        z v = lhs_entry_bc(A, c, dir, i, j, k, r, cc);
        mat[q * 125 + k + 5 * r + 25 * cc] = make_double2(v.re, v.im);
      }
}

int launch_assemble(const lns3d_params &p, const Coeffs &c, int dir, PentaPlan &plan, hipStream_t s) {
  AsmCtx A = make_asm(p, c, dir);
  long tot = (long)p.nx * p.ny;
  hipLaunchKernelGGL(k_assemble_lines, dim3(nblocks(tot, 128)), dim3(128), 0, s, A, c, dir,
                     plan.periodic ? 0 : 1, (z *)plan.lu);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int launch_assemble_ref(const lns3d_params &p, const Coeffs &c, int dir, void *mat, hipStream_t s) {
  AsmCtx A = make_asm(p, c, dir);
  long tot = (long)p.nx * p.ny;
  hipLaunchKernelGGL(k_assemble_ref, dim3(nblocks(tot, 128)), dim3(128), 0, s, A, c, dir,
                     (double2 *)mat);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// ==========================================================================
// RHS: c1 = d/dxi v first (stored), then the point operator
// one lane per (point, component), so a wave's stencil loads are contiguous
__device__ __forceinline__ void rhs_d1x(long bid, int nx, int ny, int xper, const Stencil &st, const double2 *v,
                                        double2 *c1) {
  long t = bid * blockDim.x + threadIdx.x;
  if (t >= 5L * nx * ny) return;
  const long q = t / 5;
  const int e = (int)(t - 5 * q);
  const int i = q % nx, j = q / nx;
  z g = d1<double2, z>(nx, xper, i, st.gx, st.dxinv, v + 5L * nx * j + e, 5, st.sym, e == 2 ? -1.0 : 1.0);
  c1[t] = make_double2(g.re, g.im);
}
__global__ void k_rhs_d1x(int nx, int ny, int xper, Stencil st, const double2 *v, double2 *c1) {
  rhs_d1x(xcd_tile(blockIdx.x, gridDim.x), nx, ny, xper, st, v, c1);
}

// The point operator works on tiles of RT consecutive grid points with one
// workgroup of five waves.  Phase 1 stages the tile's U and U_xi and its 17
// stencil derivatives (U_eta; U_xixi, U_xieta, U_etaeta of components 1-4) in
// LDS; phase 2 has wave m form row m of the operator for the tile, so each
// wave streams one row of every coefficient matrix ([25][P] layout, one
// contiguous run of RT values per load).  Terms are accumulated in the order
// of src/lrhs3d.f90:45-220 (U,xi and U,eta; U; U,xixi; U,xieta; U,etaeta;
// A_i/B_i; D_i), then rbe3D and rhsBC3D (impdrv3d.f90:444-470,
// src/rhsbc3d.f90:34-160, 200-215).
constexpr int RT = 64;   // the interior path's task split (320 threads, 512 tasks) assumes 64
static_assert(RT == 64, "k_rhs_point_t's interior path assumes tiles of 64 points");
constexpr int NDER = 17;

struct StepBatchItem;
__device__ __forceinline__ int step_member(const int *off, int count);
__device__ __forceinline__ void batch_point_args(const StepBatchItem *items, const int *off, int count,
                                                 lns3d_params &p, Stencil &st, Coeffs &c, const double2 *&v,
                                                 const double2 *&c1a, double2 *&r, long &tile);

// BATCH: the kernel of a staged batch (the member's deck, coefficients and
// buffers replace the arguments, from items; its workgroups are numbered
// from off[member]); otherwise one context with its tiles numbered per XCD.
// The single-context kernel keeps the arguments as kernel arguments: its
// body moved into a device function was measured 11% slower (78 to 90
// VGPRs).
template <bool BATCH>
__global__ __launch_bounds__(320) void k_rhs_point_t(lns3d_params p, Stencil st, Coeffs c, const double2 *v,
                                                    const double2 *c1a, double2 *r, int with_rbe,
                                                    const double2 *v2old, const double2 *vold,
                                                    const double2 *forcing, const StepBatchItem *items,
                                                    const int *off, int count) {
  using namespace st;
  long tile;
  if constexpr (BATCH) batch_point_args(items, off, count, p, st, c, v, c1a, r, tile);
  else tile = xcd_tile(blockIdx.x, gridDim.x);   // tiles of one XCD: consecutive rows
  __shared__ double2 sv[5 * RT], sg1[5 * RT], sd[NDER * RT], sr[5 * RT];
  const int nx = p.nx, ny = p.ny;
  const long P = (long)nx * ny;
  const long q0 = tile * RT;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int npt = (int)min((long)RT, P - q0);
  if (tid < 5 * npt) {
    sv[tid] = v[5 * q0 + tid];
    sg1[tid] = c1a[5 * q0 + tid];
  }
  // stencil derivatives into LDS, one lane per (point, component): slots
  // 0-4 U_eta, 5-8 U_xixi, 9-12 U_xieta, 13-16 U_etaeta (components 1-4).
  // A tile within one row whose points all lie inside both directions'
  // closures (no periodic direction) takes the interior stencils without
  // branches: U_eta and U_etaeta from one 7-row gather of v, then U_xixi or
  // U_xieta, then U_xieta again (threads below 192), each summed in the
  // interior formula's order (the bits of d1, d2).  Other tiles loop over
  // the tasks, whose four kinds span multiples of 64, so each wave works on
  // one kind at a time.
  // (the staged batch's kernel keeps the task loop: the interior path raises
  // it from 78 to 97 VGPRs, six to four waves per SIMD, and its small grids
  // have few such tiles)
  bool inner = false;
  if (!BATCH && !p.xper && !p.yper && npt == RT) {
    const int i0 = (int)(q0 % nx), j0 = (int)(q0 / nx);
    const int i1 = (int)((q0 + RT - 1) % nx), j1 = (int)((q0 + RT - 1) / nx);
    inner = j0 == j1 && j0 >= 3 && j0 <= ny - 4 && i0 >= 3 && i1 <= nx - 4;
  }
  if (inner) {
    const long es = 5L * nx;
    {
      const int pt = tid / 5, e = tid - 5 * pt;
      const double2 *b = v + 5 * (q0 + pt) + e;
      z x[7];
#pragma unroll
      for (int k = 0; k < 7; k++) x[k] = ld(b + (k - 3) * es);
      const z h1 = sc(st.gy[0], x[0]) + sc(st.gy[1], x[1]) + sc(st.gy[2], x[2]) + sc(st.gy[3], x[4]) +
                   sc(st.gy[4], x[5]) + sc(st.gy[5], x[6]);
      const z h3 = sc(st.dy[0], x[0]) + sc(st.dy[1], x[1]) + sc(st.dy[2], x[2]) + sc(st.dy[3], x[3]) +
                   sc(st.dy[4], x[4]) + sc(st.dy[5], x[5]) + sc(st.dy[6], x[6]);
      sd[e * RT + pt] = make_double2(h1.re, h1.im);
      if (e >= 1) sd[(12 + e) * RT + pt] = make_double2(h3.re, h3.im);
    }
#pragma unroll
    for (int sl = 0; sl < 2; sl++) {
      // task u of 512: U_xixi (u < 256), then U_xieta; components 1-4
      const int u0 = tid + 5 * 64 * sl, u = min(u0, 511);
      const bool xx = u < 256;
      const int w4 = u & 255, pt = w4 >> 2, e = 1 + (w4 & 3);
      const double2 *b = (xx ? v : c1a) + 5 * (q0 + pt) + e;
      const long s1 = xx ? 5L : es;
      z x[7];
#pragma unroll
      for (int k = 0; k < 7; k++) x[k] = ld(b + (k - 3) * s1);
      const double *wv = xx ? st.dx : st.gy;
      z h = sc(wv[0], x[0]) + sc(wv[1], x[1]) + sc(wv[2], x[2]);
      if (xx) h = h + sc(wv[3], x[3]);
      const int o = xx ? 4 : 3;
      h = h + sc(wv[o], x[4]) + sc(wv[o + 1], x[5]) + sc(wv[o + 2], x[6]);
      if (u0 < 512) sd[((xx ? 4 : 8) + e) * RT + pt] = make_double2(h.re, h.im);
    }
  } else
  for (int tsk = tid; tsk < 5 * RT + 3 * 4 * RT; tsk += 5 * 64) {
    int kind, pt, e;
    if (tsk < 5 * RT) { kind = 0; pt = tsk / 5; e = tsk - 5 * pt; }
    else { const int u = tsk - 5 * RT; kind = 1 + u / (4 * RT); const int w4 = u % (4 * RT); pt = w4 / 4; e = 1 + w4 % 4; }
    if (pt >= npt) continue;
    const long qq = q0 + pt;
    const int ii = (int)(qq % nx), jj = (int)(qq / nx);
    z h;
    if (kind == 0) {
      h = d1<double2, z>(ny, p.yper, jj, st.gy, st.dyinv, v + 5L * ii + e, 5L * nx);
      if (e == 4 && !c.bnb && (p.wallt == 1 || p.wallt == 2) && jj == 0) h = z{0, 0};   // src/cgradbc.f90:15 (Navier)
    } else if (kind == 1) {
      h = d2<double2, z>(nx, p.xper, ii, st.dx, st.dxsinv, v + 5L * nx * jj + e, 5, st.sym,
                         e == 2 ? -1.0 : 1.0);
    } else if (kind == 2) {
      h = d1<double2, z>(ny, p.yper, jj, st.gy, st.dyinv, c1a + 5L * ii + e, 5L * nx);
    } else {
      h = d2<double2, z>(ny, p.yper, jj, st.dy, st.dysinv, v + 5L * ii + e, 5L * nx);
    }
    const int sl = kind == 0 ? e : 4 * kind + e;
    sd[sl * RT + pt] = make_double2(h.re, h.im);
  }
  const long q = q0 + l;
  const int i = (int)(q % nx), j = (int)(q / nx);
  __syncthreads();
  if (l < npt) {
    const int m = w;
    auto S = [](double2 a) { return z{a.x, a.y}; };
    auto g1 = [&](int e) { return S(sg1[5 * l + e]); };
    auto g2 = [&](int e) { return S(sd[e * RT + l]); };
    auto vq = [&](int e) { return S(sv[5 * l + e]); };
    z a{0, 0};
#pragma unroll
    for (int e = 0; e < 5; e++) a = a + sc(c.Ah[(long)(m + 5 * e) * P + q], g1(e));
#pragma unroll
    for (int e = 0; e < 5; e++) a = a + sc(c.Bh[(long)(m + 5 * e) * P + q], g2(e));
#pragma unroll
    for (int e = 0; e < 5; e++) a = a + sc(c.Dh[(long)(m + 5 * e) * P + q], vq(e));
    if (m > 0) {
      // viscous terms: slots 4+e (xixi), 8+e (xieta), 12+e (etaeta)
      const double *Ws[3] = {c.V11, c.V12, c.V22};
#pragma unroll
      for (int tt = 0; tt < 3; tt++) {
        const double *W = Ws[tt];
        auto h = [&](int e) { return S(sd[(4 * tt + 4 + e) * RT + l]); };
        if (m == 1) a = a - sc(W[q], h(1)) - sc(W[4 * P + q], h(2));
        else if (m == 2) a = a - sc(W[5 * P + q], h(1)) - sc(W[P + q], h(2));
        else if (m == 3) a = a - sc(W[2 * P + q], h(3));
        else a = a - sc(W[3 * P + q], h(4));
      }
    }
    if (m >= 1 && m <= 3) {
      const double ab0 = c.ABhi[q], ab1 = c.ABhi[P + q], ab2 = c.ABhi[2 * P + q], ab3 = c.ABhi[3 * P + q];
      z ti;
      if (m == 1) ti = sc(ab0, g1(3)) + sc(ab2, g2(3));
      else if (m == 2) ti = sc(ab1, g1(3)) + sc(ab3, g2(3));
      else ti = sc(ab0, g1(1)) + sc(ab1, g1(2)) + sc(ab2, g2(1)) + sc(ab3, g2(2));
      a = a + z{ti.im, -ti.re};                       // -i * ti
    }
    {
      z b{0, 0};
#pragma unroll
      for (int e = 0; e < 5; e++) b = b + sc(c.Dhi[(long)(m + 5 * e) * P + q], vq(e));
      a = a + z{-b.im, b.re};                         // + i * b
    }
    // sponge (src/lrhs3d.f90:222-282): sp1 v, then the inflow part sp2 (v - f)
    if (c.sp1) a = a + sc(c.sp1[q], vq(m));
    if (c.sp2) {
      const z f{c.spf[2 * ((long)m * P + q)], c.spf[2 * ((long)m * P + q) + 1]};
      a = a + sc(c.sp2[q], vq(m) - f);
    }
    // smoother (src/lrhs3d.f90:293, src/smoother3d.f90): eta, then xi
    if (c.sbuf) {
      const double eb = c.sbuf[q];
      z s;
      if (smooth_at(ny, j, 0, 1.0, v + 5L * i + m, 5L * nx, s)) a = a + sc(eb, s);
      if (smooth_at(nx, i, st.sym, m == 2 ? -1.0 : 1.0, v + 5L * nx * j + m, 5, s)) a = a + sc(eb, s);
    }
    // backward Euler: r = vold - v - dtl * r with vold = v (the field already
    // satisfies the boundary conditions when the step starts)
    if (with_rbe) {
      const double dt = c.dtl[q];
      constexpr double onept33 = 1.3333333333333333333, pt33 = 3.3333333333333333333e-1;
      if (with_rbe == 2) {
        // two-step backward (src/impdrv3d.f90:473-501, rtb3D) with vold = v:
        // r = 4/3 vold - 1/3 v2old - v - 2/3 dtl r
        const z vv = vq(m), vo = ld(v2old + 5 * q + m);
        a = sc(onept33, vv) - sc(pt33, vo) - vv - sc(pt66 * dt, a);
      } else if (with_rbe == 3) {
        // backward Euler with vold given (later corrector iterations):
        // r = vold - v - dtl r (src/impdrv3d.f90:444-470)
        const z vv = vq(m), vb = ld(vold + 5 * q + m);
        a = vb - vv - sc(dt, a);
      } else if (with_rbe == 4) {
        const z vv = vq(m), vb = ld(vold + 5 * q + m), vo = ld(v2old + 5 * q + m);
        a = sc(onept33, vb) - sc(pt33, vo) - vv - sc(pt66 * dt, a);
      } else if (with_rbe == 5) {
        // forced backward Euler of the shift-invert driver (ebe3D,
        // src/si_eigdrv3d.f90:333-342): r = vold - v - dtl (r - f)
        const z vv = vq(m), vb = ld(vold + 5 * q + m), f = ld(forcing + 5 * q + m);
        a = vb - vv - sc(dt, a - f);
      } else {
        a = z{-(dt * a.re), -(dt * a.im)};
      }
      // boundary rows of the residual
      if (p.yper) {
        if (j == ny - 1) a = z{0, 0};
      } else {
        if (j == 0 && !c.bnb && i >= c.wlo && i <= c.whi) {
          // wall = 3: the bump's wall values u1w, u2w, u3w, tw, twp (src/rhsbc3d.f90:47-100),
          // whole arrays against the is:ie section: position i takes element i - is
          const int ib = i - c.wlo;
          auto W = [&](int k) { return z{c.fwall[2 * (k * nx + ib)], c.fwall[2 * (k * nx + ib) + 1]}; };
          if (m >= 1 && m <= 3) {
            const z wm = p.wall == 3 ? W(m - 1) : z{0, 0};
            const z d = vq(m) - wm;
            a = z{-d.re, -d.im};
          }
          if (m == 0 && p.wall == 3)
            a = ld(v + 5 * q) - (sc(4.0, ld(v + 5 * (q + nx))) - sc(6.0, ld(v + 5 * (q + 2L * nx))) +
                                 sc(4.0, ld(v + 5 * (q + 3L * nx))) - ld(v + 5 * (q + 4L * nx)));
          if (m == 0 && p.wall == 1)
            a = sc(gc1, ld(v + 5 * q)) + sc(gc2, ld(v + 5 * (q + nx))) + sc(gc3, ld(v + 5 * (q + 2L * nx))) +
                sc(gc4, ld(v + 5 * (q + 3L * nx))) + sc(gc5, ld(v + 5 * (q + 4L * nx)));
          if (m == 0 && p.wall == 2)
            a = ld(v + 5 * q) - sc(2.0, ld(v + 5 * (q + nx))) + ld(v + 5 * (q + 2L * nx));
          if (m == 4 && p.wallt == 0) {
            if (p.wall == 3) {
              const z d = vq(4) - W(3);
              a = z{-d.re, -d.im};
            } else {
              a = z{0, 0};
            }
          }
          if (m == 4 && p.wallt == 1) {
            a = sc(gc1, ld(v + 5 * q + 4)) + sc(gc2, ld(v + 5 * (q + nx) + 4)) +
                sc(gc3, ld(v + 5 * (q + 2L * nx) + 4)) + sc(gc4, ld(v + 5 * (q + 3L * nx) + 4)) +
                sc(gc5, ld(v + 5 * (q + 4L * nx) + 4));
            if (p.wall == 3) a = a - W(4);
          }
        }
        if (j == ny - 1) {
          if (p.top == 1) {        // src/rhsbc3d.f90:152-157: -(v - wave)
            const z f{c.ftop[2 * (m * nx + i)], c.ftop[2 * (m * nx + i) + 1]};
            const z d = vq(m) - f;
            a = z{-d.re, -d.im};
          } else {
            a = z{0, 0};
          }
        }
      }
      if (p.xper) {
        if (i == nx - 1) a = z{0, 0};
      } else {
        // src/rhsbc3d.f90:172-300: left 0 and right 0, 4 zero; left 1, 4
        // -(v - the held value); left/right 7 the operator's row
        if (i == 0 && (p.left == 1 || p.left == 4 || p.left == 5)) {
          const z f{c.fside[2 * (m * ny + j)], c.fside[2 * (m * ny + j) + 1]};
          const z d = vq(m) - f;
          a = z{-d.re, -d.im};
        }
        if (i == nx - 1 && p.right == 5) {
          const z f{c.fside[2 * ((5 + m) * ny + j)], c.fside[2 * ((5 + m) * ny + j) + 1]};
          const z d = vq(m) - f;
          a = z{-d.re, -d.im};
        }
        if (i == nx - 1 && p.right == 9) {      // src/rhsbc3d.f90:317-320
          const z d = vq(m) - extrap_log(ld(v + 5 * (q - 1) + m), ld(v + 5 * (q - 2) + m));
          a = z{-d.re, -d.im};
        }
        if ((i == 0 && p.left == 0) || (i == nx - 1 && (p.right == 0 || p.right == 4 || p.right == 8)))
          a = z{0, 0};
      }
    }
    sr[5 * l + m] = make_double2(a.re, a.im);
  }
  __syncthreads();
  if (tid < 5 * npt) {
    double2 o = sr[tid];
    // the inviscid wall rows (src/rhsbc3d.f90:108-118) on is:ie, rows 2 and 3
    // of the residual: the tangential momentum row rotated, the normal
    // velocity held at zero; the left and right columns keep their rows
    const int pt = tid / 5, m = tid - 5 * pt;
    const long qq = q0 + pt;
    const int ii = (int)(qq % nx), jj = (int)(qq / nx);
    if (c.bnb && with_rbe && !p.yper && jj == 0 && (m == 1 || m == 2) && ii >= c.wlo && ii <= c.whi) {
      const bool lr = p.xper ? ii == nx - 1
                             : ((ii == 0 && (p.left == 0 || p.left == 1 || p.left == 4 || p.left == 5)) ||
                                (ii == nx - 1 && (p.right == 0 || p.right == 4 || p.right == 5 || p.right == 8 ||
                                                  p.right == 9)));
      if (!lr) {
        const double b1 = c.bnb[ii], b2 = c.bnb[nx + ii];
        if (m == 1) {
          const z t = sc(b2, z{sr[5 * pt + 1].x, sr[5 * pt + 1].y}) - sc(b1, z{sr[5 * pt + 2].x, sr[5 * pt + 2].y});
          o = make_double2(t.re, t.im);
        } else {
          const z t = sc(b1, z{sv[5 * pt + 1].x, sv[5 * pt + 1].y}) + sc(b2, z{sv[5 * pt + 2].x, sv[5 * pt + 2].y});
          o = make_double2(-t.re, -t.im);
        }
      }
    }
    r[5 * q0 + tid] = o;
  }
}

int launch_rhs(const lns3d_params &p, const Stencil &st, const Coeffs &c, const void *v, void *r,
               void *work, int with_rbe, hipStream_t s, const void *v2old, const void *vold,
               const void *forcing) {
  long P = (long)p.nx * p.ny;
  hipLaunchKernelGGL(k_rhs_d1x, dim3(nblocks(5 * P, 256)), dim3(256), 0, s, p.nx, p.ny, p.xper, st,
                     (const double2 *)v, (double2 *)work);
  hipLaunchKernelGGL(k_rhs_point_t<false>, dim3(nblocks(P, RT)), dim3(5 * 64), 0, s, p, st, c, (const double2 *)v,
                     (const double2 *)work, (double2 *)r, with_rbe, (const double2 *)v2old,
                     (const double2 *)vold, (const double2 *)forcing, nullptr, nullptr, 0);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// ==========================================================================
// update and boundary conditions on v (src/impdrv3d.f90:340-356, itrbc3d.f90)
__global__ void k_axpy(double2 *x, const double2 *d, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double2 a = x[t], b = d[t];
  x[t] = make_double2(a.x + b.x, a.y + b.y);
}

__device__ __forceinline__ void itrbc_wall_top(long bid, const lns3d_params &p, const double *ftop,
                                               const double *fwall, int wlo, int whi, const double *bnb,
                                               double2 *v) {
  using namespace st;
  int i = (int)(bid * blockDim.x + threadIdx.x);
  const int nx = p.nx, ny = p.ny;
  if (i >= nx) return;
  auto V = [&](int d, int jj) -> double2 & { return v[5 * (i + (long)nx * jj) + d]; };
  if (p.yper) {
    for (int d = 0; d < 5; d++) V(d, ny - 1) = V(d, 0);
    return;
  }
  const bool on = i >= wlo && i <= whi;           // the wall extent is:ie
  if (bnb) {
    // inviscid (src/itrbc3d.f90:120-132): the normal velocity zeroed, the
    // tangential one kept, on is:ie
    if (on) {
      const double b1 = bnb[i], b2 = bnb[nx + i];
      const z ub = sc(b2, ld(&V(1, 0))) - sc(b1, ld(&V(2, 0)));
      const z u = sc(b2, ub), w = sc(-b1, ub);
      V(1, 0) = make_double2(u.re, u.im);
      V(2, 0) = make_double2(w.re, w.im);
    }
  } else {
    // the wall value of component d (wall = 3: the bump's, src/itrbc3d.f90:48-105):
    // its velocities on the whole row, the rest on is:ie, where the whole
    // arrays tw, twp meet the is:ie section (position i takes element i - is)
    auto W = [&](int d, int at) { return z{fwall[2 * (d * nx + at)], fwall[2 * (d * nx + at) + 1]}; };
    if (p.wall == 3)
      for (int d = 1; d <= 3; d++) {
        const z w = W(d - 1, i);
        V(d, 0) = make_double2(w.re, w.im);
      }
    if (on) {
      const int ib = i - wlo;
      if (p.wall != 3)
        for (int d = 1; d <= 3; d++) V(d, 0) = make_double2(0, 0);
      auto lin = [&](int d, double a, double b, double c2, double e) {
        z s = sc(a, ld(&V(d, 1))) + sc(b, ld(&V(d, 2))) + sc(c2, ld(&V(d, 3))) + sc(e, ld(&V(d, 4)));
        return s;
      };
      if (p.wall == 1) {
        z s = lin(0, gc2, gc3, gc4, gc5);
        V(0, 0) = make_double2(-s.re / gc1, -s.im / gc1);
      }
      if (p.wall == 2) {
        z a = ld(&V(0, 1)), b = ld(&V(0, 2));
        V(0, 0) = make_double2(2.0 * a.re - b.re, 2.0 * a.im - b.im);
      }
      if (p.wall == 3) {      // third-order extrapolation of the density
        z s = sc(4.0, ld(&V(0, 1))) - sc(6.0, ld(&V(0, 2))) + sc(4.0, ld(&V(0, 3))) - ld(&V(0, 4));
        V(0, 0) = make_double2(s.re, s.im);
      }
      if (p.wallt == 0) {
        const z w = p.wall == 3 ? W(3, ib) : z{0, 0};
        V(4, 0) = make_double2(w.re, w.im);
      }
      if (p.wallt == 1) {
        z s = lin(4, gc2, gc3, gc4, gc5);
        if (p.wall == 3) {
          const z t = W(4, ib) - s;
          V(4, 0) = make_double2(t.re / gc1, t.im / gc1);
        } else {
          V(4, 0) = make_double2(-s.re / gc1, -s.im / gc1);
        }
      }
    }
  }
  if (p.top == 1)        // src/itrbc3d.f90:187-191: the incoming wave
    for (int d = 0; d < 5; d++) V(d, ny - 1) = make_double2(ftop[2 * (d * nx + i)], ftop[2 * (d * nx + i) + 1]);
  else
    for (int d = 0; d < 5; d++) V(d, ny - 1) = make_double2(0, 0);
}
__global__ void k_itrbc_wall_top(lns3d_params p, const double *ftop, const double *fwall, int wlo, int whi,
                                 const double *bnb, double2 *v) {
  itrbc_wall_top(blockIdx.x, p, ftop, fwall, wlo, whi, bnb, v);
}

// bcfix (src/impdrv3d.f90:333-336, src/bcfix.f90): the wall row of the
// increment back from the body normal coordinates, on every wall point
__device__ __forceinline__ void bcfix_point(int i, int nx, const double *bnb, double2 *r) {
  const double b1 = bnb[i], b2 = bnb[nx + i];
  const z a = ld(r + 5 * i + 1), b = ld(r + 5 * i + 2);
  const z t = sc(b2, a) + sc(b1, b), u = sc(-b1, a) + sc(b2, b);
  r[5 * i + 1] = make_double2(t.re, t.im);
  r[5 * i + 2] = make_double2(u.re, u.im);
}
__global__ void k_bcfix(int nx, const double *bnb, double2 *r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  bcfix_point(i, nx, bnb, r);
}

// the wall normals (src/gengrid.f90:133-137) from the metric's n1, n2 on j = 0
__global__ void k_wall_normals(int nx, long P, const double *met, double *bnb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const double n1 = met[2 * P + i], n2 = met[3 * P + i];
  const double s = sqrt(n1 * n1 + n2 * n2);
  bnb[i] = -n1 / s;
  bnb[nx + i] = -n2 / s;
}

int launch_wall_normals(const lns3d_params &p, const Coeffs &c, double *bnb, hipStream_t s) {
  hipLaunchKernelGGL(k_wall_normals, dim3(nblocks(p.nx, 128)), dim3(128), 0, s, p.nx, (long)p.nx * p.ny,
                     (const double *)c.metric, bnb);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// left and right columns (src/itrbc3d.f90:200-410): periodic copy; 0 zero;
// left 1 and 4, right 4 the values in fside; right 9 the extrapolation from
// the two columns inside; 7 (symmetry) left alone
__device__ __forceinline__ void itrbc_lr(long bid, const lns3d_params &p, const double *fside, double2 *v) {
  int j = (int)(bid * blockDim.x + threadIdx.x);
  const int nx = p.nx, ny = p.ny;
  if (j >= ny) return;
  double2 *row = v + 5L * nx * j;
  for (int d = 0; d < 5; d++) {
    if (p.xper) {
      row[5L * (nx - 1) + d] = row[d];
    } else {
      if (p.left == 0) row[d] = make_double2(0, 0);
      if (p.left == 1 || p.left == 4 || p.left == 5)
        row[d] = make_double2(fside[2 * (d * ny + j)], fside[2 * (d * ny + j) + 1]);
      if (p.right == 0) row[5L * (nx - 1) + d] = make_double2(0, 0);
      if (p.right == 4 || p.right == 5)
        row[5L * (nx - 1) + d] =
            make_double2(fside[2 * ((5 + d) * ny + j)], fside[2 * ((5 + d) * ny + j) + 1]);
      if (p.right == 9) {
        const z e = extrap_log(ld(row + 5L * (nx - 2) + d), ld(row + 5L * (nx - 3) + d));
        row[5L * (nx - 1) + d] = make_double2(e.re, e.im);
      }
    }
  }
}
__global__ void k_itrbc_lr(lns3d_params p, const double *fside, double2 *v) {
  itrbc_lr(blockIdx.x, p, fside, v);
}

int launch_itrbc(const lns3d_params &p, const Coeffs &c, void *v, hipStream_t s) {
  hipLaunchKernelGGL(k_itrbc_wall_top, dim3(nblocks(p.nx, 128)), dim3(128), 0, s, p, (const double *)c.ftop,
                     (const double *)c.fwall, c.wlo, c.whi, c.bnb, (double2 *)v);
  hipLaunchKernelGGL(k_itrbc_lr, dim3(nblocks(p.ny, 128)), dim3(128), 0, s, p, (const double *)c.fside,
                     (double2 *)v);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

__global__ void k_copy2(double2 *dst, const double2 *src, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

// the shift-invert driver's residual norm terms (src/si_eigdrv3d_new.f90
// :360-372): sum of |r|^2 / dtl^2 over the points of each block
// block blk of nblk (256 threads) of sum |r|^2 / dtl^2 over P points: points
// blk 256 + t, stepping by nblk 256, then a tree over the block; every
// thread returns the block's sum
__device__ __forceinline__ double norm_block(long blk, long nblk, const double2 *r, const double *dtl, long P) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long q = blk * 256 + threadIdx.x; q < P; q += nblk * 256) {
    const double d2 = dtl[q] * dtl[q];
    for (int e = 0; e < 5; e++) {
      const double2 v = r[5 * q + e];
      acc += (v.x * v.x + v.y * v.y) / d2;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(256) void k_forced_norm(const double2 *r, const double *dtl, long P,
                                                     double *part) {
  const double t = norm_block(blockIdx.x, gridDim.x, r, dtl, P);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// src/cgrad.f90 on a complex field v(5,nx,ny): g1 = d/dxi v, g2 = d/deta v,
// one lane per (point, component), with the stencils of the residual
__global__ void k_cgrad(int nx, int ny, int xper, int yper, Stencil st, const double2 *v, double2 *g1,
                        double2 *g2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 5L * nx * ny) return;
  const int e = (int)(t % 5);
  const long q = t / 5;
  const int i = (int)(q % nx), j = (int)(q / nx);
  const z a = d1<double2, z>(nx, xper, i, st.gx, st.dxinv, v + 5L * nx * j + e, 5, st.sym, e == 2 ? -1.0 : 1.0);
  const z b = d1<double2, z>(ny, yper, j, st.gy, st.dyinv, v + 5L * i + e, 5L * nx);
  g1[t] = make_double2(a.re, a.im);
  g2[t] = make_double2(b.re, b.im);
}

// src/cgrad2.f90: g11 = d2/dxi2 v, g22 = d2/deta2 v and the cross
// derivative g12 = d/deta of the caller's g1
__global__ void k_cgrad2(int nx, int ny, int xper, int yper, Stencil st, const double2 *v, const double2 *g1,
                         double2 *g11, double2 *g12, double2 *g22) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 5L * nx * ny) return;
  const int e = (int)(t % 5);
  const long q = t / 5;
  const int i = (int)(q % nx), j = (int)(q / nx);
  const z a = d2<double2, z>(nx, xper, i, st.dx, st.dxsinv, v + 5L * nx * j + e, 5, st.sym, e == 2 ? -1.0 : 1.0);
  const z b = d1<double2, z>(ny, yper, j, st.gy, st.dyinv, g1 + 5L * i + e, 5L * nx);
  const z c = d2<double2, z>(ny, yper, j, st.dy, st.dysinv, v + 5L * i + e, 5L * nx);
  g11[t] = make_double2(a.re, a.im);
  g12[t] = make_double2(b.re, b.im);
  g22[t] = make_double2(c.re, c.im);
}

int launch_cgrad(int nx, int ny, int xper, int yper, const Stencil &st, const void *v, void *g1, void *g2,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_cgrad, dim3(nblocks(5L * nx * ny, 256)), dim3(256), 0, s, nx, ny, xper, yper, st,
                     (const double2 *)v, (double2 *)g1, (double2 *)g2);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int launch_cgrad2(int nx, int ny, int xper, int yper, const Stencil &st, const void *v, const void *g1,
                  void *g11, void *g12, void *g22, hipStream_t s) {
  hipLaunchKernelGGL(k_cgrad2, dim3(nblocks(5L * nx * ny, 256)), dim3(256), 0, s, nx, ny, xper, yper, st,
                     (const double2 *)v, (const double2 *)g1, (double2 *)g11, (double2 *)g12, (double2 *)g22);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// sqrt of the sum of nb partials, in a fixed order (one block): the same
// bits on every run
__device__ __forceinline__ double norm_sum(const double *part, int nb) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int k = threadIdx.x; k < nb; k += 256) acc += part[k];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(256) void k_norm_final(const double *part, int nb, double *out) {
  const double t = norm_sum(part, nb);
  if (threadIdx.x == 0) *out = sqrt(t);
}

int launch_residual_norm(const double *dtl, const void *r, long P, double *part, double *out, hipStream_t s) {
  const int blocks = (int)std::min<long>(nblocks(P, 256), NORM_BLOCKS);
  hipLaunchKernelGGL(k_forced_norm, dim3(blocks), dim3(256), 0, s, (const double2 *)r, dtl, P, part);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, s, (const double *)part, blocks, out);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// The per-block partials of the allocation-free residual norm: one static
// array per GPU (NORM_BLOCKS doubles, 8 KiB), held until the process ends.
// Host code never passes this name to a kernel (on the host it is only a
// stand-in for the symbol): its device address is looked up once per device
// with hipGetSymbolAddress and kept in scratch_dev.
static __device__ __align__(256) double scratch[NORM_BLOCKS];
static std::mutex scratch_mu;
static double *scratch_dev[MAX_DEVICES] = {};

int launch_residual_norm_noalloc(const double *dtl, const void *r, long P, double *out, hipStream_t s) {
  int dev = -1;
  LNS3D_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return set_error(__FILE__, __LINE__, "device index beyond MAX_DEVICES");
  double *part = nullptr;
  {
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
      void *a = nullptr;
      const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !a) return set_error(__FILE__, __LINE__, "hipGetSymbolAddress(scratch) failed");
      scratch_dev[dev] = (double *)a;
    }
    part = scratch_dev[dev];
  }
  // every partial the final sum reads is written first by k_forced_norm:
  // nothing needs clearing between calls
  return launch_residual_norm(dtl, r, P, part, out, s);
}

int launch_forced_norm(const lns3d_params &p, const Coeffs &c, const void *r, double *part, int *nb,
                       hipStream_t s) {
  const long P = (long)p.nx * p.ny;
  const int blocks = (int)std::min<long>(nblocks(P, 256), NORM_BLOCKS);
  hipLaunchKernelGGL(k_forced_norm, dim3(blocks), dim3(256), 0, s, (const double2 *)r, (const double *)c.dtl,
                     P, part);
  LNS3D_CHECK(hipGetLastError());
  *nb = blocks;
  return 0;
}

// device copy as a kernel (graph-capture friendly)
int launch_copy(void *dst, const void *src, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy2, dim3(nblocks(n, 256)), dim3(256), 0, s, (double2 *)dst,
                     (const double2 *)src, n);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int launch_axpy(void *x, const void *d, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(nblocks(n, 256)), dim3(256), 0, s, (double2 *)x, (const double2 *)d, n);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int launch_update_bc(const lns3d_params &p, const Coeffs &c, void *v, void *dv, hipStream_t s) {
  if (c.bnb && !p.yper) {      // inviscid: the increment's wall row back to x, y components
    hipLaunchKernelGGL(k_bcfix, dim3(nblocks(p.nx, 128)), dim3(128), 0, s, p.nx, c.bnb, (double2 *)dv);
    LNS3D_CHECK(hipGetLastError());
  }
  int rc = launch_axpy(v, dv, 5L * p.nx * p.ny, s);
  if (rc) return rc;
  return launch_itrbc(p, c, v, s);
}

// ==========================================================================
// approximate-factorization operator y = Lx (Ly x), blocks regenerated from
// the coefficients (no assembled matrix is stored)
__device__ __forceinline__ int slot_col(int N, int per, int pos, int k) {
  if (per) {
    int m = N - 1;
    return (((pos - 2 + k) % m) + m) % m;
  }
  if (pos == 0) { const int t[5] = {3, 4, 0, 1, 2}; return t[k]; }
  if (pos == 1) { const int t[5] = {4, 0, 1, 2, 3}; return t[k]; }
  if (pos == N - 2) { const int t[5] = {N - 4, N - 3, N - 2, N - 1, N - 5}; return t[k]; }
  if (pos == N - 1) { const int t[5] = {N - 3, N - 2, N - 1, N - 5, N - 4}; return t[k]; }
  return pos - 2 + k;
}

// One factor applied as a stencil: with the slot weights w1_k (first
// derivative) and w2_k (second derivative) of the row (interior a_k, b_k of
// src/lhs1f3d.f90; closure weights of src/lhsbt1f3d.f90 near the ends),
//   L x = x_c + dt [ X (d1 x) - V o (d2 x) - i AB o (d1 x) ] + calcd alfa dt (D + i Di) x_c
// where d1 x = sum_k w1_k x_col(k), d2 x = sum_k w2_k x_col(k), X = Ah or Bh,
// V the compact viscous entries and AB the spanwise cross entries of the
// direction; boundary-condition rows as lhs_entry.  Equal to the assembled
// block rows up to rounding.
__device__ __forceinline__ void line_op(long bid, const AsmCtx &A, const Coeffs &c, int dir, const double2 *x,
                                        double2 *y) {
  using namespace st;
  const lns3d_params &p = A.p;
  const int nx = p.nx, ny = p.ny;
  const long P = (long)nx * ny;
  long q = bid * blockDim.x + threadIdx.x;
  if (q >= P) return;
  const int i = q % nx, j = q / nx;
  const int N = dir == 1 ? nx : ny, pos = dir == 1 ? i : j;
  const int per = dir == 1 ? p.xper : p.yper;
  const int posc = (per && pos == N - 1) ? 0 : pos;       // duplicate periodic row
  const long qr = dir == 1 ? (posc + (long)nx * j) : (i + (long)nx * posc);   // row point
  double w1[5], w2[5], w1m[5], w2m[5];      // w1m, w2m: component 3 (symmetric rows)
  const bool symrow = A.sym && (((A.sym & 1) && posc <= 1) || ((A.sym & 2) && posc >= N - 2));
  if (symrow) {
    for (int k = 0; k < 5; k++) {
      sym_weights(A, posc, N, k, 0, w1[k], w2[k]);
      sym_weights(A, posc, N, k, 2, w1m[k], w2m[k]);
    }
  } else if (!per && (posc <= 1 || posc >= N - 2)) {
    double F[5], S[5];
    if (posc == 0) {
      const double f[5] = {gc4, gc5, gc1, gc2, gc3}, sd[5] = {dd4, dd5, dd1, dd2, dd3};
      for (int k = 0; k < 5; k++) { F[k] = f[k]; S[k] = sd[k]; }
    } else if (posc == 1) {
      const double f[5] = {gb5, gb1, gb2, gb3, gb4}, sd[5] = {db5, db1, db2, db3, db4};
      for (int k = 0; k < 5; k++) { F[k] = f[k]; S[k] = sd[k]; }
    } else if (posc == N - 2) {
      const double f[5] = {-gb4, -gb3, -gb2, -gb1, -gb5}, sd[5] = {db4, db3, db2, db1, db5};
      for (int k = 0; k < 5; k++) { F[k] = f[k]; S[k] = sd[k]; }
    } else {
      const double f[5] = {-gc3, -gc2, -gc1, -gc5, -gc4}, sd[5] = {dd3, dd2, dd1, dd5, dd4};
      for (int k = 0; k < 5; k++) { F[k] = f[k]; S[k] = sd[k]; }
    }
    for (int k = 0; k < 5; k++) { w1[k] = p.alfa * F[k] * A.dinv; w2[k] = p.alfa * S[k] * A.dsinv; }
  } else {
    for (int k = 0; k < 5; k++) { w1[k] = A.a[k]; w2[k] = A.b[k]; }
  }
  if (!symrow)
    for (int k = 0; k < 5; k++) { w1m[k] = w1[k]; w2m[k] = w2[k]; }
  // implicit damping weights (component 3 apart next to a symmetric end)
  double wd[5], wdm[5];
  for (int k = 0; k < 5; k++) {
    wd[k] = c.sbuf ? damp_w(posc, N, per, symrow ? A.sym : 0, k, 0) : 0.0;
    wdm[k] = c.sbuf ? damp_w(posc, N, per, symrow ? A.sym : 0, k, 2) : 0.0;
  }
  const int ir = dir == 1 ? posc : i, jr = dir == 1 ? j : posc;   // the row point
  // the inviscid wall zone (as lhs_entry_bc): rows of is:ie on the wall line
  // (xi factor) or within two rows of it (eta factor).  A block whose column
  // is rotated to the wall normal (b1, b2) acts on the column's values as the
  // unrotated block on (b2 x1 + b1 x2, -b1 x1 + b2 x2); the normal is that of
  // the non-periodic slot map's point (xi) or of the row's wall point (eta)
  const bool lrz = !p.xper && ((ir == 0 && (p.left == 0 || p.left == 4 || p.left == 5)) ||
                               (ir == nx - 1 && (p.right == 0 || p.right == 4 || p.right == 5 ||
                                                 (p.right == 8 && dir == 1))));
  const bool zone = c.bnb && !p.yper && !lrz && ir >= c.wlo && ir <= c.whi &&
                    ((dir == 1 && jr == 0) || (dir == 2 && jr <= 2));
  z d1v[5], d2v[5], xc[5], d4v[5];
  for (int e = 0; e < 5; e++) { d1v[e] = z{0, 0}; d2v[e] = z{0, 0}; d4v[e] = z{0, 0}; }
  for (int k = 0; k < 5; k++) {
    // symmetric rows: slot k is point posc + k - 2; the slots beyond the end
    // carry no weight and are skipped
    const int col = symrow ? posc - 2 + k : slot_col(N, per, posc, k);
    if (symrow && (col < 0 || col > N - 1)) continue;
    const long qc = dir == 1 ? (col + (long)nx * j) : (i + (long)nx * col);
    z xs[5];
    for (int e = 0; e < 5; e++) xs[e] = ld(x + 5 * qc + e);
    if (zone) {
      int ncol = -1;
      if (dir == 1) {
        const int t0[5] = {3, 4, 0, 1, 2}, t1[5] = {4, 0, 1, 2, 3};
        const int tm[5] = {nx - 4, nx - 3, nx - 2, nx - 1, nx - 5}, tn[5] = {nx - 3, nx - 2, nx - 1, nx - 5, nx - 4};
        ncol = ir == 0 ? t0[k] : ir == 1 ? t1[k] : ir == nx - 2 ? tm[k] : ir == nx - 1 ? tn[k] : ir - 2 + k;
      } else if (k == 2 - jr) {
        ncol = ir;
      }
      if (ncol >= 0) {
        const double b1 = c.bnb[ncol], b2 = c.bnb[nx + ncol];
        const z u = xs[1], w = xs[2];
        xs[1] = sc(b2, u) + sc(b1, w);
        xs[2] = sc(-b1, u) + sc(b2, w);
      }
    }
    for (int e = 0; e < 5; e++) {
      const z xv = xs[e];
      d1v[e] = d1v[e] + sc(e == 2 ? w1m[k] : w1[k], xv);
      d2v[e] = d2v[e] + sc(e == 2 ? w2m[k] : w2[k], xv);
      d4v[e] = d4v[e] + sc(e == 2 ? wdm[k] : wd[k], xv);
      if (k == 2) xc[e] = xv;
    }
  }
  const double dt = c.dtl[qr];
  const double *X = dir == 1 ? c.Ah : c.Bh;
  const double *V = dir == 1 ? c.V11 : c.V22;
  const double *AB = c.ABhi + (dir == 1 ? 0 : 2) * P;
  z out[5];
  for (int r = 0; r < 5; r++) {
    z a{0, 0};
    for (int cc = 0; cc < 5; cc++) a = a + sc(X[(long)(r + 5 * cc) * P + qr], d1v[cc]);
    out[r] = a;
  }
  out[1] = out[1] - sc(V[qr], d2v[1]) - sc(V[4 * P + qr], d2v[2]);
  out[2] = out[2] - sc(V[5 * P + qr], d2v[1]) - sc(V[P + qr], d2v[2]);
  out[3] = out[3] - sc(V[2 * P + qr], d2v[3]);
  out[4] = out[4] - sc(V[3 * P + qr], d2v[4]);
  {
    // symmetric rows read the cross entries through src/lhsbs1f3d.f90's
    // Ahi(6,ny,nx) shape: point posc + ny j
    const long qa = symrow ? posc + (long)ny * j : qr;
    const double ab0 = AB[qa], ab1 = AB[P + qa];
    z t1 = sc(ab0, d1v[3]), t2 = sc(ab1, d1v[3]), t3 = sc(ab0, d1v[1]) + sc(ab1, d1v[2]);
    out[1] = out[1] - z{-t1.im, t1.re};     // - i * t
    out[2] = out[2] - z{-t2.im, t2.re};
    out[3] = out[3] - z{-t3.im, t3.re};
  }
  for (int r = 0; r < 5; r++) out[r] = sc(dt, out[r]);
  if (A.calcd) {
    for (int r = 0; r < 5; r++) {
      z a{0, 0}, b{0, 0};
      for (int cc = 0; cc < 5; cc++) {
        a = a + sc(c.Dh[(long)(r + 5 * cc) * P + qr], xc[cc]);
        b = b + sc(c.Dhi[(long)(r + 5 * cc) * P + qr], xc[cc]);
      }
      const double f = p.alfa * dt;
      out[r] = out[r] + sc(f, a) + z{-f * b.im, f * b.re};
    }
  } else if (c.spl) {                  // sponge of the factor without the D terms
    const double f = p.alfa * dt * c.spl[qr];
    for (int r = 0; r < 5; r++) out[r] = out[r] + sc(f, xc[r]);
  }
  if (c.sbuf) {                        // implicit damping of the smoother
    const double f = p.alfa * dt * c.sbuf[qr];
    for (int r = 0; r < 5; r++) out[r] = out[r] + sc(f, d4v[r]);
  }
  for (int r = 0; r < 5; r++) out[r] = out[r] + xc[r];
  if (zone && jr == 0) {               // the wall row: u rotated, the normal velocity held
    const double b1 = c.bnb[ir], b2 = c.bnb[nx + ir];
    out[1] = sc(b2, out[1]) - sc(b1, out[2]);
    out[2] = ld(x + 5 * qr + 2);
  }
  // boundary-condition rows (src/lhsbc1f3d.f90, lhsbc2f3d.f90), as lhs_entry
  // (the identity rows take the unrotated values: none lies in the zone)
  for (int e = 0; e < 5; e++) xc[e] = ld(x + 5 * qr + e);
  bool ident[5] = {false, false, false, false, false};
  bool wallrow0 = false, wallrow4 = false;
  const bool lrrow = !p.xper && ((ir == 0 && (p.left == 0 || p.left == 4 || p.left == 5)) ||
                                  (ir == nx - 1 && (p.right == 0 || p.right == 4 || p.right == 5)));
  if (lrrow) for (int r = 0; r < 5; r++) ident[r] = true;
  const bool holdrow = !p.xper && dir == 1 && ir == nx - 1 && p.right == 8;   // -I
  if (!p.yper) {
    if (jr == ny - 1) for (int r = 0; r < 5; r++) ident[r] = true;
    if (jr == 0 && !c.bnb && ir >= c.wlo && ir <= c.whi) {   // the viscous wall rows
      ident[1] = ident[2] = ident[3] = true;
      if (dir == 1 && (p.wall == 1 || p.wall == 2 || p.wall == 3)) ident[0] = true;
      if (dir == 2 && (p.wall == 1 || p.wall == 3)) wallrow0 = true;
      if (p.wallt == 0 || (dir == 1 && p.wallt == 1)) ident[4] = true;
      if (dir == 2 && p.wallt == 1) wallrow4 = true;
    }
  }
  if (lrrow) wallrow0 = wallrow4 = false;
  if (!p.yper && jr == ny - 1) wallrow0 = wallrow4 = false;
  if (wallrow0 || wallrow4) {
    const double w[5] = {-gc4, -gc5, -gc1, -gc2, -gc3};
    const double w3[5] = {4.0, -1.0, -1.0, 4.0, -6.0};     // wall = 3 density row
    const double *w0 = p.wall == 3 ? w3 : w;
    z s0{0, 0}, s4{0, 0};
    for (int k = 0; k < 5; k++) {
      const int col = slot_col(N, per, posc, k);
      const long qc = dir == 1 ? (col + (long)nx * j) : (i + (long)nx * col);
      s0 = s0 + sc(w0[k], ld(x + 5 * qc + 0));
      s4 = s4 + sc(w[k], ld(x + 5 * qc + 4));
    }
    if (wallrow0 && !ident[0]) out[0] = s0;
    if (wallrow4 && !ident[4]) out[4] = s4;
  }
  // the duplicate row of a periodic line is no unknown of the factor (its
  // solve copies row 0 there, its residual is zero): no operator row
  const bool dup = per && pos == N - 1;
  for (int r = 0; r < 5; r++) {
    if (ident[r]) out[r] = xc[r];
    if (holdrow) out[r] = z{-xc[r].re, -xc[r].im};
    y[5 * q + r] = dup ? make_double2(0.0, 0.0) : make_double2(out[r].re, out[r].im);
  }
}
__global__ __launch_bounds__(256) void k_line_op(AsmCtx A, Coeffs c, int dir, const double2 *x, double2 *y) {
  line_op(xcd_tile(blockIdx.x, gridDim.x), A, c, dir, x, y);
}

int launch_af_matvec(const lns3d_params &p, const Coeffs &c, const void *x, void *y, void *tmp,
                     hipStream_t s) {
  long P = (long)p.nx * p.ny;
  AsmCtx Ax = make_asm(p, c, 1), Ay = make_asm(p, c, 2);
  hipLaunchKernelGGL(k_line_op, dim3(nblocks(P, 256)), dim3(256), 0, s, Ay, c, 2,
                     (const double2 *)x, (double2 *)tmp);
  hipLaunchKernelGGL(k_line_op, dim3(nblocks(P, 256)), dim3(256), 0, s, Ax, c, 1,
                     (const double2 *)tmp, (double2 *)y);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// y = b - y and the max norms |y|, |b| (of real and imaginary parts): one
// partial pair per block, then one block reduces the partials
constexpr int RES_BLOCKS = 512;

__global__ __launch_bounds__(256) void k_residual(const double2 *b, double2 *y, long n, double *part) {
  double ay = 0.0, abv = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    double2 bb = b[t], yy = y[t];
    double2 rr = make_double2(bb.x - yy.x, bb.y - yy.y);
    y[t] = rr;
    ay = fmax(ay, fmax(fabs(rr.x), fabs(rr.y)));
    abv = fmax(abv, fmax(fabs(bb.x), fabs(bb.y)));
  }
  __shared__ double s1[256], s2[256];
  s1[threadIdx.x] = ay;
  s2[threadIdx.x] = abv;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

__global__ __launch_bounds__(256) void k_residual_final(const double *part, int nb, double *norms) {
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    a = fmax(a, part[2 * k]);
    b = fmax(b, part[2 * k + 1]);
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    norms[0] = s1[0];
    norms[1] = s2[0];
  }
}

// norms: 2 doubles of results followed by 2*RES_BLOCKS of partials
int launch_residual_norms(const void *b, void *y, long n, double *norms, hipStream_t s) {
  const int nb = (int)std::min<long>(RES_BLOCKS, (n + 255) / 256);
  hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, s, (const double2 *)b, (double2 *)y, n,
                     norms + 2);
  hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(256), 0, s, (const double *)(norms + 2), nb,
                     norms);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// ==========================================================================
// batches: each stage of the step is one launch over the members of a batch.
// Workgroup b of a stage belongs to member k with off[k] <= b < off[k + 1]
// (the stage's own table); the member's item holds its deck, coefficients and
// buffers.  The arithmetic is that of the single-member kernels.
struct StepBatchItem {
  lns3d_params p;
  Stencil st;
  Coeffs c;
  AsmCtx ax, ay;
  double2 *buf[6];     // v, dv, rhs, t1, t2, work
  long P;
  double *norm;        // residual-norm slots (or null), as StepBatchMember
  int norm_count;
  long norm_idx;
};

__device__ __forceinline__ int step_member(const int *off, int count) {
  int lo = 0, hi = count;
  const int b = (int)blockIdx.x;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void kb_rhs_d1x(const StepBatchItem *items, const int *off, int count) {
  const int k = step_member(off, count);
  const StepBatchItem &b = items[k];
  rhs_d1x((long)blockIdx.x - off[k], b.p.nx, b.p.ny, b.p.xper, b.st, b.buf[SB_V], b.buf[SB_WORK]);
}

__device__ __forceinline__ void batch_point_args(const StepBatchItem *items, const int *off, int count,
                                                 lns3d_params &p, Stencil &st, Coeffs &c, const double2 *&v,
                                                 const double2 *&c1a, double2 *&r, long &tile) {
  const int k = step_member(off, count);
  const StepBatchItem &b = items[k];
  p = b.p;
  st = b.st;
  c = b.c;
  v = b.buf[SB_V];
  c1a = b.buf[SB_WORK];
  r = b.buf[SB_RHS];
  tile = (long)blockIdx.x - off[k];
}

// y = L_dir x with x, y chosen among the member's buffers
__global__ __launch_bounds__(256) void kb_line_op(const StepBatchItem *items, const int *off, int count, int dir,
                                                  int xs, int ys) {
  const int k = step_member(off, count);
  const StepBatchItem &b = items[k];
  line_op((long)blockIdx.x - off[k], dir == 1 ? b.ax : b.ay, b.c, dir, b.buf[xs], b.buf[ys]);
}

// elementwise stages over the 5P values of a member
__global__ void kb_vec(const StepBatchItem *items, const int *off, int count, int op) {
  const int k = step_member(off, count);
  const StepBatchItem &b = items[k];
  const long t = ((long)blockIdx.x - off[k]) * blockDim.x + threadIdx.x;
  if (t >= 5 * b.P) return;
  double2 *v = b.buf[SB_V], *dv = b.buf[SB_DV], *rhs = b.buf[SB_RHS], *t2 = b.buf[SB_T2];
  if (op == SB_COPY) {
    dv[t] = rhs[t];
  } else if (op == SB_SUB) {
    const double2 bb = rhs[t], yy = t2[t];
    t2[t] = make_double2(bb.x - yy.x, bb.y - yy.y);
  } else if (op == SB_ADD_DV || op == SB_UPDATE) {
    const double2 a = dv[t], c = t2[t];
    const double2 d = make_double2(a.x + c.x, a.y + c.y);
    dv[t] = d;
    if (op == SB_UPDATE) {
      const double2 x = v[t];
      v[t] = make_double2(x.x + d.x, x.y + d.y);
    }
  } else {   // SB_ADD_V
    const double2 x = v[t], d = dv[t];
    v[t] = make_double2(x.x + d.x, x.y + d.y);
  }
}

__global__ void kb_itrbc_wall_top(const StepBatchItem *items, const int *off, int count) {
  const int k = step_member(off, count);
  const StepBatchItem &b = items[k];
  itrbc_wall_top((long)blockIdx.x - off[k], b.p, b.c.ftop, b.c.fwall, b.c.wlo, b.c.whi, b.c.bnb, b.buf[SB_V]);
}

__global__ void kb_itrbc_lr(const StepBatchItem *items, const int *off, int count) {
  const int k = step_member(off, count);
  const StepBatchItem &b = items[k];
  itrbc_lr((long)blockIdx.x - off[k], b.p, b.c.fside, b.buf[SB_V]);
}

// the residual norm of the members with a norm output: the blocks of a
// member are those of launch_residual_norm for its P, then one block per
// member sums its partials in the order k_norm_final does
__global__ __launch_bounds__(256) void kb_norm_part(const StepBatchItem *items, const int *off, int count,
                                                    double *part) {
  const int k = step_member(off, count);
  const StepBatchItem &b = items[k];
  const double t = norm_block((long)blockIdx.x - off[k], off[k + 1] - off[k], b.buf[SB_RHS], b.c.dtl, b.P);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void kb_norm_final(const StepBatchItem *items, const int *off, const double *part,
                                                     long step) {
  const StepBatchItem &b = items[blockIdx.x];
  if (!b.norm) return;
  const double t = norm_sum(part + off[blockIdx.x], off[blockIdx.x + 1] - off[blockIdx.x]);
  if (threadIdx.x == 0) b.norm[(b.norm_idx + step) % b.norm_count] = sqrt(t);
}

// workgroups of each stage for a member
static long stage_blocks(int stage, const lns3d_params &p) {
  const long P = (long)p.nx * p.ny;
  switch (stage) {
    case SB_STAGE_D1X: return nblocks(5 * P, 256);
    case SB_STAGE_POINT: return nblocks(P, RT);
    case SB_STAGE_LINE: return nblocks(P, 256);
    case SB_STAGE_VEC: return nblocks(5 * P, 256);
    case SB_STAGE_WALL: return nblocks(p.nx, 128);
    default: return nblocks(p.ny, 128);
  }
}

int step_batch_build(const StepBatchMember *m, int count, StepBatch &out, std::vector<TableUpload> &ups) {
  step_batch_free(out);
  std::vector<StepBatchItem> items((size_t)count);
  std::vector<std::vector<int>> off(SB_NSTAGE, std::vector<int>(1, 0));
  for (int k = 0; k < count; k++) {
    const StepBatchMember &a = m[k];
    StepBatchItem &it = items[(size_t)k];
    it.p = a.p;
    it.st = a.st;
    it.c = a.c;
    it.ax = make_asm(a.p, a.c, 1);
    it.ay = make_asm(a.p, a.c, 2);
    it.buf[SB_V] = a.v;
    it.buf[SB_DV] = a.dv;
    it.buf[SB_RHS] = a.rhs;
    it.buf[SB_T1] = a.t1;
    it.buf[SB_T2] = a.t2;
    it.buf[SB_WORK] = a.work;
    it.P = (long)a.p.nx * a.p.ny;
    it.norm = a.norm;
    it.norm_count = a.norm ? a.norm_count : 0;
    it.norm_idx = a.norm_idx;
    if (a.norm && a.norm_count < 1) return set_error(__FILE__, __LINE__, "batched steps: a norm output of no slots");
    for (int sg = 0; sg < SB_NSTAGE; sg++) {
      const long nb = sg == SB_STAGE_NORM    ? (a.norm ? std::min<long>(nblocks(it.P, 256), NORM_BLOCKS) : 0)
                      : sg == SB_STAGE_BCFIX ? ((a.c.bnb && !a.p.yper) ? nblocks(a.p.nx, 128) : 0)
                                             : stage_blocks(sg, a.p);
      if ((long)off[sg].back() + nb > 0x7fffffffL) return set_error(__FILE__, __LINE__, "batch too large");
      off[sg].push_back(off[sg].back() + (int)nb);
    }
  }
  // one device block: the items, each stage's table, the norm partials
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t bytes = up(items.size() * sizeof(StepBatchItem)), at_off[SB_NSTAGE];
  for (int sg = 0; sg < SB_NSTAGE; sg++) {
    at_off[sg] = bytes;
    bytes += up(off[sg].size() * sizeof(int));
  }
  const size_t at_part = bytes;
  bytes += off[SB_STAGE_NORM].back() * sizeof(double);
  std::vector<char> host(at_part);
  memcpy(host.data(), items.data(), items.size() * sizeof(StepBatchItem));
  for (int sg = 0; sg < SB_NSTAGE; sg++) memcpy(host.data() + at_off[sg], off[sg].data(), off[sg].size() * sizeof(int));
  char *dev = nullptr;
  LNS3D_CHECK(hipMalloc(&dev, bytes));
  out.mem = dev;
  out.count = count;
  out.items = dev;
  for (int sg = 0; sg < SB_NSTAGE; sg++) {
    out.blocks[sg] = off[sg].back();
    out.off[sg] = (int *)(dev + at_off[sg]);
  }
  out.part = out.blocks[SB_STAGE_NORM] > 0 ? (double *)(dev + at_part) : nullptr;
  return stage_upload(dev, host.data(), at_part, ups);
}

void step_batch_free(StepBatch &b) {
  if (b.mem) (void)hipFree(b.mem);
  for (int sg = 0; sg < SB_NSTAGE; sg++) b.off[sg] = nullptr;
  b.mem = nullptr;
  b.items = nullptr;
  b.part = nullptr;
  b.count = 0;
}

int step_batch_rhs(const StepBatch &b, hipStream_t s) {
  const auto *it = (const StepBatchItem *)b.items;
  hipLaunchKernelGGL(kb_rhs_d1x, dim3((unsigned)b.blocks[SB_STAGE_D1X]), dim3(256), 0, s, it, b.off[SB_STAGE_D1X],
                     b.count);
  hipLaunchKernelGGL(k_rhs_point_t<true>, dim3((unsigned)b.blocks[SB_STAGE_POINT]), dim3(5 * 64), 0, s,
                     lns3d_params{}, Stencil{}, Coeffs{}, nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr,
                     it, b.off[SB_STAGE_POINT], b.count);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int step_batch_matvec(const StepBatch &b, hipStream_t s) {
  const auto *it = (const StepBatchItem *)b.items;
  const dim3 g((unsigned)b.blocks[SB_STAGE_LINE]);
  hipLaunchKernelGGL(kb_line_op, g, dim3(256), 0, s, it, b.off[SB_STAGE_LINE], b.count, 2, (int)SB_DV, (int)SB_T1);
  hipLaunchKernelGGL(kb_line_op, g, dim3(256), 0, s, it, b.off[SB_STAGE_LINE], b.count, 1, (int)SB_T1, (int)SB_T2);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// bcfix of the members with the inviscid wall rows (their wall points)
__global__ void kb_bcfix(const StepBatchItem *items, const int *off, int count) {
  const int k = step_member(off, count);
  const StepBatchItem &b = items[k];
  const int i = (int)(((long)blockIdx.x - off[k]) * blockDim.x + threadIdx.x);
  if (i >= b.p.nx) return;
  bcfix_point(i, b.p.nx, b.c.bnb, b.buf[SB_DV]);
}
int step_batch_bcfix(const StepBatch &b, hipStream_t s) {
  if (b.blocks[SB_STAGE_BCFIX] == 0) return 0;
  hipLaunchKernelGGL(kb_bcfix, dim3((unsigned)b.blocks[SB_STAGE_BCFIX]), dim3(128), 0, s,
                     (const StepBatchItem *)b.items, b.off[SB_STAGE_BCFIX], b.count);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}
int step_batch_vec(const StepBatch &b, int op, hipStream_t s) {
  hipLaunchKernelGGL(kb_vec, dim3((unsigned)b.blocks[SB_STAGE_VEC]), dim3(256), 0, s,
                     (const StepBatchItem *)b.items, b.off[SB_STAGE_VEC], b.count, op);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int step_batch_norm(const StepBatch &b, long k, hipStream_t s) {
  if (b.blocks[SB_STAGE_NORM] == 0) return 0;
  const auto *it = (const StepBatchItem *)b.items;
  hipLaunchKernelGGL(kb_norm_part, dim3((unsigned)b.blocks[SB_STAGE_NORM]), dim3(256), 0, s, it,
                     b.off[SB_STAGE_NORM], b.count, b.part);
  hipLaunchKernelGGL(kb_norm_final, dim3((unsigned)b.count), dim3(256), 0, s, it, b.off[SB_STAGE_NORM],
                     (const double *)b.part, k);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int step_batch_itrbc(const StepBatch &b, hipStream_t s) {
  const auto *it = (const StepBatchItem *)b.items;
  hipLaunchKernelGGL(kb_itrbc_wall_top, dim3((unsigned)b.blocks[SB_STAGE_WALL]), dim3(128), 0, s, it,
                     b.off[SB_STAGE_WALL], b.count);
  hipLaunchKernelGGL(kb_itrbc_lr, dim3((unsigned)b.blocks[SB_STAGE_LR]), dim3(128), 0, s, it, b.off[SB_STAGE_LR],
                     b.count);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

}  // namespace lns3d
