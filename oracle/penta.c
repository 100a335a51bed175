/*
 * Oracle: non-pivoted block LU of 5x5-block pentadiagonal lines, with the
 * reference's boundary wrap-around storage (bc) or periodic bordering (p).
 * TEST INFRASTRUCTURE ONLY (see lns3d_oracle.h).
 *
 * The reference spells each row case out (src/cpenta1bc.f:43-495 factor,
 * :499-847 solve; src/cpenta1p.f:46-654 factor, :660-1043 solve).  This
 * restatement drives the same elimination from one description of where each
 * (row, column) block is stored -- blk() below -- so every row case of the
 * reference falls out of the same loops:
 *
 *   for pivot row i:  in-block Doolittle on the diagonal block, with the row
 *                     operations applied to every block right of it
 *                     (cpenta1bc.f:54-71);
 *                     for each pivot column nc, fact = 1/U(nc,nc), then for
 *                     each lower row t holding a block in column i, store the
 *                     negated multiplier and update row t's blocks
 *                     (cpenta1bc.f:75-108).
 *   forward solve:    in-block L, then push to the lower rows (:505-700);
 *   backward solve:   rows n..1, blocks right of the diagonal in ascending
 *                     column order, then divide by U(m,m) (:705-846).
 *
 * Floating-point operation order per element matches the reference's loops.
 * The index defect at cpenta1bc.f:675 / cpenta2bc.f:676 is not reproduced
 * (the update goes to row n-1, as every other line of those files does).
 * For periodic lines the reference's factor adds the row n-5 -> row n-3
 * update of column n-1 into slot 4 (cpenta1p.f:284, :345; column n-2) rather
 * than slot 5; exact=0 reproduces that, exact=1 gives the exact LU.
 */
#include <stdlib.h>
#include <string.h>
#include "lns3d_oracle.h"

typedef struct { zdouble *p; int rs, cs; } blk_t;   /* element (r,c) = p[r*rs + c*cs] */
#define E(b, r, c) ((b).p[(r) * (b).rs + (c) * (b).cs])

typedef struct {
  int periodic, n, exact;
  zdouble *mat, *per, *per2;   /* already offset to this line */
  long ms, ps;                 /* stride (complex elements) between rows i, i+1 */
} line_t;

/* slot (0-based) of column c in row i for the non-periodic storage, or -1
 * (src/lhsbt2f3d.f90:53-59, 183-189, 337-343, 491-497) */
static int bc_slot(int n, int i, int c) {
  static const int r0[5] = {2, 3, 4, 0, 1};
  static const int r1[5] = {1, 2, 3, 4, 0};
  if (i == 0) return (c >= 0 && c <= 4) ? r0[c] : -1;
  if (i == 1) return (c >= 0 && c <= 4) ? r1[c] : -1;
  if (i == n - 2) {
    int d = c - (n - 5);              /* columns n-5 .. n-1 */
    static const int rm[5] = {4, 0, 1, 2, 3};
    return (d >= 0 && d <= 4) ? rm[d] : -1;
  }
  if (i == n - 1) {
    int d = c - (n - 5);
    static const int rn[5] = {3, 4, 0, 1, 2};
    return (d >= 0 && d <= 4) ? rn[d] : -1;
  }
  return (c - i >= -2 && c - i <= 2) ? c - i + 2 : -1;
}

static int blk(const line_t *L, int i, int c, blk_t *b) {
  int n = L->n;
  if (!L->periodic) {
    int k = bc_slot(n, i, c);
    if (k < 0) return 0;
    b->p = L->mat + (long)i * L->ms + k; b->rs = 5; b->cs = 25;
    return 1;
  }
  if (c - i >= -2 && c - i <= 2 && c >= 0 && c < n) {
    b->p = L->mat + (long)i * L->ms + (c - i + 2); b->rs = 5; b->cs = 25;
    return 1;
  }
  if ((c == n - 2 || c == n - 1) && i < c - 2) {          /* right border */
    b->p = L->per + (long)i * L->ps + (c - (n - 2)); b->rs = 2; b->cs = 10;
    return 1;
  }
  if ((i == n - 2 || i == n - 1) && c < i - 2) {          /* bottom border */
    b->p = L->per2 + (long)c * L->ps + (i - (n - 2)); b->rs = 2; b->cs = 10;
    return 1;
  }
  return 0;
}

/* destination of the pivot-row block in column c when eliminating into row t */
static void dest_blk(const line_t *L, int i, int t, int c, blk_t *b) {
  blk(L, t, c, b);
  if (L->periodic && !L->exact && i == L->n - 6 && t == L->n - 4 && c == L->n - 2) {
    b->p = L->mat + (long)t * L->ms + 3; b->rs = 5; b->cs = 25;
  }
}

/* candidate partners of row/column i, ascending: i+1..i+4, and the two
 * border indices n-2, n-1 for periodic lines */
static int candidates(const line_t *L, int i, int *cand) {
  int k = 0, n = L->n;
  for (int d = 1; d <= 4 && i + d < n; d++) cand[k++] = i + d;
  if (L->periodic)
    for (int e = n - 2; e < n; e++)
      if (e > i + 4) cand[k++] = e;
  return k;
}

static int upper_cols(const line_t *L, int i, int *cols) {
  int cand[8], nc = candidates(L, i, cand), k = 0;
  for (int a = 0; a < nc; a++) {
    blk_t b;
    if (blk(L, i, cand[a], &b)) cols[k++] = cand[a];
  }
  return k;
}

static int lower_rows(const line_t *L, int i, int *rows) {
  int cand[8], nc = candidates(L, i, cand), k = 0;
  for (int a = 0; a < nc; a++) {
    blk_t b;
    if (blk(L, cand[a], i, &b)) rows[k++] = cand[a];
  }
  return k;
}

static void factor_line(const line_t *L) {
  int n = L->n;
  int ucol[8], lrow[8];
  for (int i = 0; i < n; i++) {
    blk_t D, U[8];
    blk(L, i, i, &D);
    int nu = upper_cols(L, i, ucol);
    for (int u = 0; u < nu; u++) blk(L, i, ucol[u], &U[u]);
    for (int nc = 0; nc < 4; nc++)
      for (int l = nc + 1; l < 5; l++) {
        zdouble mult = -E(D, l, nc) / E(D, nc, nc);
        E(D, l, nc) = mult;
        for (int q = nc + 1; q < 5; q++) E(D, l, q) += mult * E(D, nc, q);
        for (int u = 0; u < nu; u++)
          for (int q = 0; q < 5; q++) E(U[u], l, q) += mult * E(U[u], nc, q);
      }
    int nl = lower_rows(L, i, lrow);
    blk_t Lb[8], Dst[8][8];
    for (int t = 0; t < nl; t++) {
      blk(L, lrow[t], i, &Lb[t]);
      for (int u = 0; u < nu; u++) dest_blk(L, i, lrow[t], ucol[u], &Dst[t][u]);
    }
    for (int nc = 0; nc < 5; nc++) {
      zdouble fact = 1.0 / E(D, nc, nc);
      for (int m = 0; m < 5; m++)
        for (int t = 0; t < nl; t++) {
          zdouble mult = -E(Lb[t], m, nc) * fact;
          E(Lb[t], m, nc) = mult;
          for (int q = nc + 1; q < 5; q++) E(Lb[t], m, q) += mult * E(D, nc, q);
          for (int u = 0; u < nu; u++)
            for (int q = 0; q < 5; q++) E(Dst[t][u], m, q) += mult * E(U[u], nc, q);
        }
    }
  }
}

static void solve_line(const line_t *L, zdouble *x, long xs) {
  int n = L->n;
  int ucol[8], lrow[8];
#define X(i, d) x[(long)(i) * xs + (d)]
  for (int i = 0; i < n; i++) {
    blk_t D;
    blk(L, i, i, &D);
    for (int nc = 0; nc < 4; nc++)
      for (int l = nc + 1; l < 5; l++) X(i, l) += E(D, l, nc) * X(i, nc);
    int nl = lower_rows(L, i, lrow);
    blk_t Lb[8];
    for (int t = 0; t < nl; t++) blk(L, lrow[t], i, &Lb[t]);
    for (int nc = 0; nc < 5; nc++)
      for (int m = 0; m < 5; m++)
        for (int t = 0; t < nl; t++) X(lrow[t], m) += E(Lb[t], m, nc) * X(i, nc);
  }
  for (int i = n - 1; i >= 0; i--) {
    blk_t D, U[8];
    blk(L, i, i, &D);
    int nu = upper_cols(L, i, ucol);
    for (int u = 0; u < nu; u++) blk(L, i, ucol[u], &U[u]);
    for (int m = 4; m >= 0; m--) {
      zdouble acc = X(i, m);
      for (int q = m + 1; q < 5; q++) acc -= X(i, q) * E(D, m, q);
      for (int q = 0; q < 5; q++)
        for (int u = 0; u < nu; u++) acc -= X(ucol[u], q) * E(U[u], m, q);
      X(i, m) = acc / E(D, m, m);
    }
  }
#undef X
}

int orc_cpenta_bc(int layout, int nsys, int n, zdouble *mat, zdouble *rhs, int code) {
  if (n < 7 || nsys < 1 || code < 0 || code > 2 || (layout != 1 && layout != 2)) return -1;
  long ms_row = layout == 1 ? 125L * nsys : 125L;
  long ms_sys = layout == 1 ? 125L : 125L * n;
  long xs_row = layout == 1 ? 5L * nsys : 5L;
  long xs_sys = layout == 1 ? 5L : 5L * n;
  for (int s = 0; s < nsys; s++) {
    line_t L = {0, n, 1, mat + s * ms_sys, NULL, NULL, ms_row, 0};
    if (code == 0 || code == 2) factor_line(&L);
    if (code == 1 || code == 2) solve_line(&L, rhs + s * xs_sys, xs_row);
  }
  return 0;
}

int orc_cpenta_p(int layout, int nsys, int np, zdouble *mat, zdouble *rhs,
                 zdouble *per, zdouble *per2, int code, int exact) {
  int n = np - 1;
  if (n < 7 || nsys < 1 || code < 0 || code > 2 || (layout != 1 && layout != 2)) return -1;
  long ms_row = layout == 1 ? 125L * nsys : 125L;
  long ms_sys = layout == 1 ? 125L : 125L * np;
  long ps_row = layout == 1 ? 50L * nsys : 50L;
  long ps_sys = layout == 1 ? 50L : 50L * np;
  long xs_row = layout == 1 ? 5L * nsys : 5L;
  long xs_sys = layout == 1 ? 5L : 5L * np;
  for (int s = 0; s < nsys; s++) {
    line_t L = {1, n, exact, mat + s * ms_sys, per + s * ps_sys, per2 + s * ps_sys,
                ms_row, ps_row};
    if (code == 0 || code == 2) {
      /* bordered fill-in storage (src/cpenta1p.f:53-78) */
      for (int i = 0; i < np; i++)
        for (int e = 0; e < 50; e++) {
          L.per[i * L.ps + e] = 0;
          L.per2[i * L.ps + e] = 0;
        }
      for (int r = 0; r < 5; r++)
        for (int c = 0; c < 5; c++) {
          long me = 5 * r + 25 * c, pe = 2 * r + 10 * c;
          L.per[0 * L.ps + pe + 0] = L.mat[0 * L.ms + me + 0];
          L.per[0 * L.ps + pe + 1] = L.mat[0 * L.ms + me + 1];
          L.per[1 * L.ps + pe + 1] = L.mat[1 * L.ms + me + 0];
          L.per2[0 * L.ps + pe + 0] = L.mat[(long)(n - 2) * L.ms + me + 4];
          L.per2[0 * L.ps + pe + 1] = L.mat[(long)(n - 1) * L.ms + me + 3];
          L.per2[1 * L.ps + pe + 1] = L.mat[(long)(n - 1) * L.ms + me + 4];
        }
      factor_line(&L);
    }
    if (code == 1 || code == 2) {
      zdouble *x = rhs + s * xs_sys;
      solve_line(&L, x, xs_row);
      for (int d = 0; d < 5; d++) x[(long)n * xs_row + d] = x[d];
    }
  }
  return 0;
}
