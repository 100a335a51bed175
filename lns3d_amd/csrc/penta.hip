// Batched complex 5x5 block-pentadiagonal line solvers for gfx950.
//
// Replaces src/cpenta1bc.f, src/cpenta2bc.f (non-periodic, boundary
// wrap-around storage) and src/cpenta1p.f, src/cpenta2p.f (periodic,
// bordered).  Three stages:
//
//  1. k_lu_lines: non-pivoted block LU of every line, one lane per line, on
//     a line-interleaved working copy ([row][slot][r][c][line]) so a wave's
//     accesses to one block element are contiguous across its 64 lines.  The
//     elimination order and the storage of negated multipliers follow the
//     reference (cpenta1bc.f:43-495, cpenta1p.f:46-654), including the
//     periodic factor's slot-4 placement at cpenta1p.f:284/345 unless exact
//     is requested.
//  2. k_records_bc: for non-periodic lines, turn row i's factors into the
//     records the fast solve streams:
//        forward   y_i = Linv_i b_i + N1_i y_{i-2} + N2_i y_{i-1}
//        backward  x_i = Uinv_i y_i + Q4_i x_{i+1} + Q5_i x_{i+2}
//     with Linv_i/Uinv_i the inverses of the in-block unit-lower / upper
//     factors, N = Linv * (negated multiplier block), Q = -Uinv * (upper
//     block).  Mathematically the reference's forward/back substitution
//     (cpenta1bc.f:499-847); the per-row critical path becomes one 5x5
//     matvec.  Boundary rows 0,1 and n-2,n-1 carry extra blocks, kept in a
//     per-line tail record.
//  3. k_solve_bc: one workgroup per line.  Waves 1-3 stream chunks of R rows
//     of records into a double-buffered LDS ring and compute the parallel
//     parts (Linv b, Uinv y); wave 0 runs the serial recurrence with lanes
//     0..4 owning one component each and the previous rows' values broadcast
//     through scalar registers.
//  Periodic lines are solved by k_solve_lines (one lane per line, the
//  reference's order) on the LU working copy.
#include "common.hpp"
#include "penta.hpp"

namespace lns3d {

// ---------------------------------------------------------------------------
// block addressing in the interleaved LU working copy
// element (r,c) of a block = p[(r*5 + c) * L]
template <typename T> struct Blk {
  cx<T> *p;
  __device__ __forceinline__ cx<T> &operator()(int r, int c) const { return p[(long)(r * 5 + c) * ldl]; }
  long ldl;
};

struct LineGeom {
  int n;         // unknown rows (np-1 for periodic)
  int periodic;
  int exact;
  long L;        // lines (interleave stride)
};

__device__ __forceinline__ int bc_slot(int n, int i, int c) {
  if (i == 0) { const int t[5] = {2, 3, 4, 0, 1}; return (c >= 0 && c <= 4) ? t[c] : -1; }
  if (i == 1) { const int t[5] = {1, 2, 3, 4, 0}; return (c >= 0 && c <= 4) ? t[c] : -1; }
  if (i == n - 2) { const int t[5] = {4, 0, 1, 2, 3}; int d = c - (n - 5); return (d >= 0 && d <= 4) ? t[d] : -1; }
  if (i == n - 1) { const int t[5] = {3, 4, 0, 1, 2}; int d = c - (n - 5); return (d >= 0 && d <= 4) ? t[d] : -1; }
  int d = c - i;
  return (d >= -2 && d <= 2) ? d + 2 : -1;
}

// where block (row i, column c) lives; false if structurally zero
template <typename T>
__device__ __forceinline__ bool blk(const LineGeom &g, cx<T> *mat, cx<T> *per, cx<T> *per2,
                                    long line, int i, int c, Blk<T> &b) {
  b.ldl = g.L;
  const int n = g.n;
  if (!g.periodic) {
    int k = bc_slot(n, i, c);
    if (k < 0) return false;
    b.p = mat + ((long)(i * 5 + k) * 25) * g.L + line;
    return true;
  }
  int d = c - i;
  if (d >= -2 && d <= 2 && c >= 0 && c < n) {
    b.p = mat + ((long)(i * 5 + d + 2) * 25) * g.L + line;
    return true;
  }
  if ((c == n - 2 || c == n - 1) && i < c - 2) {
    b.p = per + ((long)(i * 2 + (c - (n - 2))) * 25) * g.L + line;
    return true;
  }
  if ((i == n - 2 || i == n - 1) && c < i - 2) {
    b.p = per2 + ((long)(c * 2 + (i - (n - 2))) * 25) * g.L + line;
    return true;
  }
  return false;
}

__device__ __forceinline__ int partners(const LineGeom &g, int i, int *cand) {
  int k = 0;
  for (int d = 1; d <= 4 && i + d < g.n; d++) cand[k++] = i + d;
  if (g.periodic)
    for (int e = g.n - 2; e < g.n; e++)
      if (e > i + 4) cand[k++] = e;
  return k;
}

// ---------------------------------------------------------------------------
// stage 1: LU of each line (one lane per line)
template <typename T>
__global__ void k_lu_lines(LineGeom g, cx<T> *mat, cx<T> *per, cx<T> *per2) {
  long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= g.L) return;
  const int n = g.n;
  if (g.periodic) {
    // bordered fill-in storage (cpenta1p.f:53-78)
    int np = n + 1;
    for (int i = 0; i < np; i++)
      for (int e = 0; e < 50; e++) {
        per[((long)i * 50 + e) * g.L + line] = cx<T>{0, 0};
        per2[((long)i * 50 + e) * g.L + line] = cx<T>{0, 0};
      }
    for (int e = 0; e < 25; e++) {
      per[((long)(0 * 2 + 0) * 25 + e) * g.L + line] = mat[((long)(0 * 5 + 0) * 25 + e) * g.L + line];
      per[((long)(0 * 2 + 1) * 25 + e) * g.L + line] = mat[((long)(0 * 5 + 1) * 25 + e) * g.L + line];
      per[((long)(1 * 2 + 1) * 25 + e) * g.L + line] = mat[((long)(1 * 5 + 0) * 25 + e) * g.L + line];
      per2[((long)(0 * 2 + 0) * 25 + e) * g.L + line] = mat[((long)((n - 2) * 5 + 4) * 25 + e) * g.L + line];
      per2[((long)(0 * 2 + 1) * 25 + e) * g.L + line] = mat[((long)((n - 1) * 5 + 3) * 25 + e) * g.L + line];
      per2[((long)(1 * 2 + 1) * 25 + e) * g.L + line] = mat[((long)((n - 1) * 5 + 4) * 25 + e) * g.L + line];
    }
  }
  int cand[8];
  for (int i = 0; i < n; i++) {
    Blk<T> D, U[6];
    int ucol[6], nu = 0;
    blk(g, mat, per, per2, line, i, i, D);
    int nc_ = partners(g, i, cand);
    for (int a = 0; a < nc_; a++)
      if (blk(g, mat, per, per2, line, i, cand[a], U[nu])) ucol[nu++] = cand[a];
    for (int nc = 0; nc < 4; nc++)
      for (int l = nc + 1; l < 5; l++) {
        cx<T> mult = -cdiv(D(l, nc), D(nc, nc));
        D(l, nc) = mult;
        for (int q = nc + 1; q < 5; q++) D(l, q) = cfma(mult, D(nc, q), D(l, q));
        for (int u = 0; u < nu; u++)
          for (int q = 0; q < 5; q++) U[u](l, q) = cfma(mult, U[u](nc, q), U[u](l, q));
      }
    for (int a = 0; a < nc_; a++) {
      int t = cand[a];
      Blk<T> Lb;
      if (!blk(g, mat, per, per2, line, t, i, Lb)) continue;
      Blk<T> Dst[6];
      for (int u = 0; u < nu; u++) {
        blk(g, mat, per, per2, line, t, ucol[u], Dst[u]);
        if (g.periodic && !g.exact && i == n - 6 && t == n - 4 && ucol[u] == n - 2)
          Dst[u].p = mat + ((long)(t * 5 + 3) * 25) * g.L + line;
      }
      for (int nc = 0; nc < 5; nc++) {
        cx<T> fact = crecip(D(nc, nc));
        for (int m = 0; m < 5; m++) {
          cx<T> mult = -(Lb(m, nc) * fact);
          Lb(m, nc) = mult;
          for (int q = nc + 1; q < 5; q++) Lb(m, q) = cfma(mult, D(nc, q), Lb(m, q));
          for (int u = 0; u < nu; u++)
            for (int q = 0; q < 5; q++) Dst[u](m, q) = cfma(mult, U[u](nc, q), Dst[u](m, q));
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// periodic (and reference-order) solve, one lane per line, in place on the
// rhs array: element (row i, comp m) of line l at rhs[(l*sL + i*sR)*5 + m]
template <typename T>
__global__ void k_solve_lines(LineGeom g, const cx<T> *mat_, const cx<T> *per_, const cx<T> *per2_,
                              double2 *rhs, long sL, long sR) {
  long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= g.L) return;
  cx<T> *mat = const_cast<cx<T> *>(mat_), *per = const_cast<cx<T> *>(per_),
        *per2 = const_cast<cx<T> *>(per2_);
  const int n = g.n;
  auto X = [&](int i, int d) -> double2 & { return rhs[(line * sL + (long)i * sR) * 5 + d]; };
  int cand[8];
  for (int i = 0; i < n; i++) {
    Blk<T> D;
    blk(g, mat, per, per2, line, i, i, D);
    cx<T> x[5];
    for (int d = 0; d < 5; d++) x[d] = cx<T>{(T)X(i, d).x, (T)X(i, d).y};
    for (int nc = 0; nc < 4; nc++)
      for (int l = nc + 1; l < 5; l++) x[l] = cfma(D(l, nc), x[nc], x[l]);
    for (int d = 0; d < 5; d++) X(i, d) = make_double2(x[d].re, x[d].im);
    int nc_ = partners(g, i, cand);
    for (int a = 0; a < nc_; a++) {
      Blk<T> Lb;
      int t = cand[a];
      if (!blk(g, mat, per, per2, line, t, i, Lb)) continue;
      for (int m = 0; m < 5; m++) {
        cx<T> acc = cx<T>{(T)X(t, m).x, (T)X(t, m).y};
        for (int nc = 0; nc < 5; nc++) acc = cfma(Lb(m, nc), x[nc], acc);
        X(t, m) = make_double2(acc.re, acc.im);
      }
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    Blk<T> D;
    blk(g, mat, per, per2, line, i, i, D);
    cx<T> acc[5];
    for (int d = 0; d < 5; d++) acc[d] = cx<T>{(T)X(i, d).x, (T)X(i, d).y};
    int nc_ = partners(g, i, cand);
    for (int a = 0; a < nc_; a++) {
      Blk<T> U;
      int c = cand[a];
      if (!blk(g, mat, per, per2, line, i, c, U)) continue;
      cx<T> xc[5];
      for (int d = 0; d < 5; d++) xc[d] = cx<T>{(T)X(c, d).x, (T)X(c, d).y};
      for (int m = 0; m < 5; m++)
        for (int q = 0; q < 5; q++) acc[m] = cfma(-U(m, q), xc[q], acc[m]);
    }
    for (int m = 4; m >= 0; m--) {
      for (int q = m + 1; q < 5; q++) acc[m] = cfma(-D(m, q), acc[q], acc[m]);
      acc[m] = cdiv(acc[m], D(m, m));
    }
    for (int d = 0; d < 5; d++) X(i, d) = make_double2(acc[d].re, acc[d].im);
  }
  if (g.periodic)
    for (int d = 0; d < 5; d++) X(n, d) = X(0, d);
}

// ---------------------------------------------------------------------------
// stage 2: records for the fast non-periodic solve
// fwd[(line*n + i)*60 + f]: f = m*10 + q -> N2(m,q); m*10 + 5 + q -> N1(m,q);
//                           50 + k -> strict lower of Linv, k = l*(l-1)/2 + q
// bwd[(line*n + i)*66 + f] (one complex of padding): f = m*10 + q -> Q4(m,q); m*10 + 5 + q -> Q5(m,q);
//                           50 + k -> upper of Uinv (row-major, incl. diagonal)
// tail[line*150 + ...]: 0 row n-2 N(col n-5); 25 row n-1 N(col n-5);
//                       50 row n-1 N(col n-4); 75 row 0 Q(col 3);
//                       100 row 0 Q(col 4); 125 row 1 Q(col 4)   ([m][q])
template <typename TF, typename TR>
__global__ void k_records_bc(LineGeom g, const cx<TF> *mat_, cx<TR> *fwd, cx<TR> *bwd, cx<TR> *tail) {
  long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= g.L) return;
  cx<TF> *mat = const_cast<cx<TF> *>(mat_);
  const int n = g.n;
  for (int i = 0; i < n; i++) {
    Blk<TF> D;
    blk(g, mat, (cx<TF> *)nullptr, (cx<TF> *)nullptr, line, i, i, D);
    // Linv: product of the in-block elimination steps applied to I
    cx<TF> Li[5][5], Ui[5][5];
    for (int r = 0; r < 5; r++)
      for (int c = 0; c < 5; c++) Li[r][c] = cx<TF>{TF(r == c), 0};
    for (int nc = 0; nc < 4; nc++)
      for (int l = nc + 1; l < 5; l++)
        for (int c = 0; c <= nc; c++) Li[l][c] = cfma(D(l, nc), Li[nc][c], Li[l][c]);
    // Uinv by back substitution on identity columns
    for (int r = 0; r < 5; r++)
      for (int c = 0; c < 5; c++) Ui[r][c] = cx<TF>{0, 0};
    for (int c = 0; c < 5; c++)
      for (int m = c; m >= 0; m--) {
        cx<TF> acc = cx<TF>{TF(m == c), 0};
        for (int q = m + 1; q <= c; q++) acc = cfma(-D(m, q), Ui[q][c], acc);
        Ui[m][c] = cdiv(acc, D(m, m));
      }
    cx<TR> *F = fwd + ((long)line * n + i) * 60;
    cx<TR> *B = bwd + ((long)line * n + i) * 66;
    cx<TR> *Tl = tail + line * 150;
    auto Nblk = [&](int col, cx<TR> *dst, int ld) {   // Linv * M(i,col)
      Blk<TF> M;
      bool ok = blk(g, mat, (cx<TF> *)nullptr, (cx<TF> *)nullptr, line, i, col, M);
      for (int m = 0; m < 5; m++)
        for (int q = 0; q < 5; q++) {
          cx<TF> acc{0, 0};
          if (ok)
            for (int k = 0; k <= m; k++) acc = cfma(Li[m][k], M(k, q), acc);
          dst[m * ld + q] = cx<TR>{(TR)acc.re, (TR)acc.im};
        }
    };
    auto Qblk = [&](int col, cx<TR> *dst, int ld) {   // -Uinv * U(i,col)
      Blk<TF> U;
      bool ok = col < n && blk(g, mat, (cx<TF> *)nullptr, (cx<TF> *)nullptr, line, i, col, U);
      for (int m = 0; m < 5; m++)
        for (int q = 0; q < 5; q++) {
          cx<TF> acc{0, 0};
          if (ok)
            for (int k = m; k < 5; k++) acc = cfma(Ui[m][k], U(k, q), acc);
          dst[m * ld + q] = cx<TR>{(TR)-acc.re, (TR)-acc.im};
        }
    };
    Nblk(i - 1, F, 10);
    Nblk(i - 2, F + 5, 10);
    for (int l = 1, k = 0; l < 5; l++)
      for (int q = 0; q < l; q++, k++) F[50 + k] = cx<TR>{(TR)Li[l][q].re, (TR)Li[l][q].im};
    Qblk(i + 1, B, 10);
    Qblk(i + 2, B + 5, 10);
    for (int m = 0, k = 0; m < 5; m++)
      for (int q = m; q < 5; q++, k++) B[50 + k] = cx<TR>{(TR)Ui[m][q].re, (TR)Ui[m][q].im};
    if (i == n - 2) Nblk(n - 5, Tl + 0, 5);
    if (i == n - 1) { Nblk(n - 5, Tl + 25, 5); Nblk(n - 4, Tl + 50, 5); }
    if (i == 0) { Qblk(3, Tl + 75, 5); Qblk(4, Tl + 100, 5); }
    if (i == 1) Qblk(4, Tl + 125, 5);
  }
}

// ---------------------------------------------------------------------------
// stage 3: streaming solve of non-periodic lines
//
// LDS per buffer: nr record rows copied verbatim from HBM (fwd: 60, bwd: 65
// complex per row) with global_load_lds_dwordx4 (no VGPR round trip), plus
// the parallel parts c_i = Linv_i b_i / d_i = Uinv_i y_i computed by the
// loader waves into cb.
typedef __attribute__((address_space(3))) void lds_void;

template <typename T>
__device__ __forceinline__ void copy_rows_to_lds(const cx<T> *src, long src_elems_left, cx<T> *dst,
                                                 int nelem, int wave_in, int nwaves, int lane) {
  // copy nelem complex values src[0..nelem) -> dst[0..nelem) in 1 KB blocks
  // per wave instruction; lanes past the end read a clamped (valid) address
  // and write into the buffer's padding
  constexpr int PER_LANE = 16 / sizeof(cx<T>);          // complex values per 16 B
  const int nblk = (nelem * (int)sizeof(cx<T>) + 1023) / 1024;
  for (int b = wave_in; b < nblk; b += nwaves) {
    long e = (long)b * 64 * PER_LANE + (long)lane * PER_LANE;
    long ec = e < src_elems_left - PER_LANE ? e : src_elems_left - PER_LANE;
    if (ec < 0) ec = 0;
    __builtin_amdgcn_global_load_lds((const void *)(src + ec),
                                     (lds_void *)(dst + (long)b * 64 * PER_LANE), 16, 0, 0);
  }
}

template <typename T, int R>
__global__ __launch_bounds__(256) void k_solve_bc(int n, const cx<T> *__restrict__ fwd,
                                                 const cx<T> *__restrict__ bwd,
                                                 const cx<T> *__restrict__ tail, double2 *rhs,
                                                 long sL, long sR, long total_rows) {
  constexpr int RECW = 66;                                       // LDS row stride (complex)
  constexpr int BUFE = ((R * RECW * (int)sizeof(cx<T>) + 1023) / 1024) * 1024 / (int)sizeof(cx<T>);
  __shared__ __attribute__((aligned(1024))) cx<T> rec[2][BUFE];
  __shared__ cx<T> cb[2][R][5];
  __shared__ cx<T> saved[2][5];   // y_{n-5}, y_{n-4} (forward); x_3, x_4 (backward)
  const long line = blockIdx.x;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m = lane < 5 ? lane : 4;
  const int nchunk = (n + R - 1) / R;
  double2 *X = rhs + line * sL * 5;
  const long rs = sR * 5;
  const cx<T> *tl = tail + line * 150;

  // ---------------- forward: y_i = c_i + N1 y_{i-2} + N2 y_{i-1} ----------
  auto load_fwd = [&](int k) {
    int r0 = k * R, nr = min(R, n - r0);
    const cx<T> *F = fwd + ((long)line * n + r0) * 60;
    copy_rows_to_lds(F, (total_rows - ((long)line * n + r0)) * 60, rec[k & 1], nr * 60, wave - 1, 3,
                     lane);
    for (int t = tid - 64; t < nr * 5; t += 192) {
      int row = t / 5, mm = t - row * 5;
      const double2 *b = X + (long)(r0 + row) * rs;
      const cx<T> *Li = F + (long)row * 60 + 50 + mm * (mm - 1) / 2;
      double2 bm = b[mm];
      cx<T> acc{(T)bm.x, (T)bm.y};
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (q < mm) {
          double2 bq = b[q];
          acc = cfma(Li[q], cx<T>{(T)bq.x, (T)bq.y}, acc);
        }
      cb[k & 1][row][mm] = acc;
    }
    __builtin_amdgcn_s_waitcnt(0);
  };
  cx<T> y1[5], y2[5];
#pragma unroll
  for (int q = 0; q < 5; q++) { y1[q] = cx<T>{0, 0}; y2[q] = cx<T>{0, 0}; }
  if (wave != 0) load_fwd(0);
  __syncthreads();
  // one row of the recurrence on lanes 0..4 (lane m = component m):
  // out = c + W1 * ym2 + W2 * ym1, then out is broadcast into ym2 (which
  // held the row two back and is dead afterwards)
  auto row_step = [&](const cx<T> (&w)[10], cx<T> c, cx<T> (&ym2)[5], const cx<T> (&ym1)[5],
                      int i, bool fwd_pass) {
    cx<T> a0 = c, p0{0, 0}, p1{0, 0};
#pragma unroll
    for (int q = 0; q < 5; q++) a0 = cfma(w[5 + q], ym2[q], a0);
    p0 = cfma(w[0], ym1[0], p0); p1 = cfma(w[1], ym1[1], p1);
    p0 = cfma(w[2], ym1[2], p0); p1 = cfma(w[3], ym1[3], p1);
    p0 = cfma(w[4], ym1[4], p0);
    if (fwd_pass) {
      if (i >= n - 2) {      // rows n-2, n-1: blocks in columns n-5, n-4
        const cx<T> *Xa = tl + (i == n - 2 ? 0 : 25) + m * 5;
#pragma unroll
        for (int q = 0; q < 5; q++) a0 = cfma(Xa[q], saved[0][q], a0);
        if (i == n - 1)
#pragma unroll
          for (int q = 0; q < 5; q++) a0 = cfma(tl[50 + m * 5 + q], saved[1][q], a0);
      }
    } else if (i <= 1) {     // rows 1, 0: blocks in columns 3, 4
      const cx<T> *Qa = tl + (i == 1 ? 125 : 100) + m * 5;
#pragma unroll
      for (int q = 0; q < 5; q++) a0 = cfma(Qa[q], saved[1][q], a0);
      if (i == 0)
#pragma unroll
        for (int q = 0; q < 5; q++) a0 = cfma(tl[75 + m * 5 + q], saved[0][q], a0);
    }
    cx<T> acc = (a0 + p1) + p0;
    if (lane < 5) X[(long)i * rs + lane] = make_double2(acc.re, acc.im);
    if (lane < 5) {
      if (fwd_pass && (i == n - 5 || i == n - 4)) saved[i - (n - 5)][lane] = acc;
      if (!fwd_pass && (i == 3 || i == 4)) saved[i - 3][lane] = acc;
    }
    row_bcast5(acc, ym2);
  };
  auto lds_row = [&](const cx<T> *B, int stride, int row, int kk, cx<T> (&w)[10], cx<T> &c) {
    const cx<T> *p = B + row * stride + m * 10;
#pragma unroll
    for (int q = 0; q < 10; q++) w[q] = p[q];
    c = cb[kk][row][m];
  };
  cx<T> wA[10], wB[10], cA, cB;
  for (int k = 0; k < nchunk; k++) {
    if (wave == 0) {
      int r0 = k * R, nr = min(R, n - r0);
      const cx<T> *B = rec[k & 1];
      lds_row(B, 60, 0, k & 1, wA, cA);
      for (int row = 0; row < nr; row += 2) {
        // prefetches are unconditional (row index clamped into the chunk)
        lds_row(B, 60, min(row + 1, nr - 1), k & 1, wB, cB);
        row_step(wA, cA, y2, y1, r0 + row, true);             // -> y2 holds y_i
        lds_row(B, 60, min(row + 2, nr - 1), k & 1, wA, cA);
        if (row + 1 < nr) row_step(wB, cB, y1, y2, r0 + row + 1, true);
      }
      if (nr & 1) {          // odd tail: restore (older, newer) order
#pragma unroll
        for (int q = 0; q < 5; q++) { cx<T> t = y1[q]; y1[q] = y2[q]; y2[q] = t; }
      }
    } else if (k + 1 < nchunk) {
      load_fwd(k + 1);
    }
    __syncthreads();
  }

  // ---------------- backward: x_i = d_i + Q4 x_{i+1} + Q5 x_{i+2} ---------
  // chunk k holds rows rlo..rtop in ascending order in LDS
  auto load_bwd = [&](int k) {
    int rtop = n - 1 - k * R, nr = min(R, rtop + 1), rlo = rtop - nr + 1;
    const cx<T> *Bw = bwd + ((long)line * n + rlo) * 66;
    copy_rows_to_lds(Bw, (total_rows - ((long)line * n + rlo)) * 66, rec[k & 1], nr * 66, wave - 1, 3,
                     lane);
    for (int t = tid - 64; t < nr * 5; t += 192) {
      int row = t / 5, mm = t - row * 5;     // row counted from rlo
      const double2 *y = X + (long)(rlo + row) * rs;
      const cx<T> *Ui = Bw + (long)row * 66 + 50 + mm * 5 - mm * (mm - 1) / 2;
      cx<T> acc{0, 0};
#pragma unroll
      for (int q = 0; q < 5; q++)
        if (q < 5 - mm) {
          double2 yq = y[mm + q];
          acc = cfma(Ui[q], cx<T>{(T)yq.x, (T)yq.y}, acc);
        }
      cb[k & 1][row][mm] = acc;
    }
    __builtin_amdgcn_s_waitcnt(0);
  };
#pragma unroll
  for (int q = 0; q < 5; q++) { y1[q] = cx<T>{0, 0}; y2[q] = cx<T>{0, 0}; }
  if (wave != 0) load_bwd(0);
  __syncthreads();
  for (int k = 0; k < nchunk; k++) {
    if (wave == 0) {
      int rtop = n - 1 - k * R, nr = min(R, rtop + 1), rlo = rtop - nr + 1;
      const cx<T> *B = rec[k & 1];
      lds_row(B, 66, nr - 1, k & 1, wA, cA);
      for (int j = 0; j < nr; j += 2) {       // j counts rows from the top of the chunk
        int row = nr - 1 - j;
        lds_row(B, 66, max(row - 1, 0), k & 1, wB, cB);
        row_step(wA, cA, y2, y1, rlo + row, false);
        lds_row(B, 66, max(row - 2, 0), k & 1, wA, cA);
        if (j + 1 < nr) row_step(wB, cB, y1, y2, rlo + row - 1, false);
      }
      if (nr & 1) {
#pragma unroll
        for (int q = 0; q < 5; q++) { cx<T> t = y1[q]; y1[q] = y2[q]; y2[q] = t; }
      }
    } else if (k + 1 < nchunk) {
      load_bwd(k + 1);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// layout conversions between the reference's arrays and the working copy
// ref mat (5,5,5,...) element (slot k, r, c) of (line l, row i) at
//   ((i*sRm + l*sLm) * 125 + k + 5r + 25c)
__global__ void k_import_mat(const double2 *ref, cx<double> *lu, int n, long L, long sLm, long sRm) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;   // over (i, k, r, c, line)
  long total = (long)n * 125 * L;
  if (idx >= total) return;
  long line = idx % L;
  long e = idx / L;            // (i*5 + k)*25 + r*5 + c
  int c = e % 5, r = (e / 5) % 5, k = (e / 25) % 5;
  long i = e / 125;
  double2 v = ref[(i * sRm + line * sLm) * 125 + k + 5 * r + 25 * c];
  lu[idx] = cx<double>{v.x, v.y};
}

// ---------------------------------------------------------------------------
// host side
static int grid_for(long n, int b) { return (int)((n + b - 1) / b); }

int penta_plan_create(PentaPlan *pl, int periodic, int exact, int n_rows, long lines, int prec) {
  pl->periodic = periodic;
  pl->exact = exact;
  pl->np = n_rows;
  pl->n = periodic ? n_rows - 1 : n_rows;
  pl->L = lines;
  pl->prec = prec;
  if (pl->n < 7 || lines < 1) return set_error(__FILE__, __LINE__, "line length must be >= 7 unknowns");
  size_t ne = (size_t)pl->np * 125 * lines;
  LNS3D_CHECK(hipMalloc(&pl->lu, ne * sizeof(cx<double>)));
  pl->per = pl->per2 = nullptr;
  pl->fwd = pl->bwd = pl->tail = nullptr;
  if (periodic) {
    LNS3D_CHECK(hipMalloc(&pl->per, (size_t)pl->np * 50 * lines * sizeof(cx<double>)));
    LNS3D_CHECK(hipMalloc(&pl->per2, (size_t)pl->np * 50 * lines * sizeof(cx<double>)));
  } else {
    size_t es = prec ? sizeof(cx<float>) : sizeof(cx<double>);
    LNS3D_CHECK(hipMalloc(&pl->fwd, (size_t)pl->n * 60 * lines * es));
    LNS3D_CHECK(hipMalloc(&pl->bwd, (size_t)pl->n * 66 * lines * es));
    LNS3D_CHECK(hipMalloc(&pl->tail, (size_t)150 * lines * es));
  }
  return 0;
}

void penta_plan_destroy(PentaPlan *pl) {
  if (pl->lu) (void)hipFree(pl->lu);
  if (pl->per) (void)hipFree(pl->per);
  if (pl->per2) (void)hipFree(pl->per2);
  if (pl->fwd) (void)hipFree(pl->fwd);
  if (pl->bwd) (void)hipFree(pl->bwd);
  if (pl->tail) (void)hipFree(pl->tail);
  pl->lu = pl->per = pl->per2 = pl->fwd = pl->bwd = pl->tail = nullptr;
}

int penta_import(PentaPlan *pl, const void *ref_mat, long sLm, long sRm, hipStream_t s) {
  long total = (long)pl->np * 125 * pl->L;
  hipLaunchKernelGGL(k_import_mat, dim3(grid_for(total, 256)), dim3(256), 0, s,
                     (const double2 *)ref_mat, (cx<double> *)pl->lu, pl->np, pl->L, sLm, sRm);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int penta_factor(PentaPlan *pl, hipStream_t s) {
  LineGeom g{pl->n, pl->periodic, pl->exact, pl->L};
  hipLaunchKernelGGL(k_lu_lines<double>, dim3(grid_for(pl->L, 64)), dim3(64), 0, s, g,
                     (cx<double> *)pl->lu, (cx<double> *)pl->per, (cx<double> *)pl->per2);
  LNS3D_CHECK(hipGetLastError());
  if (!pl->periodic) {
    if (pl->prec)
      hipLaunchKernelGGL((k_records_bc<double, float>), dim3(grid_for(pl->L, 64)), dim3(64), 0, s, g,
                         (const cx<double> *)pl->lu, (cx<float> *)pl->fwd, (cx<float> *)pl->bwd,
                         (cx<float> *)pl->tail);
    else
      hipLaunchKernelGGL((k_records_bc<double, double>), dim3(grid_for(pl->L, 64)), dim3(64), 0, s, g,
                         (const cx<double> *)pl->lu, (cx<double> *)pl->fwd, (cx<double> *)pl->bwd,
                         (cx<double> *)pl->tail);
    LNS3D_CHECK(hipGetLastError());
  }
  return 0;
}

int penta_factor_records_only(PentaPlan *pl, hipStream_t s) {
  if (pl->periodic) return set_error(__FILE__, __LINE__, "records are for non-periodic lines");
  LineGeom g{pl->n, pl->periodic, pl->exact, pl->L};
  if (pl->prec)
    hipLaunchKernelGGL((k_records_bc<double, float>), dim3(grid_for(pl->L, 64)), dim3(64), 0, s, g,
                       (const cx<double> *)pl->lu, (cx<float> *)pl->fwd, (cx<float> *)pl->bwd,
                       (cx<float> *)pl->tail);
  else
    hipLaunchKernelGGL((k_records_bc<double, double>), dim3(grid_for(pl->L, 64)), dim3(64), 0, s, g,
                       (const cx<double> *)pl->lu, (cx<double> *)pl->fwd, (cx<double> *)pl->bwd,
                       (cx<double> *)pl->tail);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int penta_solve(PentaPlan *pl, void *rhs, long sL, long sR, hipStream_t s) {
  LineGeom g{pl->n, pl->periodic, pl->exact, pl->L};
  if (pl->periodic) {
    hipLaunchKernelGGL(k_solve_lines<double>, dim3(grid_for(pl->L, 64)), dim3(64), 0, s, g,
                       (const cx<double> *)pl->lu, (const cx<double> *)pl->per,
                       (const cx<double> *)pl->per2, (double2 *)rhs, sL, sR);
  } else {
    long tot = (long)pl->n * pl->L;
    bool many = pl->L >= 512;    // several lines per CU: smaller LDS ring
    if (pl->prec) {
      auto *f = (const cx<float> *)pl->fwd; auto *b = (const cx<float> *)pl->bwd;
      auto *t = (const cx<float> *)pl->tail;
      if (many) hipLaunchKernelGGL((k_solve_bc<float, 28>), dim3((unsigned)pl->L), dim3(256), 0, s, pl->n, f, b, t, (double2 *)rhs, sL, sR, tot);
      else hipLaunchKernelGGL((k_solve_bc<float, 56>), dim3((unsigned)pl->L), dim3(256), 0, s, pl->n, f, b, t, (double2 *)rhs, sL, sR, tot);
    } else {
      auto *f = (const cx<double> *)pl->fwd; auto *b = (const cx<double> *)pl->bwd;
      auto *t = (const cx<double> *)pl->tail;
      if (many) hipLaunchKernelGGL((k_solve_bc<double, 14>), dim3((unsigned)pl->L), dim3(256), 0, s, pl->n, f, b, t, (double2 *)rhs, sL, sR, tot);
      else hipLaunchKernelGGL((k_solve_bc<double, 28>), dim3((unsigned)pl->L), dim3(256), 0, s, pl->n, f, b, t, (double2 *)rhs, sL, sR, tot);
    }
  }
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int penta_solve_reference_order(PentaPlan *pl, void *rhs, long sL, long sR, hipStream_t s) {
  LineGeom g{pl->n, pl->periodic, pl->exact, pl->L};
  hipLaunchKernelGGL(k_solve_lines<double>, dim3(grid_for(pl->L, 64)), dim3(64), 0, s, g,
                     (const cx<double> *)pl->lu, (const cx<double> *)pl->per,
                     (const cx<double> *)pl->per2, (double2 *)rhs, sL, sR);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

}  // namespace lns3d
