// Batched complex 5x5 block-pentadiagonal line solvers for gfx950.
//
// Replaces src/cpenta1bc.f, src/cpenta2bc.f (non-periodic, boundary
// wrap-around storage) and src/cpenta1p.f, src/cpenta2p.f (periodic,
// bordered).  Three stages:
//
//  1. Non-pivoted block LU.  Non-periodic lines: k_factor_chains, one wave
//     per chain of rows with the rows being updated in an LDS window (lines
//     of 16+ rows, when there are few lines, use a twisted order: two chains
//     per line meeting in the middle).  Periodic lines: k_lu_per, one wave
//     per line on a line-interleaved copy.  The storage of negated
//     multipliers and the in-pivot order follow the reference
//     (cpenta1bc.f:43-495, cpenta1p.f:46-654), including the periodic
//     factor's slot-4 placement at cpenta1p.f:284/345 unless exact is
//     requested.
//  2. k_records_rows: for non-periodic lines, turn row i's factors into the
//     records the fast solve streams:
//        forward   y_i = Linv_i b_i + N1_i y_{i-2} + N2_i y_{i-1}
//        backward  x_i = Uinv_i y_i + Q4_i x_{i+1} + Q5_i x_{i+2}
//     with Linv_i/Uinv_i the inverses of the in-block unit-lower / upper
//     factors, N = Linv * (negated multiplier block), Q = -Uinv * (upper
//     block).  Mathematically the reference's forward/back substitution
//     (cpenta1bc.f:499-847); the per-row critical path becomes one 5x5
//     matvec.  Boundary rows 0,1 and n-2,n-1 carry extra blocks, kept in a
//     per-line tail record.
//  3. k_solve_tw (twisted lines) / k_solve_bc: waves 1-3 stream chunks of
//     records into a double-buffered LDS ring and compute the parallel parts
//     (Linv b, Uinv y); wave 0 runs the serial recurrences, one line or one
//     chain per 16-lane DPP row, lanes 0..4 owning one component each.
//  Periodic lines: k_records_per forms the same records from the bordered
//  LU plus two border blocks per row (p2: rows n-2, n-1 in column i; qp:
//  row i in columns n-2, n-1), and k_solve_per streams them like k_solve_bc
//  with the border sums formed by the loader waves.  k_solve_lines (one lane
//  per line, the reference's substitution order on the LU working copy) is
//  kept as the order-exact solve behind lns3d_penta_solve_reference_order.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "penta.hpp"

namespace lns3d {

// ---------------------------------------------------------------------------
// block addressing in the LU working copy.  Periodic lines: line-interleaved
// ([row][slot][r][c][line], one wave per line in k_lu_per); non-periodic
// lines: line-major ([line][row][slot][r][c], one wave per chain in
// k_factor_chains).  Element (r,c) of a block = p[(r*5 + c) * ldl].
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
  int h = -1;    // twisted factor (non-periodic): rows 0..h+1 top-down after
                 // rows n-1..h+2 bottom-up; -1: rows in order 0..n-1
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
    b.p = mat + ((line * n + i) * 5 + k) * 25;
    b.ldl = 1;
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

// rows (and columns) that a pivot on row i eliminates from / updates: the
// not yet pivoted neighbours in the pivot direction (dir +1: below, -1: above)
__device__ __forceinline__ int partners(const LineGeom &g, int i, int dir, int *cand) {
  int k = 0;
  const int hi = (g.h >= 0 && dir > 0) ? g.h + 1 : g.n - 1;   // twisted top phase stops at h+1
  for (int d = 1; d <= 4; d++) {
    int t = i + dir * d;
    if (t >= 0 && t <= hi) cand[k++] = t;
  }
  if (g.periodic)
    for (int e = g.n - 2; e < g.n; e++)
      if (e > i + 4) cand[k++] = e;
  return k;
}


// ---------------------------------------------------------------------------
// stage 1 for periodic lines, one wave per line: the bordered block LU in
// the reference's order (cpenta1p.f:46-654): pivots 0..n-1, each with its
// partner rows i+1..i+4 and the border rows n-2, n-1 (partners()), the
// storage of blk() and the fill-in placement of the shipped factor unless
// exact is requested; every element receives its operations in the
// reference's order.  Per pivot, the blocks it touches
// -- the pivot row's diagonal and right blocks, and each partner row's
// blocks in the pivot column and in the right blocks' columns (at most 49)
// -- are brought into LDS, eliminated there (the in-block steps and the
// partner steps each in two phases: multipliers, then one element per
// task), and written back.
constexpr int LUP_SLOTS = 49;    // 0 D, 1..6 right blocks, 7 + 7a + b partner a (b 0: pivot column)
__global__ __launch_bounds__(64) void k_lu_per(LineGeom g, cx<double> *mat, cx<double> *per, cx<double> *per2) {
  using T = double;
  __shared__ cx<T> blkv[LUP_SLOTS][25];
  __shared__ cx<T> *bptr[LUP_SLOTS];      // the slots' blocks in HBM (null: no such block)
  __shared__ cx<T> mlt[6][5];
  const long line = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = g.n, np = n + 1;
  const long L = g.L;
  // bordered fill-in storage (cpenta1p.f:53-78)
  for (long e = lane; e < (long)np * 50; e += 64) {
    per[e * L + line] = cx<T>{0, 0};
    per2[e * L + line] = cx<T>{0, 0};
  }
  __builtin_amdgcn_s_waitcnt(0);
  wave_sync();
  if (lane < 25) {
    const int e = lane;
    per[((long)(0 * 2 + 0) * 25 + e) * L + line] = mat[((long)(0 * 5 + 0) * 25 + e) * L + line];
    per[((long)(0 * 2 + 1) * 25 + e) * L + line] = mat[((long)(0 * 5 + 1) * 25 + e) * L + line];
    per[((long)(1 * 2 + 1) * 25 + e) * L + line] = mat[((long)(1 * 5 + 0) * 25 + e) * L + line];
    per2[((long)(0 * 2 + 0) * 25 + e) * L + line] = mat[((long)((n - 2) * 5 + 4) * 25 + e) * L + line];
    per2[((long)(0 * 2 + 1) * 25 + e) * L + line] = mat[((long)((n - 1) * 5 + 3) * 25 + e) * L + line];
    per2[((long)(1 * 2 + 1) * 25 + e) * L + line] = mat[((long)((n - 1) * 5 + 4) * 25 + e) * L + line];
  }
  __builtin_amdgcn_s_waitcnt(0);
  wave_sync();
  for (int i = 0; i < n; i++) {
    // candidates a = 0..5: rows / columns i+1..i+4 in the line, then the
    // border rows / columns n-2, n-1 beyond i+4 (partners(), dir +1)
    auto cand = [&](int a) {
      const int t = a < 4 ? i + 1 + a : n - 6 + a;
      const bool ok = a < 4 ? t <= n - 1 : t > i + 4;
      return ok ? t : -1;
    };
    // where slot k lives in HBM (false: no such block); uniform over the wave
    auto where = [&](int k, Blk<T> &B) -> bool {
      if (k == 0) return blk(g, mat, per, per2, line, i, i, B);
      if (k <= 6) {
        const int c = cand(k - 1);
        return c >= 0 && blk(g, mat, per, per2, line, i, c, B);
      }
      const int a = (k - 7) / 7, bb = (k - 7) % 7, t = cand(a);
      Blk<T> Lb;
      if (t < 0 || !blk(g, mat, per, per2, line, t, i, Lb)) return false;   // not a partner
      if (bb == 0) { B = Lb; return true; }
      Blk<T> U;
      const int c = cand(bb - 1);
      if (c < 0 || !blk(g, mat, per, per2, line, i, c, U)) return false;    // no right block there
      if (!blk(g, mat, per, per2, line, t, c, B)) return false;
      if (!g.exact && i == n - 6 && t == n - 4 && c == n - 2)             // cpenta1p.f:284, 345
        B.p = mat + ((long)(t * 5 + 3) * 25) * L + line;
      return true;
    };
    // the slots' addresses, one lane per slot, then the blocks into LDS (all
    // loads in flight at once)
    if (lane < LUP_SLOTS) {
      Blk<T> B;
      bptr[lane] = where(lane, B) ? B.p : nullptr;
    }
    wave_sync();
    constexpr int PER_LANE = (LUP_SLOTS * 25 + 63) / 64;
    cx<T> tmp[PER_LANE];
#pragma unroll
    for (int r = 0; r < PER_LANE; r++) {
      const int e = r * 64 + lane, k = e / 25, el = e % 25;
      const cx<T> *bp = e < LUP_SLOTS * 25 ? bptr[k] : nullptr;
      tmp[r] = bp ? bp[(long)el * L] : cx<T>{0, 0};
    }
#pragma unroll
    for (int r = 0; r < PER_LANE; r++) {
      const int e = r * 64 + lane;
      if (e < LUP_SLOTS * 25) blkv[e / 25][e % 25] = tmp[r];
    }
    wave_sync();
    // slot existence, uniform: right blocks and partners
    unsigned uok = 0, pok = 0;
#pragma unroll
    for (int u = 0; u < 6; u++) {
      if (bptr[1 + u]) uok |= 1u << u;
      if (bptr[7 + 7 * u]) pok |= 1u << u;
    }
    cx<T> *D = blkv[0];
    // in-block LU of D, its row operations applied to the right blocks
    for (int nc = 0; nc < 4; nc++) {
      if (lane < 4 - nc) {
        const int l = nc + 1 + lane;
        const cx<T> mult = -cdiv(D[l * 5 + nc], D[nc * 5 + nc]);
        mlt[0][l] = mult;
        D[l * 5 + nc] = mult;
      }
      wave_sync();
      const int nt = (4 - nc) * 5 * 7;          // tasks (l, q, b): b 0 D (q > nc), 1 + u right block u
      for (int tk = lane; tk < nt; tk += 64) {
        const int l = nc + 1 + tk / 35, q = (tk / 7) % 5, b = tk % 7;
        const cx<T> mult = mlt[0][l];
        if (b == 0) {
          if (q > nc) D[l * 5 + q] = cfma(mult, D[nc * 5 + q], D[l * 5 + q]);
        } else if (uok >> (b - 1) & 1) {
          cx<T> *U = blkv[b];
          U[l * 5 + q] = cfma(mult, U[nc * 5 + q], U[l * 5 + q]);
        }
      }
      wave_sync();
    }
    // partner rows: the block in the pivot column becomes the multipliers,
    // the partner's blocks in the right blocks' columns are updated
    for (int nc = 0; nc < 5; nc++) {
      if (lane < 30) {
        const int a = lane / 5, m = lane % 5;
        if (pok >> a & 1) {
          cx<T> *Lb = blkv[7 + 7 * a];
          const cx<T> fact = crecip(D[nc * 5 + nc]);
          const cx<T> mult = -(Lb[m * 5 + nc] * fact);
          mlt[a][m] = mult;
          Lb[m * 5 + nc] = mult;
        }
      }
      wave_sync();
      const int nt = 6 * 5 * 5 * 7;             // tasks (a, m, q, b)
      for (int tk = lane; tk < nt; tk += 64) {
        const int a = tk / 175, m = (tk / 35) % 5, q = (tk / 7) % 5, b = tk % 7;
        if (!(pok >> a & 1)) continue;
        const cx<T> mult = mlt[a][m];
        cx<T> *Lb = blkv[7 + 7 * a];
        if (b == 0) {
          if (q > nc) Lb[m * 5 + q] = cfma(mult, D[nc * 5 + q], Lb[m * 5 + q]);
        } else if (uok >> (b - 1) & 1) {
          cx<T> *Dst = blkv[7 + 7 * a + b];
          Dst[m * 5 + q] = cfma(mult, blkv[b][nc * 5 + q], Dst[m * 5 + q]);
        }
      }
      wave_sync();
    }
    // write the blocks back; the next pivot reads them after the stores
#pragma unroll
    for (int r = 0; r < PER_LANE; r++) {
      const int e = r * 64 + lane, k = e / 25, el = e % 25;
      cx<T> *bp = e < LUP_SLOTS * 25 ? bptr[k] : nullptr;
      if (bp) bp[(long)el * L] = blkv[k][el];
    }
    __builtin_amdgcn_s_waitcnt(0);
    wave_sync();
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
    int nc_ = partners(g, i, 1, cand);
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
    int nc_ = partners(g, i, 1, cand);
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
// Records of (line, chain row)s: eight per workgroup, 32 lanes each (a row
// and its inverses in LDS; lane (m, q) = (l / 5, l % 5) of lanes 0..24 forms
// entry (m, q) of a block).  The row's five LU blocks are read with
// consecutive lanes on consecutive entries.
template <typename TF, typename TR>
__global__ __launch_bounds__(256) void k_records_rows(LineGeom g, const cx<TF> *mat, cx<TR> *fwd, cx<TR> *bwd,
                                                      cx<TR> *tail) {
  __shared__ cx<TF> rowb[8][125];
  __shared__ cx<TF> Lis[8][25], Uis[8][25];
  const int n = g.n;
  const int grp = threadIdx.x >> 5, gl = threadIdx.x & 31;
  const long idx0 = (long)blockIdx.x * 8 + grp;
  const bool valid = idx0 < g.L * n;
  const long idx = valid ? idx0 : g.L * n - 1;        // idle groups repeat the last row, store nothing
  const long line = idx / n;
  const int jj = (int)(idx - line * n);
  const bool tw = g.h >= 0;
  const int nchain = tw ? 2 : 1;
  // chain t: rows in elimination order; t = 1 runs bottom-up (row n-1-j)
  const int t = (tw && jj >= g.h + 2) ? 1 : 0;
  const int j = t == 0 ? jj : jj - (g.h + 2);
  const int len = tw ? (t == 0 ? g.h + 2 : n - g.h - 2) : n;
  const int dir = t == 0 ? 1 : -1;
  const int i = t == 0 ? j : n - 1 - j;
  cx<TF> *rb = rowb[grp];
  const cx<TF> *src = mat + ((line * n + i) * 125);
  for (int e = gl; e < 125; e += 32) rb[e] = src[e];
  const int m = gl / 5, q = gl % 5;
  const bool act = gl < 25;
  const int sd = bc_slot(n, i, i) * 25;
  if (act) Lis[grp][gl] = cx<TF>{TF(m == q), 0};
  wave_sync();
  // Linv: the in-block elimination steps applied to I
#pragma unroll
  for (int nc = 0; nc < 4; nc++) {
    if (act && m > nc && q <= nc) Lis[grp][gl] = cfma(rb[sd + m * 5 + nc], Lis[grp][nc * 5 + q], Lis[grp][gl]);
    wave_sync();
  }
  // Uinv by back substitution on identity columns (lane c < 5: column c)
  if (gl < 5) {
    const int c = gl;
    cx<TF> u[5];
#pragma unroll
    for (int r = 0; r < 5; r++) u[r] = cx<TF>{0, 0};
#pragma unroll
    for (int mm = 4; mm >= 0; mm--)
      if (mm <= c) {
        cx<TF> acc = cx<TF>{TF(mm == c), 0};
#pragma unroll
        for (int k = mm + 1; k < 5; k++)
          if (k <= c) acc = cfma(-rb[sd + mm * 5 + k], u[k], acc);
        u[mm] = cdiv(acc, rb[sd + mm * 5 + mm]);
      }
#pragma unroll
    for (int r = 0; r < 5; r++) Uis[grp][r * 5 + c] = u[r];
  }
  wave_sync();
  const cx<TF> *Li = Lis[grp], *Ui = Uis[grp];
  // entry (m, q) of Linv * M(i, col) (a multiplier block), zero if col has none
  auto Nent = [&](int col) {
    cx<TF> acc{0, 0};
    const int sl = (col >= 0 && col < n) ? bc_slot(n, i, col) : -1;
    if (sl >= 0)
      for (int k = 0; k <= m; k++) acc = cfma(Li[m * 5 + k], rb[sl * 25 + k * 5 + q], acc);
    return cx<TR>{(TR)acc.re, (TR)acc.im};
  };
  // entry (m, q) of -Uinv * U(i, col), col pivoted after i
  auto Qent = [&](int col) {
    cx<TF> acc{0, 0};
    const bool later = !(tw && t == 0 && col >= g.h + 2);
    const int sl = (later && col >= 0 && col < n) ? bc_slot(n, i, col) : -1;
    if (sl >= 0)
      for (int k = m; k < 5; k++) acc = cfma(Ui[m * 5 + k], rb[sl * 25 + k * 5 + q], acc);
    return cx<TR>{(TR)-acc.re, (TR)-acc.im};
  };
  cx<TR> *F = fwd + idx * 60;
  cx<TR> *B = bwd + idx * 66;
  cx<TR> *Tl = tail + (line * nchain + t) * 150;
  if (valid && act) {
    F[m * 10 + q] = Nent(i - dir);
    F[m * 10 + 5 + q] = Nent(i - 2 * dir);
    B[m * 10 + q] = Qent(i + dir);
    B[m * 10 + 5 + q] = Qent(i + 2 * dir);
    // forward blocks of the chain's last two rows outside the band: columns
    // n-5, n-4 (plain) or the bottom chain's last rows h+2, h+3 (twisted);
    // the bottom chain of a twisted factor has none (zeros)
    const bool fw_tail = !(tw && t == 1);
    const int x0 = !fw_tail ? -1 : (tw ? g.h + 2 : n - 5), x1 = !fw_tail ? -1 : (tw ? g.h + 3 : n - 4);
    if (j == (fw_tail ? len - 2 : 0)) Tl[m * 5 + q] = Nent(x0);
    if (j == (fw_tail ? len - 1 : 0)) { Tl[25 + m * 5 + q] = Nent(x0); Tl[50 + m * 5 + q] = Nent(x1); }
    // backward blocks of the chain's first two rows: columns 3, 4 counted
    // along the chain
    const int c3 = t == 0 ? 3 : n - 4, c4 = t == 0 ? 4 : n - 5;
    if (j == 0) { Tl[75 + m * 5 + q] = Qent(c3); Tl[100 + m * 5 + q] = Qent(c4); }
    if (j == 1) Tl[125 + m * 5 + q] = Qent(c4);
  }
  if (valid && gl < 10) {            // strict lower of Linv, row-major
    const int l = gl < 1 ? 1 : gl < 3 ? 2 : gl < 6 ? 3 : 4, c = gl - l * (l - 1) / 2;
    F[50 + gl] = cx<TR>{(TR)Li[l * 5 + c].re, (TR)Li[l * 5 + c].im};
  }
  if (valid && gl < 16) {            // upper of Uinv incl. diagonal, row-major; then the pad
    if (gl < 15) {
      const int r = gl < 5 ? 0 : gl < 9 ? 1 : gl < 12 ? 2 : gl < 14 ? 3 : 4;
      const int c = r + gl - (5 * r - r * (r - 1) / 2);
      B[50 + gl] = cx<TR>{(TR)Ui[r * 5 + c].re, (TR)Ui[r * 5 + c].im};
    } else {
      B[65] = cx<TR>{0, 0};
    }
  }
}

// ---------------------------------------------------------------------------
// Twisted lines take their chain steps in pairs where no hand-off or boundary
// block is involved.  In processing order a step is u_s = c_s + W2_s u_{s-1}
// + W1_s u_{s-2} (forward W2 = N2, W1 = N1; backward W2 = Q4, W1 = Q5).  The
// second step of a pair written in the same inputs as the first,
//   u_{s+1} = c' + W2' u_{s-1} + W1' u_{s-2},
//   W2' = W1_{s+1} + W2_{s+1} W2_s,  W1' = W2_{s+1} W1_s,  c' = c_{s+1} + W2_{s+1} c_s,
// lets the solve form both at once (lanes 0-4 and 5-9 of a DPP row).  The
// factor rewrites the second step's record with W2', W1' and keeps
// M = W2_{s+1} Linv_s (forward) / W2_{s+1} Uinv_s (backward) in a side array
// (xf / xb, 25 complex per pair), so that the solve's loader waves form
// c' = Linv_{s+1} b_{s+1} + M b_s (forward; backward with Uinv and the
// forward result) from the staged right-hand sides in one pass.
//
// Pairs (s, s+1), s even, of chunk k (steps [kR, min(kR+R, len0))) lie in
// [p_end, m_end); the same rule on both sides, as it depends on the solve's
// chunk size R.
__host__ __device__ inline void tw_pair_range(int R, int k, bool fw, int len0, int len1, int &p_end,
                                              int &m_end) {
  const int p0 = fw ? 0 : 4;                                   // even, after the backward hand-off
  const int main_hi = fw ? len1 - 2 : (len0 - 5 < len1 - 3 ? len0 - 5 : len1 - 3);
  const int s_begin = k * R, s_end = (k * R + R < len0) ? k * R + R : len0;
  p_end = s_begin > (p0 < s_end ? p0 : s_end) ? s_begin : (p0 < s_end ? p0 : s_end);
  const int hi = main_hi < s_end ? main_hi : s_end;
  m_end = p_end + ((hi > p_end ? hi - p_end : 0) / 6) * 6;
}
// chunk size of the twisted solve for a record precision
constexpr int tw_chunk(bool single) { return single ? 12 : 6; }

// the pair rewrite at factor time: one 32-lane group per (line, chain, pass,
// pair), lane (m, q) of lanes 0..24 forms entry (m, q) of W2', W1' and M
template <typename TR>
__global__ __launch_bounds__(256) void k_pair_records(int n, int h, long L, int R, cx<TR> *fwd, cx<TR> *bwd,
                                                      cx<TR> *xf, cx<TR> *xb) {
  const int len0 = h + 2, len1 = n - h - 2, npair = (len0 + 1) / 2;
  const int grp = threadIdx.x >> 5, gl = threadIdx.x & 31;
  const long id = (long)blockIdx.x * 8 + grp;     // ((line * 2 + chain) * 2 + pass) * npair + pair
  const long total = L * 4 * npair;
  const long idc = id < total ? id : total - 1;
  const int q2 = (int)(idc % npair), pass = (int)((idc / npair) % 2), cc = (int)((idc / npair / 2) % 2);
  const long line = idc / npair / 4;
  const bool fw = pass == 0;
  const int s = 2 * q2;
  int p_end, m_end;
  tw_pair_range(R, s / R, fw, len0, len1, p_end, m_end);
  const bool is_pair = id < total && s >= p_end && s + 1 < m_end;
  // chain rows of steps s, s+1 and their records
  const int coff = cc * len0;
  const int j0 = fw ? s : len0 - 1 - s + cc * (len1 + 2 - len0);
  const int j1 = fw ? j0 + 1 : j0 - 1;
  const int stride = fw ? 60 : 66;
  cx<TR> *base = fw ? fwd : bwd;
  cx<TR> *r0 = base + (line * n + coff + (is_pair ? j0 : 0)) * stride;
  cx<TR> *r1 = base + (line * n + coff + (is_pair ? j1 : 0)) * stride;
  cx<TR> *x = (fw ? xf : xb) + ((line * 2 + cc) * npair + q2) * 25;
  const int m = gl / 5, q = gl % 5;
  cx<double> w2n{0, 0}, w1n{0, 0}, mx{0, 0};
  if (is_pair && gl < 25) {
    // M(m, q) = (W2_{s+1} Linv_s)(m, q) forward (Linv: unit lower, strict
    // part at 50 + l(l-1)/2 + c), (W2_{s+1} Uinv_s)(m, q) backward (Uinv:
    // upper incl. diagonal at 50 + 5k - k(k-1)/2 + (c - k))
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const cx<double> a{(double)r1[m * 10 + k].re, (double)r1[m * 10 + k].im};
      cx<TR> lk{0, 0};
      if (fw) lk = k == q ? cx<TR>{1, 0} : (k > q ? r0[50 + k * (k - 1) / 2 + q] : cx<TR>{0, 0});
      else if (k <= q) lk = r0[50 + 5 * k - k * (k - 1) / 2 + (q - k)];
      mx = cfma(a, cx<double>{(double)lk.re, (double)lk.im}, mx);
    }
    w2n = cx<double>{(double)r1[m * 10 + 5 + q].re, (double)r1[m * 10 + 5 + q].im};
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const cx<double> a{(double)r1[m * 10 + k].re, (double)r1[m * 10 + k].im};
      const cx<double> b2{(double)r0[k * 10 + q].re, (double)r0[k * 10 + q].im};
      const cx<double> b1{(double)r0[k * 10 + 5 + q].re, (double)r0[k * 10 + 5 + q].im};
      w2n = cfma(a, b2, w2n);
      w1n = cfma(a, b1, w1n);
    }
  }
  wave_sync();
  if (is_pair && gl < 25) {
    x[m * 5 + q] = cx<TR>{(TR)mx.re, (TR)mx.im};
    r1[m * 10 + q] = cx<TR>{(TR)w2n.re, (TR)w2n.im};
    r1[m * 10 + 5 + q] = cx<TR>{(TR)w1n.re, (TR)w1n.im};
  }
}

// ---------------------------------------------------------------------------
// stage 1 for non-periodic lines: block LU with one wave per chain, the
// rows v..v+4 of the chain in an LDS window.  A plain factor is one chain
// (rows 0..n-1).  The twisted factor runs two chains concurrently, rows
// 0..h-1 top-down and rows n-1..h+2 bottom-up (the bottom one in reversed
// indices, where the band looks the same); both update the 2x2 block of
// rows/columns h, h+1, the top chain in place and the bottom chain into a
// scratch copy that starts at zero.  Once both are done the two are summed
// and rows h, h+1 are eliminated.  Inside a pivot the elimination follows
// cpenta1bc.f:43-495: in-block LU with negated multipliers stored below the
// diagonal, its row operations applied to the pivot row's right blocks, then
// each partner row's block in the pivot column turned into multipliers and
// the partner's blocks updated.
constexpr int FW = 5;

__global__ __launch_bounds__(128) void k_factor_chains(LineGeom g, cx<double> *lu, cx<double> *mid) {
  using T = double;
  __shared__ cx<T> win[2][FW][125];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long line = blockIdx.x;
  const int n = g.n, h = g.h;
  const bool tw = h >= 0;
  cx<T> *lul = lu + line * n * 125;
  cx<T> *midl = mid ? mid + line * 4 * 25 : nullptr;
  auto wsync = [] { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); __builtin_amdgcn_wave_barrier(); __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); };
  // one chain on window w, chain direction tt, pivots [v0, v1), rows above
  // lim never touched
  auto chain = [&](int w, int tt, int v0, int v1, int lim, bool preloaded) {
    auto real = [&](int v) { return tt == 0 ? v : n - 1 - v; };
    auto slot = [&](int vr, int vc) {
      if (vr < 0 || vc < 0 || vr >= n || vc >= n) return -1;
      return bc_slot(n, real(vr), real(vc));
    };
    auto B = [&](int vr, int sl) -> cx<T> * { return &win[w][vr % FW][sl * 25]; };
    auto middle = [&](int vr, int vc) {
      int r = real(vr), c = real(vc);
      return tw && (r == h || r == h + 1) && (c == h || c == h + 1);
    };
    // a row into the window: entries e = lane, lane + 64 (x0, x1) of row vr
    auto put_row = [&](int vr, cx<T> x0, cx<T> x1) {
#pragma unroll
      for (int h2 = 0; h2 < 2; h2++) {
        const int e = lane + 64 * h2;
        if (e >= 125) continue;
        cx<T> x = h2 == 0 ? x0 : x1;
        if (tt == 1 && tw && (real(vr) == h || real(vr) == h + 1)) {   // bottom chain: middle blocks start at zero
          int sl = e / 25;
          for (int vc = vr - 4; vc <= vr + 4; vc++)
            if (slot(vr, vc) == sl && middle(vr, vc)) x = cx<T>{0, 0};
        }
        win[w][vr % FW][e] = x;
      }
    };
    auto load_row = [&](int vr) {
      const cx<T> *src = lul + (long)real(vr) * 125;
      put_row(vr, src[lane], lane + 64 < 125 ? src[lane + 64] : cx<T>{0, 0});
    };
    // write back a row: own = all slots (pivot rows) or the chain's share
    auto store_row = [&](int vr, bool all) {
      cx<T> *dst = lul + (long)real(vr) * 125;
      if (all) {
        for (int e = lane; e < 125; e += 64) dst[e] = win[w][vr % FW][e];
        return;
      }
      for (int e = lane; e < 125; e += 64) {
        const int sl = e / 25;
        int vcol = -1;
        for (int vc = vr - 4; vc <= vr + 4; vc++)
          if (slot(vr, vc) == sl) vcol = vc;
        if (vcol < 0) { if (all) dst[e] = win[w][vr % FW][e]; continue; }
        const int c = real(vcol);
        if (all) { dst[e] = win[w][vr % FW][e]; continue; }
        if (middle(vr, vcol)) {
          if (tt == 0) dst[e] = win[w][vr % FW][e];
          else midl[((real(vr) - h) * 2 + (c - h)) * 25 + (e % 25)] = win[w][vr % FW][e];
        } else if ((tt == 0 && c <= h + 1) || (tt == 1 && c >= h + 2)) {
          dst[e] = win[w][vr % FW][e];
        }
      }
    };
    if (!preloaded)
      for (int vr = v0; vr <= min(v0 + FW - 1, lim); vr++) load_row(vr);
    wsync();
    for (int v = v0; v < v1; v++) {
      // row v+5 comes into the window when row v leaves it; its loads are
      // issued now, so that they arrive while row v is eliminated
      const bool pf_ok = v + FW <= lim;
      cx<T> pf0, pf1;
      {
        const cx<T> *src = lul + (long)real(pf_ok ? v + FW : v) * 125;
        pf0 = src[lane];
        pf1 = lane + 64 < 125 ? src[lane + 64] : cx<T>{0, 0};
      }
      const int sd = slot(v, v);
      cx<T> *D = B(v, sd);
      // Interior pivot: rows v..v+4 (those up to lim) are interior rows of
      // the line, and v+2 <= lim.  Then the right blocks are columns v+1,
      // v+2 and the partners rows v+1, v+2, at slots fixed by the chain
      // direction -- the generic path below with its lists in constants
      // (the same operations on every element, in the same order).
      const int sg = tt == 0 ? 1 : -1;
      const bool fast = v >= 2 && v + 2 <= lim && min(v + 4, lim) <= n - 3;
      if (fast) {
        cx<T> *U0 = B(v, 2 + sg), *U1 = B(v, 2 + 2 * sg);
#pragma unroll
        for (int nc = 0; nc < 4; nc++) {
          if (lane < 4 - nc) {
            const int l = nc + 1 + lane;
            D[l * 5 + nc] = -cdiv(D[l * 5 + nc], D[nc * 5 + nc]);
          }
          wsync();
          const int nd = (4 - nc) * (4 - nc), ntask = nd + 2 * (4 - nc) * 5;
          for (int k = lane; k < ntask; k += 64) {
            if (k < nd) {
              const int l = nc + 1 + k / (4 - nc), q = nc + 1 + k % (4 - nc);
              D[l * 5 + q] = cfma(D[l * 5 + nc], D[nc * 5 + q], D[l * 5 + q]);
            } else {
              const int kk = k - nd, u = kk / ((4 - nc) * 5), rem = kk % ((4 - nc) * 5);
              const int l = nc + 1 + rem / 5, q = rem % 5;
              cx<T> *U = u ? U1 : U0;
              U[l * 5 + q] = cfma(D[l * 5 + nc], U[nc * 5 + q], U[l * 5 + q]);
            }
          }
          wsync();
        }
        cx<T> *L0 = B(v + 1, 2 - sg), *L1 = B(v + 2, 2 - 2 * sg);
#pragma unroll
        for (int nc = 0; nc < 5; nc++) {
          const cx<T> fact = crecip(D[nc * 5 + nc]);
          if (lane < 10) {
            cx<T> *Lb = lane < 5 ? L0 : L1;
            const int mm = lane % 5;
            Lb[mm * 5 + nc] = -(Lb[mm * 5 + nc] * fact);
          }
          wsync();
          const int nl = 2 * 5 * (4 - nc), ntask = nl + 2 * 2 * 25;
          for (int k = lane; k < ntask; k += 64) {
            if (k < nl) {
              const int pp = k / (5 * (4 - nc)), rem = k % (5 * (4 - nc));
              const int mm = rem / (4 - nc), q = nc + 1 + rem % (4 - nc);
              cx<T> *Lb = pp ? L1 : L0;
              Lb[mm * 5 + q] = cfma(Lb[mm * 5 + nc], D[nc * 5 + q], Lb[mm * 5 + q]);
            } else {
              const int kk = k - nl, pp = kk / 50, rem = kk % 50;
              const int u = rem / 25, mm = (rem % 25) / 5, q = rem % 5;
              cx<T> *Lb = pp ? L1 : L0;
              cx<T> *Dst = B(v + 1 + pp, 2 + sg * (u - pp));
              const cx<T> *U = u ? U1 : U0;
              Dst[mm * 5 + q] = cfma(Lb[mm * 5 + nc], U[nc * 5 + q], Dst[mm * 5 + q]);
            }
          }
          wsync();
        }
      } else {
        // blocks right of the pivot (columns uc, slots us) and partner rows
        // (pr); filled and read with constant indices only, so that they stay
        // in registers (a run-time index puts a local array in scratch)
        int uc[4] = {0, 0, 0, 0}, us[4] = {0, 0, 0, 0}, pr[4] = {0, 0, 0, 0}, nu = 0, nparts = 0;
#pragma unroll
        for (int d = 1; d <= 4; d++) {
          const int vc = v + d;
          const int su = vc <= lim ? slot(v, vc) : -1;
          const bool hp = vc <= lim && slot(vc, v) >= 0;
#pragma unroll
          for (int i = 0; i < 4; i++) {
            if (su >= 0 && nu == i) { uc[i] = vc; us[i] = su; }
            if (hp && nparts == i) pr[i] = vc;
          }
          nu += su >= 0;
          nparts += hp;
        }
        auto pick = [](const int (&a)[4], int i) { return i == 0 ? a[0] : i == 1 ? a[1] : i == 2 ? a[2] : a[3]; };
        // in-block LU of D, its row operations applied to the right blocks
        // (loops over nc unrolled: the task decodes divide by constants)
#pragma unroll
        for (int nc = 0; nc < 4; nc++) {
          if (lane < 4 - nc) {
            const int l = nc + 1 + lane;
            D[l * 5 + nc] = -cdiv(D[l * 5 + nc], D[nc * 5 + nc]);
          }
          wsync();
          const int nd = (4 - nc) * (4 - nc), ntask = nd + nu * (4 - nc) * 5;
          for (int k = lane; k < ntask; k += 64) {
            if (k < nd) {
              const int l = nc + 1 + k / (4 - nc), q = nc + 1 + k % (4 - nc);
              D[l * 5 + q] = cfma(D[l * 5 + nc], D[nc * 5 + q], D[l * 5 + q]);
            } else {
              const int kk = k - nd, u = kk / ((4 - nc) * 5), rem = kk % ((4 - nc) * 5);
              const int l = nc + 1 + rem / 5, q = rem % 5;
              cx<T> *U = B(v, pick(us, u));
              U[l * 5 + q] = cfma(D[l * 5 + nc], U[nc * 5 + q], U[l * 5 + q]);
            }
          }
          wsync();
        }
        // partner rows: the block in the pivot column becomes the multipliers
#pragma unroll
        for (int nc = 0; nc < 5; nc++) {
          const cx<T> fact = crecip(D[nc * 5 + nc]);
          if (lane < nparts * 5) {
            cx<T> *Lb = B(pick(pr, lane / 5), slot(pick(pr, lane / 5), v));
            const int mm = lane % 5;
            Lb[mm * 5 + nc] = -(Lb[mm * 5 + nc] * fact);
          }
          wsync();
          const int nl = nparts * 5 * (4 - nc), ntask = nl + nparts * nu * 25;
          for (int k = lane; k < ntask; k += 64) {
            if (k < nl) {
              const int pp = k / (5 * (4 - nc)), rem = k % (5 * (4 - nc));
              const int mm = rem / (4 - nc), q = nc + 1 + rem % (4 - nc);
              cx<T> *Lb = B(pick(pr, pp), slot(pick(pr, pp), v));
              Lb[mm * 5 + q] = cfma(Lb[mm * 5 + nc], D[nc * 5 + q], Lb[mm * 5 + q]);
            } else {
              const int kk = k - nl, pp = kk / (nu * 25), rem = kk % (nu * 25);
              const int u = rem / 25, mm = (rem % 25) / 5, q = rem % 5;
              const int sdst = slot(pick(pr, pp), pick(uc, u));
              if (sdst < 0) continue;      // no such block (does not occur in the band)
              cx<T> *Lb = B(pick(pr, pp), slot(pick(pr, pp), v));
              cx<T> *Dst = B(pick(pr, pp), sdst);
              const cx<T> *U = B(v, pick(us, u));
              Dst[mm * 5 + q] = cfma(Lb[mm * 5 + nc], U[nc * 5 + q], Dst[mm * 5 + q]);
            }
          }
          wsync();
        }
      }
      // row v is final: write it back and bring in row v+5
      store_row(v, true);
      wsync();
      if (pf_ok) put_row(v + FW, pf0, pf1);
      wsync();
    }
    // rows the chain updated but did not pivot
    for (int vr = v1; vr <= min(v1 + FW - 1, lim); vr++) store_row(vr, !tw);
  };
  if (!tw) {
    chain(0, 0, 0, n, n - 1, false);
    return;
  }
  if (wave == 0) chain(0, 0, 0, h, h + 1, false);
  else chain(1, 1, 0, n - h - 2, n - 1 - h, false);
  __syncthreads();
  if (wave != 0) return;
  // rows h, h+1 with both chains' updates of their 2x2 middle block summed,
  // then their elimination
  for (int rr = h; rr <= h + 1; rr++)
    for (int e = lane; e < 125; e += 64) win[0][rr % FW][e] = lul[(long)rr * 125 + e];
  wsync();
  for (int e = lane; e < 100; e += 64) {
    const int b = e / 25, r = h + b / 2, c = h + b % 2;
    cx<T> &x = win[0][r % FW][bc_slot(n, r, c) * 25 + e % 25];
    x = x + midl[e];
  }
  chain(0, 0, h, h + 2, h + 1, true);
}

// ---------------------------------------------------------------------------
// stage 3: streaming solve of non-periodic lines
//
// A workgroup of four waves.  Wave 0 runs the recurrences; waves 1-3 stream
// the records through a ring of NBUF = 4 LDS buffers.  Chunk k (R rows of
// each line or chain of the workgroup) belongs to loader wave 1 + k % 3,
// which
//   * in iteration k - 3 starts its copies: the record rows verbatim
//     (global_load_lds_dwordx4, no VGPR round trip; fwd 60, bwd 66 complex
//     per row) and the right-hand-side rows gathered into a staging buffer;
//   * in iteration k - 1 waits for them (the only copies it has in flight)
//     and replaces each staged row by its parallel part c_i = Linv_i b_i
//     (forward) or d_i = Uinv_i y_i (backward).
// A chunk's copies so have two iterations to arrive while the compute wave
// works through the chunks before it.
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

// one wave copies nelem complex values src[0..nelem) into LDS at dst, 1 KB
// (16 B per lane) per instruction; lanes past the end are masked off (nelem
// is even for complex float, so no lane copies past the end)
template <typename T>
__device__ __forceinline__ void dma_rows(const cx<T> *src, cx<T> *dst, int nelem, int lane) {
  constexpr int PER_LANE = 16 / sizeof(cx<T>);
  for (int b = 0; b * 64 * PER_LANE < nelem; b++) {
    const int e = (b * 64 + lane) * PER_LANE;
    if (e < nelem)
      __builtin_amdgcn_global_load_lds((const void *)(src + e), (lds_void *)(dst + b * 64 * PER_LANE),
                                       16, 0, 0);
  }
}

// the parallel part of one row, in place: bp[0..5) holds the row's
// right-hand side (forward) or forward result (backward) in complex double,
// rp the row's record.  Lane mm of the row overwrites bp[mm] with component
// mm of c = Linv b (forward) or d = Uinv y (backward), as cx<T>.  The five
// lanes of a row sit in one wave, so all of them have read the row before
// any writes to it.
template <typename T>
__device__ __forceinline__ void parallel_part(double2 *bp, const cx<T> *rp, int mm, bool fw) {
  cx<T> acc{0, 0};
  if (fw) {
    const cx<T> *Li = rp + 50 + mm * (mm - 1) / 2;
    const double2 bm = bp[mm];
    acc = cx<T>{(T)bm.x, (T)bm.y};
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (q < mm) {
        const double2 bq = bp[q];
        acc = cfma(Li[q], cx<T>{(T)bq.x, (T)bq.y}, acc);
      }
  } else {
    const cx<T> *Ui = rp + 50 + mm * 5 - mm * (mm - 1) / 2;
#pragma unroll
    for (int q = 0; q < 5; q++)
      if (q < 5 - mm) {
        const double2 yq = bp[mm + q];
        acc = cfma(Ui[q], cx<T>{(T)yq.x, (T)yq.y}, acc);
      }
  }
  *reinterpret_cast<cx<T> *>(bp + mm) = acc;
}

// The workgroup barrier of the loader waves: their LDS writes are complete
// (lgkmcnt 0) but the copies they have in flight are not waited for, as a
// __syncthreads fence would (it treats global_load_lds as pending LDS
// writes).  The empty asm statements keep the compiler from moving memory
// accesses across the barrier.
__device__ __forceinline__ void loader_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);     // lgkmcnt(0); vmcnt, expcnt not waited for
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// the parallel part of a row, plus M b_prev for the second step of a pair
// (M: the pair's side block, b_prev: the staged right-hand side of the first
// step).  One straight-line path for every lane of a round, so that all of
// them read before any writes.
template <typename T>
__device__ __forceinline__ void parallel_part_x(double2 *bp, const cx<T> *rp, const cx<T> *M,
                                                const double2 *bprev, int mm, bool fw, bool second) {
  cx<T> acc{0, 0};
  double2 bz[5];
  cx<T> mz[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    bz[q] = bprev[q];
    mz[q] = second ? M[mm * 5 + q] : cx<T>{0, 0};
  }
  if (fw) {
    const int base = 50 + mm * (mm - 1) / 2;
    const double2 bm = bp[mm];
    double2 bq[4];
    cx<T> l[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      bq[q] = bp[q];
      l[q] = rp[min(base + q, 59)];
    }
    acc = cx<T>{(T)bm.x, (T)bm.y};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const cx<T> w = q < mm ? l[q] : cx<T>{0, 0};
      acc = cfma(w, cx<T>{(T)bq[q].x, (T)bq[q].y}, acc);
    }
  } else {
    const int base = 50 + mm * 5 - mm * (mm - 1) / 2;
    double2 yq[5];
    cx<T> u[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
      yq[q] = bp[min(mm + q, 4)];
      u[q] = rp[min(base + q, 64)];
    }
#pragma unroll
    for (int q = 0; q < 5; q++) {
      const cx<T> w = q < 5 - mm ? u[q] : cx<T>{0, 0};
      acc = cfma(w, cx<T>{(T)yq[q].x, (T)yq[q].y}, acc);
    }
  }
#pragma unroll
  for (int q = 0; q < 5; q++) acc = cfma(mz[q], cx<T>{(T)bz[q].x, (T)bz[q].y}, acc);
  *reinterpret_cast<cx<T> *>(bp + mm) = acc;
}

// The loader side of one pass over nchunk chunks (loader wave w = 0, 1, 2)
// with a ring of NBUF (4 or 5) buffers.  Chunk k belongs to wave k % 3,
// which starts its copies in iteration k - (NBUF - 1) and, in iteration
// k - 1, waits for them and prepares the chunk before it starts any further
// copies.  So a wave has one chunk's copies in flight when it waits, and
// vmcnt(0) waits for exactly those (with NBUF = 5 the wave preparing chunk
// k + 1 is the one that then starts chunk k + 4; in the prologue wave 0
// waits for chunks 0 and 3).  One barrier after the prologue and one per
// chunk, as in compute_pass.
template <int NBUF, typename Issue, typename Prep>
__device__ __forceinline__ void stream_pass(int nchunk, int w, Issue issue, Prep prep) {
  static_assert(NBUF == 4 || NBUF == 5, "one chunk in flight per wave when it waits");
  for (int c = w; c < NBUF - 1 && c < nchunk; c += 3) issue(c);
  if (w == 0) {
    __builtin_amdgcn_s_waitcnt(0);
    prep(0);
  }
  loader_barrier();
  for (int k = 0; k < nchunk; k++) {
    if (k + 1 < nchunk && (k + 1) % 3 == w) {
      __builtin_amdgcn_s_waitcnt(0);      // this wave's copies of chunk k + 1
      prep(k + 1);
    }
    if (k + NBUF - 1 < nchunk && (k + NBUF - 1) % 3 == w) issue(k + NBUF - 1);   // buffer of chunk k - 1
    loader_barrier();
  }
}

template <typename Body>
__device__ __forceinline__ void compute_pass(int nchunk, Body body) {
  __syncthreads();
  for (int k = 0; k < nchunk; k++) {
    body(k);
    __syncthreads();
  }
}

// the solve of workgroup bid (lines bid G .. bid G + cnt - 1, cnt <= G)
template <typename T, int R, int G>
__device__ __forceinline__ void solve_bc(long bid, int n, long L, const cx<T> *__restrict__ fwd,
                                         const cx<T> *__restrict__ bwd, const cx<T> *__restrict__ tail,
                                         double2 *rhs, long sL, long sR) {
  constexpr int NBUF = 4, BUFE = R * 66;                        // LDS record rows of one line
  static_assert(R % 6 == 0, "a chunk is whole groups of six rows");
  __shared__ __attribute__((aligned(16))) cx<T> rec[NBUF][G][BUFE];
  __shared__ double2 bst[NBUF][G][R][5];    // rhs rows, then c (or d) in place
  const long line0 = bid * G;
  const int cnt = (int)min((long)G, L - line0);
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  // cnt lines in this workgroup; DPP row r (16 lanes) of the compute wave
  // runs the recurrence of line g = r mod cnt (rows past the last line repeat
  // one and store the same values).  Lane m of a row owns component m (lanes
  // past 4 duplicate component 4).  LDS slots past cnt are not loaded.
  const int g = (lane >> 4) % cnt;
  const int m = min(lane & 15, 4);
  auto line_of = [&](int gg) { return line0 + min(gg, cnt - 1); };
  const long line = line_of(g);
  const int nchunk = (n + R - 1) / R;
  double2 *X = rhs + line * sL * 5;
  const long rs = sR * 5;
  const cx<T> *tl = tail + line * 150;
  // rows of chunk k, ascending as they sit in LDS: [lo, lo + nr); forward
  // chunks run up from row 0, backward chunks down from row n-1
  auto chunk = [&](int k, bool fw, int &lo, int &nr) {
    if (fw) {
      lo = k * R;
      nr = min(R, n - lo);
    } else {
      const int top = n - 1 - k * R;
      nr = min(R, top + 1);
      lo = top - nr + 1;
    }
  };

  if (wave != 0) {
    // ---- loader waves ----
    auto issue = [&](int k, bool fw) {
      const int buf = k % NBUF;
      int lo, nr;
      chunk(k, fw, lo, nr);
#pragma unroll
      for (int gg = 0; gg < G; gg++) {
        if (gg >= cnt) break;
        const long r0 = line_of(gg) * n + lo;
        if (fw) dma_rows(fwd + r0 * 60, rec[buf][gg], nr * 60, lane);
        else dma_rows(bwd + r0 * 66, rec[buf][gg], nr * 66, lane);
      }
      double2 *dst = &bst[buf][0][0][0];
#pragma unroll
      for (int b = 0; b * 64 < G * R * 5; b++) {
        const int e = b * 64 + lane, gg = e / (R * 5), j = (e / 5) % R, q = e % 5;
        if (e < G * R * 5 && gg < cnt && j < nr)
          __builtin_amdgcn_global_load_lds(
              (const void *)(rhs + line_of(gg) * sL * 5 + (long)(lo + j) * rs + q),
              (lds_void *)(dst + b * 64), 16, 0, 0);
      }
    };
    auto prep = [&](int k, bool fw) {
      const int buf = k % NBUF;
      int lo, nr;
      chunk(k, fw, lo, nr);
      for (int r0 = 0; r0 < G * R; r0 += 12) {     // 12 rows (60 lanes) at a time
        const int rr = r0 + lane / 5, gg = rr / R, j = rr % R;
        if (lane < 60 && rr < G * R && gg < cnt && j < nr)
          parallel_part(bst[buf][gg][j], rec[buf][gg] + j * (fw ? 60 : 66), lane % 5, fw);
      }
    };
    stream_pass<NBUF>(nchunk, wave - 1, [&](int k) { issue(k, true); }, [&](int k) { prep(k, true); });
    stream_pass<NBUF>(nchunk, wave - 1, [&](int k) { issue(k, false); }, [&](int k) { prep(k, false); });
    return;
  }

  // ---- compute wave: the recurrence on lanes 0..4 (lane m = component m)
  // Row i is split into the part that needs the newest value,
  //   y_i = a0_i + W2_i * y_{i-1}          (products summed as a tree)
  // and the part formed one row ahead, off the critical path,
  //   a0_i = c_i + W1_i * y_{i-2} (+ boundary-row blocks).
  // Record rows are read from LDS two rows ahead of use into three rotating
  // register sets (W2 parts and W1+c parts separately).  The newest two
  // values live in yA/yB, broadcast to every lane, with roles alternating.
  // The recurrence wave issues first on its SIMD: where it shares the SIMD
  // with loader waves (two workgroups on a CU) they would split its issue
  // slots (1-4% of the solve time).
  __builtin_amdgcn_s_setprio(3);
  cx<T> yA[5], yB[5];
  cx<T> w2s[3][5], w1s[3][5], cs[3];
  // rows kept for the boundary-row blocks: y_{n-5}, y_{n-4} (forward),
  // x_3, x_4 (backward), as broadcast values
  cx<T> s0[5], s1[5];
  // the term in the newest value: two FMA accumulators (four independent
  // chains over re and im), 22 instructions
  auto crit = [&](const cx<T> (&w)[5], const cx<T> (&yn)[5], cx<T> a0) -> cx<T> {
    cx<T> a = cfma(w[0], yn[0], a0), b = cmul(w[1], yn[1]);
    a = cfma(w[2], yn[2], a);
    b = cfma(w[3], yn[3], b);
    a = cfma(w[4], yn[4], a);
    return a + b;
  };
  auto pre0 = [&](const cx<T> (&w)[5], cx<T> c, const cx<T> (&yo)[5]) -> cx<T> {
    cx<T> a = cfma(w[0], yo[0], c), b = cmul(w[1], yo[1]);
    a = cfma(w[2], yo[2], a);
    b = cfma(w[3], yo[3], b);
    a = cfma(w[4], yo[4], a);
    return a + b;
  };
  // blocks of the boundary rows outside the two-row band
  auto bnd = [&](int i, bool fwd_pass) -> cx<T> {
    cx<T> b{0, 0};
    if (fwd_pass) {
      if (i >= n - 2) {      // rows n-2, n-1: blocks in columns n-5, n-4
        const cx<T> *Xa = tl + (i == n - 2 ? 0 : 25) + m * 5;
#pragma unroll
        for (int q = 0; q < 5; q++) b = cfma(Xa[q], s0[q], b);
        if (i == n - 1)
#pragma unroll
          for (int q = 0; q < 5; q++) b = cfma(tl[50 + m * 5 + q], s1[q], b);
      }
    } else if (i <= 1) {     // rows 1, 0: blocks in columns 3, 4
      const cx<T> *Qa = tl + (i == 1 ? 125 : 100) + m * 5;
#pragma unroll
      for (int q = 0; q < 5; q++) b = cfma(Qa[q], s1[q], b);
      if (i == 0)
#pragma unroll
        for (int q = 0; q < 5; q++) b = cfma(tl[75 + m * 5 + q], s0[q], b);
    }
    return b;
  };
  auto keep = [&](const cx<T> (&y)[5], int i, bool fwd_pass) {
    if (i == (fwd_pass ? n - 5 : 3))
#pragma unroll
      for (int q = 0; q < 5; q++) s0[q] = y[q];
    if (i == (fwd_pass ? n - 4 : 4))
#pragma unroll
      for (int q = 0; q < 5; q++) s1[q] = y[q];
  };
  // one chunk: rows of buffer kk in processing order j = 0..nr-1 (row in
  // the buffer = j forward, nr-1-j backward; grid row = ibase +/- j).  Rows
  // are processed in branch-free groups of six; the remainder and the last
  // five rows of a pass (which carry the boundary-row blocks) take a plain
  // one-row-at-a-time path.
  auto run_chunk = [&](const cx<T> *B, int stride, int kk, int nr, int ibase, bool fwd_pass) {
    auto rix = [&](int j) { return fwd_pass ? j : nr - 1 - j; };     // j < nr at every call
    auto gi = [&](int j) { return fwd_pass ? ibase + j : ibase - j; };
    auto ldW2 = [&](int j, cx<T> (&w)[5]) {
      const cx<T> *p = B + rix(j) * stride + m * 10;
#pragma unroll
      for (int q = 0; q < 5; q++) w[q] = p[q];
    };
    auto ldW1 = [&](int j, cx<T> (&w)[5], cx<T> &c) {
      const int r = rix(j);
      const cx<T> *p = B + r * stride + m * 10 + 5;
#pragma unroll
      for (int q = 0; q < 5; q++) w[q] = p[q];
      c = *reinterpret_cast<const cx<T> *>(&bst[kk][g][r][m]);
    };
    int before_tail = fwd_pass ? n - 5 - ibase : ibase - 4;
    before_tail = max(0, min(nr, before_tail));
    const int nmain = (before_tail / 6) * 6;
    double2 *xp = X + (long)gi(0) * rs + m;
    const long xstep = fwd_pass ? rs : -rs;
    if (nmain > 0) {
      ldW2(0, w2s[0]);
      ldW2(1, w2s[1]);
      ldW1(0, w1s[0], cs[0]);
      ldW1(1, w1s[1], cs[1]);
      ldW1(2, w1s[2], cs[2]);
      cx<T> a0 = pre0(w1s[0], cs[0], yA);
      // rows two and three ahead are read only while they lie in this
      // chunk's groups (the next chunk starts by reading its own first rows)
      auto sub = [&](int j, bool more, auto U) {
        constexpr int u = decltype(U)::value;
        constexpr int u0 = u % 3, u1 = (u + 1) % 3, u2 = (u + 2) % 3;
        cx<T>(&yn)[5] = (u % 2 == 0) ? yB : yA;      // newest value, y_{i-1}
        cx<T>(&yd)[5] = (u % 2 == 0) ? yA : yB;      // receives y_i
        cx<T> y = crit(w2s[u0], yn, a0);
        a0 = pre0(w1s[u1], cs[u1], yn);
        *xp = make_double2(y.re, y.im);    // lanes of one component store the same value
        xp += xstep;
        if (u + 2 < 6 || more) ldW2(j + 2, w2s[u2]);
        if (u + 3 < 6 || more) ldW1(j + 3, w1s[u0], cs[u0]);
        row_bcast5(y, yd);
      };
      for (int j = 0; j < nmain; j += 6) {
        const bool more = j + 6 < nmain;
        sub(j, more, std::integral_constant<int, 0>());
        sub(j + 1, more, std::integral_constant<int, 1>());
        sub(j + 2, more, std::integral_constant<int, 2>());
        sub(j + 3, more, std::integral_constant<int, 3>());
        sub(j + 4, more, std::integral_constant<int, 4>());
        sub(j + 5, more, std::integral_constant<int, 5>());
      }
    }
    for (int j = nmain; j < nr; j++) {
      const int i = gi(j);
      cx<T> w1[5], w2[5], c;
      ldW1(j, w1, c);
      ldW2(j, w2);
      cx<T> y = crit(w2, yB, pre0(w1, c, yA) + bnd(i, fwd_pass));
      *xp = make_double2(y.re, y.im);
      xp += xstep;
      row_bcast5(y, yA);                 // y_{i-2} is dead
#pragma unroll
      for (int q = 0; q < 5; q++) { cx<T> t = yA[q]; yA[q] = yB[q]; yB[q] = t; }
      keep(yB, i, fwd_pass);
    }
  };

  // forward: y_i = c_i + N1 y_{i-2} + N2 y_{i-1}; backward: x_i = d_i +
  // Q4 x_{i+1} + Q5 x_{i+2}
#pragma unroll
  for (int q = 0; q < 5; q++) { yA[q] = cx<T>{0, 0}; yB[q] = cx<T>{0, 0}; }
  compute_pass(nchunk, [&](int k) {
    int lo, nr;
    chunk(k, true, lo, nr);
    run_chunk(rec[k % NBUF][g], 60, k % NBUF, nr, lo, true);
  });
#pragma unroll
  for (int q = 0; q < 5; q++) { yA[q] = cx<T>{0, 0}; yB[q] = cx<T>{0, 0}; }
  compute_pass(nchunk, [&](int k) {
    int lo, nr;
    chunk(k, false, lo, nr);
    run_chunk(rec[k % NBUF][g], 66, k % NBUF, nr, lo + nr - 1, false);
  });
}

template <typename T, int R, int G>
__global__ __launch_bounds__(256) void k_solve_bc(int n, long L, const cx<T> *__restrict__ fwd,
                                                 const cx<T> *__restrict__ bwd,
                                                 const cx<T> *__restrict__ tail, double2 *rhs,
                                                 long sL, long sR) {
  solve_bc<T, R, G>(blockIdx.x, n, L, fwd, bwd, tail, rhs, sL, sR);
}

// ---------------------------------------------------------------------------
// stage 3, twisted lines: the factor eliminated rows n-1..h+2 bottom-up and
// then rows 0..h+1 top-down, so each line is two recurrences of about n/2
// rows ("chains"): chain 0 runs rows 0..h+1, chain 1 rows n-1..h+2.  They
// run in lockstep on two DPP rows of the compute wave (G lines per
// workgroup, 2G chains).  Forward: chain 1 finishes first and hands its last
// two values (rows h+2, h+3) to chain 0, whose last two rows carry those
// columns.  Backward: chain 0 starts at row h+1; chain 1 starts two steps
// later from chain 0's x_{h+1}, x_h.
// HALF (one line per workgroup, single steps): the DPP rows that would
// repeat the two chains instead take half of each row's columns (below).
template <typename T, int R, int G, bool PAIRS, int NBUF, bool HALF>
__device__ __forceinline__ void solve_tw(long bid, int n, int h, long L, const cx<T> *__restrict__ fwd,
                                         const cx<T> *__restrict__ bwd, const cx<T> *__restrict__ tail,
                                         const cx<T> *__restrict__ xf, const cx<T> *__restrict__ xb,
                                         double2 *rhs, long sL, long sR) {
  constexpr int C = 2 * G;
  constexpr int BUFE = R * 66;
  static_assert(R % 6 == 0, "a chunk is whole groups of six rows");
  static_assert(!HALF || (G == 1 && !PAIRS), "column halves: one line per workgroup, single steps");
  __shared__ __attribute__((aligned(16))) cx<T> rec[NBUF][C][BUFE];
  __shared__ double2 bst[NBUF][C][R][5];   // rhs rows, then c (or d) in place
  constexpr int XE = PAIRS ? (R / 2) * 25 + 1 : 1;   // side blocks of a chunk's pairs (+1: copy padding)
  __shared__ __attribute__((aligned(16))) cx<T> xtr[NBUF][C][XE];
  __shared__ cx<T> xch[G][2][5];     // values handed between the chains of a line
  __shared__ cx<T> kept[C][2][5];    // x at chain rows 3, 4 (backward boundary blocks)
  // zero rows standing in for the staged c rows in the column half without c
  __shared__ double2 zrows[HALF ? R : 1][5];
  const long line0 = bid * G;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int ch = (lane >> 4) % C;    // chain of this DPP row (rows past 2G repeat)
  const int g = ch >> 1, t = ch & 1;
  const int m = min(lane & 15, 4);
  const int hq = lane >> 5;          // column half (HALF): DPP rows 0, 1 and rows 2, 3
  auto line_of = [&](int gg) { return min(line0 + gg, L - 1); };
  const long line = line_of(g);
  const int len0 = h + 2, len1 = n - h - 2;          // len0 >= len1 + 2
  // chain tt (0 or 1) in arithmetic form: a select here becomes a per-lane
  // table in scratch, whose loads would wait on the loaders' copies
  auto clen = [&](int tt) { return len0 + tt * (len1 - len0); };
  auto coff = [&](int tt) { return tt * len0; };
  auto real_row = [&](int tt, int j) { return j + tt * (n - 1 - 2 * j); };
  const int len = clen(t);
  const long rs = sR * 5;
  double2 *X = rhs + line * sL * 5;
  const cx<T> *tl = tail + (line * 2 + t) * 150;
  const int nchunk = (len0 + R - 1) / R;             // steps per pass: len0
  const int npair = (len0 + 1) / 2;                  // pair slots per chain in xf / xb
  // chain row at step s of the backward pass
  auto bj = [&](int tt, int s) { return len0 - 1 - s + tt * (len1 + 2 - len0); };
  // rows of chunk k held in the LDS buffer for chain tt: [jlo, jlo + nr)
  auto chunk_rows = [&](int tt, int k, bool fw, int &jlo, int &nr) {
    const int l = clen(tt);
    if (fw) {
      jlo = k * R;
      nr = max(0, min(R, l - jlo));
    } else {
      int jhi = min(l - 1, bj(tt, k * R));
      jlo = max(0, bj(tt, k * R + R - 1));
      nr = max(0, jhi - jlo + 1);
    }
  };

  if (wave != 0) {
    // ---- loader waves ----
    auto issue = [&](int k, bool fw) {
      const int buf = k % NBUF;
#pragma unroll
      for (int cc = 0; cc < C; cc++) {
        int jlo, nr;
        chunk_rows(cc & 1, k, fw, jlo, nr);
        const long r0 = line_of(cc >> 1) * n + coff(cc & 1) + jlo;
        if (fw) dma_rows(fwd + r0 * 60, rec[buf][cc], nr * 60, lane);
        else dma_rows(bwd + r0 * 66, rec[buf][cc], nr * 66, lane);
      }
      int p_end, m_end;
      tw_pair_range(R, k, fw, len0, len1, p_end, m_end);
      if (PAIRS && m_end > p_end) {
        const int ne = (m_end - p_end) / 2 * 25;
#pragma unroll
        for (int cc = 0; cc < C; cc++)
          dma_rows((fw ? xf : xb) + ((line_of(cc >> 1) * 2 + (cc & 1)) * npair + p_end / 2) * 25, xtr[buf][cc],
                   (ne + 1) & ~1, lane);
      }
      double2 *dst = &bst[buf][0][0][0];
#pragma unroll
      for (int b = 0; b * 64 < C * R * 5; b++) {
        const int e = b * 64 + lane, cc = e / (R * 5), j = (e / 5) % R, q = e % 5;
        int jlo, nr;
        chunk_rows(cc & 1, k, fw, jlo, nr);
        if (e < C * R * 5 && j < nr)
          __builtin_amdgcn_global_load_lds(
              (const void *)(rhs + line_of(cc >> 1) * sL * 5 + (long)real_row(cc & 1, jlo + j) * rs + q),
              (lds_void *)(dst + b * 64), 16, 0, 0);
      }
    };
    auto prep = [&](int k, bool fw) {
      const int buf = k % NBUF;
      int p_end, m_end;
      tw_pair_range(R, k, fw, len0, len1, p_end, m_end);
      for (int r0 = 0; r0 < C * R; r0 += 12) {     // 12 rows (60 lanes) at a time (a chain's rows
        const int rr = r0 + lane / 5, cc = rr / R, j = rr % R;   // never straddle two rounds)
        int jlo, nr;
        chunk_rows(cc & 1, k, fw, jlo, nr);
        // the step of LDS row j, and whether it is the second of a pair
        const int st = fw ? k * R + j : bj(cc & 1, 0) - (jlo + j);
        const bool second = PAIRS && (st & 1) && st - 1 >= p_end && st < m_end;
        const int jp = second ? (fw ? j - 1 : j + 1) : j;   // LDS row of step st - 1
        if (lane < 60 && rr < C * R && j < nr) {
          if constexpr (PAIRS)
            parallel_part_x(bst[buf][cc][j], rec[buf][cc] + j * (fw ? 60 : 66),
                            xtr[buf][cc] + (second ? (st - 1 - p_end) / 2 : 0) * 25, bst[buf][cc][jp], lane % 5,
                            fw, second);
          else
            parallel_part(bst[buf][cc][j], rec[buf][cc] + j * (fw ? 60 : 66), lane % 5, fw);
        }
      }
    };
    stream_pass<NBUF>(nchunk, wave - 1, [&](int k) { issue(k, true); }, [&](int k) { prep(k, true); });
    stream_pass<NBUF>(nchunk, wave - 1, [&](int k) { issue(k, false); }, [&](int k) { prep(k, false); });
    return;
  }

  // ---- compute wave (same recurrence as k_solve_bc, per chain) ----
  __builtin_amdgcn_s_setprio(3);   // as in solve_bc
  cx<T> yA[5], yB[5];
  cx<T> w2s[3][5], w1s[3][5], cs[3];
  cx<T> s0[5], s1[5];
  // HALF: the two newest values (lane m: component m), always current, and
  // their split broadcasts, current while `split` holds
  cx<T> vA{0, 0}, vB{0, 0}, SA[3], SB[3];
  bool split = false;
  if (HALF) {
    for (int e = lane; e < R * 5; e += 64) zrows[e / 5][e % 5] = make_double2(0.0, 0.0);
    wave_sync();
  }
  // split broadcast of v into S, in place: slot k = component 3 hq + k of
  // the chain; slot 2 of the upper half is never written and holds zero
  auto split_bcast = [&](cx<T> v, cx<T> (&S)[3]) {
    S[0] = row_bcast_rows<3, 0xC>(row_bcast_rows<0, 0x3>(S[0], v), v);
    S[1] = row_bcast_rows<4, 0xC>(row_bcast_rows<1, 0x3>(S[1], v), v);
    S[2] = row_bcast_rows<2, 0x3>(S[2], v);
  };
  // the term in the newest value: two FMA accumulators (four independent
  // chains over re and im), 22 instructions
  auto crit = [&](const cx<T> (&w)[5], const cx<T> (&yn)[5], cx<T> a0) -> cx<T> {
    cx<T> a = cfma(w[0], yn[0], a0), b = cmul(w[1], yn[1]);
    a = cfma(w[2], yn[2], a);
    b = cfma(w[3], yn[3], b);
    a = cfma(w[4], yn[4], a);
    return a + b;
  };
  auto pre0 = [&](const cx<T> (&w)[5], cx<T> c, const cx<T> (&yo)[5]) -> cx<T> {
    cx<T> a = cfma(w[0], yo[0], c), b = cmul(w[1], yo[1]);
    a = cfma(w[2], yo[2], a);
    b = cfma(w[3], yo[3], b);
    a = cfma(w[4], yo[4], a);
    return a + b;
  };
  // one step of a pair: c + W2 yn + W1 yo, in four independent FMA chains
  auto pair_eval = [&](const cx<T> (&w2)[5], const cx<T> (&w1)[5], cx<T> c, const cx<T> (&yn)[5],
                       const cx<T> (&yo)[5]) -> cx<T> {
    cx<T> a = cfma(w1[0], yo[0], c), b = cmul(w2[0], yn[0]);
#pragma unroll
    for (int q = 1; q < 5; q++) {
      a = cfma(w1[q], yo[q], a);
      b = cfma(w2[q], yn[q], b);
    }
    return a + b;
  };
  auto run_pass = [&](bool fw) {
    const int stride = fw ? 60 : 66;
#pragma unroll
    for (int q = 0; q < 5; q++) { yA[q] = cx<T>{0, 0}; yB[q] = cx<T>{0, 0}; }
    vA = vB = cx<T>{0, 0};
    split = false;
    compute_pass(nchunk, [&](int k) {
      {
        const int kk = k % NBUF;
        const cx<T> *B = rec[kk][ch];
        int jlo, nr;
        chunk_rows(t, k, fw, jlo, nr);
        const int s_begin = k * R, s_end = min(k * R + R, len0);
        // buffer row of step s (clamped into the chunk's rows)
        auto brow = [&](int s) {
          int b = fw ? s - s_begin : bj(t, s) - jlo;
          return max(0, min(max(nr, 1) - 1, b));
        };
        // the same inside the main range, where every step's row is in the
        // buffer for both chains
        auto brow_m = [&](int s) { return fw ? s - s_begin : bj(t, s) - jlo; };
        auto ldW2 = [&](int s, cx<T> (&w)[5]) {
          const cx<T> *pp = B + brow(s) * stride + m * 10;
#pragma unroll
          for (int q = 0; q < 5; q++) w[q] = pp[q];
        };
        auto ldW1 = [&](int s, cx<T> (&w)[5], cx<T> &c) {
          const int r = brow(s);
          const cx<T> *pp = B + r * stride + m * 10 + 5;
#pragma unroll
          for (int q = 0; q < 5; q++) w[q] = pp[q];
          c = *reinterpret_cast<const cx<T> *>(&bst[kk][ch][r][m]);
        };
        auto jrow = [&](int s) { return fw ? s : bj(t, s); };
        // one step on the plain path: hand-offs, boundary blocks, inactive chains
        auto plain = [&](int s) {
          const int j = jrow(s);
          const bool act = j >= 0 && j < len;
          if (HALF) {          // vA, vB are the state; yA, yB live only within a plain step
            row_bcast5(vA, yA);
            row_bcast5(vB, yB);
            split = false;
          }
          if (!fw && t == 1 && s == 2) {            // chain 1 starts from x_h, x_{h+1}
#pragma unroll
            for (int q = 0; q < 5; q++) { yA[q] = xch[g][0][q]; yB[q] = xch[g][1][q]; }
            if (HALF) { vA = xch[g][0][m]; vB = xch[g][1][m]; }
          }
          if (fw && t == 0 && j == len0 - 2) {      // columns h+2, h+3 from chain 1
#pragma unroll
            for (int q = 0; q < 5; q++) { s0[q] = xch[g][0][q]; s1[q] = xch[g][1][q]; }
          }
          if (!fw && j == 1) {
#pragma unroll
            for (int q = 0; q < 5; q++) { s0[q] = kept[ch][0][q]; s1[q] = kept[ch][1][q]; }
          }
          cx<T> w1[5], w2[5], c;
          ldW1(s, w1, c);
          ldW2(s, w2);
          cx<T> b{0, 0};
          if (fw) {
            if (t == 0 && j >= len0 - 2) {
              const cx<T> *Xa = tl + (j == len0 - 2 ? 0 : 25) + m * 5;
#pragma unroll
              for (int q = 0; q < 5; q++) b = cfma(Xa[q], s0[q], b);
              if (j == len0 - 1)
#pragma unroll
                for (int q = 0; q < 5; q++) b = cfma(tl[50 + m * 5 + q], s1[q], b);
            }
          } else if (j >= 0 && j <= 1) {
            const cx<T> *Qa = tl + (j == 1 ? 125 : 100) + m * 5;
#pragma unroll
            for (int q = 0; q < 5; q++) b = cfma(Qa[q], s1[q], b);
            if (j == 0)
#pragma unroll
              for (int q = 0; q < 5; q++) b = cfma(tl[75 + m * 5 + q], s0[q], b);
          }
          cx<T> y = crit(w2, yB, pre0(w1, c, yA) + b);
          if (act) X[(long)real_row(t, j) * rs + m] = make_double2(y.re, y.im);
          row_bcast5(y, yA);
#pragma unroll
          for (int q = 0; q < 5; q++) { cx<T> tt = yA[q]; yA[q] = yB[q]; yB[q] = tt; }
          if (HALF) { vA = vB; vB = y; }
          if (act) {
            if (fw && t == 1 && j == len1 - 1) xch[g][0][m] = y;     // z_{h+2}
            if (fw && t == 1 && j == len1 - 2) xch[g][1][m] = y;     // z_{h+3}
            if (!fw && t == 0 && j == len0 - 1) xch[g][1][m] = y;    // x_{h+1}
            if (!fw && t == 0 && j == len0 - 2) xch[g][0][m] = y;    // x_h
            if (!fw && (j == 3 || j == 4)) kept[ch][j - 3][m] = y;
          }
        };
        int p_end, m_end;
        tw_pair_range(R, k, fw, len0, len1, p_end, m_end);
        for (int s = s_begin; s < p_end; s++) plain(s);
        if (PAIRS && m_end > p_end) {
          // pairs of steps (s, s+1): lanes 0-4 step s, lanes 5-9 step s+1
          // (its record and parallel part rewritten for the pair), lanes
          // 10-15 repeat lane 9; both from y_{s-2} (yA), y_{s-1} (yB).
          // Coefficients two pairs ahead in three rotating register sets.
          const int pl = min(lane & 15, 9), prow = pl >= 5 ? 1 : 0, pm = pl - 5 * prow;
          auto ldP = [&](int st, cx<T> (&w2)[5], cx<T> (&w1)[5], cx<T> &c) {
            const int r = brow(st + prow);
            const cx<T> *pp = B + r * stride + pm * 10;
#pragma unroll
            for (int q = 0; q < 5; q++) { w2[q] = pp[q]; w1[q] = pp[5 + q]; }
            c = *reinterpret_cast<const cx<T> *>(&bst[kk][ch][r][pm]);
          };
          const long xstep = ((t == 0) == fw) ? rs : -rs;
          double2 *xq = X + (long)real_row(t, jrow(p_end + prow)) * rs + pm;
          ldP(p_end, w2s[0], w1s[0], cs[0]);
          ldP(p_end + 2, w2s[1], w1s[1], cs[1]);
          auto psub = [&](int st, auto U) {
            constexpr int u = decltype(U)::value, u2 = (u + 2) % 3;
            const cx<T> y = pair_eval(w2s[u], w1s[u], cs[u], yB, yA);
            *xq = make_double2(y.re, y.im);
            xq += 2 * xstep;
            ldP(st + 4, w2s[u2], w1s[u2], cs[u2]);
            row_bcast5_at<0>(y, yA);
            row_bcast5_at<5>(y, yB);
          };
          for (int st = p_end; st < m_end; st += 6) {
            psub(st, std::integral_constant<int, 0>());
            psub(st + 2, std::integral_constant<int, 1>());
            psub(st + 4, std::integral_constant<int, 2>());
          }
        }
        if (HALF && m_end > p_end) {
          // Columns split over the wave's halves: half hq of a chain sums
          // columns 3 hq + k (k = 0, 1, 2) of each block row, slot 2 of the
          // upper half meeting a zero broadcast; the upper half adds c.  The
          // halves' partial rows are added across the wave (half_sum), so both
          // halves hold y and broadcast it.  Steps two and three ahead are
          // read only while they lie in the main range.
          if (!split) {
            SA[2] = SB[2] = cx<T>{0, 0};
            split_bcast(vA, SA);
            split_bcast(vB, SB);
            split = true;
          }
          const int qb = 3 * hq;
          const double2 *cb = hq ? &bst[kk][ch][0][m] : &zrows[0][m];   // row r: cb + 5 r
          auto ldW2h = [&](int s, cx<T> (&w)[3]) {
            const cx<T> *pp = B + brow_m(s) * stride + m * 10 + qb;
#pragma unroll
            for (int k = 0; k < 3; k++) w[k] = pp[k];
          };
          auto ldW1h = [&](int s, cx<T> (&w)[3], cx<T> &c) {
            const int r = brow_m(s);
            const cx<T> *pp = B + r * stride + m * 10 + 5 + qb;
#pragma unroll
            for (int k = 0; k < 3; k++) w[k] = pp[k];
            c = *reinterpret_cast<const cx<T> *>(cb + 5 * r);
          };
          // a + b: the half's part of c + W1 y_{i-2}, formed a step ahead
          auto pre0h = [&](const cx<T> (&w)[3], cx<T> c, const cx<T> (&S)[3], cx<T> &a, cx<T> &b) {
            a = cfma(w[0], S[0], c);
            b = cmul(w[1], S[1]);
            a = cfma(w[2], S[2], a);
          };
          auto crith = [&](const cx<T> (&w)[3], const cx<T> (&S)[3], cx<T> a, cx<T> b) -> cx<T> {
            a = cfma(w[0], S[0], a);
            b = cfma(w[1], S[1], b);
            a = cfma(w[2], S[2], a);
            return half_sum(a + b);
          };
          cx<T> w2h[3][3], w1h[3][3], chh[3];
          double2 *xp = X + (long)real_row(t, jrow(p_end)) * rs + m;
          const long xstep = ((t == 0) == fw) ? rs : -rs;
          ldW2h(p_end, w2h[0]);
          ldW2h(p_end + 1, w2h[1]);
          ldW1h(p_end, w1h[0], chh[0]);
          ldW1h(p_end + 1, w1h[1], chh[1]);
          ldW1h(p_end + 2, w1h[2], chh[2]);
          cx<T> pa, pb;
          pre0h(w1h[0], chh[0], SA, pa, pb);
          auto subh = [&](int s, bool more, auto U) {
            constexpr int u = decltype(U)::value;
            constexpr int u0 = u % 3, u1 = (u + 1) % 3, u2 = (u + 2) % 3;
            cx<T>(&Sn)[3] = (u % 2 == 0) ? SB : SA;      // y_{i-1}
            cx<T>(&Sd)[3] = (u % 2 == 0) ? SA : SB;      // receives y_i
            const cx<T> y = crith(w2h[u0], Sn, pa, pb);
            pre0h(w1h[u1], chh[u1], Sn, pa, pb);
            *xp = make_double2(y.re, y.im);
            xp += xstep;
            if (u + 2 < 6 || more) ldW2h(s + 2, w2h[u2]);
            if (u + 3 < 6 || more) ldW1h(s + 3, w1h[u0], chh[u0]);
            split_bcast(y, Sd);
            if constexpr (u == 4) vA = y;
            if constexpr (u == 5) vB = y;
          };
          for (int s = p_end; s < m_end; s += 6) {
            const bool more = s + 6 < m_end;
            subh(s, more, std::integral_constant<int, 0>());
            subh(s + 1, more, std::integral_constant<int, 1>());
            subh(s + 2, more, std::integral_constant<int, 2>());
            subh(s + 3, more, std::integral_constant<int, 3>());
            subh(s + 4, more, std::integral_constant<int, 4>());
            subh(s + 5, more, std::integral_constant<int, 5>());
          }
        }
        if (!PAIRS && !HALF && m_end > p_end) {
          // one step at a time in groups of six (no pair rewrite); steps two
          // and three ahead are read only while they lie in the main range
          auto ldW2m = [&](int s, cx<T> (&w)[5]) {
            const cx<T> *pp = B + brow_m(s) * stride + m * 10;
#pragma unroll
            for (int q = 0; q < 5; q++) w[q] = pp[q];
          };
          auto ldW1m = [&](int s, cx<T> (&w)[5], cx<T> &c) {
            const int r = brow_m(s);
            const cx<T> *pp = B + r * stride + m * 10 + 5;
#pragma unroll
            for (int q = 0; q < 5; q++) w[q] = pp[q];
            c = *reinterpret_cast<const cx<T> *>(&bst[kk][ch][r][m]);
          };
          int j0 = jrow(p_end);
          double2 *xp = X + (long)real_row(t, j0) * rs + m;
          const long xstep = ((t == 0) == fw) ? rs : -rs;
          ldW2m(p_end, w2s[0]);
          ldW2m(p_end + 1, w2s[1]);
          ldW1m(p_end, w1s[0], cs[0]);
          ldW1m(p_end + 1, w1s[1], cs[1]);
          ldW1m(p_end + 2, w1s[2], cs[2]);
          cx<T> a0 = pre0(w1s[0], cs[0], yA);
          auto sub = [&](int s, bool more, auto U) {
            constexpr int u = decltype(U)::value;
            constexpr int u0 = u % 3, u1 = (u + 1) % 3, u2 = (u + 2) % 3;
            cx<T>(&yn)[5] = (u % 2 == 0) ? yB : yA;
            cx<T>(&yd)[5] = (u % 2 == 0) ? yA : yB;
            cx<T> y = crit(w2s[u0], yn, a0);
            a0 = pre0(w1s[u1], cs[u1], yn);
            *xp = make_double2(y.re, y.im);
            xp += xstep;
            if (u + 2 < 6 || more) ldW2m(s + 2, w2s[u2]);
            if (u + 3 < 6 || more) ldW1m(s + 3, w1s[u0], cs[u0]);
            row_bcast5(y, yd);
          };
          for (int s = p_end; s < m_end; s += 6) {
            const bool more = s + 6 < m_end;
            sub(s, more, std::integral_constant<int, 0>());
            sub(s + 1, more, std::integral_constant<int, 1>());
            sub(s + 2, more, std::integral_constant<int, 2>());
            sub(s + 3, more, std::integral_constant<int, 3>());
            sub(s + 4, more, std::integral_constant<int, 4>());
            sub(s + 5, more, std::integral_constant<int, 5>());
          }
        }
        for (int s = m_end; s < s_end; s++) plain(s);
      }
    });
  };
  run_pass(true);
  run_pass(false);
}

template <typename T, int R, int G, bool PAIRS, int NBUF, bool HALF = false>
__global__ __launch_bounds__(256, 2) void k_solve_tw(int n, int h, long L, const cx<T> *__restrict__ fwd,
                                                 const cx<T> *__restrict__ bwd,
                                                 const cx<T> *__restrict__ tail,
                                                 const cx<T> *__restrict__ xf,
                                                 const cx<T> *__restrict__ xb, double2 *rhs,
                                                 long sL, long sR) {
  solve_tw<T, R, G, PAIRS, NBUF, HALF>(blockIdx.x, n, h, L, fwd, bwd, tail, xf, xb, rhs, sL, sR);
}

// ---------------------------------------------------------------------------
// periodic lines (cpenta1p.f / cpenta2p.f): records from the bordered LU of
// k_lu_per and a streaming solve.  The LU's band gives the same recurrence
// as non-periodic lines; the border (rows n-2, n-1 in every column: per2;
// columns n-2, n-1 in every row: per) becomes two more per-row records:
//   p2[row c] = (Linv_{n-2} per2(n-2,c), Linv_{n-1} per2(n-1,c)): the forward
//     contributions of y_c to the last two rows, summed by the loader waves;
//   qp[row i] = (-Uinv_i per(i,n-2), -Uinv_i per(i,n-1)): the backward terms
//     in x_{n-2}, x_{n-1}, which are known first, folded into d_i.
__device__ __forceinline__ void inblock_inverses(const Blk<double> &D, cx<double> (&Li)[5][5],
                                                 cx<double> (&Ui)[5][5]) {
  for (int r = 0; r < 5; r++)
    for (int c = 0; c < 5; c++) { Li[r][c] = cx<double>{double(r == c), 0}; Ui[r][c] = cx<double>{0, 0}; }
  for (int nc = 0; nc < 4; nc++)
    for (int l = nc + 1; l < 5; l++)
      for (int c = 0; c <= nc; c++) Li[l][c] = cfma(D(l, nc), Li[nc][c], Li[l][c]);
  for (int c = 0; c < 5; c++)
    for (int m = c; m >= 0; m--) {
      cx<double> acc = cx<double>{double(m == c), 0};
      for (int q = m + 1; q <= c; q++) acc = cfma(-D(m, q), Ui[q][c], acc);
      Ui[m][c] = cdiv(acc, D(m, m));
    }
}

__device__ __noinline__ void records_per_row(LineGeom g, cx<double> *mat, cx<double> *per, cx<double> *per2,
                                             cx<double> *fwd, cx<double> *bwd, cx<double> *p2,
                                             cx<double> *qp, long idx) {
  const int n = g.n;
  const long line = idx / n;
  const int i = (int)(idx - line * n);
  Blk<double> D, M;
  cx<double> Li[5][5], Ui[5][5];
  blk(g, mat, per, per2, line, i, i, D);
  inblock_inverses(D, Li, Ui);
  cx<double> *F = fwd + idx * 60, *B = bwd + idx * 66, *P2 = p2 + idx * 50, *QP = qp + idx * 50;
  auto lmul = [&](const cx<double> (&A)[5][5], int row, int col, cx<double> *dst, int ld) {
    bool ok = col >= 0 && col < n && blk(g, mat, per, per2, line, row, col, M);
    for (int m = 0; m < 5; m++)
      for (int q = 0; q < 5; q++) {
        cx<double> acc{0, 0};
        if (ok)
          for (int k = 0; k <= m; k++) acc = cfma(A[m][k], M(k, q), acc);
        dst[m * ld + q] = acc;
      }
  };
  auto umul = [&](int col, cx<double> *dst, int ld) {
    bool ok = col >= 0 && col < n && blk(g, mat, per, per2, line, i, col, M);
    for (int m = 0; m < 5; m++)
      for (int q = 0; q < 5; q++) {
        cx<double> acc{0, 0};
        if (ok)
          for (int k = m; k < 5; k++) acc = cfma(Ui[m][k], M(k, q), acc);
        dst[m * ld + q] = cx<double>{-acc.re, -acc.im};
      }
  };
  // band blocks only (columns i-1, i-2 / i+1, i+2 within the line)
  if (i >= 1) lmul(Li, i, i - 1, F, 10); else lmul(Li, i, -1, F, 10);
  if (i >= 2) lmul(Li, i, i - 2, F + 5, 10); else lmul(Li, i, -1, F + 5, 10);
  for (int l = 1, k = 0; l < 5; l++)
    for (int q = 0; q < l; q++, k++) F[50 + k] = Li[l][q];
  umul(i + 1 < n ? i + 1 : -1, B, 10);
  umul(i + 2 < n ? i + 2 : -1, B + 5, 10);
  for (int m = 0, k = 0; m < 5; m++)
    for (int q = m; q < 5; q++, k++) B[50 + k] = Ui[m][q];
  B[65] = cx<double>{0, 0};
  // border columns of row i (per exists for i < col-2)
  umul(i < n - 4 ? n - 2 : -1, QP, 5);
  umul(i < n - 3 ? n - 1 : -1, QP + 25, 5);
  // border rows in column i (per2 exists for i < row-2), with those rows' Linv
  cx<double> Lb[5][5], Ub[5][5];
  blk(g, mat, per, per2, line, n - 2, n - 2, D);
  inblock_inverses(D, Lb, Ub);
  lmul(Lb, n - 2, i < n - 4 ? i : -1, P2, 5);
  blk(g, mat, per, per2, line, n - 1, n - 1, D);
  inblock_inverses(D, Lb, Ub);
  lmul(Lb, n - 1, i < n - 3 ? i : -1, P2 + 25, 5);
}

__global__ __launch_bounds__(128) void k_records_per(LineGeom g, cx<double> *mat, cx<double> *per, cx<double> *per2,
                              cx<double> *fwd, cx<double> *bwd, cx<double> *p2, cx<double> *qp) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < g.L * g.n) records_per_row(g, mat, per, per2, fwd, bwd, p2, qp, idx);
}

// streaming solve of periodic lines: G lines per workgroup, one per DPP row
// of the compute wave; rows 0..n-3 in chunks of R through the LDS ring.
//   forward: rows 0..n-3 on the band, then rows n-2, n-1 with the border sums
//            the loaders formed from p2 and the computed y;
//   backward: rows n-1, n-2 first, then rows n-3..0 with d_i including
//            qp_i (x_{n-2}, x_{n-1}); finally x(np) = x(1) (cpenta1p.f:1037).
// T: the records' precision (complex float records are the complex-double
// ones rounded; the right-hand side and solution stay complex double).
// Workgroup bid takes lines bid*G .. bid*G + G - 1.
template <typename T, int R, int G>
__device__ __forceinline__ void solve_per(long bid, int n, long L, const cx<T> *__restrict__ fwd,
                                          const cx<T> *__restrict__ bwd, const cx<T> *__restrict__ p2,
                                          const cx<T> *__restrict__ qp, double2 *rhs, long sL, long sR) {
  constexpr int RECW = 66;
  constexpr int BUFE = ((R * RECW * (int)sizeof(cx<T>) + 1023) / 1024) * 1024 / (int)sizeof(cx<T>);
  __shared__ __attribute__((aligned(1024))) cx<T> rec[2][G][BUFE];
  __shared__ cx<T> cb[2][G][R][5];
  __shared__ cx<T> bsum[G][2][5];     // forward border sums into rows n-2, n-1
  constexpr int NSUB = 192 / (10 * G);  // loader threads per border sum
  __shared__ cx<T> bpart[NSUB][10 * G];
  __shared__ cx<T> xb[G][2][5];       // x_{n-2}, x_{n-1}
  const long line0 = bid * G;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int g = (lane >> 4) % G;
  const int m = min(lane & 15, 4);
  auto line_of = [&](int gg) { return min(line0 + gg, L - 1); };
  const long line = line_of(g);
  const long rs = sR * 5;
  double2 *X = rhs + line * sL * 5;
  const int nreg = n - 2;                         // rows 0..n-3: band recurrence only
  const int nck = (nreg + R - 1) / R;
  auto zx = [](double2 a) { return cx<T>{(T)a.x, (T)a.y}; };

  // ---- loaders: record rows [r0, r0+nr) of pass fw into buffer kk, with
  // c = Linv b (forward) or d = Uinv y + qp (x_{n-2}, x_{n-1}) (backward)
  auto load = [&](int kk, int r0, int nr, bool fw, bool border) {
    const int stride = fw ? 60 : 66;
    const cx<T> *base = fw ? fwd : bwd;
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
      const long r = line_of(gg) * n + r0;
      copy_rows_to_lds(base + r * stride, (L * n - r) * stride, rec[kk][gg], nr * stride, wave - 1, 3, lane);
    }
    for (int t = tid - 64; t < G * nr * 5; t += 192) {
      const int gg = t / (nr * 5), rem = t - gg * nr * 5, row = rem / 5, mm = rem - row * 5;
      const long ln = line_of(gg), ri = ln * n + r0 + row;
      const double2 *xr = rhs + ln * sL * 5 + (long)(r0 + row) * rs;
      cx<T> acc{0, 0};
      if (fw) {
        const cx<T> *Li = fwd + ri * 60 + 50 + mm * (mm - 1) / 2;
        acc = zx(xr[mm]);
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (q < mm) acc = cfma(Li[q], zx(xr[q]), acc);
      } else {
        const cx<T> *Ui = bwd + ri * 66 + 50 + mm * 5 - mm * (mm - 1) / 2;
#pragma unroll
        for (int q = 0; q < 5; q++)
          if (q < 5 - mm) acc = cfma(Ui[q], zx(xr[mm + q]), acc);
        if (border) {
          const cx<T> *Q = qp + ri * 50 + mm * 5;
#pragma unroll
          for (int q = 0; q < 5; q++) acc = cfma(Q[q], xb[gg][0][q], acc);
#pragma unroll
          for (int q = 0; q < 5; q++) acc = cfma(Q[25 + q], xb[gg][1][q], acc);
        }
      }
      cb[kk][gg][row][mm] = acc;
    }
    __builtin_amdgcn_s_waitcnt(0);
  };
  // loader threads keep partial forward border sums of (line gg, border row
  // bt, component mm) over the rows the compute wave finished: NSUB threads
  // per sum, each taking every NSUB-th row of a chunk
  const int at = tid - 64;
  const bool acc_thread = at >= 0 && at < NSUB * 10 * G;
  const int aid = acc_thread ? at % (10 * G) : 0, asub = acc_thread ? at / (10 * G) : 0;
  const int agg = aid / 10, abt = (aid / 5) % 2, amm = aid % 5;
  cx<T> bacc{0, 0};
  auto accumulate = [&](int r0, int nr) {
    if (!acc_thread) return;
    const long ln = line_of(agg);
    for (int row = asub; row < nr; row += NSUB) {
      const long ri = ln * n + r0 + row;
      const double2 *yr = rhs + ln * sL * 5 + (long)(r0 + row) * rs;
      const cx<T> *P = p2 + ri * 50 + abt * 25 + amm * 5;
#pragma unroll
      for (int q = 0; q < 5; q++) bacc = cfma(P[q], zx(yr[q]), bacc);
    }
  };

  // ---- compute wave
  cx<T> yA[5], yB[5];
  cx<T> w2s[3][5], w1s[3][5], cs[3];
  // the term in the newest value: two FMA accumulators (four independent
  // chains over re and im), 22 instructions
  auto crit = [&](const cx<T> (&w)[5], const cx<T> (&yn)[5], cx<T> a0) -> cx<T> {
    cx<T> a = cfma(w[0], yn[0], a0), b = cmul(w[1], yn[1]);
    a = cfma(w[2], yn[2], a);
    b = cfma(w[3], yn[3], b);
    a = cfma(w[4], yn[4], a);
    return a + b;
  };
  auto pre0 = [&](const cx<T> (&w)[5], cx<T> c, const cx<T> (&yo)[5]) -> cx<T> {
    cx<T> a = cfma(w[0], yo[0], c), b = cmul(w[1], yo[1]);
    a = cfma(w[2], yo[2], a);
    b = cfma(w[3], yo[3], b);
    a = cfma(w[4], yo[4], a);
    return a + b;
  };
  // rows of buffer kk in processing order j (buffer row = j forward,
  // nr-1-j backward); extra(j) is added to the row's known part
  // two rows on the plain path: the border rows of either pass; extra adds
  // the forward border sums
  auto border_rows = [&](int kk, int ibase, bool fw, bool extra) {
    const int stride = fw ? 60 : 66;
    const cx<T> *B = rec[kk][g];
    for (int j = 0; j < 2; j++) {
      const int r = fw ? j : 1 - j;
      cx<T> w1[5], w2[5];
#pragma unroll
      for (int q = 0; q < 5; q++) { w2[q] = B[r * stride + m * 10 + q]; w1[q] = B[r * stride + m * 10 + 5 + q]; }
      cx<T> a = pre0(w1, cb[kk][g][r][m], yA);
      if (extra) a = a + bsum[g][j][m];
      cx<T> y = crit(w2, yB, a);
      X[(long)(fw ? ibase + j : ibase - j) * rs + m] = make_double2(y.re, y.im);
      row_bcast5(y, yA);
#pragma unroll
      for (int q = 0; q < 5; q++) { cx<T> t = yA[q]; yA[q] = yB[q]; yB[q] = t; }
    }
  };
  auto run_chunk = [&](int kk, int nr, int ibase, bool fw) {
    const int stride = fw ? 60 : 66;
    const cx<T> *B = rec[kk][g];
    auto rix = [&](int j) { j = min(j, nr - 1); return fw ? j : nr - 1 - j; };
    auto ldW2 = [&](int j, cx<T> (&w)[5]) {
      const cx<T> *pp = B + rix(j) * stride + m * 10;
#pragma unroll
      for (int q = 0; q < 5; q++) w[q] = pp[q];
    };
    auto ldW1 = [&](int j, cx<T> (&w)[5], cx<T> &c) {
      const int r = rix(j);
      const cx<T> *pp = B + r * stride + m * 10 + 5;
#pragma unroll
      for (int q = 0; q < 5; q++) w[q] = pp[q];
      c = cb[kk][g][r][m];
    };
    const long xstep = fw ? rs : -rs;
    double2 *xp = X + (long)(fw ? ibase : ibase) * rs + m;
    const int nmain = (nr / 6) * 6;
    if (nmain > 0) {
      ldW2(0, w2s[0]);
      ldW2(1, w2s[1]);
      ldW1(0, w1s[0], cs[0]);
      ldW1(1, w1s[1], cs[1]);
      ldW1(2, w1s[2], cs[2]);
      cx<T> a0 = pre0(w1s[0], cs[0], yA);
      auto sub = [&](int j, auto U) {
        constexpr int u = decltype(U)::value;
        constexpr int u0 = u % 3, u1 = (u + 1) % 3, u2 = (u + 2) % 3;
        cx<T>(&yn)[5] = (u % 2 == 0) ? yB : yA;
        cx<T>(&yd)[5] = (u % 2 == 0) ? yA : yB;
        cx<T> y = crit(w2s[u0], yn, a0);
        a0 = pre0(w1s[u1], cs[u1], yn);
        *xp = make_double2(y.re, y.im);
        xp += xstep;
        ldW2(j + 2, w2s[u2]);
        ldW1(j + 3, w1s[u0], cs[u0]);
        row_bcast5(y, yd);
      };
      for (int j = 0; j < nmain; j += 6) {
        sub(j, std::integral_constant<int, 0>());
        sub(j + 1, std::integral_constant<int, 1>());
        sub(j + 2, std::integral_constant<int, 2>());
        sub(j + 3, std::integral_constant<int, 3>());
        sub(j + 4, std::integral_constant<int, 4>());
        sub(j + 5, std::integral_constant<int, 5>());
      }
    }
    for (int j = nmain; j < nr; j++) {
      cx<T> w1[5], w2[5], c;
      ldW1(j, w1, c);
      ldW2(j, w2);
      cx<T> y = crit(w2, yB, pre0(w1, c, yA));
      *xp = make_double2(y.re, y.im);
      xp += xstep;
      row_bcast5(y, yA);
#pragma unroll
      for (int q = 0; q < 5; q++) { cx<T> t = yA[q]; yA[q] = yB[q]; yB[q] = t; }
    }
  };

  // ================= forward
#pragma unroll
  for (int q = 0; q < 5; q++) { yA[q] = cx<T>{0, 0}; yB[q] = cx<T>{0, 0}; }
  if (wave != 0) load(0, 0, min(R, nreg), true, false);
  __syncthreads();
  for (int k = 0; k < nck; k++) {
    const int r0 = k * R, nr = min(R, nreg - r0);
    if (wave == 0) {
      run_chunk(k & 1, nr, r0, true);
    } else {
      if (k >= 1) accumulate(r0 - R, R);
      if (k + 1 < nck) load((k + 1) & 1, r0 + R, min(R, nreg - r0 - R), true, false);
    }
    __syncthreads();
  }
  // the last regular chunk's sums, the records of rows n-2, n-1
  if (wave != 0) {
    const int r0 = (nck - 1) * R;
    accumulate(r0, nreg - r0);
    if (acc_thread) bpart[asub][aid] = bacc;
    load(nck & 1, nreg, 2, true, false);
  }
  __syncthreads();
  if (tid >= 64 && tid < 64 + 10 * G) {
    const int id = tid - 64;
    cx<T> t{0, 0};
    for (int k = 0; k < NSUB; k++) t = t + bpart[k][id];
    bsum[id / 10][(id / 5) % 2][id % 5] = t;
  }
  __syncthreads();
  if (wave == 0) border_rows(nck & 1, nreg, true, true);
  __syncthreads();

  // ================= backward
#pragma unroll
  for (int q = 0; q < 5; q++) { yA[q] = cx<T>{0, 0}; yB[q] = cx<T>{0, 0}; }
  if (wave != 0) load(0, nreg, 2, false, false);
  __syncthreads();
  if (wave == 0) {
    border_rows(0, n - 1, false, false);
    // x_{n-1} is yA, x_{n-2} is yB (all components on every lane); pick
    // component m with constant indices (a runtime index would move the
    // arrays out of registers)
    cx<T> xa = yB[0], xc = yA[0];
#pragma unroll
    for (int q = 1; q < 5; q++)
      if (m == q) { xa = yB[q]; xc = yA[q]; }
    if ((lane & 15) < 5) { xb[g][0][m] = xa; xb[g][1][m] = xc; }
  }
  __syncthreads();
  if (wave != 0) {
    const int rtop = nreg - 1, nr = min(R, rtop + 1);
    load(1, rtop - nr + 1, nr, false, true);
  }
  __syncthreads();
  for (int k = 1; k <= nck; k++) {
    const int rtop = nreg - 1 - (k - 1) * R, nr = min(R, rtop + 1);
    if (wave == 0) {
      run_chunk(k & 1, nr, rtop, false);
    } else if (k < nck) {
      const int rt2 = rtop - nr, nr2 = min(R, rt2 + 1);
      load((k + 1) & 1, rt2 - nr2 + 1, nr2, false, true);
    }
    __syncthreads();
  }
  // x(np) = x(1)
  if (wave == 0 && (lane & 15) < 5) X[(long)n * rs + m] = X[m];
}
template <typename T, int R, int G>
__global__ __launch_bounds__(256) void k_solve_per(int n, long L, const cx<T> *__restrict__ fwd,
                                                  const cx<T> *__restrict__ bwd,
                                                  const cx<T> *__restrict__ p2,
                                                  const cx<T> *__restrict__ qp, double2 *rhs,
                                                  long sL, long sR) {
  solve_per<T, R, G>(blockIdx.x, n, L, fwd, bwd, p2, qp, rhs, sL, sR);
}

// ---------------------------------------------------------------------------
// layout conversions between the reference's arrays and the working copy
// ref mat (5,5,5,...) element (slot k, r, c) of (line l, row i) at
//   ((i*sRm + l*sLm) * 125 + k + 5r + 25c)
__global__ void k_import_mat(const double2 *ref, cx<double> *lu, int n, long L, long sLm, long sRm,
                             int line_major) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;   // over (i, k, r, c, line)
  long total = (long)n * 125 * L;
  if (idx >= total) return;
  long line = idx % L;
  long e = idx / L;            // (i*5 + k)*25 + r*5 + c
  int c = e % 5, r = (e / 5) % 5, k = (e / 25) % 5;
  long i = e / 125;
  double2 v = ref[(i * sRm + line * sLm) * 125 + k + 5 * r + 25 * c];
  lu[line_major ? line * n * 125 + e : idx] = cx<double>{v.x, v.y};
}

// Factor health (the reference has no pivoting and no check: a zero pivot
// spreads Inf/NaN through its line, src/cpenta1bc.f:80-82 fact = one/pivot).
// One flag per line: set when an entry of the line's factors is not finite
// or a pivot (diagonal of the slot-3 block) is zero.  a holds total complex
// values; line-major arrays hold per_line values per line, interleaved ones
// put the line index fastest (stride L).
template <typename T>
__global__ void k_check_factor(const cx<T> *a, long total, long L, int line_major, long per_line,
                               int pivots, int *flag) {
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const cx<T> x = a[idx];
    const long line = line_major ? idx / per_line : idx % L;
    const long e = line_major ? idx % per_line : idx / L;     // ((row*5 + slot)*25 + 5r + c)
    bool bad = !(isfinite(x.re) && isfinite(x.im));
    if (pivots && (e / 25) % 5 == 2 && (e % 25) % 6 == 0 && x.re == T(0) && x.im == T(0)) bad = true;
    if (bad) flag[line] = 1;
  }
}
__global__ void k_zero_i32(int *x, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

// ---------------------------------------------------------------------------
// host side
static int grid_for(long n, int b) { return (int)((n + b - 1) / b); }

static int check_launch_grid(long total) { return (int)std::min<long>(grid_for(total, 256), 4096); }

int penta_check_launch(PentaPlan *pl, hipStream_t s) {
  if (!pl->flags) LNS3D_CHECK(hipMalloc(&pl->flags, (size_t)pl->L * sizeof(int)));
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(pl->L, 256)), dim3(256), 0, s, (int *)pl->flags, pl->L);
  const long n = pl->n, L = pl->L;
  if (pl->periodic) {
    const long tl = (long)pl->np * 125 * L, tp = (long)pl->np * 50 * L;
    hipLaunchKernelGGL(k_check_factor<double>, dim3(check_launch_grid(tl)), dim3(256), 0, s,
                       (const cx<double> *)pl->lu, tl, L, 0, 0L, 1, (int *)pl->flags);
    hipLaunchKernelGGL(k_check_factor<double>, dim3(check_launch_grid(tp)), dim3(256), 0, s,
                       (const cx<double> *)pl->per, tp, L, 0, 0L, 0, (int *)pl->flags);
    hipLaunchKernelGGL(k_check_factor<double>, dim3(check_launch_grid(tp)), dim3(256), 0, s,
                       (const cx<double> *)pl->per2, tp, L, 0, 0L, 0, (int *)pl->flags);
  } else {
    const long tl = n * 125 * L;
    hipLaunchKernelGGL(k_check_factor<double>, dim3(check_launch_grid(tl)), dim3(256), 0, s,
                       (const cx<double> *)pl->lu, tl, L, 1, n * 125, 1, (int *)pl->flags);
  }
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// records of a plan whose LU another plan holds (the complex-float records
// of the refined step): their entries, per line
int penta_check_records_launch(PentaPlan *pl, hipStream_t s) {
  if (!pl->prec) return 0;
  if (!pl->flags) LNS3D_CHECK(hipMalloc(&pl->flags, (size_t)pl->L * sizeof(int)));
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(pl->L, 256)), dim3(256), 0, s, (int *)pl->flags, pl->L);
  const long n = pl->n, L = pl->L;
  const long tf = n * 60 * L, tb = n * 66 * L;
  hipLaunchKernelGGL(k_check_factor<float>, dim3(check_launch_grid(tf)), dim3(256), 0, s,
                     (const cx<float> *)pl->fwd, tf, L, 1, n * 60, 0, (int *)pl->flags);
  hipLaunchKernelGGL(k_check_factor<float>, dim3(check_launch_grid(tb)), dim3(256), 0, s,
                     (const cx<float> *)pl->bwd, tb, L, 1, n * 66, 0, (int *)pl->flags);
  if (pl->periodic) {         // the border records, rounded too
    const long tq = n * 50 * L;
    hipLaunchKernelGGL(k_check_factor<float>, dim3(check_launch_grid(tq)), dim3(256), 0, s,
                       (const cx<float> *)pl->p2, tq, L, 1, n * 50, 0, (int *)pl->flags);
    hipLaunchKernelGGL(k_check_factor<float>, dim3(check_launch_grid(tq)), dim3(256), 0, s,
                       (const cx<float> *)pl->qp, tq, L, 1, n * 50, 0, (int *)pl->flags);
  }
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int penta_check_result(PentaPlan *pl, hipStream_t s, long *bad_lines, long *first_bad) {
  *bad_lines = 0;
  *first_bad = -1;
  if (!pl->flags) return 0;
  std::vector<int> f((size_t)pl->L);
  LNS3D_CHECK(hipMemcpyAsync(f.data(), pl->flags, f.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  LNS3D_CHECK(hipStreamSynchronize(s));
  for (long l = 0; l < pl->L; l++)
    if (f[l]) {
      if (*first_bad < 0) *first_bad = l;
      ++*bad_lines;
    }
  return 0;
}

static std::mutex g_tune_mu;
static std::map<std::tuple<int, long, int>, int> g_tuned;

int penta_variant_for(int n, long lines, int prec) {
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple(n, lines, prec));
    if (it != g_tuned.end()) return it->second;
  }
  if (n >= 16 && lines < 512) return prec ? PENTA_TWISTED_PAIRS : PENTA_TWISTED_HALF;
  return lines >= 512 ? PENTA_PLAIN3 : PENTA_PLAIN1;
}

int penta_tune_set(int n, long lines, int prec, int variant) {
  if (variant < -1 || variant >= PENTA_NVARIANT)
    return set_error(__FILE__, __LINE__, "line-solve variant out of range");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (variant < 0) g_tuned.erase(std::make_tuple(n, lines, prec));
  else g_tuned[std::make_tuple(n, lines, prec)] = variant;
  return 0;
}

int penta_plan_create(PentaPlan *pl, int periodic, int exact, int n_rows, long lines, int prec) {
  const int n = periodic ? n_rows - 1 : n_rows;
  return penta_plan_create_variant(pl, periodic, exact, n_rows, lines, prec,
                                   periodic ? PENTA_PLAIN1 : penta_variant_for(n, lines, prec));
}

int penta_plan_create_variant(PentaPlan *pl, int periodic, int exact, int n_rows, long lines, int prec,
                              int variant) {
  pl->periodic = periodic;
  pl->exact = exact;
  pl->np = n_rows;
  pl->n = periodic ? n_rows - 1 : n_rows;
  pl->L = lines;
  pl->prec = prec;
  if (pl->n < 7 || lines < 1) return set_error(__FILE__, __LINE__, "line length must be >= 7 unknowns");
  if (penta_twisted(variant) && pl->n < 16) variant = PENTA_PLAIN1;
  pl->variant = periodic ? PENTA_PLAIN1 : variant;
  pl->h = (!periodic && penta_twisted(pl->variant)) ? (pl->n - 1) / 2 : -1;
  size_t ne = (size_t)pl->np * 125 * lines;
  LNS3D_CHECK(hipMalloc(&pl->lu, ne * sizeof(cx<double>)));
  pl->per = pl->per2 = nullptr;
  pl->fwd = pl->bwd = pl->tail = pl->mid = pl->p2 = pl->qp = pl->xf = pl->xb = pl->flags = nullptr;
  if (periodic) {
    // complex-float plans take their records from a complex-double plan
    // (penta_records_narrow): no bordered LU of their own
    const size_t z = prec ? sizeof(cx<float>) : sizeof(cx<double>);
    if (!prec) {
      LNS3D_CHECK(hipMalloc(&pl->per, (size_t)pl->np * 50 * lines * z));
      LNS3D_CHECK(hipMalloc(&pl->per2, (size_t)pl->np * 50 * lines * z));
    }
    LNS3D_CHECK(hipMalloc(&pl->fwd, (size_t)pl->n * 60 * lines * z));
    LNS3D_CHECK(hipMalloc(&pl->bwd, (size_t)pl->n * 66 * lines * z));
    LNS3D_CHECK(hipMalloc(&pl->p2, (size_t)pl->n * 50 * lines * z));
    LNS3D_CHECK(hipMalloc(&pl->qp, (size_t)pl->n * 50 * lines * z));
  } else {
    size_t es = prec ? sizeof(cx<float>) : sizeof(cx<double>);
    LNS3D_CHECK(hipMalloc(&pl->fwd, (size_t)pl->n * 60 * lines * es));
    LNS3D_CHECK(hipMalloc(&pl->bwd, (size_t)pl->n * 66 * lines * es));
    LNS3D_CHECK(hipMalloc(&pl->tail, (size_t)150 * lines * (pl->h >= 0 ? 2 : 1) * es));
    if (pl->h >= 0) LNS3D_CHECK(hipMalloc(&pl->mid, (size_t)100 * lines * sizeof(cx<double>)));
    if (pl->variant == PENTA_TWISTED_PAIRS) {
      // 25 per (line, chain, pair), plus one complex past the end (a copy of
      // an odd count of complex float moves 16 B at a time)
      const size_t nx = ((size_t)lines * 2 * ((pl->h + 3) / 2) * 25 + 2) * es;
      LNS3D_CHECK(hipMalloc(&pl->xf, nx));
      LNS3D_CHECK(hipMalloc(&pl->xb, nx));
    }
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
  if (pl->mid) (void)hipFree(pl->mid);
  if (pl->p2) (void)hipFree(pl->p2);
  if (pl->qp) (void)hipFree(pl->qp);
  if (pl->xf) (void)hipFree(pl->xf);
  if (pl->xb) (void)hipFree(pl->xb);
  if (pl->flags) (void)hipFree(pl->flags);
  pl->lu = pl->per = pl->per2 = pl->fwd = pl->bwd = pl->tail = pl->mid = pl->p2 = pl->qp = nullptr;
  pl->xf = pl->xb = pl->flags = nullptr;
}

int penta_import(PentaPlan *pl, const void *ref_mat, long sLm, long sRm, hipStream_t s) {
  long total = (long)pl->np * 125 * pl->L;
  hipLaunchKernelGGL(k_import_mat, dim3(grid_for(total, 256)), dim3(256), 0, s,
                     (const double2 *)ref_mat, (cx<double> *)pl->lu, pl->np, pl->L, sLm, sRm,
                     pl->periodic ? 0 : 1);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int penta_factor_records_only(PentaPlan *pl, hipStream_t s) {
  if (pl->periodic) return set_error(__FILE__, __LINE__, "records are for non-periodic lines");
  LineGeom g{pl->n, pl->periodic, pl->exact, pl->L, pl->h};
  const int nb = grid_for(pl->L * (long)pl->n, 8);
  if (pl->prec)
    hipLaunchKernelGGL((k_records_rows<double, float>), dim3(nb), dim3(256), 0, s, g,
                       (const cx<double> *)pl->lu, (cx<float> *)pl->fwd, (cx<float> *)pl->bwd,
                       (cx<float> *)pl->tail);
  else
    hipLaunchKernelGGL((k_records_rows<double, double>), dim3(nb), dim3(256), 0, s, g,
                       (const cx<double> *)pl->lu, (cx<double> *)pl->fwd, (cx<double> *)pl->bwd,
                       (cx<double> *)pl->tail);
  LNS3D_CHECK(hipGetLastError());
  if (pl->variant == PENTA_TWISTED_PAIRS) {   // the pair rewrite for the solve's chunk size
    const long groups = pl->L * 4 * ((pl->h + 3) / 2);
    if (pl->prec)
      hipLaunchKernelGGL((k_pair_records<float>), dim3(grid_for(groups, 8)), dim3(256), 0, s, pl->n, pl->h, pl->L,
                         tw_chunk(true), (cx<float> *)pl->fwd, (cx<float> *)pl->bwd, (cx<float> *)pl->xf,
                         (cx<float> *)pl->xb);
    else
      hipLaunchKernelGGL((k_pair_records<double>), dim3(grid_for(groups, 8)), dim3(256), 0, s, pl->n, pl->h, pl->L,
                         tw_chunk(false), (cx<double> *)pl->fwd, (cx<double> *)pl->bwd, (cx<double> *)pl->xf,
                         (cx<double> *)pl->xb);
    LNS3D_CHECK(hipGetLastError());
  }
  return 0;
}

__global__ void k_narrow(const cx<double> *src, cx<float> *dst, long count) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const cx<double> a = src[t];
  dst[t] = cx<float>{(float)a.re, (float)a.im};
}

int penta_records_narrow(PentaPlan *dst, const PentaPlan *src, hipStream_t s) {
  if (!dst->periodic || !src->periodic || !dst->prec || src->prec || dst->n != src->n || dst->L != src->L)
    return set_error(__FILE__, __LINE__, "penta_records_narrow: a complex-float and a complex-double periodic plan of one shape");
  const long n = src->n, L = src->L;
  const void *from[4] = {src->fwd, src->bwd, src->p2, src->qp};
  void *to[4] = {dst->fwd, dst->bwd, dst->p2, dst->qp};
  const long count[4] = {n * 60 * L, n * 66 * L, n * 50 * L, n * 50 * L};
  for (int k = 0; k < 4; k++)
    hipLaunchKernelGGL(k_narrow, dim3(grid_for(count[k], 256)), dim3(256), 0, s, (const cx<double> *)from[k],
                       (cx<float> *)to[k], count[k]);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

int penta_factor(PentaPlan *pl, hipStream_t s) {
  LineGeom g{pl->n, pl->periodic, pl->exact, pl->L, pl->h};
  if (pl->periodic && pl->prec)
    return set_error(__FILE__, __LINE__, "periodic complex-float plans take their records from a complex-double plan");
  if (pl->periodic) {
    hipLaunchKernelGGL(k_lu_per, dim3((unsigned)pl->L), dim3(64), 0, s, g,
                       (cx<double> *)pl->lu, (cx<double> *)pl->per, (cx<double> *)pl->per2);
    LNS3D_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_records_per, dim3(grid_for(pl->L * (long)pl->n, 128)), dim3(128), 0, s, g,
                       (cx<double> *)pl->lu, (cx<double> *)pl->per, (cx<double> *)pl->per2,
                       (cx<double> *)pl->fwd, (cx<double> *)pl->bwd, (cx<double> *)pl->p2,
                       (cx<double> *)pl->qp);
    LNS3D_CHECK(hipGetLastError());
    return 0;
  }
  hipLaunchKernelGGL(k_factor_chains, dim3((unsigned)pl->L), dim3(pl->h >= 0 ? 128 : 64), 0, s, g,
                     (cx<double> *)pl->lu, (cx<double> *)pl->mid);
  LNS3D_CHECK(hipGetLastError());
  return penta_factor_records_only(pl, s);
}

// non-periodic solve kernels by variant.  Chunks of R rows, four LDS
// buffers: twisted lines and one line per workgroup use longer chunks (one
// workgroup per line holds more rows); two, three or four lines per
// workgroup take chunks of six rows so that two or more workgroups fit a CU.
template <typename T>
static void launch_solve_bc(const PentaPlan *pl, void *rhs, long sL, long sR, hipStream_t s) {
  constexpr bool D = sizeof(T) == 8;
  const auto *f = (const cx<T> *)pl->fwd;
  const auto *b = (const cx<T> *)pl->bwd;
  const auto *t = (const cx<T> *)pl->tail;
  auto *x = (double2 *)rhs;
  auto wgs = [&](int G) { return dim3((unsigned)((pl->L + G - 1) / G)); };
  switch (pl->variant) {
    // twisted lines, a ring of four buffers (five measured slower: the wave
    // preparing a chunk then also starts the next copies).  The kernel is
    // bounded to 256 VGPRs per wave: above that only one workgroup fits a CU,
    // and the 257th line of a 257-line solve would run alone after the rest.
    case PENTA_TWISTED:        // one step at a time
      hipLaunchKernelGGL((k_solve_tw<T, D ? 6 : 12, 1, false, 4>), wgs(1), dim3(256), 0, s, pl->n, pl->h, pl->L, f,
                         b, t, (const cx<T> *)pl->xf, (const cx<T> *)pl->xb, x, sL, sR);
      break;
    case PENTA_TWISTED_HALF:   // one step at a time, columns split over the wave's halves
      hipLaunchKernelGGL((k_solve_tw<T, D ? 6 : 12, 1, false, 4, true>), wgs(1), dim3(256), 0, s, pl->n, pl->h,
                         pl->L, f, b, t, (const cx<T> *)pl->xf, (const cx<T> *)pl->xb, x, sL, sR);
      break;
    case PENTA_TWISTED_PAIRS:  // pairs of steps (factor-time rewrite)
      static_assert(tw_chunk(!D) == (D ? 6 : 12), "the pair rewrite assumes the solve's chunk size");
      hipLaunchKernelGGL((k_solve_tw<T, D ? 6 : 12, 1, true, 4>), wgs(1), dim3(256), 0, s, pl->n, pl->h, pl->L, f, b,
                         t, (const cx<T> *)pl->xf, (const cx<T> *)pl->xb, x, sL, sR);
      break;
    case PENTA_PLAIN2:
      hipLaunchKernelGGL((k_solve_bc<T, D ? 6 : 12, 2>), wgs(2), dim3(256), 0, s, pl->n, pl->L, f, b, t, x, sL, sR);
      break;
    case PENTA_PLAIN3:
      hipLaunchKernelGGL((k_solve_bc<T, 6, 3>), wgs(3), dim3(256), 0, s, pl->n, pl->L, f, b, t, x, sL, sR);
      break;
    case PENTA_PLAIN4:
      hipLaunchKernelGGL((k_solve_bc<T, 6, 4>), wgs(4), dim3(256), 0, s, pl->n, pl->L, f, b, t, x, sL, sR);
      break;
    default:
      hipLaunchKernelGGL((k_solve_bc<T, D ? 12 : 24, 1>), wgs(1), dim3(256), 0, s, pl->n, pl->L, f, b, t, x, sL, sR);
  }
}

int penta_solve(PentaPlan *pl, void *rhs, long sL, long sR, hipStream_t s) {
  LineGeom g{pl->n, pl->periodic, pl->exact, pl->L};
  if (pl->periodic) {
    (void)g;
    auto per = [&](auto tag) {
      using T = decltype(tag);
      auto *f = (const cx<T> *)pl->fwd; auto *b = (const cx<T> *)pl->bwd;
      auto *p2 = (const cx<T> *)pl->p2; auto *qp = (const cx<T> *)pl->qp;
      // one line per workgroup while two workgroups per CU hold every line
      if (pl->L > 512)
        hipLaunchKernelGGL((k_solve_per<T, 6, 4>), dim3((unsigned)((pl->L + 3) / 4)), dim3(256), 0, s, pl->n, pl->L,
                           f, b, p2, qp, (double2 *)rhs, sL, sR);
      else
        hipLaunchKernelGGL((k_solve_per<T, 30, 1>), dim3((unsigned)pl->L), dim3(256), 0, s, pl->n, pl->L, f, b, p2,
                           qp, (double2 *)rhs, sL, sR);
    };
    if (pl->prec) per(float{});
    else per(double{});
  } else {
    if (pl->prec) launch_solve_bc<float>(pl, rhs, sL, sR, s);
    else launch_solve_bc<double>(pl, rhs, sL, sR, s);
  }
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// batches: one launch solves the lines of many plans of one solve kernel.
// Workgroup b belongs to member k with off[k] <= b < off[k + 1]; the member's
// item holds its plan's records and its two right-hand sides.
__device__ __forceinline__ int batch_member(const int *off, int count) {
  int lo = 0, hi = count;
  const int b = (int)blockIdx.x;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <typename T, int R, int G>
__global__ __launch_bounds__(256) void kb_solve_bc(const PentaBatchItem *items, const int *off, int count,
                                                  int which) {
  const int k = batch_member(off, count);
  const PentaBatchItem &b = items[k];
  solve_bc<T, R, G>((long)blockIdx.x - off[k], b.n, b.L, (const cx<T> *)b.fwd, (const cx<T> *)b.bwd,
                    (const cx<T> *)b.tail, b.rhs[which], b.sL, b.sR);
}

template <typename T, int R, int G, bool PAIRS, int NBUF, bool HALF = false>
__global__ __launch_bounds__(256, 2) void kb_solve_tw(const PentaBatchItem *items, const int *off, int count,
                                                     int which) {
  const int k = batch_member(off, count);
  const PentaBatchItem &b = items[k];
  solve_tw<T, R, G, PAIRS, NBUF, HALF>((long)blockIdx.x - off[k], b.n, b.h, b.L, (const cx<T> *)b.fwd,
                                       (const cx<T> *)b.bwd, (const cx<T> *)b.tail, (const cx<T> *)b.xf,
                                       (const cx<T> *)b.xb, b.rhs[which], b.sL, b.sR);
}

template <typename T, int R, int G>
__global__ __launch_bounds__(256) void kb_solve_per(const PentaBatchItem *items, const int *off, int count,
                                                   int which) {
  const int k = batch_member(off, count);
  const PentaBatchItem &b = items[k];
  solve_per<T, R, G>((long)blockIdx.x - off[k], b.n, b.L, (const cx<T> *)b.fwd, (const cx<T> *)b.bwd,
                     (const cx<T> *)b.p2, (const cx<T> *)b.qp, b.rhs[which], b.sL, b.sR);
}

// batch groups of periodic lines, by their solve kernel (as penta_solve
// chooses it: four lines per workgroup above 512 lines, else one)
constexpr int BATCH_PERIODIC1 = 1001, BATCH_PERIODIC4 = 1004;

// lines per workgroup of a batch group's solve kernel
static int variant_lines(int v) {
  return v == PENTA_PLAIN2 ? 2 : v == PENTA_PLAIN3 ? 3 : (v == PENTA_PLAIN4 || v == BATCH_PERIODIC4) ? 4 : 1;
}

int stage_upload(void *dev, const void *src, size_t bytes, std::vector<TableUpload> &ups) {
  void *h = nullptr;
  LNS3D_CHECK(hipHostMalloc(&h, bytes, hipHostMallocDefault));
  memcpy(h, src, bytes);
  ups.push_back(TableUpload{dev, h, bytes});
  return 0;
}

int penta_batch_build(const PentaPlan *const *plans, double2 *const *rhs0, double2 *const *rhs1, const long *sL,
                      const long *sR, int count, std::vector<PentaBatchGroup> &out,
                      std::vector<TableUpload> &ups) {
  out.clear();
  std::map<std::pair<int, int>, std::vector<int>> by;   // (variant, precision) -> members
  for (int k = 0; k < count; k++) {
    const PentaPlan *pl = plans[k];
    const int v = pl->periodic ? (pl->L > 512 ? BATCH_PERIODIC4 : BATCH_PERIODIC1) : pl->variant;
    by[{v, pl->prec}].push_back(k);
  }
  for (auto &kv : by) {
    // longest lines first: their workgroups start in the first wave and the
    // short ones fill in behind them
    std::stable_sort(kv.second.begin(), kv.second.end(),
                     [&](int a, int b) { return plans[a]->n > plans[b]->n; });
    PentaBatchGroup g;
    g.variant = kv.first.first;
    g.prec = kv.first.second;
    g.count = (int)kv.second.size();
    std::vector<PentaBatchItem> items;
    std::vector<int> off(1, 0);
    for (int k : kv.second) {
      const PentaPlan *pl = plans[k];
      PentaBatchItem it;
      it.n = pl->n;
      it.h = pl->h;
      it.L = pl->L;
      it.fwd = pl->fwd;
      it.bwd = pl->bwd;
      it.tail = pl->tail;
      it.xf = pl->xf;
      it.xb = pl->xb;
      it.p2 = pl->p2;
      it.qp = pl->qp;
      it.rhs[0] = rhs0[k];
      it.rhs[1] = rhs1[k];
      it.sL = sL[k];
      it.sR = sR[k];
      items.push_back(it);
      const long wg = (pl->L + variant_lines(g.variant) - 1) / variant_lines(g.variant);
      if ((long)off.back() + wg > 0x7fffffffL) return set_error(__FILE__, __LINE__, "batch too large");
      off.push_back(off.back() + (int)wg);
    }
    g.blocks = off.back();
    LNS3D_CHECK(hipMalloc(&g.items, items.size() * sizeof(PentaBatchItem)));
    out.push_back(g);             // freed with the groups from here on
    LNS3D_CHECK(hipMalloc(&out.back().off, off.size() * sizeof(int)));
    int rc = stage_upload(g.items, items.data(), items.size() * sizeof(PentaBatchItem), ups);
    if (!rc) rc = stage_upload(out.back().off, off.data(), off.size() * sizeof(int), ups);
    if (rc) return rc;
  }
  return 0;
}

void penta_batch_free(std::vector<PentaBatchGroup> &groups) {
  for (auto &g : groups) {
    if (g.items) (void)hipFree(g.items);
    if (g.off) (void)hipFree(g.off);
  }
  groups.clear();
}

template <typename T>
static void launch_batch_group(const PentaBatchGroup &g, int which, hipStream_t s) {
  constexpr bool D = sizeof(T) == 8;
  const auto *it = (const PentaBatchItem *)g.items;
  const dim3 grid((unsigned)g.blocks), blk(256);
  switch (g.variant) {
    case PENTA_TWISTED:
      hipLaunchKernelGGL((kb_solve_tw<T, D ? 6 : 12, 1, false, 4>), grid, blk, 0, s, it, g.off, g.count, which);
      break;
    case PENTA_TWISTED_HALF:
      hipLaunchKernelGGL((kb_solve_tw<T, D ? 6 : 12, 1, false, 4, true>), grid, blk, 0, s, it, g.off, g.count,
                         which);
      break;
    case PENTA_TWISTED_PAIRS:
      hipLaunchKernelGGL((kb_solve_tw<T, D ? 6 : 12, 1, true, 4>), grid, blk, 0, s, it, g.off, g.count, which);
      break;
    case PENTA_PLAIN2:
      hipLaunchKernelGGL((kb_solve_bc<T, D ? 6 : 12, 2>), grid, blk, 0, s, it, g.off, g.count, which);
      break;
    case PENTA_PLAIN3:
      hipLaunchKernelGGL((kb_solve_bc<T, 6, 3>), grid, blk, 0, s, it, g.off, g.count, which);
      break;
    case PENTA_PLAIN4:
      hipLaunchKernelGGL((kb_solve_bc<T, 6, 4>), grid, blk, 0, s, it, g.off, g.count, which);
      break;
    case BATCH_PERIODIC1:
      hipLaunchKernelGGL((kb_solve_per<T, 30, 1>), grid, blk, 0, s, it, g.off, g.count, which);
      break;
    case BATCH_PERIODIC4:
      hipLaunchKernelGGL((kb_solve_per<T, 6, 4>), grid, blk, 0, s, it, g.off, g.count, which);
      break;
    default:
      hipLaunchKernelGGL((kb_solve_bc<T, D ? 12 : 24, 1>), grid, blk, 0, s, it, g.off, g.count, which);
  }
}

int penta_batch_solve(const std::vector<PentaBatchGroup> &groups, int which, hipStream_t s) {
  for (const auto &g : groups) {
    if (g.prec) launch_batch_group<float>(g, which, s);
    else launch_batch_group<double>(g, which, s);
    LNS3D_CHECK(hipGetLastError());
  }
  return 0;
}

int penta_solve_reference_order(PentaPlan *pl, void *rhs, long sL, long sR, hipStream_t s) {
  if (pl->h >= 0)
    return set_error(__FILE__, __LINE__, "reference-order solve needs factors in row order (this plan is twisted)");
  LineGeom g{pl->n, pl->periodic, pl->exact, pl->L};
  hipLaunchKernelGGL(k_solve_lines<double>, dim3(grid_for(pl->L, 64)), dim3(64), 0, s, g,
                     (const cx<double> *)pl->lu, (const cx<double> *)pl->per,
                     (const cx<double> *)pl->per2, (double2 *)rhs, sL, sR);
  LNS3D_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// autotuner: synthetic diagonally dominant lines of the requested shape,
// every solve variant factored and timed, the fastest kept for the shape
__global__ void k_fill_synthetic(cx<double> *lu, long total, unsigned seed) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  unsigned x = (unsigned)(idx * 2654435761u) ^ seed;
  x ^= x >> 13; x *= 0x5bd1e995u; x ^= x >> 15;
  double a = ((x & 0xffff) / 65536.0 - 0.5) * 0.1, b = ((x >> 16) / 65536.0 - 0.5) * 0.1;
  const int e = (int)(idx % 125), k = e / 25, r = (e / 5) % 5, c = e % 5;
  if (k == 2 && r == c) a += 3.0;
  lu[idx] = cx<double>{a, b};
}

int penta_tune(int n, long lines, int prec, int *variant, float *times_ms) {
  float best = 1e30f;
  int bestv = -1;
  hipStream_t s;
  LNS3D_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  LNS3D_CHECK(hipEventCreate(&e0));
  LNS3D_CHECK(hipEventCreate(&e1));
  void *rhs = nullptr;
  LNS3D_CHECK(hipMalloc(&rhs, (size_t)n * lines * 5 * sizeof(double2)));
  LNS3D_CHECK(hipMemsetAsync(rhs, 0, (size_t)n * lines * 5 * sizeof(double2), s));
  int rc = 0;
  for (int v = 0; v < PENTA_NVARIANT && !rc; v++) {
    if (times_ms) times_ms[v] = -1.0f;
    if (penta_twisted(v) && n < 16) continue;
    PentaPlan pl;
    rc = penta_plan_create_variant(&pl, 0, 0, n, lines, prec, v);
    if (!rc) {
      long total = (long)n * 125 * lines;
      hipLaunchKernelGGL(k_fill_synthetic, dim3(grid_for(total, 256)), dim3(256), 0, s,
                         (cx<double> *)pl.lu, total, 12345u);
      rc = penta_factor(&pl, s);
    }
    if (!rc) rc = penta_solve(&pl, rhs, n, 1, s);      // warm-up
    if (!rc) {
      (void)hipEventRecord(e0, s);
      for (int k = 0; k < 3 && !rc; k++) rc = penta_solve(&pl, rhs, n, 1, s);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms /= 3;
      if (times_ms) times_ms[v] = ms;
      if (ms < best) { best = ms; bestv = v; }
    }
    (void)hipStreamSynchronize(s);
    penta_plan_destroy(&pl);
  }
  (void)hipFree(rhs);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[std::make_tuple(n, lines, prec)] = bestv;
  }
  if (variant) *variant = bestv;
  return 0;
}

}  // namespace lns3d
