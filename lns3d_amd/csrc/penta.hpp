// Host interface of the block-pentadiagonal line solvers (penta.hip).
#pragma once
#include <vector>
#include <hip/hip_runtime.h>

namespace lns3d {

// A device table of a batch graph, held in pinned host memory until the
// graph's first launch copies it on the launch stream (building a graph then
// puts no work on a device queue, where it could wait behind other streams'
// kernels).  stage_upload allocates the pinned copy; the graph frees it.
struct TableUpload {
  void *dev = nullptr;
  void *host = nullptr;
  size_t bytes = 0;
};
int stage_upload(void *dev, const void *src, size_t bytes, std::vector<TableUpload> &ups);

// One set of independent lines of equal length, factored once, solved many
// times.  Device buffers:
//   lu, per, per2 : LU working copy (np rows): line-interleaved for periodic
//                   lines, line-major for the others
//   fwd, bwd, tail: solve records of non-periodic lines (prec 0: complex
//                   double, 1: complex float)
struct PentaPlan {
  int periodic = 0, exact = 0;
  int n = 0, np = 0;   // unknown rows, stored rows (np = n+1 when periodic)
  int h = -1;          // twisted factor split (rows 0..h+1 | h+2..n-1), -1: plain
  int variant = 0;     // solve kernel (non-periodic): see penta_variant_* below
  long L = 0;          // number of lines
  int prec = 0;
  void *lu = nullptr, *per = nullptr, *per2 = nullptr;
  void *fwd = nullptr, *bwd = nullptr, *tail = nullptr;
  void *mid = nullptr;   // twisted factor: the bottom chain's share of the middle block
  void *p2 = nullptr, *qp = nullptr;   // periodic lines: border records (complex double)
  void *xf = nullptr, *xb = nullptr;   // twisted lines: per step pair, the second row's
                                       // original N2 (forward) / Q4 (backward) block
  void *flags = nullptr;   // per line: 1 when its factors hold a non-finite entry or a zero pivot
};

// one member of a batched line solve: a plan's records (non-periodic lines)
// and two right-hand sides, chosen per launch
struct PentaBatchItem {
  int n, h;
  long L;
  const void *fwd, *bwd, *tail, *xf, *xb;
  const void *p2, *qp;     // periodic lines: the border records
  double2 *rhs[2];
  long sL, sR;
};
// the members of a batch that share a solve kernel; device tables
struct PentaBatchGroup {
  int variant = 0, prec = 0, count = 0;
  long blocks = 0;
  void *items = nullptr;   // PentaBatchItem[count]
  int *off = nullptr;      // first workgroup of each member, [count + 1]
};

// solve-kernel variants of non-periodic lines: one line per workgroup with a
// deep LDS ring, two, three or four lines per workgroup, or twisted lines
// (two chains): single steps, steps in pairs, or single steps with each
// row's columns split over the two halves of the wave
enum { PENTA_PLAIN1 = 0, PENTA_PLAIN4 = 1, PENTA_TWISTED = 2, PENTA_PLAIN2 = 3, PENTA_PLAIN3 = 4,
       PENTA_TWISTED_PAIRS = 5, PENTA_TWISTED_HALF = 6, PENTA_NVARIANT = 7 };
inline bool penta_twisted(int v) {
  return v == PENTA_TWISTED || v == PENTA_TWISTED_PAIRS || v == PENTA_TWISTED_HALF;
}
// the variant a plan of this shape uses: the autotuned one when
// penta_tune has measured the shape, else a fixed rule
int penta_variant_for(int n, long lines, int prec);
// time every variant on synthetic lines of this shape (n rows, lines,
// precision) and record the fastest; returns the chosen variant in *variant
int penta_tune(int n, long lines, int prec, int *variant, float *times_ms);
// record a variant for the shape without timing (later plans use it);
// variant -1 drops the record (back to the fixed rule)
int penta_tune_set(int n, long lines, int prec, int variant);
int penta_plan_create(PentaPlan *pl, int periodic, int exact, int n_rows, long lines, int prec);
int penta_plan_create_variant(PentaPlan *pl, int periodic, int exact, int n_rows, long lines, int prec,
                              int variant);
void penta_plan_destroy(PentaPlan *pl);
// copy a device matrix in the reference's layout mat(5,5,5,...) into the
// working copy; block (line l, row i) starts at complex offset (i*sRm + l*sLm)*125
int penta_import(PentaPlan *pl, const void *ref_mat, long sLm, long sRm, hipStream_t s);
// assembled working copy -> factors (+ records for non-periodic lines)
int penta_factor(PentaPlan *pl, hipStream_t s);
// records only, from an LU already in pl->lu (non-periodic lines)
int penta_factor_records_only(PentaPlan *pl, hipStream_t s);
// periodic lines: the complex-float records of dst, rounded from the
// complex-double records of src (same shape, factored)
int penta_records_narrow(PentaPlan *dst, const PentaPlan *src, hipStream_t s);
// in-place solve of rhs (complex double); component d of (line l, row i) at
// complex offset (l*sL + i*sR)*5 + d
int penta_solve(PentaPlan *pl, void *rhs, long sL, long sR, hipStream_t s);
int penta_solve_reference_order(PentaPlan *pl, void *rhs, long sL, long sR, hipStream_t s);
// batched solves: the plans grouped by solve kernel, with right-hand sides
// rhs0[k], rhs1[k] of strides sL[k], sR[k]; penta_batch_solve launches one
// kernel per group on right-hand side `which` (0 or 1) of every member
int penta_batch_build(const PentaPlan *const *plans, double2 *const *rhs0, double2 *const *rhs1, const long *sL,
                      const long *sR, int count, std::vector<PentaBatchGroup> &out,
                      std::vector<TableUpload> &ups);
int penta_batch_solve(const std::vector<PentaBatchGroup> &groups, int which, hipStream_t s);
void penta_batch_free(std::vector<PentaBatchGroup> &groups);
// factor health: flag the lines whose LU (and border fill) holds a
// non-finite entry or a zero pivot (launch after penta_factor), or whose
// complex-float records hold a non-finite entry (after the records); then
// count them (synchronises s)
int penta_check_launch(PentaPlan *pl, hipStream_t s);
int penta_check_records_launch(PentaPlan *pl, hipStream_t s);
int penta_check_result(PentaPlan *pl, hipStream_t s, long *bad_lines, long *first_bad);

}  // namespace lns3d
