// Device-side pieces of one implicit step (step.hip) and their host launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/lns3d_amd.h"
#include "penta.hpp"

namespace lns3d {

// finite-difference weights derived from the scheme selectors and spacing
struct Stencil {
  double gx[6], gy[6];   // interior first derivative (premultiplied by 1/d)
  double dx[7], dy[7];   // interior second derivative (premultiplied by 1/d^2)
  double dxinv, dyinv, dxsinv, dysinv;
  double dxi, deta;
  int sym;               // symmetric xi ends (left/right = 7): bit 0 left, bit 1 right
};
Stencil make_stencil(const lns3d_params &p);
// the same from explicit spacings and symmetric-end bits (bit 0 left, 1 right)
Stencil make_stencil_spacing(int optx, int opty, double dxi, double deta, int sym);
// src/cgrad.f90 and src/cgrad2.f90 on a device field v (5,nx,ny) complex:
// g1, g2 first derivatives; g11, g12 (= d/deta of g1), g22 second
int launch_cgrad(int nx, int ny, int xper, int yper, const Stencil &st, const void *v, void *g1, void *g2,
                 hipStream_t s);
int launch_cgrad2(int nx, int ny, int xper, int yper, const Stencil &st, const void *v, const void *g1,
                  void *g11, void *g12, void *g22, hipStream_t s);

// coefficients on the device, structure of arrays over the P = nx*ny points
struct Coeffs {
  double *Ah, *Bh, *Dh, *Dhi;   // [25][P], entry (r,c) at index r + 5c
  double *V11, *V12, *V22;      // [6][P]
  double *ABhi;                 // [4][P]
  double *dtl;                  // [P]
  double *vm;                   // [5][P] mean flow (AoS (5,nx,ny) as passed)
  double *metric;               // [10][P]
  // sponge (src/lrhs3d.f90:222-282, src/lhs2f3d.f90:207-225); null: none
  double *sp1;                  // [P] on v in the residual: spg (ispg 1, 2, 3) or spg + spg2 (5)
  double *sp2;                  // [P] inflow sponge spg2 on (v - forcing) (ispg 2, 3)
  double *spf;                  // [5][P] complex inflow forcing, (re, im) pairs (ispg 2, 3)
  double *spl;                  // [P] eta-factor diagonal: spg (ispg 1) or spg + spg2
  // top = 1 (src/rhsbc3d.f90:129-161, src/itrbc3d.f90:145-195): the incoming
  // wave on the top boundary, [5][nx] complex as (re, im) pairs
  double *ftop;
  // left = 1, 4 and right = 4 (src/rhsbc3d.f90:176-246, src/itrbc3d.f90
  // :219-376): the values the left and right columns are held at, [2][5][ny]
  // complex as (re, im) pairs (side 0 left, side 1 right)
  double *fside;
  // smoother (src/smoother3d.f90 and the implicit damping of the factors):
  // eps_e buff per point, [P]; null: eps_e = 0
  double *sbuf;
  // the bump wall (wall = 3, src/rhsbc3d.f90:47-100, src/itrbc3d.f90:48-105):
  // u1w, u2w, u3w, tw, twp, [5][nx] complex as (re, im) pairs
  double *fwall;
  double S0;                    // Sutherland's constant (mattyp = 1), 110.4 K / Tref
  // the wall rows of the factors, the residual and the field (src/lhsbc1f3d.f90,
  // lhsbc2f3d.f90, rhsbc3d.f90, itrbc3d.f90) cover points wlo..whi of the wall
  // (the &extras is:ie, 0-based inclusive)
  int wlo = 0, whi = 0;
  // the inviscid deck (Navier = .false.): the wall normals bnb, [2][nx]
  // (src/gengrid.f90:133-137); null for the viscous rows
  const double *bnb = nullptr;
  // the xi factor carries the D terms (1; the eta factor the sponge), or the
  // eta factor does (0: useCalcd, src/lhs1f3d.f90:27-37)
  int calcd_xi = 1;
};

int launch_genmtrx(const lns3d_params &p, const Stencil &st, Coeffs &c, double *work, hipStream_t s);
int launch_timestep(const lns3d_params &p, Coeffs &c, double *work, hipStream_t s);
// inflow forcing of the sponge (ispg 2 or 3) into c.spf from the mean flow c.vm,
// the physical x (P, device) and wamp (nx, device; ispg 3)
int launch_sponge_forcing(const lns3d_params &p, int ispg, const Coeffs &c, const double *x,
                          const double *wamp, double x0, hipStream_t s);
// the top = 1 wave into c.ftop from the mean flow c.vm, the top row of the
// physical x (nx, device) and wamp (nx, device; null for 1)
// the acoustic wave of the left (side 0) or right (side 1) column into
// c.fside, damped (left = 1) or not (left/right = 5), from the mean flow and
// the x value each line reads (ny, device)
int launch_side_wave(const lns3d_params &p, const Coeffs &c, const double *xcol, double x0, int side,
                     int damped, hipStream_t s);
int launch_top_forcing(const lns3d_params &p, const Coeffs &c, const double *xtop, const double *wamp,
                       double x0, hipStream_t s);
// assemble the LHS of direction dir (1: xi, 2: eta) into the line-interleaved
// working copy of plan (lines: y index for dir 1, x index for dir 2)
int launch_assemble(const lns3d_params &p, const Coeffs &c, int dir, PentaPlan &plan, hipStream_t s);
// assembled LHS in the reference layout mat(5,5,5,nx,ny) (testing)
int launch_assemble_ref(const lns3d_params &p, const Coeffs &c, int dir, void *mat, hipStream_t s);
// r = -dtl * L(v) with the boundary rows of rhsBC3D; work holds 5P complex
// with_rbe: 0 the operator only, 1 backward Euler (rbe3D), 2 two-step
// backward (rtb3D) with v2old, the field at the start of the previous step,
// both with vold = v; 3 and 4 the same with vold given (corrector iterations);
// 5 the forced form vold - v - dtl (L(v) - f) of the shift-invert driver
int launch_rhs(const lns3d_params &p, const Stencil &st, const Coeffs &c, const void *v, void *r,
               void *work, int with_rbe, hipStream_t s, const void *v2old = nullptr,
               const void *vold = nullptr, const void *forcing = nullptr);
// the residual norm sqrt(sum |r|^2 / dtl^2) into *out (device), its
// per-block partials in part (NORM_BLOCKS doubles); no host synchronisation
constexpr int NORM_BLOCKS = 1024;
int launch_residual_norm(const double *dtl, const void *r, long P, double *part, double *out, hipStream_t s);
// the same with the partials in the library's static device array (no
// allocation, calls on one GPU share it); on the current device
constexpr int MAX_DEVICES = 64;
int launch_residual_norm_noalloc(const double *dtl, const void *r, long P, double *out, hipStream_t s);
// sum over the points of |r|^2 / dtl^2, per block into part (at most
// NORM_BLOCKS doubles); the caller adds them
int launch_forced_norm(const lns3d_params &p, const Coeffs &c, const void *r, double *part, int *nb,
                       hipStream_t s);
// ftop: the top = 1 wave (Coeffs::ftop), null for top = 0
// v += dv and itrBC3D; for the inviscid deck bcfix turns dv's wall row back
// to x, y components first (src/impdrv3d.f90:333-336)
int launch_update_bc(const lns3d_params &p, const Coeffs &c, void *v, void *dv, hipStream_t s);
int launch_itrbc(const lns3d_params &p, const Coeffs &c, void *v, hipStream_t s);
// the wall normals bnb [2][nx] from the metric's n1, n2 on the wall row
int launch_wall_normals(const lns3d_params &p, const Coeffs &c, double *bnb, hipStream_t s);
// y = M x with M = Lx * Ly the approximate-factorization operator (FP64,
// blocks regenerated from the coefficients); tmp holds 5P complex
int launch_af_matvec(const lns3d_params &p, const Coeffs &c, const void *x, void *y, void *tmp,
                     hipStream_t s);
// y = b - y (complex, n values) and the max norms of y and b into norms[0..1];
// norms holds 2 + 2*512 doubles (the rest is per-block partials)
int launch_residual_norms(const void *b, void *y, long n, double *norms, hipStream_t s);
int launch_axpy(void *x, const void *d, long n, hipStream_t s);   // x += d
int launch_copy(void *dst, const void *src, long n, hipStream_t s);   // complex n values

// batched steps: one launch per stage over the members of a batch (device
// tables built once).  Buffers of a member: v, dv, rhs (the step's
// residual), t1, t2 (the correction), work (U_xi)
enum { SB_V = 0, SB_DV, SB_RHS, SB_T1, SB_T2, SB_WORK };
// elementwise stages: dv = rhs; t2 = rhs - t2; dv += t2; dv += t2 then
// v += dv; v += dv
enum { SB_COPY = 0, SB_SUB, SB_ADD_DV, SB_UPDATE, SB_ADD_V };
enum { SB_STAGE_D1X = 0, SB_STAGE_POINT, SB_STAGE_LINE, SB_STAGE_VEC, SB_STAGE_WALL, SB_STAGE_LR, SB_STAGE_NORM,
       SB_STAGE_BCFIX, SB_NSTAGE };
// norm (optional): the member's residual-norm output of norm_count slots;
// step k of a captured sequence writes slot (norm_idx + k) mod norm_count
struct StepBatchMember {
  lns3d_params p;
  Stencil st;
  Coeffs c;
  double2 *v, *dv, *rhs, *t1, *t2, *work;
  double *norm = nullptr;
  int norm_count = 0;
  long norm_idx = 0;
};
struct StepBatch {
  int count = 0;
  void *mem = nullptr;       // the one device block the tables live in
  void *items = nullptr;
  int *off[SB_NSTAGE] = {};
  long blocks[SB_NSTAGE] = {};
  double *part = nullptr;    // per-block partial sums of the norm stage
};
// the device tables (their contents staged in ups for the first launch)
int step_batch_build(const StepBatchMember *m, int count, StepBatch &out, std::vector<TableUpload> &ups);
void step_batch_free(StepBatch &b);
// rhs = the backward-Euler residual with the rhsBC3D rows (launch_rhs mode 1)
int step_batch_rhs(const StepBatch &b, hipStream_t s);
// t2 = Lx (Ly dv) through t1 (launch_af_matvec)
int step_batch_matvec(const StepBatch &b, hipStream_t s);
int step_batch_vec(const StepBatch &b, int op, hipStream_t s);
// itrBC3D of v (launch_itrbc)
int step_batch_itrbc(const StepBatch &b, hipStream_t s);
// the residual norm sqrt(sum |rhs|^2 / dtl^2) of every member with a norm
// output, into its slot for step k (the bits of launch_residual_norm)
int step_batch_norm(const StepBatch &b, long k, hipStream_t s);
// bcfix on the increments of the members with the inviscid wall rows (no
// launch when none has them); between dv += t2 and v += dv
int step_batch_bcfix(const StepBatch &b, hipStream_t s);

}  // namespace lns3d
