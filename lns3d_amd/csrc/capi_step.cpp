// C-ABI of the device-resident implicit step: one context per problem
// (grid, mean flow, coefficients, factors, work arrays), all in HBM.
#include <hip/hip_runtime.h>
#include <math.h>
#include <cmath>
#include <stdio.h>
#include <string.h>
#include <string>
#include <complex>
#include <utility>
#include <vector>

#include "../../include/lns3d_amd.h"
#include "common.hpp"
#include "ctx.hpp"
#include "step.hpp"

using namespace lns3d;

namespace {
int fail(int code, const char *msg) {
  lns3d::set_error("capi_step", 0, msg);
  return code;
}
}  // namespace


static int check_params(const lns3d_params &p) {
  if (p.nx < 7 || p.ny < 7) return fail(LNS3D_EINVAL, "nx and ny must be >= 7");
  if (!std::isfinite(p.Ma) || !std::isfinite(p.Re) || !std::isfinite(p.Pr) || !std::isfinite(p.gamma) ||
      !std::isfinite(p.kz) || !std::isfinite(p.omega))
    return fail(LNS3D_EINVAL, "Ma, Re, Pr, gamma, kz and omega must be finite");
  if (p.Re < 0.0) return fail(LNS3D_EINVAL, "Re must be >= 0");
  // Re = 0 is the reference's inviscid deck for mattyp = 0 (src/input.f90:31-36)
  if (p.Re == 0.0 && p.mattyp != 0)
    return fail(LNS3D_EINVAL, "Re = 0 (inviscid) is defined for mattyp = 0 only (src/input.f90:31-36)");
  if (p.wall_is < 0 || p.wall_ie < 0 || p.wall_is > p.nx || p.wall_ie > p.nx ||
      ((p.wall_is ? p.wall_is : 1) > (p.wall_ie ? p.wall_ie : p.nx)))
    return fail(LNS3D_EINVAL, "wall extent: 1 <= wall_is <= wall_ie <= nx (0 for the default)");
  // idamp (src/input.f90:268, src/damper.f90): 1 damps the outflow column's
  // mean xi second derivatives and Vh11 (src/genmtrx.f90:87-96, 877-881);
  // 2 only forms damp, which this path never reads
  if (p.idamp < 0 || p.idamp > 2) return fail(LNS3D_EINVAL, "idamp must be 0, 1 or 2 (src/damper.f90:26-47)");
  if ((p.use_calcd != 0 && p.use_calcd != 1) || (p.update_lhs != 0 && p.update_lhs != 1))
    return fail(LNS3D_EINVAL, "use_calcd and update_lhs are 0 or 1");
  if (p.wall < 0 || p.wall > 3 || p.wallt < 0 || p.wallt > 2 || (p.top != 0 && p.top != 1) ||
      (p.left != 0 && p.left != 1 && p.left != 4 && p.left != 5 && p.left != 7) ||
      (p.right != 0 && p.right != 4 && p.right != 5 && p.right != 7 && p.right != 8 && p.right != 9) ||
      (p.linear != 1 && p.linear != 2) ||
      (p.mattyp < 0 || p.mattyp > 3) || (p.loctime != 0 && p.loctime != 1))
    return fail(LNS3D_EUNSUPPORTED, "unsupported boundary/physics selector (see lns3d_params)");
  // right = 4 reads the eigenfunction profiles, which the reference reads
  // only for left = 4 (src/gendist.f90:17)
  if (!p.xper && p.right == 4 && p.left != 4)
    return fail(LNS3D_EUNSUPPORTED, "right = 4 needs left = 4 (the eigenfunction inflow)");
  // right = 5 reads x(j,nx) of x(nx,ny) (src/rhsbc3d.f90:296): on the grid
  // only for nx <= ny
  if (!p.xper && p.right == 5 && p.ny - 1 + (long)p.nx * (p.nx - 1) >= (long)p.nx * p.ny)
    return fail(LNS3D_EUNSUPPORTED, "right = 5 needs nx <= ny (the reference reads x(j,nx) off the grid otherwise)");
  // symmetric ends: src/lhsbs1f3d.f90:17 reads the spanwise cross entries of
  // row (i, j) at point i + ny j (its Ahi(6,ny,nx) shape), which lies on the
  // grid only when nx >= ny (as in the reference's own symmetric case)
  const long P = (long)p.nx * p.ny;
  if (!p.xper && ((p.left == 7 && 1 + (long)p.ny * (p.ny - 1) >= P) ||
                  (p.right == 7 && p.nx - 1 + (long)p.ny * (p.ny - 1) >= P)))
    return fail(LNS3D_EUNSUPPORTED, "left/right = 7 needs nx >= ny (the reference reads off the grid otherwise)");
  if (p.precision != 0 && p.precision != 1) return fail(LNS3D_EINVAL, "precision must be 0 or 1");
  return 0;
}

// the deck into the context: the inviscid deck holds Re = 1e99 (src/input.f90
// :31-36); the wall extent and the wall normals of its rows
static int apply_params(lns3d_ctx *c, const lns3d_params &p) {
  c->p = p;
  c->navier = !(p.Re == 0.0 && p.mattyp == 0);
  if (!c->navier) c->p.Re = 1.0e99;
  c->c.wlo = p.wall_is ? p.wall_is - 1 : 0;
  c->c.whi = p.wall_ie ? p.wall_ie - 1 : p.nx - 1;
  if (!c->navier && !c->bnb_block) LNS3D_CHECK(hipMalloc(&c->bnb_block, (size_t)2 * p.nx * sizeof(double)));
  c->c.bnb = c->navier ? nullptr : c->bnb_block;
  c->bnb_ready = false;
  c->c.calcd_xi = 1;
  c->cur_iter = 1;
  return 0;
}

extern "C" int lns3d_ctx_create(const lns3d_params *p, lns3d_ctx **out) {
  if (!p || !out) return fail(LNS3D_EINVAL, "lns3d_ctx_create: null argument");
  int rc = check_params(*p);
  if (rc) return rc;
  auto *c = new lns3d_ctx;
  if ((rc = apply_params(c, *p))) {
    delete c;
    return rc;
  }
  c->st = make_stencil(*p);
  const long P = (long)p->nx * p->ny;
  size_t ncoef = (size_t)P * (25 * 4 + 6 * 3 + 4 + 1 + 5 + 10);
  auto bail = [&](int code) {
    lns3d_ctx_destroy(c);
    return code;
  };
  if (hipMalloc(&c->coef_block, ncoef * sizeof(double)) != hipSuccess) return bail(fail(LNS3D_EHIP, "hipMalloc coefficients"));
  double *b = c->coef_block;
  c->c.Ah = b; b += 25 * P;
  c->c.Bh = b; b += 25 * P;
  c->c.Dh = b; b += 25 * P;
  c->c.Dhi = b; b += 25 * P;
  c->c.V11 = b; b += 6 * P;
  c->c.V12 = b; b += 6 * P;
  c->c.V22 = b; b += 6 * P;
  c->c.ABhi = b; b += 4 * P;
  c->c.dtl = b; b += P;
  c->c.vm = b; b += 5 * P;
  c->c.metric = b;
  if (hipMalloc(&c->work, (size_t)P * 5 * 16) != hipSuccess ||
      hipMalloc(&c->r, (size_t)P * 5 * 16) != hipSuccess ||
      hipMalloc(&c->scal, 4096 * sizeof(double)) != hipSuccess)
    return bail(fail(LNS3D_EHIP, "hipMalloc work"));
  rc = penta_plan_create(&c->px, p->xper, p->periodic_exact, p->nx, p->ny, 0);
  if (!rc) rc = penta_plan_create(&c->py, p->yper, p->periodic_exact, p->ny, p->nx, 0);
  if (rc) return bail(rc);
  *out = c;
  return 0;
}

extern "C" int lns3d_ctx_destroy(lns3d_ctx *c) {
  if (!c) return 0;
  penta_plan_destroy(&c->px);
  penta_plan_destroy(&c->py);
  penta_plan_destroy(&c->px32);
  penta_plan_destroy(&c->py32);
  if (c->coef_block) (void)hipFree(c->coef_block);
  if (c->work) (void)hipFree(c->work);
  if (c->r) (void)hipFree(c->r);
  if (c->t1) (void)hipFree(c->t1);
  if (c->t2) (void)hipFree(c->t2);
  if (c->t3) (void)hipFree(c->t3);
  if (c->scal) (void)hipFree(c->scal);
  if (c->vstate) (void)hipFree(c->vstate);
  if (c->sponge_block) (void)hipFree(c->sponge_block);
  if (c->top_block) (void)hipFree(c->top_block);
  if (c->side_block) (void)hipFree(c->side_block);
  if (c->smooth_block) (void)hipFree(c->smooth_block);
  if (c->wall_block) (void)hipFree(c->wall_block);
  if (c->bnb_block) (void)hipFree(c->bnb_block);
  if (c->vhist) (void)hipFree(c->vhist);
  if (c->vstart) (void)hipFree(c->vstart);
  delete c;
  return 0;
}

// new physics for an existing context (same grid size, periodicity and
// precision): the coefficients and factors must be formed again
// (lns3d_setup, lns3d_factor)
extern "C" int lns3d_ctx_set_params(lns3d_ctx *c, const lns3d_params *p) {
  if (!c || !p) return fail(LNS3D_EINVAL, "lns3d_ctx_set_params: null argument");
  if (p->nx != c->p.nx || p->ny != c->p.ny || p->xper != c->p.xper || p->yper != c->p.yper ||
      p->precision != c->p.precision)
    return fail(LNS3D_EINVAL, "lns3d_ctx_set_params: grid size, periodicity and precision are fixed");
  int rc = check_params(*p);
  if (rc) return rc;
  if ((rc = apply_params(c, *p))) return rc;
  c->st = make_stencil(*p);
  c->px.exact = c->py.exact = p->periodic_exact;
  c->coeffs_ready = false;
  c->vm_ready = false;
  c->factored = false;
  // a new run: the time scheme starts again; the sponge and the top wave,
  // formed from the old physics, are set again by the caller
  c->lstep = 0;
  if (c->impl != 1) c->p.alfa = c->impl == 3 ? 0.5 : 1.0;
  c->c.sp1 = c->c.sp2 = c->c.spf = c->c.spl = nullptr;
  c->c.ftop = nullptr;
  c->c.fside = nullptr;
  c->c.sbuf = nullptr;
  c->c.fwall = nullptr;
  c->eig.clear();
  return 0;
}

extern "C" int lns3d_ctx_set_stream(lns3d_ctx *c, void *stream) {
  if (!c) return fail(LNS3D_EINVAL, "null context");
  c->stream = (hipStream_t)stream;
  return 0;
}

// metric (nx,ny,10) and vm (5,nx,ny), host arrays; coefficients and the time
// step are then generated on the device (src/genmtrx.f90, src/temporal.f90)
extern "C" int lns3d_setup(lns3d_ctx *c, const double *metric, const double *vm) {
  if (!c || !metric || !vm) return fail(LNS3D_EINVAL, "lns3d_setup: null argument");
  if (c->p.mattyp == 1 && !c->s0_set)
    return fail(LNS3D_EINVAL, "lns3d_setup: mattyp = 1 needs Sutherland's constant (lns3d_set_sutherland)");
  const long P = (long)c->p.nx * c->p.ny;
  LNS3D_CHECK(hipMemcpyAsync(c->c.metric, metric, (size_t)P * 10 * 8, hipMemcpyHostToDevice, c->stream));
  LNS3D_CHECK(hipMemcpyAsync(c->c.vm, vm, (size_t)P * 5 * 8, hipMemcpyHostToDevice, c->stream));
  int rc = 0;
  if (!c->navier) {
    rc = launch_wall_normals(c->p, c->c, c->bnb_block, c->stream);
    if (rc) return rc;
    c->bnb_ready = true;
  }
  rc = launch_genmtrx(c->p, c->st, c->c, (double *)c->work, c->stream);
  if (!rc) rc = launch_timestep(c->p, c->c, c->scal, c->stream);
  if (rc) return rc;
  c->coeffs_ready = true;
  c->vm_ready = true;
  c->factored = false;
  return 0;
}

// Sponge (src/global.f90:132-141 module state: ispg, spg, spg2, wamp, x0;
// terms of src/lrhs3d.f90:222-282 and src/lhs2f3d.f90:207-225).  The
// residual's sponge terms and the eta factor's diagonal are formed from the
// profiles here; the inflow forcing (ispg 2, 3) from the mean flow, x, wamp,
// x0 and the context's omega, Ma, Re, Pr, gamma.
extern "C" int lns3d_set_sponge(lns3d_ctx *c, int ispg, const double *spg, const double *spg2,
                                const double *x, const double *wamp, double x0) {
  if (!c) return fail(LNS3D_EINVAL, "lns3d_set_sponge: null context");
  if (ispg < 0 || ispg > 5) return fail(LNS3D_EINVAL, "lns3d_set_sponge: ispg must be 0..5");
  if (ispg == 4 && (c->p.left != 4 || c->eig.empty() || c->p.omega == 0.0))
    return fail(LNS3D_EUNSUPPORTED, "lns3d_set_sponge: ispg = 4 (src/sponge.f90:536, cspg_it) needs left = 4 "
                                    "with its profiles (lns3d_set_inflow first) and omega != 0");
  c->factored = false;
  if (ispg == 0) {
    c->c.sp1 = c->c.sp2 = c->c.spf = c->c.spl = nullptr;
    return 0;
  }
  const bool inflow = ispg == 2 || ispg == 3;
  if (!spg || (ispg >= 2 && !spg2) || ((inflow || ispg == 4) && !x) || (ispg == 3 && !wamp))
    return fail(LNS3D_EINVAL, "lns3d_set_sponge: a profile this ispg needs is null");
  if (inflow && !c->vm_ready)
    return fail(LNS3D_EINVAL, "lns3d_set_sponge: the inflow forcing needs the mean flow (lns3d_setup first)");
  const long P = (long)c->p.nx * c->p.ny;
  if (!c->sponge_block)
    LNS3D_CHECK(hipMalloc(&c->sponge_block, (size_t)(14 * P + c->p.nx) * sizeof(double)));
  double *sp1 = c->sponge_block, *sp2 = sp1 + P, *spl = sp2 + P, *spf = spl + P, *xd = spf + 10 * P,
         *wd = xd + P;
  std::vector<double> h1((size_t)P), hl((size_t)P);
  for (long q = 0; q < P; q++) {
    h1[q] = ispg == 5 ? spg[q] + spg2[q] : spg[q];
    hl[q] = ispg == 1 ? spg[q] : spg[q] + spg2[q];
  }
  LNS3D_CHECK(hipMemcpyAsync(sp1, h1.data(), (size_t)P * 8, hipMemcpyHostToDevice, c->stream));
  LNS3D_CHECK(hipMemcpyAsync(spl, hl.data(), (size_t)P * 8, hipMemcpyHostToDevice, c->stream));
  c->c.sp1 = sp1;
  c->c.spl = spl;
  c->c.sp2 = c->c.spf = nullptr;
  std::vector<double> hf;
  if (ispg == 4) {
    // cspg_it (src/sponge.f90:536-606): (spg + spg2) (v - cvic), cvic =
    // profile(j) exp(i lalpha x(i,j)); the sp2 (v - f) form with sp2 = hl
    const int nx = c->p.nx;
    const std::complex<double> la(c->lalpha[0], c->lalpha[1]), I(0.0, 1.0);
    hf.assign((size_t)10 * P, 0.0);
    for (long q = 0; q < P; q++) {
      const long j = q / nx;
      const std::complex<double> e = std::exp(I * la * x[q]);
      for (int m = 0; m < 5; m++) {
        const std::complex<double> f = std::complex<double>(c->eig[10 * j + 2 * m], c->eig[10 * j + 2 * m + 1]) * e;
        hf[2 * ((size_t)m * P + q)] = f.real();
        hf[2 * ((size_t)m * P + q) + 1] = f.imag();
      }
    }
    LNS3D_CHECK(hipMemcpyAsync(sp2, hl.data(), (size_t)P * 8, hipMemcpyHostToDevice, c->stream));
    LNS3D_CHECK(hipMemcpyAsync(spf, hf.data(), (size_t)P * 80, hipMemcpyHostToDevice, c->stream));
    c->c.sp1 = nullptr;
    c->c.sp2 = sp2;
    c->c.spf = spf;
  }
  if (inflow) {
    LNS3D_CHECK(hipMemcpyAsync(sp2, spg2, (size_t)P * 8, hipMemcpyHostToDevice, c->stream));
    LNS3D_CHECK(hipMemcpyAsync(xd, x, (size_t)P * 8, hipMemcpyHostToDevice, c->stream));
    if (ispg == 3)
      LNS3D_CHECK(hipMemcpyAsync(wd, wamp, (size_t)c->p.nx * 8, hipMemcpyHostToDevice, c->stream));
    c->c.sp2 = sp2;
    c->c.spf = spf;
    int rc = launch_sponge_forcing(c->p, ispg, c->c, xd, wd, x0, c->stream);
    if (rc) return rc;
  }
  LNS3D_CHECK(hipStreamSynchronize(c->stream));     // the host staging vectors end here
  return 0;
}

// top = 1: the incoming acoustic wave on the top boundary
// (src/rhsbc3d.f90:129-161, src/itrbc3d.f90:145-195) from the mean flow, the
// top row of the physical x, wamp (NULL: 1, as src/itrbc3d.f90:154-158) and x0
extern "C" int lns3d_set_top_forcing(lns3d_ctx *c, const double *x, const double *wamp, double x0) {
  if (!c || !x) return fail(LNS3D_EINVAL, "lns3d_set_top_forcing: null argument");
  if (c->p.top != 1) return fail(LNS3D_EINVAL, "lns3d_set_top_forcing: the context has top != 1");
  if (!c->vm_ready)
    return fail(LNS3D_EINVAL, "lns3d_set_top_forcing: the wave needs the mean flow (lns3d_setup first)");
  const int nx = c->p.nx, ny = c->p.ny;
  if (!c->top_block) LNS3D_CHECK(hipMalloc(&c->top_block, (size_t)12 * nx * sizeof(double)));
  double *ftop = c->top_block, *xd = ftop + 10 * nx, *wd = xd + nx;
  std::vector<double> xt((size_t)nx);
  for (int i = 0; i < nx; i++) xt[i] = x[i + (long)nx * (ny - 1)];
  LNS3D_CHECK(hipMemcpyAsync(xd, xt.data(), (size_t)nx * 8, hipMemcpyHostToDevice, c->stream));
  if (wamp) LNS3D_CHECK(hipMemcpyAsync(wd, wamp, (size_t)nx * 8, hipMemcpyHostToDevice, c->stream));
  c->c.ftop = ftop;
  int rc = launch_top_forcing(c->p, c->c, xd, wamp ? wd : nullptr, x0, c->stream);
  if (rc) return rc;
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// Sutherland's law (mattyp = 1, src/getmat.f90:28-40): the nondimensional
// constant S0 = 110.4 K / Tref as src/input.f90:41-47 forms it from the deck.
// Before lns3d_setup (the coefficients use it).
extern "C" int lns3d_set_sutherland(lns3d_ctx *c, double S0) {
  if (!c || !(S0 > 0.0)) return fail(LNS3D_EINVAL, "lns3d_set_sutherland: needs a context and S0 > 0");
  c->c.S0 = S0;
  c->s0_set = true;
  c->coeffs_ready = false;
  c->factored = false;
  return 0;
}

// The bump wall (wall = 3): the wall values src/genbump.f90:74-131 forms from
// the bump shape, uw host (nx,5) complex as (re, im) pairs: u1w, u2w, u3w, tw,
// twp (src/rhsbc3d.f90:47-100, src/itrbc3d.f90:48-105).  Before lns3d_factor.
extern "C" int lns3d_set_wall(lns3d_ctx *c, const double *uw) {
  if (!c || !uw) return fail(LNS3D_EINVAL, "lns3d_set_wall: null argument");
  if (c->p.wall != 3 || c->p.yper) return fail(LNS3D_EINVAL, "lns3d_set_wall: the context has no wall = 3");
  const int nx = c->p.nx;
  if (!c->wall_block) LNS3D_CHECK(hipMalloc(&c->wall_block, (size_t)10 * nx * sizeof(double)));
  std::vector<double> h((size_t)10 * nx);
  for (int k = 0; k < 5; k++)
    for (int i = 0; i < nx; i++) {              // (nx,5) complex -> [5][nx] (re, im)
      h[2 * ((size_t)k * nx + i)] = uw[2 * ((size_t)k * nx + i)];
      h[2 * ((size_t)k * nx + i) + 1] = uw[2 * ((size_t)k * nx + i) + 1];
    }
  LNS3D_CHECK(hipMemcpyAsync(c->wall_block, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  c->c.fwall = c->wall_block;
  c->factored = false;
  return 0;
}

// Smoother (the deck's eps_e; src/smoother3d.f90 in the residual, the implicit
// damping of the factors): eps_e and the damping profile buff (host (nx,ny)),
// which src/buffer.f90 (init_buffer) forms from ibuff; eps_e = 0 turns it off.
// Before lns3d_factor; again after lns3d_ctx_set_params.
extern "C" int lns3d_set_smoother(lns3d_ctx *c, double eps_e, const double *buff) {
  if (!c) return fail(LNS3D_EINVAL, "lns3d_set_smoother: null context");
  c->factored = false;
  if (eps_e == 0.0) {
    c->c.sbuf = nullptr;
    return 0;
  }
  if (!buff) return fail(LNS3D_EINVAL, "lns3d_set_smoother: null buff");
  const long P = (long)c->p.nx * c->p.ny;
  if (!c->smooth_block) LNS3D_CHECK(hipMalloc(&c->smooth_block, (size_t)P * sizeof(double)));
  std::vector<double> eb((size_t)P);
  for (long q = 0; q < P; q++) eb[q] = eps_e * buff[q];
  LNS3D_CHECK(hipMemcpyAsync(c->smooth_block, eb.data(), (size_t)P * 8, hipMemcpyHostToDevice, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  c->c.sbuf = c->smooth_block;
  return 0;
}

// Left and right boundary values (src/rhsbc3d.f90:176-246, src/itrbc3d.f90
// :219-376): left = 1 the acoustic wave on the left column (from the mean flow,
// x (host (nx,ny)) and x0, as the top wave with amplitude 1); left = 4 the
// eigenfunction profiles eig (host (10,ny): rhor, rhoi, ur, ui, vr, vi, wr,
// wi, tr, ti, the columns of src/gendist.f90's inflow file); right = 4 the
// same profiles times exp(i lalpha x(nx,j)), lalpha = (re, im).
extern "C" int lns3d_set_inflow(lns3d_ctx *c, const double *x, double x0, const double *eig,
                                const double *lalpha) {
  if (!c) return fail(LNS3D_EINVAL, "lns3d_set_inflow: null context");
  const lns3d_params &p = c->p;
  const bool l1 = !p.xper && p.left == 1, l4 = !p.xper && p.left == 4, r4 = !p.xper && p.right == 4;
  const bool l5 = !p.xper && p.left == 5, r5 = !p.xper && p.right == 5;
  if (!l1 && !l4 && !r4 && !l5 && !r5)
    return fail(LNS3D_EINVAL, "lns3d_set_inflow: the context has no left = 1, 4, 5 or right = 4, 5");
  if (((l1 || r4 || l5 || r5) && !x) || ((l4 || r4) && !eig))
    return fail(LNS3D_EINVAL, "lns3d_set_inflow: null argument");
  if (l4 && p.omega == 0.0)
    return fail(LNS3D_EUNSUPPORTED, "lns3d_set_inflow: left = 4 with omega = 0 varies in time in the reference "
                                    "(its transient eigenfunction), not implemented");
  if ((l1 || l5 || r5) && !c->vm_ready)
    return fail(LNS3D_EINVAL, "lns3d_set_inflow: the waves need the mean flow (lns3d_setup first)");
  const int nx = p.nx, ny = p.ny;
  if (!c->side_block) LNS3D_CHECK(hipMalloc(&c->side_block, (size_t)22 * ny * sizeof(double)));
  std::vector<double> hs((size_t)20 * ny, 0.0);
  if (l4)
    for (int j = 0; j < ny; j++)
      for (int d = 0; d < 5; d++) {
        hs[2 * ((size_t)d * ny + j)] = eig[10 * (size_t)j + 2 * d];
        hs[2 * ((size_t)d * ny + j) + 1] = eig[10 * (size_t)j + 2 * d + 1];
      }
  if (r4) {
    const std::complex<double> la(lalpha ? lalpha[0] : 0.0, lalpha ? lalpha[1] : 0.0);
    for (int j = 0; j < ny; j++) {
      const std::complex<double> e = std::exp(std::complex<double>(0.0, 1.0) * la * x[nx - 1 + (long)nx * j]);
      for (int d = 0; d < 5; d++) {
        const std::complex<double> f =
            std::complex<double>(eig[10 * (size_t)j + 2 * d], eig[10 * (size_t)j + 2 * d + 1]) * e;
        hs[2 * ((size_t)(5 + d) * ny + j)] = f.real();
        hs[2 * ((size_t)(5 + d) * ny + j) + 1] = f.imag();
      }
    }
  }
  LNS3D_CHECK(hipMemcpyAsync(c->side_block, hs.data(), hs.size() * 8, hipMemcpyHostToDevice, c->stream));
  c->c.fside = c->side_block;
  if (l4) {                          // kept for the ispg = 4 sponge
    c->eig.assign(eig, eig + 10 * (size_t)ny);
    c->lalpha[0] = lalpha ? lalpha[0] : 0.0;
    c->lalpha[1] = lalpha ? lalpha[1] : 0.0;
  }
  // the waves: the x value line j reads is x(1,j) for left = 1, and the
  // reference's transposed x(j,1) and x(j,nx) (memory offsets j and
  // j + nx (nx - 1) of x(nx,ny)) for left = 5 and right = 5
  std::vector<double> xl((size_t)2 * ny);
  for (int j = 0; j < ny; j++) {
    xl[j] = l1 ? x[(long)nx * j] : l5 ? x[j] : 0.0;
    xl[ny + j] = r5 ? x[j + (long)nx * (nx - 1)] : 0.0;
  }
  double *xd = c->side_block + 20 * ny;
  if (l1 || l5 || r5)
    LNS3D_CHECK(hipMemcpyAsync(xd, xl.data(), (size_t)2 * ny * 8, hipMemcpyHostToDevice, c->stream));
  int rc = 0;
  if (l1 || l5) rc = launch_side_wave(p, c->c, xd, x0, 0, l1 ? 1 : 0, c->stream);
  if (!rc && r5) rc = launch_side_wave(p, c->c, xd + ny, x0, 1, 0, c->stream);
  if (rc) return rc;
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// Coefficients in the reference's layouts (host): Ah,Bh,Dh,Dhi (5,5,nx,ny),
// Vh11,Vh12,Vh22,ABhi (6,nx,ny), dtl (nx,ny) -- the arrays src/genmtrx.f90
// and src/temporal.f90 leave in modules local2d/local3d.
int lns3d::ctx_put_soa(lns3d_ctx *c, double *dst, const double *src, int ne, int stride) {
  const long P = (long)c->p.nx * c->p.ny;
  std::vector<double> soa((size_t)P * ne);
  for (long q = 0; q < P; q++)
    for (int e = 0; e < ne; e++) soa[(size_t)e * P + q] = src[q * stride + e];
  LNS3D_CHECK(hipMemcpyAsync(dst, soa.data(), (size_t)ne * P * 8, hipMemcpyHostToDevice, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));     // the host staging vector ends here
  return 0;
}

extern "C" int lns3d_set_coeffs(lns3d_ctx *c, const double *Ah, const double *Bh, const double *Dh,
                                const double *Dhi, const double *Vh11, const double *Vh12,
                                const double *Vh22, const double *ABhi, const double *dtl) {
  if (!c || !Ah || !Bh || !Dh || !Dhi || !Vh11 || !Vh12 || !Vh22 || !ABhi || !dtl)
    return fail(LNS3D_EINVAL, "lns3d_set_coeffs: null argument");
  int rc = ctx_put_soa(c, c->c.Ah, Ah, 25, 25);
  if (!rc) rc = ctx_put_soa(c, c->c.Bh, Bh, 25, 25);
  if (!rc) rc = ctx_put_soa(c, c->c.Dh, Dh, 25, 25);
  if (!rc) rc = ctx_put_soa(c, c->c.Dhi, Dhi, 25, 25);
  if (!rc) rc = ctx_put_soa(c, c->c.V11, Vh11, 6, 6);
  if (!rc) rc = ctx_put_soa(c, c->c.V12, Vh12, 6, 6);
  if (!rc) rc = ctx_put_soa(c, c->c.V22, Vh22, 6, 6);
  if (!rc) rc = ctx_put_soa(c, c->c.ABhi, ABhi, 4, 6);
  if (!rc) rc = ctx_put_soa(c, c->c.dtl, dtl, 1, 1);
  if (rc) return rc;
  c->coeffs_ready = true;
  c->factored = false;
  return 0;
}

extern "C" int lns3d_get_coeffs(lns3d_ctx *c, double *Ah, double *Bh, double *Dh, double *Dhi,
                                double *Vh11, double *Vh12, double *Vh22, double *ABhi, double *dtl) {
  if (!c || !c->coeffs_ready) return fail(LNS3D_EINVAL, "lns3d_get_coeffs: no coefficients");
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  const long P = (long)c->p.nx * c->p.ny;
  std::vector<double> soa((size_t)P * 25);
  auto get = [&](double *dst, int ne, int stride, const double *src) -> int {
    if (!dst) return 0;
    if (hipMemcpy(soa.data(), src, (size_t)ne * P * 8, hipMemcpyDeviceToHost) != hipSuccess)
      return fail(LNS3D_EHIP, "hipMemcpy coefficients");
    for (long q = 0; q < P; q++)
      for (int e = 0; e < stride; e++) dst[q * stride + e] = e < ne ? soa[(size_t)e * P + q] : 0.0;
    return 0;
  };
  int rc = get(Ah, 25, 25, c->c.Ah);
  if (!rc) rc = get(Bh, 25, 25, c->c.Bh);
  if (!rc) rc = get(Dh, 25, 25, c->c.Dh);
  if (!rc) rc = get(Dhi, 25, 25, c->c.Dhi);
  if (!rc) rc = get(Vh11, 6, 6, c->c.V11);
  if (!rc) rc = get(Vh12, 6, 6, c->c.V12);
  if (!rc) rc = get(Vh22, 6, 6, c->c.V22);
  if (!rc) rc = get(ABhi, 4, 6, c->c.ABhi);
  if (!rc) rc = get(dtl, 1, 1, c->c.dtl);
  return rc;
}

// assemble (src/lhs1f3d.f90 / lhs2f3d.f90 with closures and BC rows) and
// factor both directions; precision 1 also builds complex-float factors
extern "C" int lns3d_factor(lns3d_ctx *c) {
  if (!c || !c->coeffs_ready) return fail(LNS3D_EINVAL, "lns3d_factor: no coefficients");
  if (c->p.top == 1 && !c->c.ftop)
    return fail(LNS3D_EINVAL, "lns3d_factor: top = 1 needs its incoming wave (lns3d_set_top_forcing)");
  if (!c->p.xper && (c->p.left == 1 || c->p.left == 4 || c->p.left == 5 || c->p.right == 4 || c->p.right == 5) &&
      !c->c.fside)
    return fail(LNS3D_EINVAL, "lns3d_factor: left = 1, 4, 5 and right = 4, 5 need their values (lns3d_set_inflow)");
  if (!c->p.yper && c->p.wall == 3 && !c->c.fwall)
    return fail(LNS3D_EINVAL, "lns3d_factor: wall = 3 needs its wall values (lns3d_set_wall)");
  if (!c->navier && !c->bnb_ready)
    return fail(LNS3D_EINVAL, "lns3d_factor: the inviscid rows need the wall normals (lns3d_setup with the metric)");
  // useCalcd (src/lhs1f3d.f90:27-37): the D terms in the xi factor when
  // istep - 1 + iter is even, istep counting this run's steps from 1
  c->c.calcd_xi = c->p.use_calcd ? (int)((c->lstep + c->cur_iter) % 2 == 0) : 1;
  int rc = launch_assemble(c->p, c->c, 1, c->px, c->stream);
  if (!rc) rc = penta_factor(&c->px, c->stream);
  if (!rc) rc = launch_assemble(c->p, c->c, 2, c->py, c->stream);
  if (!rc) rc = penta_factor(&c->py, c->stream);
  if (rc) return rc;
  if (c->p.precision == 1) {
    if (!c->have32) {
      // The complex-float records decode the complex-double LU, whose layout
      // depends on its family (a twisted LU eliminates the bottom chain
      // bottom-up): take the tuned complex-float variant when it is of the
      // same family, else the family's default.
      auto family_variant = [](const PentaPlan &d, int n, long lines) {
        const int v = penta_variant_for(n, lines, 1);
        if (penta_twisted(v) == (d.h >= 0)) return v;
        return d.h >= 0 ? (int)PENTA_TWISTED_PAIRS : (lines >= 512 ? (int)PENTA_PLAIN3 : (int)PENTA_PLAIN1);
      };
      // periodic lines: complex-float records rounded from the complex-double ones
      rc = penta_plan_create_variant(&c->px32, c->p.xper, 0, c->p.nx, c->p.ny, 1,
                                     c->p.xper ? 0 : family_variant(c->px, c->p.nx, c->p.ny));
      if (!rc)
        rc = penta_plan_create_variant(&c->py32, c->p.yper, 0, c->p.ny, c->p.nx, 1,
                                       c->p.yper ? 0 : family_variant(c->py, c->p.ny, c->p.nx));
      if (!rc && (c->px32.h != c->px.h || c->py32.h != c->py.h))
        rc = fail(LNS3D_EHIP, "lns3d_factor: complex-float plans split their lines unlike the complex-double LU");
      const long P = (long)c->p.nx * c->p.ny;
      if (!rc && (hipMalloc(&c->t1, (size_t)P * 80) != hipSuccess || hipMalloc(&c->t2, (size_t)P * 80) != hipSuccess ||
                  hipMalloc(&c->t3, (size_t)P * 80) != hipSuccess))
        rc = fail(LNS3D_EHIP, "hipMalloc refinement buffers");
      // the complex-float plans borrow the complex-double LU working copy
      if (!rc) {
        (void)hipFree(c->px32.lu); c->px32.lu = nullptr;
        (void)hipFree(c->py32.lu); c->py32.lu = nullptr;
      }
      if (rc) return rc;
      c->have32 = true;
    }
    // the complex-float records come from the same complex-double LU (or,
    // periodic lines, from its records)
    auto records32 = [&](PentaPlan &d32, PentaPlan &d64) {
      if (d32.periodic) return penta_records_narrow(&d32, &d64, c->stream);
      std::swap(d32.lu, d64.lu);
      const int r = penta_factor_records_only(&d32, c->stream);
      std::swap(d32.lu, d64.lu);
      return r;
    };
    rc = records32(c->px32, c->px);
    if (!rc) rc = records32(c->py32, c->py);
    if (rc) return rc;
  }
  // factor health, both directions (and the complex-float records)
  rc = penta_check_launch(&c->px, c->stream);
  if (!rc) rc = penta_check_launch(&c->py, c->stream);
  long bad[4] = {0, 0, 0, 0}, first[4] = {-1, -1, -1, -1};
  if (!rc) rc = penta_check_result(&c->px, c->stream, &bad[0], &first[0]);
  if (!rc) rc = penta_check_result(&c->py, c->stream, &bad[1], &first[1]);
  if (!rc && c->p.precision == 1) {
    rc = penta_check_records_launch(&c->px32, c->stream);
    if (!rc) rc = penta_check_records_launch(&c->py32, c->stream);
    if (!rc) rc = penta_check_result(&c->px32, c->stream, &bad[2], &first[2]);
    if (!rc) rc = penta_check_result(&c->py32, c->stream, &bad[3], &first[3]);
  }
  if (rc) return rc;
  c->factored = true;
  c->factored_calcd = c->c.calcd_xi;
  for (int k = 0; k < 4; k++)
    if (bad[k]) {
      char msg[200];
      snprintf(msg, sizeof msg, "lns3d_factor: %s factor: %ld line(s) with a zero pivot or a non-finite entry "
               "(first: %s line %ld)", k % 2 ? "eta" : "xi", bad[k], k % 2 ? "eta" : "xi", first[k] + 1);
      c->factored = false;
      return fail(LNS3D_ENONFINITE, msg);
    }
  return 0;
}

extern "C" int lns3d_assemble(lns3d_ctx *c, int dir, void *mat_dev) {
  if (!c || !c->coeffs_ready || (dir != 1 && dir != 2) || !mat_dev)
    return fail(LNS3D_EINVAL, "lns3d_assemble: bad arguments");
  return launch_assemble_ref(c->p, c->c, dir, mat_dev, c->stream);
}

extern "C" int lns3d_lrhs3d(lns3d_ctx *c, const void *v_dev, void *r_dev) {
  if (!c || !c->coeffs_ready || !v_dev || !r_dev) return fail(LNS3D_EINVAL, "lns3d_lrhs3d: bad arguments");
  return launch_rhs(c->p, c->st, c->c, v_dev, r_dev, c->work, 0, c->stream);
}

extern "C" int lns3d_itrbc3d(lns3d_ctx *c, void *v_dev) {
  if (!c || !v_dev) return fail(LNS3D_EINVAL, "lns3d_itrbc3d: bad arguments");
  // the boundary values these selectors hold the field at must be formed first
  if (!c->p.yper && c->p.top == 1 && !c->c.ftop)
    return fail(LNS3D_EINVAL, "lns3d_itrbc3d: top = 1 needs its incoming wave (lns3d_set_top_forcing)");
  if (!c->p.xper && (c->p.left == 1 || c->p.left == 4 || c->p.left == 5 || c->p.right == 4 || c->p.right == 5) &&
      !c->c.fside)
    return fail(LNS3D_EINVAL, "lns3d_itrbc3d: left = 1, 4, 5 and right = 4, 5 need their values (lns3d_set_inflow)");
  if (!c->p.yper && c->p.wall == 3 && !c->c.fwall)
    return fail(LNS3D_EINVAL, "lns3d_itrbc3d: wall = 3 needs its wall values (lns3d_set_wall)");
  if (!c->navier && !c->bnb_ready)
    return fail(LNS3D_EINVAL, "lns3d_itrbc3d: the inviscid rows need the wall normals (lns3d_setup with the metric)");
  return launch_itrbc(c->p, c->c, v_dev, c->stream);
}

// the residual norm of a step into the next slot of the context's output
// (lns3d_ctx_set_residual_output), if one is set
static int record_norm(lns3d_ctx *c, const void *r) {
  if (!c->norm_out) return 0;
  double *out = c->norm_out + (c->norm_idx++ % c->norm_count);
  return launch_residual_norm(c->c.dtl, r, (long)c->p.nx * c->p.ny, c->scal + 2048, out, c->stream);
}

extern "C" int lns3d_residual_norm(lns3d_ctx *c, const void *r_dev, double *norm_dev) {
  if (!c || !c->coeffs_ready || !r_dev || !norm_dev) return fail(LNS3D_EINVAL, "lns3d_residual_norm: bad arguments");
  return launch_residual_norm(c->c.dtl, r_dev, (long)c->p.nx * c->p.ny, c->scal + 2048, norm_dev, c->stream);
}

extern "C" int lns3d_residual_norm_noalloc(const void *r_dev, const double *dtl_dev, long npts, double *norm_dev,
                                           void *stream) {
  if (!r_dev || !dtl_dev || !norm_dev) return fail(LNS3D_EINVAL, "lns3d_residual_norm_noalloc: null argument");
  if (npts < 1 || npts > LNS3D_NOALLOC_MAX_POINTS)
    return fail(LNS3D_EINVAL, "lns3d_residual_norm_noalloc: npts must be 1..LNS3D_NOALLOC_MAX_POINTS (2^28)");
  return launch_residual_norm_noalloc(dtl_dev, r_dev, npts, norm_dev, (hipStream_t)stream);
}

extern "C" int lns3d_ctx_set_residual_output(lns3d_ctx *c, double *norms_dev, int count) {
  if (!c || (norms_dev && count < 1)) return fail(LNS3D_EINVAL, "lns3d_ctx_set_residual_output: bad arguments");
  c->norm_out = norms_dev;
  c->norm_count = norms_dev ? count : 0;
  c->norm_idx = 0;
  return 0;
}

static int solve_xy(lns3d_ctx *c, PentaPlan &px, PentaPlan &py, void *r) {
  // xi lines: line j, row i at (i + nx*j); eta lines: line i, row j
  int rc = penta_solve(&px, r, c->p.nx, 1, c->stream);
  if (!rc) rc = penta_solve(&py, r, 1, c->p.nx, c->stream);
  return rc;
}

extern "C" int lns3d_solve(lns3d_ctx *c, void *r_dev) {
  if (!c || !c->factored || !r_dev) return fail(LNS3D_ENOTFACTORED, "lns3d_solve: not factored");
  return solve_xy(c, c->px, c->py, r_dev);
}

extern "C" int lns3d_solve_dir(lns3d_ctx *c, int dir, void *r_dev) {
  if (!c || !c->factored || !r_dev || (dir != 1 && dir != 2))
    return fail(LNS3D_EINVAL, "lns3d_solve_dir: bad arguments or not factored");
  if (dir == 1) return penta_solve(&c->px, r_dev, c->p.nx, 1, c->stream);
  return penta_solve(&c->py, r_dev, 1, c->p.nx, c->stream);
}

// updateLHS: the factors formed again at every iteration (src/impdrv3d.f90
// :248,310).  They depend on the step and iteration only through useCalcd's
// D-term placement, so they are formed again when that placement changes.
static int refactor_for(lns3d_ctx *c, int iter) {
  if (!c->p.update_lhs) return 0;
  const int want = c->p.use_calcd ? (int)((c->lstep + iter) % 2 == 0) : 1;
  if (want == c->factored_calcd) return 0;
  c->cur_iter = iter;
  const int e = lns3d_factor(c);
  c->cur_iter = 1;
  return e;
}

// One implicit step (src/impdrv3d.f90:146-356, impl = 1, niter = 1):
// r = -dtl * lrhs3D(v) with rhsBC3D rows, xi solve, eta solve, v += r,
// itrBC3D.  v must already satisfy the boundary conditions (call
// lns3d_itrbc3d once before the first step, as src/impdrv3d.f90:126 does).
// dv_dev (optional) receives the increment r.
//
// impl = 2 (two-step backward, src/impdrv3d.f90:134-153, 187-190): the first
// step is backward Euler; from the second on the residual is rtb3D with the
// field at the start of the previous step, and at the second the factors are
// formed again with alfa = 2/3 (src/impdrv3d.f90:229-231).
extern "C" int lns3d_step(lns3d_ctx *c, void *v_dev, void *dv_dev) {
  if (!c || !c->factored || !v_dev) return fail(LNS3D_ENOTFACTORED, "lns3d_step: not factored");
  if (c->impl == 3 && c->niter > 1)
    return fail(LNS3D_EUNSUPPORTED, "lns3d_step: corrector iterations with impl = 3 are not implemented");
  void *r = dv_dev ? dv_dev : c->r;
  const long n = 5L * c->p.nx * c->p.ny;
  const bool two = c->impl == 2 && c->lstep >= 1;
  int rc = 0;
  if (two && c->p.alfa != 6.6666666666666666666e-1) {
    c->p.alfa = 6.6666666666666666666e-1;
    rc = lns3d_factor(c);
    if (rc) return rc;
  }
  auto refactor = [&](int iter) { return refactor_for(c, iter); };
  if (c->niter == 1) {
    rc = refactor(1);
    if (!rc) rc = launch_rhs(c->p, c->st, c->c, v_dev, r, c->work, two ? 2 : 1, c->stream, c->vhist);
    if (!rc) rc = record_norm(c, r);
    if (!rc && c->impl == 2) rc = launch_copy(c->vhist, v_dev, n, c->stream);
    if (!rc) rc = solve_xy(c, c->px, c->py, r);
    if (!rc) rc = launch_update_bc(c->p, c->c, v_dev, r, c->stream);
  } else {
    // corrector iterations (src/impdrv3d.f90:173-358): vold stays the field at
    // the start of the step; the factors are those of the first iteration
    rc = launch_copy(c->vstart, v_dev, n, c->stream);
    for (int it = 0; it < c->niter && !rc; it++) {
      const int mode = it == 0 ? (two ? 2 : 1) : (two ? 4 : 3);
      rc = refactor(it + 1);
      if (!rc) rc = launch_rhs(c->p, c->st, c->c, v_dev, r, c->work, mode, c->stream, c->vhist, c->vstart);
      if (!rc && it == 0) rc = record_norm(c, r);
      if (!rc) rc = solve_xy(c, c->px, c->py, r);
      if (!rc) rc = launch_update_bc(c->p, c->c, v_dev, r, c->stream);
    }
    if (!rc && c->impl == 2) std::swap(c->vhist, c->vstart);   // this step's start is the next v2old
  }
  if (!rc) c->lstep++;
  return rc;
}

// One pseudo-time corrector of the shift-invert driver (src/si_eigdrv3d_new
// .f90:327-358): r = rhsBC3D(vold - v - dtl (lrhs3D(v) - f)), the two line
// solves with the factors, v += r, itrBC3D.  vold NULL means v (the first
// corrector of a step); dv_dev NULL keeps the increment in the context.
extern "C" int lns3d_step_forced(lns3d_ctx *c, void *v_dev, const void *vold_dev, const void *f_dev,
                                 void *dv_dev) {
  if (!c || !c->factored || !v_dev || !f_dev)
    return fail(LNS3D_EINVAL, "lns3d_step_forced: needs a factored context, v and f");
  if (c->impl != 1)
    return fail(LNS3D_EUNSUPPORTED, "lns3d_step_forced: the shift-invert steps are backward Euler (impl = 1)");
  void *r = dv_dev ? dv_dev : c->r;
  int rc = launch_rhs(c->p, c->st, c->c, v_dev, r, c->work, 5, c->stream, nullptr,
                      vold_dev ? vold_dev : v_dev, f_dev);
  if (!rc) rc = solve_xy(c, c->px, c->py, r);
  if (!rc) rc = launch_update_bc(c->p, c->c, v_dev, r, c->stream);
  return rc;
}

// The shift-invert driver's residual after a corrector
// (src/si_eigdrv3d_new.f90:359-374): r = rhsBC3D(vold - v - dtl (lrhs3D(v) -
// f)) into r_dev and sqrt(sum |r|^2 / dtl^2) into *norm_out (synchronises).
extern "C" int lns3d_forced_residual(lns3d_ctx *c, const void *v_dev, const void *vold_dev, const void *f_dev,
                                     void *r_dev, double *norm_out) {
  if (!c || !c->coeffs_ready || !v_dev || !vold_dev || !f_dev || !r_dev || !norm_out)
    return fail(LNS3D_EINVAL, "lns3d_forced_residual: bad arguments");
  int rc = launch_rhs(c->p, c->st, c->c, v_dev, r_dev, c->work, 5, c->stream, nullptr, vold_dev, f_dev);
  int nb = 0;
  if (!rc) rc = launch_forced_norm(c->p, c->c, r_dev, c->scal + 16, &nb, c->stream);
  if (rc) return rc;
  std::vector<double> part((size_t)nb);
  LNS3D_CHECK(hipMemcpyAsync(part.data(), c->scal + 16, (size_t)nb * 8, hipMemcpyDeviceToHost, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += part[b];
  *norm_out = sqrt(s);
  return 0;
}

// corrector iterations per step, the deck's niter (src/impdrv3d.f90:173)
extern "C" int lns3d_set_corrector_iterations(lns3d_ctx *c, int niter) {
  if (!c || niter < 1) return fail(LNS3D_EINVAL, "lns3d_set_corrector_iterations: niter must be >= 1");
  c->niter = niter;
  if (niter > 1 && !c->vstart) LNS3D_CHECK(hipMalloc(&c->vstart, 5ul * c->p.nx * c->p.ny * 16));
  return 0;
}

// Time scheme (the deck's impl, src/impdrv3d.f90:179-199): 1 backward Euler
// (alfa 1), 2 two-step backward (alfa 1 for the first step, then 2/3), 3
// alfa 1/2 (its residual lrhs3D(0.5 (v + vold)) equals lrhs3D(v) at the
// start of a step, where vold = v).  Starts a new run (no earlier step);
// factor (again) afterwards.
extern "C" int lns3d_set_time_scheme(lns3d_ctx *c, int impl) {
  if (!c || impl < 1 || impl > 3) return fail(LNS3D_EINVAL, "lns3d_set_time_scheme: impl must be 1, 2 or 3");
  c->impl = impl;
  c->lstep = 0;
  c->p.alfa = impl == 3 ? 0.5 : 1.0;
  c->factored = false;
  if (impl == 2 && !c->vhist) LNS3D_CHECK(hipMalloc(&c->vhist, 5ul * c->p.nx * c->p.ny * 16));
  return 0;
}

// impl = 2 runs resumed from a restart (src/genini3d.f90:55-57 reads vold):
// vold_dev (device, complex (5,nx,ny)) is the field at the start of the last
// step taken, lstep the number of steps taken
extern "C" int lns3d_set_time_history(lns3d_ctx *c, const void *vold_dev, long lstep) {
  if (!c || c->impl != 2 || !vold_dev || lstep < 0)
    return fail(LNS3D_EINVAL, "lns3d_set_time_history: needs impl = 2, a field and lstep >= 0");
  int rc = launch_copy(c->vhist, vold_dev, 5L * c->p.nx * c->p.ny, c->stream);
  if (rc) return rc;
  c->lstep = lstep;
  return 0;
}

// The same step with the line solves in complex float used as a
// defect-correction preconditioner for the complex-double AF system
// M dv = r, M = Lx Ly (M applied with blocks regenerated in double):
//   dv = 0; repeat { res = r - M dv; dv += M32^{-1} res } until
//   |res|_max <= tol |r|_max or iters reach max_iters.
extern "C" int lns3d_step_refined(lns3d_ctx *c, void *v_dev, void *dv_dev, double tol, int max_iters,
                                  int *iters_out, double *resid_out) {
  if (c && (c->impl == 2 || c->niter > 1))
    return fail(LNS3D_EUNSUPPORTED, "lns3d_step_refined: impl = 2 and corrector iterations take lns3d_step");
  if (!c || !c->factored || !c->have32 || !v_dev)
    return fail(LNS3D_ENOTFACTORED, "lns3d_step_refined: needs precision = 1 and lns3d_factor");
  const long n = 5L * c->p.nx * c->p.ny;
  void *dv = dv_dev ? dv_dev : c->r;
  void *rhs = c->t3;            // the residual r of the step
  int rc = refactor_for(c, 1);
  if (!rc) rc = launch_rhs(c->p, c->st, c->c, v_dev, rhs, c->work, 1, c->stream);
  if (rc) return rc;
  rc = launch_copy(dv, rhs, n, c->stream);
  if (rc) return rc;
  rc = solve_xy(c, c->px32, c->py32, dv);
  if (rc) return rc;
  // max_iters <= 0: no caller limit; the loop still stops after 64 solves
  const int cap = max_iters > 0 ? max_iters : 64;
  int it = 1;
  double nr[2] = {0, 0};
  for (;; it++) {
    rc = launch_af_matvec(c->p, c->c, dv, c->t2, c->t1, c->stream);       // t2 = M dv
    if (!rc) rc = launch_residual_norms(rhs, c->t2, n, c->scal + 8, c->stream);   // t2 = rhs - M dv
    if (rc) return rc;
    LNS3D_CHECK(hipMemcpyAsync(nr, c->scal + 8, 16, hipMemcpyDeviceToHost, c->stream));
    LNS3D_CHECK(hipStreamSynchronize(c->stream));
    if (nr[0] <= tol * nr[1] || it >= cap || !(nr[0] == nr[0])) break;
    rc = solve_xy(c, c->px32, c->py32, c->t2);
    if (!rc) rc = launch_axpy(dv, c->t2, n, c->stream);
    if (rc) return rc;
  }
  c->last_iters = it;
  c->last_resid = nr[1] > 0 ? nr[0] / nr[1] : 0.0;
  if (iters_out) *iters_out = it;
  if (resid_out) *resid_out = c->last_resid;
  if (!(c->last_resid == c->last_resid)) return fail(LNS3D_ENONFINITE, "refinement produced NaN");
  rc = launch_update_bc(c->p, c->c, v_dev, dv, c->stream);
  if (!rc) c->lstep++;
  return rc;
}

// the same step, but only the complex-float solve (no refinement): the
// pseudo-time iteration itself corrects the defect in steady marching
extern "C" int lns3d_step_fp32(lns3d_ctx *c, void *v_dev, void *dv_dev) {
  if (c && (c->impl == 2 || c->niter > 1))
    return fail(LNS3D_EUNSUPPORTED, "lns3d_step_fp32: impl = 2 and corrector iterations take lns3d_step");
  if (!c || !c->factored || !c->have32 || !v_dev)
    return fail(LNS3D_ENOTFACTORED, "lns3d_step_fp32: needs precision = 1 and lns3d_factor");
  void *r = dv_dev ? dv_dev : c->r;
  int rc = refactor_for(c, 1);
  if (!rc) rc = launch_rhs(c->p, c->st, c->c, v_dev, r, c->work, 1, c->stream);
  if (!rc) rc = record_norm(c, r);
  if (!rc) rc = solve_xy(c, c->px32, c->py32, r);
  if (!rc) rc = launch_update_bc(c->p, c->c, v_dev, r, c->stream);
  if (!rc) c->lstep++;
  return rc;
}

// The refined step with a fixed number of corrections and no host
// synchronisation, so that it can be captured in a graph.  With check set,
// one more operator application leaves |res|_max / |r|_max of the final dv
// for lns3d_last_residual.
extern "C" int lns3d_step_refined_fixed(lns3d_ctx *c, void *v_dev, void *dv_dev, int corrections,
                                        int check) {
  if (c && (c->impl == 2 || c->niter > 1))
    return fail(LNS3D_EUNSUPPORTED, "lns3d_step_refined_fixed: impl = 2 and corrector iterations take lns3d_step");
  if (!c || !c->factored || !c->have32 || !v_dev || corrections < 0)
    return fail(LNS3D_ENOTFACTORED, "lns3d_step_refined_fixed: needs precision = 1 and lns3d_factor");
  const long n = 5L * c->p.nx * c->p.ny;
  void *dv = dv_dev ? dv_dev : c->r;
  void *rhs = c->t3;
  int rc = refactor_for(c, 1);
  if (!rc) rc = launch_rhs(c->p, c->st, c->c, v_dev, rhs, c->work, 1, c->stream);
  if (!rc) rc = record_norm(c, rhs);
  if (!rc) rc = launch_copy(dv, rhs, n, c->stream);
  if (!rc) rc = solve_xy(c, c->px32, c->py32, dv);
  for (int k = 0; k < corrections && !rc; k++) {
    rc = launch_af_matvec(c->p, c->c, dv, c->t2, c->t1, c->stream);
    if (!rc) rc = launch_residual_norms(rhs, c->t2, n, c->scal + 8, c->stream);
    if (!rc) rc = solve_xy(c, c->px32, c->py32, c->t2);
    if (!rc) rc = launch_axpy(dv, c->t2, n, c->stream);
  }
  if (!rc && check) {
    rc = launch_af_matvec(c->p, c->c, dv, c->t2, c->t1, c->stream);
    if (!rc) rc = launch_residual_norms(rhs, c->t2, n, c->scal + 8, c->stream);
  }
  if (!rc) rc = launch_update_bc(c->p, c->c, v_dev, dv, c->stream);
  if (!rc) c->lstep++;
  return rc;
}

extern "C" int lns3d_last_residual(lns3d_ctx *c, double *rel) {
  if (!c || !rel) return fail(LNS3D_EINVAL, "lns3d_last_residual: null argument");
  double nr[2];
  LNS3D_CHECK(hipMemcpyAsync(nr, c->scal + 8, 16, hipMemcpyDeviceToHost, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  *rel = nr[1] > 0 ? nr[0] / nr[1] : 0.0;
  return 0;
}

// bytes of device memory a context of these parameters holds once factored
extern "C" int lns3d_ctx_bytes(const lns3d_params *p, size_t *bytes) {
  if (!p || !bytes) return fail(LNS3D_EINVAL, "lns3d_ctx_bytes: null argument");
  const size_t P = (size_t)p->nx * p->ny;
  size_t b = P * (25 * 4 + 6 * 3 + 4 + 1 + 5 + 10) * 8 + P * 160 + 512;   // coefficients, work, r
  auto plan = [&](int periodic, size_t es) {
    size_t x = P * 125 * 16;                                  // LU working copy
    if (periodic) x += 2 * P * 50 * 16;
    else x += P * 126 * es + P * 2 * 150 * es / 32 + P * 100 * 16 / 32 + P * 26 * es;   // + pair side blocks
    return x;
  };
  b += plan(p->xper, 16) + plan(p->yper, 16);
  if (p->precision == 1) b += 2 * P * 126 * 8 + 3 * P * 80;
  *bytes = b;
  return 0;
}

// ---------------------------------------------------------------------------
// Graphs: nsteps steps of count contexts captured once (each context on its
// own stream branch, forked from and joined to one capture stream) and
// replayed with one launch.  mode 0: complex-double step; 1: complex-float
// factors, no refinement; 2: complex-float factors with `corrections`
// defect corrections.
struct lns3d_graph {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  hipStream_t s = nullptr;
  // mode 3: the device tables of the batched stages, copied from their
  // pinned staging on the first launch's stream (up_ev marks the copies)
  StepBatch sb;
  std::vector<PentaBatchGroup> gx, gy;
  std::vector<TableUpload> ups;
  bool uploaded = false;
  hipEvent_t up_ev = nullptr;
  hipStream_t up_stream = nullptr;   // the stream the copies were made on
};

static void graph_free(lns3d_graph *gr) {
  if (gr->up_ev) {
    (void)hipEventSynchronize(gr->up_ev);
    (void)hipEventDestroy(gr->up_ev);
  }
  for (auto &u : gr->ups) (void)hipHostFree(u.host);
  if (gr->x) (void)hipGraphExecDestroy(gr->x);
  if (gr->g) (void)hipGraphDestroy(gr->g);
  if (gr->s) (void)hipStreamDestroy(gr->s);
  step_batch_free(gr->sb);
  penta_batch_free(gr->gx);
  penta_batch_free(gr->gy);
  delete gr;
}

// mode 3: the refined steps of every member with one launch per stage: the
// residual, the two line solves, the defect corrections and the update each
// run over all members at once (the arithmetic of lns3d_step_refined_fixed)
static int graph_create_batched(lns3d_ctx **ctxs, void **v_dev, void **dv_dev, int count, int nsteps,
                                int corrections, lns3d_graph **out) {
  auto *gr = new lns3d_graph;
  std::vector<StepBatchMember> m((size_t)count);
  std::vector<const PentaPlan *> px((size_t)count), py((size_t)count);
  std::vector<double2 *> r0((size_t)count), r1((size_t)count);
  std::vector<long> sx1((size_t)count), sx2((size_t)count), sy1((size_t)count), sy2((size_t)count);
  for (int i = 0; i < count; i++) {
    lns3d_ctx *c = ctxs[i];
    StepBatchMember &a = m[(size_t)i];
    a.p = c->p;
    a.st = c->st;
    a.c = c->c;
    a.v = (double2 *)v_dev[i];
    a.dv = (double2 *)(dv_dev && dv_dev[i] ? dv_dev[i] : c->r);
    a.rhs = (double2 *)c->t3;
    a.t1 = (double2 *)c->t1;
    a.t2 = (double2 *)c->t2;
    a.work = (double2 *)c->work;
    a.norm = c->norm_out;
    a.norm_count = c->norm_count;
    a.norm_idx = c->norm_idx;
    px[(size_t)i] = &c->px32;
    py[(size_t)i] = &c->py32;
    r0[(size_t)i] = a.dv;
    r1[(size_t)i] = a.t2;
    sx1[(size_t)i] = c->p.nx;   // xi lines: line j, row i at (i + nx j)
    sx2[(size_t)i] = 1;
    sy1[(size_t)i] = 1;         // eta lines: line i, row j
    sy2[(size_t)i] = c->p.nx;
  }
  int rc = hipStreamCreateWithFlags(&gr->s, hipStreamNonBlocking) != hipSuccess ? fail(LNS3D_EHIP, "stream") : 0;
  if (!rc) rc = step_batch_build(m.data(), count, gr->sb, gr->ups);
  if (!rc) rc = penta_batch_build(px.data(), r0.data(), r1.data(), sx1.data(), sx2.data(), count, gr->gx, gr->ups);
  if (!rc) rc = penta_batch_build(py.data(), r0.data(), r1.data(), sy1.data(), sy2.data(), count, gr->gy, gr->ups);
  if (!rc && hipEventCreateWithFlags(&gr->up_ev, hipEventDisableTiming) != hipSuccess) rc = fail(LNS3D_EHIP, "event");
  if (!rc && hipStreamBeginCapture(gr->s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    rc = fail(LNS3D_EHIP, "hipStreamBeginCapture");
  const bool capturing = !rc;
  const hipStream_t s = gr->s;
  // members with the inviscid wall rows take bcfix on the final increment
  // (as launch_update_bc): dv += t2, bcfix, v += dv instead of the fused update
  const bool fix = !rc && gr->sb.blocks[SB_STAGE_BCFIX] > 0;
  for (int k = 0; k < nsteps && !rc; k++) {
    rc = step_batch_rhs(gr->sb, s);
    if (!rc) rc = step_batch_norm(gr->sb, k, s);
    if (!rc) rc = step_batch_vec(gr->sb, SB_COPY, s);
    if (!rc) rc = penta_batch_solve(gr->gx, 0, s);
    if (!rc) rc = penta_batch_solve(gr->gy, 0, s);
    for (int q = 0; q < corrections && !rc; q++) {
      rc = step_batch_matvec(gr->sb, s);
      if (!rc) rc = step_batch_vec(gr->sb, SB_SUB, s);
      if (!rc) rc = penta_batch_solve(gr->gx, 1, s);
      if (!rc) rc = penta_batch_solve(gr->gy, 1, s);
      if (!rc) rc = step_batch_vec(gr->sb, q == corrections - 1 && !fix ? SB_UPDATE : SB_ADD_DV, s);
    }
    if (!rc && fix) rc = step_batch_bcfix(gr->sb, s);
    if (!rc && (corrections == 0 || fix)) rc = step_batch_vec(gr->sb, SB_ADD_V, s);
    if (!rc) rc = step_batch_itrbc(gr->sb, s);
  }
  if (capturing) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (!rc && e != hipSuccess) rc = fail(LNS3D_EHIP, "hipStreamEndCapture");
    gr->g = g;
  }
  if (!rc && hipGraphInstantiate(&gr->x, gr->g, nullptr, nullptr, 0) != hipSuccess)
    rc = fail(LNS3D_EHIP, "hipGraphInstantiate");
  if (rc) {
    graph_free(gr);
    return rc;
  }
  // the graph's steps took the next nsteps norm slots, as mode 2's do
  for (int i = 0; i < count; i++)
    if (ctxs[i]->norm_out) ctxs[i]->norm_idx += nsteps;
  *out = gr;
  return 0;
}

extern "C" int lns3d_graph_create(lns3d_ctx **ctxs, void **v_dev, void **dv_dev, int count, int nsteps,
                                  int mode, int corrections, lns3d_graph **out) {
  if (!ctxs || !v_dev || count < 1 || nsteps < 1 || mode < 0 || mode > 3 || corrections < 0 || !out)
    return fail(LNS3D_EINVAL, "lns3d_graph_create: bad arguments");
  for (int i = 0; i < count; i++) {
    if (!ctxs[i] || !ctxs[i]->factored || !v_dev[i] || (mode > 0 && !ctxs[i]->have32))
      return fail(LNS3D_EINVAL, "lns3d_graph_create: every context must be factored (precision 1 for modes 1, 2)");
    const lns3d_ctx *c = ctxs[i];
    // the captured steps must all be alike: updateLHS with useCalcd forms
    // factors on the host between steps; the complex-float steps refuse
    // impl = 2 and corrector iterations
    if ((c->p.update_lhs && c->p.use_calcd) || (mode > 0 && (c->impl == 2 || c->niter > 1)))
      return fail(LNS3D_EUNSUPPORTED, "lns3d_graph_create: updateLHS with useCalcd, and impl = 2 or corrector "
                                      "iterations outside mode 0, take lns3d_step");
    // impl = 2: from the step after the first two (the history exists and
    // the factors hold alfa = 2/3); with corrector iterations the history
    // buffers trade places each step, so the graph holds an even count
    if (c->impl == 2 && (c->lstep < 2 || c->p.alfa != 6.6666666666666666666e-1))
      return fail(LNS3D_EUNSUPPORTED, "lns3d_graph_create: impl = 2 graphs start after the first two steps");
    if (c->impl == 2 && c->niter > 1 && nsteps % 2)
      return fail(LNS3D_EUNSUPPORTED, "lns3d_graph_create: impl = 2 with corrector iterations takes an even step count");
  }
  if (mode == 3) return graph_create_batched(ctxs, v_dev, dv_dev, count, nsteps, corrections, out);
  auto *gr = new lns3d_graph;
  std::vector<hipStream_t> st(count, nullptr);
  std::vector<hipEvent_t> ev(count + 1, nullptr);
  std::vector<hipStream_t> saved(count);
  int rc = 0;
  auto cleanup = [&]() {
    for (auto s2 : st) if (s2) (void)hipStreamDestroy(s2);
    for (auto e : ev) if (e) (void)hipEventDestroy(e);
  };
  if (hipStreamCreateWithFlags(&gr->s, hipStreamNonBlocking) != hipSuccess) rc = fail(LNS3D_EHIP, "stream");
  for (int i = 0; i < count && !rc; i++)
    if (hipStreamCreateWithFlags(&st[i], hipStreamNonBlocking) != hipSuccess) rc = fail(LNS3D_EHIP, "stream");
  for (int i = 0; i <= count && !rc; i++)
    if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) rc = fail(LNS3D_EHIP, "event");
  if (!rc && hipStreamBeginCapture(gr->s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    rc = fail(LNS3D_EHIP, "hipStreamBeginCapture");
  bool capturing = !rc;
  if (!rc && hipEventRecord(ev[count], gr->s) != hipSuccess) rc = fail(LNS3D_EHIP, "fork event");
  for (int i = 0; i < count && !rc; i++) {
    lns3d_ctx *c = ctxs[i];
    if (hipStreamWaitEvent(st[i], ev[count], 0) != hipSuccess) { rc = fail(LNS3D_EHIP, "fork wait"); break; }
    saved[i] = c->stream;
    c->stream = st[i];
    void *dv = dv_dev ? dv_dev[i] : nullptr;
    for (int k = 0; k < nsteps && !rc; k++) {
      if (mode == 0) rc = lns3d_step(c, v_dev[i], dv);
      else if (mode == 1) rc = lns3d_step_fp32(c, v_dev[i], dv);
      else rc = lns3d_step_refined_fixed(c, v_dev[i], dv, corrections, 0);
    }
    c->stream = saved[i];
    if (!rc && hipEventRecord(ev[i], st[i]) != hipSuccess) rc = fail(LNS3D_EHIP, "join event");
    if (!rc && hipStreamWaitEvent(gr->s, ev[i], 0) != hipSuccess) rc = fail(LNS3D_EHIP, "join wait");
  }
  if (capturing) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(gr->s, &g);
    if (!rc && e != hipSuccess) rc = fail(LNS3D_EHIP, "hipStreamEndCapture");
    gr->g = g;
  }
  if (!rc && hipGraphInstantiate(&gr->x, gr->g, nullptr, nullptr, 0) != hipSuccess)
    rc = fail(LNS3D_EHIP, "hipGraphInstantiate");
  cleanup();
  if (rc) {
    graph_free(gr);
    return rc;
  }
  *out = gr;
  return 0;
}

extern "C" int lns3d_graph_launch(lns3d_graph *gr, void *stream) {
  if (!gr) return fail(LNS3D_EINVAL, "lns3d_graph_launch: null graph");
  const hipStream_t s = stream ? (hipStream_t)stream : gr->s;
  if (!gr->uploaded && !gr->ups.empty()) {
    for (auto &u : gr->ups) LNS3D_CHECK(hipMemcpyAsync(u.dev, u.host, u.bytes, hipMemcpyHostToDevice, s));
    LNS3D_CHECK(hipEventRecord(gr->up_ev, s));
    gr->uploaded = true;
    gr->up_stream = s;
  } else if (gr->uploaded && s != gr->up_stream) {
    LNS3D_CHECK(hipStreamWaitEvent(s, gr->up_ev, 0));   // launched on another stream than the copies
  }
  LNS3D_CHECK(hipGraphLaunch(gr->x, s));
  return 0;
}

extern "C" int lns3d_graph_synchronize(lns3d_graph *gr) {
  if (!gr) return fail(LNS3D_EINVAL, "lns3d_graph_synchronize: null graph");
  LNS3D_CHECK(hipStreamSynchronize(gr->s));
  return 0;
}

extern "C" int lns3d_graph_destroy(lns3d_graph *gr) {
  if (!gr) return 0;
  (void)hipStreamSynchronize(gr->s);
  graph_free(gr);
  return 0;
}

static_assert(PENTA_NVARIANT == LNS3D_TUNE_VARIANTS, "the header's variant count");

// autotuner of the line-solve kernels: measures every variant on synthetic
// lines of n rows x lines and keeps the fastest for plans of that shape
// created afterwards; times_ms (may be NULL) receives LNS3D_TUNE_VARIANTS
// timings (-1: n/a)
extern "C" int lns3d_tune_lines(int n, long lines, int precision, int *variant, float *times_ms) {
  if (n < 7 || lines < 1 || (precision != 0 && precision != 1))
    return fail(LNS3D_EINVAL, "lns3d_tune_lines: bad shape");
  return penta_tune(n, lines, precision, variant, times_ms);
}

extern "C" int lns3d_tune_set(int n, long lines, int precision, int variant) {
  if (n < 7 || lines < 1 || (precision != 0 && precision != 1) || variant < -1 || variant >= LNS3D_TUNE_VARIANTS)
    return fail(LNS3D_EINVAL, "lns3d_tune_set: bad shape or variant");
  return penta_tune_set(n, lines, precision, variant);
}

// Host-array driving of the time loop (src/impdrv3d.f90:126-406): the
// context keeps the state v in HBM between calls.  The boundary conditions
// are applied once to a freshly uploaded state before its first step, as
// src/impdrv3d.f90:126 does before the loop.
extern "C" int lns3d_state_upload(lns3d_ctx *c, const void *v_host) {
  if (!c || !v_host) return fail(LNS3D_EINVAL, "lns3d_state_upload: null argument");
  const size_t bytes = 5ul * c->p.nx * c->p.ny * 16;
  if (!c->vstate) LNS3D_CHECK(hipMalloc(&c->vstate, bytes));
  LNS3D_CHECK(hipMemcpyAsync(c->vstate, v_host, bytes, hipMemcpyHostToDevice, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  c->state_fresh = true;
  return 0;
}

extern "C" int lns3d_state_download(lns3d_ctx *c, void *v_host) {
  if (!c || !v_host || !c->vstate) return fail(LNS3D_EINVAL, "lns3d_state_download: no state");
  const size_t bytes = 5ul * c->p.nx * c->p.ny * 16;
  LNS3D_CHECK(hipMemcpyAsync(v_host, c->vstate, bytes, hipMemcpyDeviceToHost, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int lns3d_advance(lns3d_ctx *c, int nsteps, void *dv_host) {
  if (!c || !c->vstate || nsteps < 0) return fail(LNS3D_EINVAL, "lns3d_advance: no state");
  if (!c->factored) return fail(LNS3D_ENOTFACTORED, "lns3d_advance: call lns3d_factor first");
  int rc = 0;
  if (c->state_fresh) {
    rc = launch_itrbc(c->p, c->c, c->vstate, c->stream);
    c->state_fresh = false;
  }
  for (int k = 0; k < nsteps && !rc; k++) rc = lns3d_step(c, c->vstate, c->r);
  if (rc) return rc;
  if (dv_host && nsteps > 0)
    LNS3D_CHECK(hipMemcpyAsync(dv_host, c->r, 5ul * c->p.nx * c->p.ny * 16, hipMemcpyDeviceToHost,
                               c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int lns3d_ctx_synchronize(lns3d_ctx *c) {
  if (!c) return fail(LNS3D_EINVAL, "null context");
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}
