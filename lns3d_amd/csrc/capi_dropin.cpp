// C-ABI drop-ins with the reference's argument lists for the RHS (lrhs3D),
// the LHS assembly (lhs1f3D, lhs2f3D) and the derivative routines (cgrad,
// cgrad2): host arrays in the reference's layouts, the arithmetic on the
// device.  The routines that read module state in the reference take it from
// lns3d_set_module_state, which keeps one device context for them.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/lns3d_amd.h"
#include "common.hpp"
#include "ctx.hpp"
#include "step.hpp"

using namespace lns3d;

namespace {
int fail(int code, const char *msg) {
  lns3d::set_error("capi_dropin", 0, msg);
  return code;
}

std::mutex g_mu;
lns3d_ctx *g_ctx = nullptr;        // the context of the forwarded module state
void *g_mat = nullptr;             // assembled LHS staging (125 P complex)
size_t g_mat_bytes = 0;
void *g_buf = nullptr;             // field staging for cgrad / cgrad2 (5 arrays of 5 P complex)
size_t g_buf_bytes = 0;
const int *g_istep = nullptr;      // the caller's step and iteration counters (useCalcd)
const int *g_iter = nullptr;

int grow(void **p, size_t *have, size_t need) {
  if (*have >= need) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  LNS3D_CHECK(hipMalloc(p, need));
  *have = need;
  return 0;
}

bool close_to(double a, double b) { return fabs(a - b) <= 1e-14 * fabs(b); }
}  // namespace

extern "C" int lns3d_set_module_state(const lns3d_module_state *m) {
  if (!m) return fail(LNS3D_EINVAL, "lns3d_set_module_state: null state");
  if (m->ndof != 5) return fail(LNS3D_EUNSUPPORTED, "lns3d_set_module_state: ndof must be 5");
  if (m->nx < 7 || m->ny < 7) return fail(LNS3D_EINVAL, "lns3d_set_module_state: nx and ny must be >= 7");
  if (!close_to(m->dxi, 1.0 / (m->nx - 1)) || !close_to(m->deta, 1.0 / (m->ny - 1)))
    return fail(LNS3D_EINVAL, "lns3d_set_module_state: dxi, deta must be 1/(nx-1), 1/(ny-1) (src/gengrid.f90:36,44)");
  if (m->carp) return fail(LNS3D_EUNSUPPORTED, "lns3d_set_module_state: the Carpenter closure (carp) is not implemented");
  if (m->bsym || m->tsym) return fail(LNS3D_EUNSUPPORTED, "lns3d_set_module_state: bsym/tsym are not implemented");
  if (m->lsym != (!m->xper && m->left == 7) || m->rsym != (!m->xper && m->right == 7))
    return fail(LNS3D_EINVAL, "lns3d_set_module_state: lsym/rsym disagree with left/right = 7 (src/input.f90:235-243)");
  if (!close_to(m->gamma1, m->gamma - 1.0)) return fail(LNS3D_EINVAL, "lns3d_set_module_state: gamma1 != gamma - 1");
  if (!m->navier && (m->mattyp != 0 || !m->bnb))
    return fail(LNS3D_EINVAL, "lns3d_set_module_state: Navier = .false. comes with mattyp = 0 and the wall normals bnb");
  if (!m->vm || !m->Ah || !m->Bh || !m->Dh || !m->Dhi || !m->Vh11 || !m->Vh12 || !m->Vh22 || !m->ABhi || !m->dtl)
    return fail(LNS3D_EINVAL, "lns3d_set_module_state: the mean flow and the coefficient arrays are required");
  lns3d_params p{};
  p.nx = m->nx; p.ny = m->ny; p.optx = m->optx; p.opty = m->opty; p.xper = m->xper; p.yper = m->yper;
  p.wall = m->wall; p.wallt = m->wallt; p.top = m->top; p.left = m->left; p.right = m->right;
  p.linear = m->linear; p.mattyp = m->mattyp; p.loctime = m->loctime;
  p.Ma = m->Ma; p.Pr = m->Pr; p.gamma = m->gamma; p.kz = m->kz; p.omega = m->omega;
  // the reference's inviscid deck holds Re = 1e99 with Navier = .false.
  // (src/input.f90:31-36); here that deck is Re = 0
  p.Re = m->navier ? m->Re : 0.0;
  p.cflmax = m->cflmax; p.dtmax = m->dtmax; p.alfa = m->alfa;
  p.periodic_exact = 0; p.precision = 0;
  p.wall_is = m->is; p.wall_ie = m->ie; p.idamp = m->idamp;
  p.use_calcd = m->use_calcd; p.update_lhs = m->update_lhs;
  if (p.wall_is == 1 && p.wall_ie == p.nx) p.wall_is = p.wall_ie = 0;   // the default extent

  std::lock_guard<std::mutex> lk(g_mu);
  int rc = 0;
  if (g_ctx && (g_ctx->p.nx != p.nx || g_ctx->p.ny != p.ny || g_ctx->p.xper != p.xper || g_ctx->p.yper != p.yper)) {
    lns3d_ctx_destroy(g_ctx);
    g_ctx = nullptr;
  }
  if (g_ctx) rc = lns3d_ctx_set_params(g_ctx, &p);
  else rc = lns3d_ctx_create(&p, &g_ctx);
  if (rc) return rc;
  lns3d_ctx *c = g_ctx;
  const long P = (long)p.nx * p.ny;
  if (p.mattyp == 1 && (rc = lns3d_set_sutherland(c, m->S0))) return rc;
  rc = lns3d_set_coeffs(c, m->Ah, m->Bh, m->Dh, m->Dhi, m->Vh11, m->Vh12, m->Vh22, m->ABhi, m->dtl);
  if (rc) return rc;
  LNS3D_CHECK(hipMemcpyAsync(c->c.vm, m->vm, (size_t)P * 5 * 8, hipMemcpyHostToDevice, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  c->vm_ready = true;
  if (m->ispg && (rc = lns3d_set_sponge(c, m->ispg, m->spg, m->spg2, m->x, m->wamp, m->x0))) return rc;
  if (m->eps_e != 0.0 && (rc = lns3d_set_smoother(c, m->eps_e, m->buff))) return rc;
  if (!p.yper && p.top == 1 && (rc = lns3d_set_top_forcing(c, m->x, m->wamp, m->x0))) return rc;
  if (!p.xper && (p.left == 1 || p.left == 4 || p.left == 5 || p.right == 4 || p.right == 5) &&
      (rc = lns3d_set_inflow(c, m->x, m->x0, m->eig, m->lalpha)))
    return rc;
  if (!p.yper && p.wall == 3 && (rc = lns3d_set_wall(c, m->bump))) return rc;
  if (!c->navier) {
    // bnb(nx,2) is the [2][nx] layout the inviscid rows read
    LNS3D_CHECK(hipMemcpyAsync(c->bnb_block, m->bnb, (size_t)2 * p.nx * 8, hipMemcpyHostToDevice, c->stream));
    LNS3D_CHECK(hipStreamSynchronize(c->stream));
    c->bnb_ready = true;
  }
  g_istep = m->istep;
  g_iter = m->iter;
  return 0;
}

static int need_ctx(const char *who) {
  if (!g_ctx || !g_ctx->coeffs_ready) {
    char msg[160];
    snprintf(msg, sizeof msg, "%s: no module state (lns3d_set_module_state first)", who);
    return fail(LNS3D_EINVAL, msg);
  }
  return 0;
}

extern "C" int lns3d_drop_lrhs3d(const void *vl, void *rl, const double *vml, const double *xl, const double *yl) {
  (void)vml; (void)xl; (void)yl;
  if (!vl || !rl) return fail(LNS3D_EINVAL, "lrhs3d: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = need_ctx("lrhs3d");
  if (rc) return rc;
  lns3d_ctx *c = g_ctx;
  const size_t bytes = 5ul * c->p.nx * c->p.ny * 16;
  if (!c->vstate) LNS3D_CHECK(hipMalloc(&c->vstate, bytes));
  LNS3D_CHECK(hipMemcpyAsync(c->vstate, vl, bytes, hipMemcpyHostToDevice, c->stream));
  rc = lns3d_lrhs3d(c, c->vstate, c->r);
  if (rc) return rc;
  LNS3D_CHECK(hipMemcpyAsync(rl, c->r, bytes, hipMemcpyDeviceToHost, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

static int drop_lhs(int dir, void *mat, const double *AB, const double *Dh, const double *Dhi, const double *V,
                    const double *ABhi, const double *dtl) {
  const char *who = dir == 1 ? "lhs1f3d" : "lhs2f3d";
  if (!mat || !AB || !Dh || !Dhi || !V || !ABhi || !dtl) return fail(LNS3D_EINVAL, "lhs: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = need_ctx(who);
  if (rc) return rc;
  lns3d_ctx *c = g_ctx;
  // the coefficient arrays the reference passes as arguments
  rc = ctx_put_soa(c, dir == 1 ? c->c.Ah : c->c.Bh, AB, 25, 25);
  if (!rc) rc = ctx_put_soa(c, c->c.Dh, Dh, 25, 25);
  if (!rc) rc = ctx_put_soa(c, c->c.Dhi, Dhi, 25, 25);
  if (!rc) rc = ctx_put_soa(c, dir == 1 ? c->c.V11 : c->c.V22, V, 6, 6);
  if (!rc) rc = ctx_put_soa(c, c->c.ABhi, ABhi, 4, 6);
  if (!rc) rc = ctx_put_soa(c, c->c.dtl, dtl, 1, 1);
  if (rc) return rc;
  c->factored = false;
  // useCalcd (src/lhs1f3d.f90:27-37, lhs2f3d.f90:27-37): the xi factor
  // holds the D terms when istep - 1 + iter is even, the eta factor otherwise
  const int istep = g_istep ? *g_istep : 1, iter = g_iter ? *g_iter : 1;
  c->c.calcd_xi = c->p.use_calcd ? (int)((istep - 1 + iter) % 2 == 0) : 1;
  const size_t bytes = 125ul * c->p.nx * c->p.ny * 16;
  if ((rc = grow(&g_mat, &g_mat_bytes, bytes))) return rc;
  rc = lns3d_assemble(c, dir, g_mat);
  if (rc) return rc;
  LNS3D_CHECK(hipMemcpyAsync(mat, g_mat, bytes, hipMemcpyDeviceToHost, c->stream));
  LNS3D_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int lns3d_drop_lhs1f3d(void *mat, const double *Ah, const double *Dh, const double *Dhi,
                                  const double *Vh11, const double *ABhi, const double *dtl, const void *vl) {
  (void)vl;
  return drop_lhs(1, mat, Ah, Dh, Dhi, Vh11, ABhi, dtl);
}

extern "C" int lns3d_drop_lhs2f3d(void *mat, const double *Bh, const double *Dh, const double *Dhi,
                                  const double *Vh22, const double *ABhi, const double *dtl, const void *vl) {
  (void)vl;
  return drop_lhs(2, mat, Bh, Dh, Dhi, Vh22, ABhi, dtl);
}

static int check_grad_args(int ndof, int nx, int ny, double dx, double dy, int optx, int opty, int bsym, int tsym,
                           int carp) {
  if (ndof != 5) return fail(LNS3D_EUNSUPPORTED, "cgrad: ndof must be 5");
  if (nx < 7 || ny < 7 || !(dx > 0.0) || !(dy > 0.0)) return fail(LNS3D_EINVAL, "cgrad: nx, ny >= 7 and dx, dy > 0");
  if (bsym || tsym || carp) return fail(LNS3D_EUNSUPPORTED, "cgrad: bsym, tsym and carp are not implemented");
  (void)optx; (void)opty;
  return 0;
}

extern "C" int lns3d_drop_cgrad(int ndof, int nx, int ny, const void *v, void *g1v, void *g2v, double dx, double dy,
                                int optx, int opty, int xper, int yper, int lsym, int rsym, int bsym, int tsym,
                                int carp) {
  if (!v || !g1v || !g2v) return fail(LNS3D_EINVAL, "cgrad: null argument");
  int rc = check_grad_args(ndof, nx, ny, dx, dy, optx, opty, bsym, tsym, carp);
  if (rc) return rc;
  const Stencil st = make_stencil_spacing(optx, opty, dx, dy, xper ? 0 : ((lsym ? 1 : 0) | (rsym ? 2 : 0)));
  const size_t f = 5ul * nx * ny * 16;
  std::lock_guard<std::mutex> lk(g_mu);
  if ((rc = grow(&g_buf, &g_buf_bytes, 5 * f))) return rc;
  char *b = (char *)g_buf;
  LNS3D_CHECK(hipMemcpy(b, v, f, hipMemcpyHostToDevice));
  rc = launch_cgrad(nx, ny, xper, yper, st, b, b + f, b + 2 * f, nullptr);
  if (rc) return rc;
  LNS3D_CHECK(hipMemcpy(g1v, b + f, f, hipMemcpyDeviceToHost));
  LNS3D_CHECK(hipMemcpy(g2v, b + 2 * f, f, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int lns3d_drop_cgrad2(int ndof, int nx, int ny, const void *v, const void *g1v, void *g11v, void *g12v,
                                 void *g22v, double dx, double dy, int optx, int opty, int xper, int yper, int lsym,
                                 int rsym, int bsym, int tsym, int carp) {
  if (!v || !g1v || !g11v || !g12v || !g22v) return fail(LNS3D_EINVAL, "cgrad2: null argument");
  int rc = check_grad_args(ndof, nx, ny, dx, dy, optx, opty, bsym, tsym, carp);
  if (rc) return rc;
  const Stencil st = make_stencil_spacing(optx, opty, dx, dy, xper ? 0 : ((lsym ? 1 : 0) | (rsym ? 2 : 0)));
  const size_t f = 5ul * nx * ny * 16;
  std::lock_guard<std::mutex> lk(g_mu);
  if ((rc = grow(&g_buf, &g_buf_bytes, 5 * f))) return rc;
  char *b = (char *)g_buf;
  LNS3D_CHECK(hipMemcpy(b, v, f, hipMemcpyHostToDevice));
  LNS3D_CHECK(hipMemcpy(b + f, g1v, f, hipMemcpyHostToDevice));
  rc = launch_cgrad2(nx, ny, xper, yper, st, b, b + f, b + 2 * f, b + 3 * f, b + 4 * f, nullptr);
  if (rc) return rc;
  LNS3D_CHECK(hipMemcpy(g11v, b + 2 * f, f, hipMemcpyDeviceToHost));
  LNS3D_CHECK(hipMemcpy(g12v, b + 3 * f, f, hipMemcpyDeviceToHost));
  LNS3D_CHECK(hipMemcpy(g22v, b + 4 * f, f, hipMemcpyDeviceToHost));
  return 0;
}
