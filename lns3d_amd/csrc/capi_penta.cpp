// C-ABI of the block-pentadiagonal line solvers: the drop-in entry points
// with the reference's Fortran argument meaning (host arrays in its layout)
// and the device-resident handle API.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <map>
#include <mutex>
#include <algorithm>
#include <string>
#include <tuple>

#include "../../include/lns3d_amd.h"
#include "common.hpp"
#include "penta.hpp"

namespace lns3d {

static thread_local std::string g_err;

int set_error(const char *file, int line, const char *msg) {
  char buf[512];
  const char *base = strrchr(file, '/');
  snprintf(buf, sizeof buf, "%s:%d: %s", base ? base + 1 : file, line, msg);
  g_err = buf;
  return LNS3D_EHIP;
}
static int fail(int code, const char *msg) {
  g_err = msg;
  return code;
}

}  // namespace lns3d

using namespace lns3d;

struct lns3d_penta {
  PentaPlan plan;
  // periodic lines in complex float: the complex-double bordered LU, whose
  // records are rounded into plan
  PentaPlan plan64;
  bool has64 = false;
  int layout = 1;
  int nsys = 0;
  int nrows = 0;   // n (bc) or np (periodic), as the caller passes it
  int factored = 0;
};

extern "C" const char *lns3d_last_error(void) { return g_err.c_str(); }

extern "C" int lns3d_penta_create(int kind, int layout, int nsys, int nrows, int precision,
                                  int exact, lns3d_penta **out) {
  if (!out || (kind != 0 && kind != 1) || (layout != 1 && layout != 2) || nsys < 1 ||
      (precision != 0 && precision != 1))
    return fail(LNS3D_EINVAL, "lns3d_penta_create: bad arguments");
  int n_unknown = kind == 1 ? nrows - 1 : nrows;
  if (n_unknown < 7) return fail(LNS3D_EINVAL, "lns3d_penta_create: lines need >= 7 unknowns");
  auto *h = new lns3d_penta;
  h->layout = layout;
  h->nsys = nsys;
  h->nrows = nrows;
  int rc = penta_plan_create(&h->plan, kind, exact, nrows, nsys, precision);
  if (!rc && kind == 1 && precision == 1) {
    h->has64 = true;
    rc = penta_plan_create(&h->plan64, kind, exact, nrows, nsys, 0);
  }
  if (rc) {
    penta_plan_destroy(&h->plan);
    if (h->has64) penta_plan_destroy(&h->plan64);
    delete h;
    return rc;
  }
  *out = h;
  return 0;
}

extern "C" int lns3d_penta_destroy(lns3d_penta *h) {
  if (!h) return 0;
  penta_plan_destroy(&h->plan);
  if (h->has64) penta_plan_destroy(&h->plan64);
  delete h;
  return 0;
}

static void ref_strides(const lns3d_penta *h, long *sL, long *sR) {
  if (h->layout == 1) { *sL = 1; *sR = h->nsys; }   // (.., nsys, n)
  else { *sL = h->nrows; *sR = 1; }                 // (.., n, nsys)
}

extern "C" int lns3d_penta_factor(lns3d_penta *h, const void *mat_dev, void *stream) {
  if (!h || !mat_dev) return fail(LNS3D_EINVAL, "lns3d_penta_factor: null argument");
  long sL, sR;
  ref_strides(h, &sL, &sR);
  hipStream_t s = (hipStream_t)stream;
  PentaPlan &lu = h->has64 ? h->plan64 : h->plan;   // the plan holding the complex-double LU
  int rc = penta_import(&lu, mat_dev, sL, sR, s);
  if (!rc) rc = penta_factor(&lu, s);
  if (!rc) rc = penta_check_launch(&lu, s);
  if (!rc && h->has64) rc = penta_records_narrow(&h->plan, &lu, s);
  if (!rc && h->plan.prec) rc = penta_check_records_launch(&h->plan, s);
  long bad = 0, first = -1;
  if (!rc) rc = penta_check_result(&h->plan, s, &bad, &first);
  if (!rc && h->has64) {
    long bad64 = 0, first64 = -1;
    rc = penta_check_result(&lu, s, &bad64, &first64);
    if (bad64 && (first < 0 || first64 < first)) first = first64;
    bad = std::max(bad, bad64);
  }
  if (rc) return rc;
  h->factored = 1;
  if (bad) {
    char msg[160];
    snprintf(msg, sizeof msg, "lns3d_penta_factor: %ld line(s) with a zero pivot or a non-finite factor (first: line %ld)",
             bad, first + 1);
    return fail(LNS3D_ENONFINITE, msg);
  }
  return 0;
}

extern "C" int lns3d_penta_solve(lns3d_penta *h, void *rhs_dev, void *stream) {
  if (!h || !rhs_dev) return fail(LNS3D_EINVAL, "lns3d_penta_solve: null argument");
  if (!h->factored) return fail(LNS3D_ENOTFACTORED, "lns3d_penta_solve: not factored");
  long sL, sR;
  ref_strides(h, &sL, &sR);
  return penta_solve(&h->plan, rhs_dev, sL, sR, (hipStream_t)stream);
}

extern "C" int lns3d_penta_solve_reference_order(lns3d_penta *h, void *rhs_dev, void *stream) {
  if (!h || !rhs_dev) return fail(LNS3D_EINVAL, "lns3d_penta_solve_reference_order: null argument");
  if (!h->factored) return fail(LNS3D_ENOTFACTORED, "not factored");
  long sL, sR;
  ref_strides(h, &sL, &sR);
  return penta_solve_reference_order(h->has64 ? &h->plan64 : &h->plan, rhs_dev, sL, sR, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// drop-ins with host arrays.  Factors stay on the device, keyed by the host
// address of mat and the shape: the reference's driver never reads mat after
// factoring (src/impdrv3d.f90:269,295,319,327), so mat is left as passed.
namespace {
struct DropKey {
  const void *mat;
  int kind, layout, nsys, nrows;
  bool operator<(const DropKey &o) const {
    return std::tie(mat, kind, layout, nsys, nrows) < std::tie(o.mat, o.kind, o.layout, o.nsys, o.nrows);
  }
};
struct DropEntry {
  lns3d_penta *h = nullptr;
  void *dmat = nullptr, *drhs = nullptr;
};
std::map<DropKey, DropEntry> g_drop;
std::mutex g_drop_mu;
int g_periodic_exact = 0;
}  // namespace

extern "C" int lns3d_set_periodic_exact(int exact) {
  g_periodic_exact = exact ? 1 : 0;
  return 0;
}

static int dropin(int kind, int layout, int nsys, int nrows, int nblk, void *mat, void *rhs, int code) {
  if (nblk != 5) return fail(LNS3D_EUNSUPPORTED, "cpenta: only nblk = 5 is supported");
  if (code < 0 || code > 2 || !mat || (code != 0 && !rhs))
    return fail(LNS3D_EINVAL, "cpenta: bad arguments");
  std::lock_guard<std::mutex> lk(g_drop_mu);
  DropKey key{mat, kind, layout, nsys, nrows};
  size_t nmat = (size_t)nsys * nrows * 125 * 16, nrhs = (size_t)nsys * nrows * 5 * 16;
  auto it = g_drop.find(key);
  int factor_rc = 0;
  if (code == 0 || code == 2) {
    if (it == g_drop.end()) {
      DropEntry e;
      int rc = lns3d_penta_create(kind, layout, nsys, nrows, 0, g_periodic_exact, &e.h);
      if (rc) return rc;
      if (hipMalloc(&e.dmat, nmat) != hipSuccess || hipMalloc(&e.drhs, nrhs) != hipSuccess) {
        lns3d_penta_destroy(e.h);
        return fail(LNS3D_EHIP, "cpenta: device allocation failed");
      }
      it = g_drop.emplace(key, e).first;
    }
    it->second.h->plan.exact = g_periodic_exact;
    LNS3D_CHECK(hipMemcpy(it->second.dmat, mat, nmat, hipMemcpyHostToDevice));
    // a zero pivot or a non-finite factor is reported (LNS3D_ENONFINITE) and,
    // as in the reference, the solve of code 2 still runs and spreads it
    factor_rc = lns3d_penta_factor(it->second.h, it->second.dmat, nullptr);
    if (factor_rc && factor_rc != LNS3D_ENONFINITE) return factor_rc;
  } else if (it == g_drop.end()) {
    return fail(LNS3D_ENOTFACTORED, "cpenta: solve (code 1) before factor (code 0) for this mat");
  }
  if (code == 1 || code == 2) {
    const std::string msg = lns3d_last_error();
    LNS3D_CHECK(hipMemcpy(it->second.drhs, rhs, nrhs, hipMemcpyHostToDevice));
    int rc = lns3d_penta_solve(it->second.h, it->second.drhs, nullptr);
    if (rc) return rc;
    LNS3D_CHECK(hipMemcpy(rhs, it->second.drhs, nrhs, hipMemcpyDeviceToHost));
    if (factor_rc) g_err = msg;
  }
  return factor_rc;
}

extern "C" int lns3d_cpenta1bc(int nsys, int n, int nblk, void *mat, void *rhs, int code) {
  return dropin(0, 1, nsys, n, nblk, mat, rhs, code);
}
extern "C" int lns3d_cpenta2bc(int n, int nsys, int nblk, void *mat, void *rhs, int code) {
  return dropin(0, 2, nsys, n, nblk, mat, rhs, code);
}
extern "C" int lns3d_cpenta1p(int nsys, int np, int nblk, void *mat, void *rhs, void *per,
                              void *per2, int code) {
  (void)per; (void)per2;
  return dropin(1, 1, nsys, np, nblk, mat, rhs, code);
}
extern "C" int lns3d_cpenta2p(int np, int nsys, int nblk, void *mat, void *rhs, void *per,
                              void *per2, int code) {
  (void)per; (void)per2;
  return dropin(1, 2, nsys, np, nblk, mat, rhs, code);
}

extern "C" int lns3d_release_factors(void) {
  std::lock_guard<std::mutex> lk(g_drop_mu);
  for (auto &kv : g_drop) {
    lns3d_penta_destroy(kv.second.h);
    (void)hipFree(kv.second.dmat);
    (void)hipFree(kv.second.drhs);
  }
  g_drop.clear();
  return 0;
}
