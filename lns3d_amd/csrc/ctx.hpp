// The device-resident context of one problem (coefficients, factors, work
// arrays, optional boundary and forcing data), shared by the C-ABI sources.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/lns3d_amd.h"
#include "common.hpp"
#include "step.hpp"

using lns3d::Coeffs;
using lns3d::PentaPlan;
using lns3d::Stencil;

struct lns3d_ctx {
  lns3d_params p{};
  Stencil st{};
  Coeffs c{};
  hipStream_t stream = nullptr;
  PentaPlan px, py;          // xi lines (ny of length nx), eta lines (nx of length ny)
  PentaPlan px32, py32;      // complex-float factors for the refined step
  bool have32 = false;
  void *work = nullptr;      // 5P complex
  void *r = nullptr;         // 5P complex
  void *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;   // 5P complex each (refinement)
  double *scal = nullptr;    // device scalars: [0,8) time step, [8,2048) refinement norms and
                             // partials, [2048,3072) step residual-norm partials
  double *norm_out = nullptr;  // per-step residual norms (lns3d_ctx_set_residual_output)
  int norm_count = 0;
  long norm_idx = 0;
  double *coef_block = nullptr;
  double *sponge_block = nullptr;   // sp1, sp2, spl (P each), spf (10P), x and wamp staging
  double *top_block = nullptr;      // ftop (10 nx), x top row and wamp staging (nx each)
  double *side_block = nullptr;     // fside (20 ny), x left column staging (ny)
  double *smooth_block = nullptr;   // eps_e buff (P)
  double *wall_block = nullptr;     // fwall (10 nx)
  std::vector<double> eig;          // left = 4 profiles (10 ny), kept for ispg = 4
  double lalpha[2] = {0.0, 0.0};
  bool coeffs_ready = false, factored = false;
  bool vm_ready = false;            // the mean flow is on the device (lns3d_setup)
  bool s0_set = false;              // Sutherland's constant given (mattyp = 1)
  int impl = 1;                     // time scheme (src/impdrv3d.f90:179-199)
  long lstep = 0;                   // steps taken with this scheme
  void *vhist = nullptr;            // impl = 2: 5P complex, v at the start of the last step
  int niter = 1;                    // corrector iterations per step (src/impdrv3d.f90:173)
  void *vstart = nullptr;           // niter > 1: 5P complex, v at the start of this step
  void *vstate = nullptr;    // 5P complex: the state of the host-array driver
  bool state_fresh = false;  // uploaded, boundary conditions not yet applied
  int last_iters = 0;
  double last_resid = 0.0;
  bool navier = true;               // false: the inviscid deck (Re = 0, src/input.f90:31-36)
  double *bnb_block = nullptr;      // wall normals [2][nx] of the inviscid rows
  bool bnb_ready = false;
  int cur_iter = 1;                 // corrector iteration being taken (useCalcd's parity)
  int factored_calcd = -1;          // the D-term placement the factors hold
};

namespace lns3d {
// host array (P points, `stride` values per point, the first `ne` used) into
// a structure-of-arrays coefficient block [ne][P] on the device
int ctx_put_soa(lns3d_ctx *c, double *dst, const double *src, int ne, int stride);
}  // namespace lns3d
