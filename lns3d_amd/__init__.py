"""MI355X (gfx950) implementation of the LNS3D implicit approximate-
factorization step: block-pentadiagonal line solvers, finite-difference RHS,
LHS assembly and boundary closures, behind the C ABI of include/lns3d_amd.h.
"""
from .config import Params, make_grid, make_mean, make_disturbance  # noqa: F401
