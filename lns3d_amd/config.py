"""Synthetic problem definitions for the implicit step.

Everything here is input generation on the host (grid, metrics, mean flow,
initial disturbance), the part the reference reads from files at start-up
(src/gengrid.f90:17-80, src/genmean.f90:28-33, src/genini3d.f90:26-57).

Grid: x uniform on [xmin, xmax]; y tanh-stretched on [0, ymax] with the
formula and constants of util/genmesh.f90:72-102 (the ``Yi < 0`` branch),
evaluated in double precision.  Metrics are the analytic ones from the same
file (m1 = 1/(xmax-xmin), n2 and n22 from the stretching, others zero).

Mean flow: a parallel boundary layer u = tanh(y), v = w = 0, temperature from
the Crocco-Busemann relation for an adiabatic wall with unit recovery factor,
T = 1 + (gamma-1)/2 Ma^2 (1 - u^2), and density rho = 1/T (uniform pressure).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Params:
    """Scheme, physics and boundary selectors (the module state of
    src/global.f90 that the step reads).  Defaults: the benchmark case."""
    nx: int = 129
    ny: int = 65
    optx: int = -1
    opty: int = -1
    xper: int = 0
    yper: int = 0
    wall: int = 0
    wallt: int = 0
    top: int = 0
    left: int = 0
    right: int = 0
    linear: int = 2
    mattyp: int = 0
    loctime: int = 0
    Ma: float = 0.5
    Re: float = 1000.0
    Pr: float = 0.72
    gamma: float = 1.4
    kz: float = 0.2          # spanwise wavenumber beta
    omega: float = 0.0
    cflmax: float = 5.0
    dtmax: float = 1.0e30
    alfa: float = 1.0
    periodic_exact: int = 0
    wall_is: int = 0         # &extras is, ie (1-based wall extent); 0: 1..nx
    wall_ie: int = 0
    idamp: int = 0
    use_calcd: int = 0       # &extras useCalcd, updateLHS
    update_lhs: int = 0
    xmin: float = 0.0
    xmax: float = 100.0
    ymax: float = 40.0

    @property
    def npts(self) -> int:
        return self.nx * self.ny


def make_grid(p: Params):
    """x, y (nx, ny) and the 10 metric arrays (10, nx, ny), Fortran order."""
    nx, ny = p.nx, p.ny
    deta = 1.0 / (ny - 1)
    eta = np.arange(ny) * deta
    ds1, ds2, dd = 0.0005, 5.0, 5.36966703089523
    ymax = p.ymax
    r = math.sqrt(ds2 / ds1)
    s = 0.5 + 1.0 / math.tanh(dd / 2.0) * np.tanh(dd * (-0.5 + eta)) / 2.0
    y1 = ymax * s / (r + (1.0 - r) * s)
    n2 = (np.cosh(dd * (-0.5 + eta)) / math.cosh(dd / 2.0) *
          (r * np.sinh(dd * (1.0 - eta)) + np.sinh(dd * eta)) ** 2 /
          (dd * r * ymax * (np.sinh(dd * (1.0 - eta)) + np.sinh(dd * eta))))
    n22 = (ds1 * np.cosh(dd * (-0.5 + eta)) *
           (np.cosh(dd * (1.0 - 3.0 * eta)) - r * np.cosh(dd * (2.0 - 3.0 * eta)) +
            np.cosh(dd * (1.0 - eta)) - 2.0 * r * np.cosh(dd * (1.0 - eta)) +
            2.0 * np.cosh(dd * eta) - r * np.cosh(dd * eta)) /
           math.sinh(dd / 2.0) / math.cosh(dd / 2.0) ** 2 /
           np.cosh(dd / 2.0 - dd * eta) ** 2 *
           (r * np.sinh(dd * (1.0 - eta)) + np.sinh(dd * eta)) ** 3 /
           (4.0 * dd * ds2 * ymax ** 2 *
            (np.sinh(dd * (1.0 - eta)) + np.sinh(dd * eta))))
    dx = (p.xmax - p.xmin) / (nx - 1)
    x1 = p.xmin + np.arange(nx) * dx
    x = np.asfortranarray(np.repeat(x1[:, None], ny, axis=1))
    y = np.asfortranarray(np.repeat(y1[None, :], nx, axis=0))
    met = np.zeros((nx, ny, 10), order="F")
    met[:, :, 0] = 1.0 / (p.xmax - p.xmin)   # m1
    met[:, :, 3] = n2[None, :]               # n2
    met[:, :, 9] = n22[None, :]              # n22
    return x, y, met


def make_mean(p: Params, y: np.ndarray) -> np.ndarray:
    """Mean flow vm(5, nx, ny): rho, u, v, w, T."""
    u = np.tanh(y)
    T = 1.0 + 0.5 * (p.gamma - 1.0) * p.Ma ** 2 * (1.0 - u ** 2)
    vm = np.zeros((5, p.nx, p.ny), order="F")
    vm[0] = 1.0 / T
    vm[1] = u
    vm[4] = T
    return vm


def make_disturbance(p: Params, x: np.ndarray, seed: int = 0, amp: float = 1.0e-6,
                     x_c: float = 30.0, sigma: float = 5.0) -> np.ndarray:
    """Initial disturbance v(5, nx, ny): seeded complex N(0, amp^2) times a
    Gaussian bump in x centred at x_c."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((5, p.nx, p.ny)) + 1j * rng.standard_normal((5, p.nx, p.ny))
    bump = np.exp(-0.5 * ((x - x_c) / sigma) ** 2)
    return np.asfortranarray(amp * z * bump[None, :, :])


def disturbance_envelope(x: np.ndarray, x_c: float = 30.0, sigma: float = 5.0) -> np.ndarray:
    """the Gaussian bump in x of make_disturbance, (nx, ny)"""
    return np.exp(-0.5 * ((x - x_c) / sigma) ** 2)


def make_disturbance_device(envelope_dev, seed: int, amp: float = 1.0e-6):
    """The same kind of initial disturbance drawn on the device (torch's
    generator, seeded; the current stream): complex N(0, amp^2) in each of
    the real and imaginary parts times the envelope, as the flat (5, nx, ny)
    state in the reference's order.  envelope_dev: the envelope flattened in
    Fortran order (nx, ny), complex double, on the device.  Reproducible for
    a seed, but not the same numbers as make_disturbance."""
    import torch
    gen = torch.Generator(device=envelope_dev.device)
    gen.manual_seed(int(seed))
    P = envelope_dev.numel()
    # torch's complex normal has variance 1/2 in each part
    z = torch.randn((P, 5), dtype=torch.complex128, generator=gen, device=envelope_dev.device)
    return (z * (amp * math.sqrt(2.0)) * envelope_dev[:, None]).reshape(-1)
