"""Forward source forms.

* The reference's: FORWARD_SOURCE's time series (driver.f90:1120-1145) with
  DETERMINE_STATUS's sinc-Gaussian x profile (driver.f90:536-553), injected
  on the single row e_rad as vr * fwdsource(time+1) * delta_width / rho0
  (physics2d.f90:1056-1059).  It is built inside the C ABI from the source
  positions (``Sparc.forward``).
* A vertical force Gaussian in x and z with a derivative-of-Gaussian time
  profile (:func:`gaussian_force`), the source of the 128 x 150 plain forward
  configuration.  It goes through the separable form of the C ABI
  (``sparc_set_forward_profile``): RHS v_z += vr(x) * series(t) * zprof(z).
"""
from __future__ import annotations

import math

import numpy as np

from .background import grid_x_mm, grid_z_mm
from .params import RSUN, SparcParams

DIMRHO = 1e-3   # initialize.f90:753-776


def derivative_of_gaussian(timestep: float, nt: int, nu0_hz: float = 3e-3, amp: float = 1e-5,
                           delay_widths: float = 5.0) -> np.ndarray:
    """s(t) = -amp (t - t0)/tau exp(1/2 - (t - t0)^2 / (2 tau^2)) at t = k * timestep,
    k = 0..nt-1: the time derivative of a Gaussian, whose amplitude spectrum
    nu exp(-2 pi^2 nu^2 tau^2) peaks at nu0 = 1/(2 pi tau); |s| peaks at amp,
    and t0 = delay_widths * tau puts the start of the run in its quiet tail."""
    tau = 1.0 / (2.0 * math.pi * nu0_hz)
    t = np.arange(nt) * timestep - delay_widths * tau
    return -amp * (t / tau) * np.exp(0.5 - t * t / (2.0 * tau * tau))


def gaussian_force(p: SparcParams, model: np.ndarray, xsrc_mm, maxtime: int, sigma_mm: float = 0.5,
                   z0_mm: float = -0.15, nu0_hz: float = 3e-3, amp: float = 1e-5, cut: float = 1e-12):
    """Vertical force density f = amp * exp(-dx^2 / 2 s^2) * g(z) * s(t) / rho0 for
    sources at xsrc_mm (Mm from the box centre, periodic distance), with
    g(z) = exp(-(z - z0)^2 / 2 s^2) / (s sqrt(2 pi)) (unit integral over the
    non-dimensional z) cut to 0 below `cut` of its peak.

    Returns (vr [nsrc][nx], zprof [nz], series [maxtime + 8]) for
    ``Sparc.forward_profile`` / the oracle's ``set_forward_profile``; the
    series has the time-index convention of FORWARD_SOURCE (value k used by
    every stage of step k)."""
    xs = np.atleast_1d(np.asarray(xsrc_mm, dtype=np.float64))
    x = grid_x_mm(p.nx, p.xlength)
    L = p.xlength * 1e-8
    d = (x[None, :] - xs[:, None] + 0.5 * L) % L - 0.5 * L
    vr = np.exp(-d * d / (2.0 * sigma_mm ** 2))
    z = grid_z_mm(model[: p.nz])
    gz = np.exp(-(z - z0_mm) ** 2 / (2.0 * sigma_mm ** 2))
    gz[gz < cut] = 0.0
    s_nd = sigma_mm * 1e8 / RSUN
    rho0 = model[: p.nz, 2] / DIMRHO
    zprof = gz / (s_nd * math.sqrt(2.0 * math.pi)) / rho0
    series = derivative_of_gaussian(p.timestep, maxtime + 8, nu0_hz, amp)
    return vr, zprof, series
