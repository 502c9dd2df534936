"""Windowed cross-correlation travel-time misfit of the vz traces and its
adjoint source.

The reference measures travel times between observed and synthetic vz with
the Gizon & Birch (2002) formula inside ADJOINT_SOURCE_FILT
(driver.f90:1148-1754; ridge filters and window picking around it) and turns
them into the adjoint source written to adjoint_srcSS/source.fits.  Here the
same linearised formula is written in torch on the GPU, over the receivers
5-60 Mm from each source and a tapered time window (the bands of params.0),
and the adjoint source is the time-reversed gradient of the misfit with
respect to the traces, as a force density (see autograd.py for why that
scaling makes the kernels gradients).
"""
from __future__ import annotations

import math

import torch


def window_weights(nt: int, cadence_s: float, window_s: float, device=None) -> torch.Tensor:
    """1 over the first window_s seconds, cosine taper over its last 10 %."""
    t = torch.arange(nt, dtype=torch.float64, device=device) * cadence_s
    w = torch.ones(nt, dtype=torch.float64, device=device)
    taper = 0.1 * window_s
    w = torch.where(t > window_s, torch.zeros_like(w), w)
    ramp = (t > window_s - taper) & (t <= window_s)
    return torch.where(ramp, 0.5 * (1 + torch.cos(math.pi * (t - (window_s - taper)) / taper)), w)


def travel_times(vz: torch.Tensor, vz_ref: torch.Tensor, mask: torch.Tensor, w: torch.Tensor,
                 cadence_s: float, floor: float = 1e-2) -> tuple[torch.Tensor, torch.Tensor]:
    """Linearised cross-correlation travel times (Gizon & Birch 2002, eq. 13):
    tau = -int W dref/dt (vz - vref) dt / int W (dref/dt)^2 dt, so that
    vz(t) = vref(t - tau) gives tau > 0.  vz, vz_ref: [nsrc][nt][nx];
    mask: [nsrc][nx] receivers; w: [nt] window.  Receivers whose windowed
    reference energy is below `floor` times the source's largest are dropped
    (tau is ill-conditioned where the wave has not arrived).
    Returns (tau [nsrc][nx], the receiver mask used)."""
    dref = torch.gradient(vz_ref, spacing=cadence_s, dim=1)[0]
    num = (w[None, :, None] * dref * (vz - vz_ref)).sum(1)
    den = (w[None, :, None] * dref * dref).sum(1)
    m = mask * (den * mask > floor * (den * mask).amax(dim=1, keepdim=True))
    return -num / torch.where(m > 0, den, torch.ones_like(den)) * m, m


def travel_time_misfit(vz, vz_ref, tau_obs, mask, w, cadence_s, floor: float = 1e-2):
    """0.5 * sum over sources and receivers of (tau - tau_obs)^2."""
    tau, m = travel_times(vz, vz_ref, mask, w, cadence_s, floor)
    return 0.5 * ((tau - tau_obs) * m).pow(2).sum()


def travel_time_adjoint_source(vz: torch.Tensor, vz_ref: torch.Tensor, tau_obs: torch.Tensor,
                               mask: torch.Tensor, w: torch.Tensor, cadence_s: float, dT: float,
                               dx: float, floor: float = 1e-2):
    """Misfit 0.5 * sum (tau - tau_obs)^2 and its adjoint source
    vr = flip_t(dL/dvz) / (dT * dx) [nsrc][nt][nx] (time reversed, per unit
    non-dimensional time and length; dT = steps * deltat, dx the grid step)."""
    v = vz.detach().to(torch.float64).requires_grad_(True)
    with torch.enable_grad():
        L = travel_time_misfit(v, vz_ref, tau_obs, mask, w, cadence_s, floor)
        g, = torch.autograd.grad(L, v)
    return L.detach(), (torch.flip(g, dims=[1]) / (dT * dx)).contiguous()
