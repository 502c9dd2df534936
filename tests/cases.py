"""Small seeded problem set-ups shared by the tests."""
from __future__ import annotations

import numpy as np

from sparc_flows_amd import background as B
from sparc_flows_amd.params import SparcParams


def small_case(nx=64, nz=60, npml=6, nph=5, flow=True, dc=True, xlength_mm=200.0):
    """Polytrope on nz points from -20 to +1 Mm, Courant-0.25 time step."""
    bg = B.polytrope(nz)
    dt = B.stable_timestep(bg)
    p = SparcParams(nx=nx, nz=nz, xlength=xlength_mm * 1e8, timestep=dt, npmlbot=npml,
                    npmltop=npml, npmlhor=nph)
    psi = B.supergranule_psi(nx, p.xlength, bg, sigx_mm=15.0) if flow else None
    c2d = B.sound_speed_bump(nx, p.xlength, bg) if dc else None
    return p, bg, c2d, psi


def adjoint_source(p: SparcParams, nt: int, nx: int, seed=0) -> np.ndarray:
    """A smooth seeded adjoint source [nt][nx] (stands in for source.fits)."""
    rng = np.random.default_rng(seed)
    t = np.arange(nt)[:, None]
    x = np.arange(nx)[None, :]
    a = rng.uniform(0.5, 1.5, 3)
    s = (np.sin(2 * np.pi * t / max(nt, 2) * a[0]) * np.exp(-((x - nx * 0.4) / (nx * 0.08)) ** 2)
         + 0.5 * np.cos(2 * np.pi * t / max(nt, 2) * a[1]) * np.exp(-((x - nx * 0.7) / (nx * 0.05)) ** 2))
    return 1e-6 * s


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
