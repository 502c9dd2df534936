"""Background models and model fields in the reference's formats.

* Model files: ``nz`` text rows of six columns, z/R_sun, c, rho, p, g, Gamma_1
  in cgs, read by ``solar_data`` (initialize.f90:743-747).
* The polytrope of the benchmark configurations (adiabatic, Gamma = 5/3,
  index m = 1.5, g = 274 m/s^2, depth d = z0 - z with z0 = +1.2 Mm,
  rho = rho_ref (d / 1 Mm)^m, p from hydrostatic balance p = rho g d / (m+1),
  c^2 = Gamma p / rho), on nz uniform points from z_bot to z_top.
* Flow stream function psi (Mm units as in model_psi_lsLL.fits) of the
  supergranule-like cell rho v = curl(psi_phys y), and the sound-speed bump.
"""
from __future__ import annotations

import numpy as np

from .params import RSUN

# Density at 1 Mm depth for the synthetic polytrope (g/cm^3).  The benchmark
# configurations leave it free; every result here scales out of it except the
# kernel amplitudes, which scale with rho0.
RHO_REF = 1.0e-6
G_SURF = 274.0e2          # cm/s^2
GAMMA = 5.0 / 3.0
POLY_M = 1.5
Z0_MM = 1.2


def polytrope(nz: int, z_bot_mm: float = -20.0, z_top_mm: float = 1.0,
              rho_ref: float = RHO_REF) -> np.ndarray:
    """Rows (z/R_sun, c, rho, p, g, Gamma) of the adiabatic polytrope."""
    zmm = np.linspace(z_bot_mm, z_top_mm, nz)
    d_cm = (Z0_MM - zmm) * 1e8
    rho = rho_ref * (d_cm / 1e8) ** POLY_M
    p = rho * G_SURF * d_cm / (POLY_M + 1.0)
    c = np.sqrt(GAMMA * p / rho)
    out = np.empty((nz, 6))
    out[:, 0] = 1.0 + zmm * 1e8 / RSUN
    out[:, 1] = c
    out[:, 2] = rho
    out[:, 3] = p
    out[:, 4] = G_SURF
    out[:, 5] = GAMMA
    return out


def read_model_file(path, nz: int) -> np.ndarray:
    """First nz rows of a six-column model file (initialize.f90:743-747)."""
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) < 6:
                continue
            rows.append([float(t.replace("D", "E").replace("d", "e")) for t in parts[:6]])
            if len(rows) == nz:
                break
    if len(rows) < nz:
        raise ValueError(f"{path}: {len(rows)} rows, need {nz}")
    return np.asarray(rows, dtype=np.float64)


def write_model_file(path, bg: np.ndarray) -> None:
    np.savetxt(path, bg, fmt="%.17e")


def stable_timestep(bg: np.ndarray, courant: float = 0.25) -> float:
    """dt = courant * min(dz) / max(c) in seconds (the benchmark configurations' rule)."""
    zc = bg[:, 0] * RSUN
    return courant * float(np.min(np.diff(zc))) / float(np.max(bg[:, 1]))


def cadence_aligned_timestep(dt_max: float, cadence: float = 30.0) -> float:
    """Largest dt <= dt_max for which floor(cadence/dt) * dt equals cadence to
    rounding, so that the output samples (every floor(outputcad/dt) steps,
    initialize.f90:998-1001) and the adjoint-source nodes (multiples of
    cadforcing, physics.f90:684-697) fall on the same times."""
    n = int(np.ceil(cadence / dt_max - 1e-12))
    dt = cadence / n
    while int(np.floor(cadence / dt)) != n:     # keep floor() at n despite rounding
        dt = np.nextafter(dt, 0.0)
    return float(dt)


def grid_x_mm(nx: int, xlength_cm: float) -> np.ndarray:
    """Horizontal positions (Mm) relative to the box centre, x = (i-1)/nx (initialize.f90:194-198)."""
    return (np.arange(nx) / nx - 0.5) * xlength_cm * 1e-8


def grid_z_mm(bg: np.ndarray) -> np.ndarray:
    return (bg[:, 0] - 1.0) * RSUN * 1e-8


def supergranule_psi(nx: int, xlength_cm: float, bg: np.ndarray, vmax_ms: float = 240.0,
                     sigx_mm: float = 15.0, sigz_mm: float = 0.9, z0_mm: float = -2.3) -> np.ndarray:
    """psi (Mm units, [nz][nx]) whose flow rho v = curl(psi_phys y) peaks at vmax.

    The physical stream function is psi_phys = A exp(-x^2/2sx^2) exp(-(z-z0)^2/2sz^2);
    the reference expects psi_ref with psi_phys = rho c psi_ref (driver.f90:108), so
    psi_ref = psi_phys / (rho c).  A is fixed by the peak |v| evaluated on the grid.
    """
    x = grid_x_mm(nx, xlength_cm)
    z = grid_z_mm(bg)
    shape = np.exp(-(z[:, None] - z0_mm) ** 2 / (2 * sigz_mm ** 2)) * \
        np.exp(-x[None, :] ** 2 / (2 * sigx_mm ** 2))
    rho = bg[:, 2][:, None]
    c = bg[:, 1][:, None]
    # v = curl(psi y)/rho in cm/s with x, z in cm: vx = -d_z psi / rho, vz = d_x psi / rho
    zc = z * 1e8
    xc = x * 1e8
    dpsidz = np.gradient(shape, zc, axis=0)
    dpsidx = np.gradient(shape, xc, axis=1)
    vmag = np.sqrt((dpsidz / rho) ** 2 + (dpsidx / rho) ** 2)
    amp = vmax_ms * 1e2 / float(np.max(vmag))
    psi_phys = amp * shape                      # g cm^-1 s^-1 (rho v L)
    # driver.f90:108: psivar = rho0*c*psi/(diml/1e8) in nondimensional units; with
    # rho0 = rho/1e-3, c = c/1e6 and derivatives in 1/R_sun this maps psi (Mm) to
    # psi_phys = rho c psi_ref * 1e8 (cm per Mm).
    return psi_phys / (rho * c * 1e8)


def sound_speed_bump(nx: int, xlength_cm: float, bg: np.ndarray, amp: float = 0.05,
                     sigx_mm: float = 10.0, sigz_mm: float = 1.5, z0_mm: float = -3.0) -> np.ndarray:
    """c(x,z) in cm/s with dc/c = amp exp(-x^2/2sx^2) exp(-(z-z0)^2/2sz^2)."""
    x = grid_x_mm(nx, xlength_cm)
    z = grid_z_mm(bg)
    bump = amp * np.exp(-x[None, :] ** 2 / (2 * sigx_mm ** 2)) * \
        np.exp(-(z[:, None] - z0_mm) ** 2 / (2 * sigz_mm ** 2))
    return bg[:, 1][:, None] * (1.0 + bump)


def supergranule_flow_3d(nx: int, ny: int, xlength_cm: float, ylength_cm: float, bg: np.ndarray,
                          vmax_ms: float = 240.0, sigr_mm: float = 15.0, sigz_mm: float = 0.9,
                          z0_mm: float = -2.3):
    """Axisymmetric supergranule: the 2D stream function rotated about the z axis,
    rho v = curl(psi phi_hat) with psi = A exp(-r^2/2sr^2) exp(-(z-z0)^2/2sz^2);
    returns (vx, vy, vz) in cm/s, [nz][ny][nx] each, peak |v| = vmax."""
    x = grid_x_mm(nx, xlength_cm) * 1e8
    y = grid_x_mm(ny, ylength_cm) * 1e8
    z = grid_z_mm(bg) * 1e8
    rho = bg[:, 2][:, None, None]
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X ** 2 + Y ** 2)
    sr, sz, z0 = sigr_mm * 1e8, sigz_mm * 1e8, z0_mm * 1e8
    gz = np.exp(-(z - z0) ** 2 / (2 * sz ** 2))[:, None, None]
    dgz = (-(z - z0) / sz ** 2)[:, None, None] * gz
    gr = np.exp(-r ** 2 / (2 * sr ** 2))[None]
    # v_r = -(1/rho) d psi/dz, v_z = (1/rho) (1/r) d(r psi)/dr
    vr = -(gr * dgz) / rho
    vz = (gz * gr * (2.0 - r[None] ** 2 / sr ** 2)) / rho
    with np.errstate(invalid="ignore", divide="ignore"):
        cx = np.where(r > 0, X / r, 0.0)[None]
        cy = np.where(r > 0, Y / r, 0.0)[None]
    vx, vy = vr * cx, vr * cy
    amp = vmax_ms * 1e2 / float(np.sqrt(vx ** 2 + vy ** 2 + vz ** 2).max())
    return amp * vx, amp * vy, amp * vz


def sound_speed_bump_3d(nx: int, ny: int, xlength_cm: float, ylength_cm: float, bg: np.ndarray,
                        amp: float = 0.05, sigr_mm: float = 10.0, sigz_mm: float = 1.5,
                        z0_mm: float = -3.0) -> np.ndarray:
    """c(x, y, z) in cm/s with the radially symmetric bump dc/c = amp exp(-r^2/2sr^2) exp(-(z-z0)^2/2sz^2)."""
    x = grid_x_mm(nx, xlength_cm)
    y = grid_x_mm(ny, ylength_cm)
    z = grid_z_mm(bg)
    r2 = x[None, None, :] ** 2 + y[None, :, None] ** 2
    bump = amp * np.exp(-r2 / (2 * sigr_mm ** 2)) * np.exp(-(z[:, None, None] - z0_mm) ** 2 / (2 * sigz_mm ** 2))
    return bg[:, 1][:, None, None] * (1.0 + bump)


def source_positions(nsrc: int, seed: int = 0, lo: float = -60.0, hi: float = 60.0) -> np.ndarray:
    """Source x positions (Mm) drawn uniformly and rounded to %.1f like master.pixels
    (generate_master_pixels.py:8-14)."""
    rng = np.random.default_rng(seed)
    xs = np.sort(rng.uniform(lo, hi, nsrc))
    return np.array([float(f"{v:.1f}") for v in xs])
