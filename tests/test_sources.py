"""The Gaussian force source (sparc_flows_amd.sources) and the oracle's
separable source form, on the CPU.

* the time series is the derivative of a Gaussian: its amplitude spectrum
  peaks at 3 mHz, its peak magnitude is amp and it starts in its quiet tail;
* rho0 * zprof integrates to 1 over the non-dimensional z and peaks at
  z0 = -0.15 Mm; the x profile is a periodic Gaussian of the right width
  centred on each source;
* the oracle's separable form with zprof = delta_width / rho0 on row e_rad
  and series = FORWARD_SOURCE's reproduces the reference's one-row form.
"""
from math import erf

import numpy as np

import oracle
from sparc_flows_amd import background as B
from sparc_flows_amd.params import SparcParams
from sparc_flows_amd.sources import derivative_of_gaussian, gaussian_force
from tests.cases import rel_l2, small_case


def test_time_series_spectrum_peaks_at_3_mhz():
    dt = 0.5
    s = derivative_of_gaussian(dt, 8000, nu0_hz=3e-3, amp=1e-5)
    assert abs(np.abs(s).max() - 1e-5) < 1e-9
    assert abs(s[0]) < 1e-3 * 1e-5
    nu = np.fft.rfftfreq(s.size, dt)
    S = np.abs(np.fft.rfft(s))
    k = int(np.argmax(S))
    assert abs(nu[k] - 3e-3) <= nu[1]


def test_spatial_profiles():
    bg = B.polytrope(150)
    p = SparcParams(nx=128, nz=150, xlength=200e8, timestep=B.stable_timestep(bg), npmlbot=12,
                    npmltop=12, npmlhor=9)
    xs = [0.0, 97.0]                       # the second wraps across the periodic edge
    vr, zp, ser = gaussian_force(p, bg, xs, 500)
    assert vr.shape == (2, 128) and zp.shape == (150,) and ser.shape == (508,)
    x = B.grid_x_mm(p.nx, p.xlength)
    assert abs(x[np.argmax(vr[0])] - 0.0) < 200.0 / 128
    d = (x - 97.0 + 100.0) % 200.0 - 100.0
    assert np.allclose(vr[1], np.exp(-d * d / (2 * 0.25)), rtol=0, atol=1e-15)
    z_nd = bg[:, 0]
    rho0 = bg[:, 2] / 1e-3
    integral = float(np.sum(0.5 * (zp[1:] * rho0[1:] + zp[:-1] * rho0[:-1]) * np.diff(z_nd)))
    zmm = B.grid_z_mm(bg)
    # unit integral over all z; the top of the box (+1 Mm) cuts the upper tail
    inside = 0.5 * (erf((zmm[-1] + 0.15) / (0.5 * np.sqrt(2))) - erf((zmm[0] + 0.15) / (0.5 * np.sqrt(2))))
    assert 0.98 < inside < 0.995 and abs(integral - inside) < 2e-3   # trapezoid rule, 0.14 Mm rows
    assert abs(zmm[np.argmax(zp * rho0)] + 0.15) < 0.5 * (zmm[1] - zmm[0]) + 1e-9


def test_separable_form_reproduces_the_one_row_source():
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    maxtime = 300
    o1 = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    o1.set_forward_source(4.0, maxtime)
    ref = o1.run(maxtime)
    e_rad = o1.info("e_rad")
    zp = np.zeros(p.nz)
    zp[e_rad] = o1.info("delta_width") / o1.array("rho0").reshape(p.nz, p.nx)[e_rad, 0]
    o2 = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    o2.set_forward_profile(o1.array("vr"), zp, o1.array("fwdsource"), maxtime)
    got = o2.run(maxtime)
    assert np.abs(ref).max() > 0 and rel_l2(got, ref) < 1e-14
