"""CPU restatement of SPARC's 3D solver with flows, option 16
(MP_FLOWS_PML_3D_DISPL, displacement.f90:377-477), for small grids.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg,
where it is the checker or the CPU baseline.  The product never imports it.

Numpy for the pointwise algebra and the periodic x/y derivatives (the FFT
product of ddxyz, derivatives.f90:175-305), the C oracle's dbyd2 for the
compact z derivative.  Followed as written:

- background and z-PML: mp_Initialize_RHS (physics.f90:34-404) -- e_rad and
  o_rad by the last tie of abs(z - target) (:44-60), source_dep (:105), d rho0/dz
  and d p0/dz by ddz with ibc 1 (:207-208), the PML profiles with 2*3*3 c(x,y,1)
  at the bottom and 3*3 c(x,y,nz) at the top (:265-290); g takes the PML factor
  of column (1,1) (:276, :287);
- right-hand side: displacement.f90:395-475 -- ddxyz with bc 1 for xi, bc 0 for
  p, bc 1 for every velocity component, the z-PML on d xi_z/dz and d p/dz only,
  the full advection -2 v0 . grad v, the source forcing * source_dep(k) on
  rows |k - e_rad| <= 4, and rows 1 and nz of every field's RHS zeroed;
- step: step.f90:43-119 with HORIZONTAL_PMLS (params.i:86): the RK stages as
  in 2D, then FILTER_VARS_3D_HORIZ_PML (physics.f90:836-882) on all six fields
  after every step with mod(time+1, 2) == 0.

Three points where the reference cannot run this path as written are fixed
here, and the GPU implementation follows the same fixes:
- the forward source: kernel-mode forward runs never set `forcing`
  (driver.f90:358-360); here forcing(x, y) = vr(x, y) * fwdsource(time+1),
  with vr as DETERMINE_STATUS builds it (driver.f90:536-553: the x profile
  repeated along y, a line source) or, for a point source at (xsrc, ysrc),
  the same profile in the horizontal distance, and fwdsource from
  FORWARD_SOURCE (driver.f90:1120-1145);
- the flow: the reference derives only v0_x, v0_z from a 2D psi
  (driver.f90:94-126); here (v0_x, v0_y, v0_z) are given (in cm/s), as the
  compute_data branch gives them (driver.f90:65-73);
- the x/y grid lengths are both runtime values (params.i fixes them).

Parity pins (no reference output exists for this path): the derivative
closed forms, the x <-> y exchange symmetry, linearity, y-invariance and
zero-in/zero-out, in tests/test_oracle3d.py.
"""
from __future__ import annotations

import dataclasses
import math
import struct

import numpy as np

from . import dbyd2 as _dbyd2
from . import filter_var_z as _filter_var_z
from . import forward_source_series

RSUN = 6.95989467700e10      # initialize.f90:37
PI = 3.141592653589793


def f32(v: float) -> float:
    return struct.unpack("f", struct.pack("f", float(v)))[0]


@dataclasses.dataclass
class Params3D:
    nx: int = 32
    ny: int = 32
    nz: int = 60
    xlength: float = 200.0e8
    ylength: float = 200.0e8
    timestep: float = 2.0
    cadforcing: float = 30.0
    outputcad: float = 30.0
    obsheight: float = 2.0e7
    excitdep: float = -1.5e7
    npmlbot: int = 10
    npmltop: int = 10
    nupeak: float = f32(0.0017)
    nuwidth: float = f32(0.0055)

    @property
    def steps(self) -> int:
        return int(math.floor(self.outputcad / self.timestep))


def _deriv_fft(f: np.ndarray, axis: int, stretch: float) -> np.ndarray:
    """d/dx of periodic rows on the unit interval, Nyquist bin dropped (FFTW
    c2r ignores the imaginary part of i k F at k = n/2)."""
    n = f.shape[axis]
    F = np.fft.rfft(f, axis=axis)
    k = 2.0 * PI * np.arange(n // 2 + 1)
    if n % 2 == 0:
        k[-1] = 0.0
    shape = [1] * f.ndim
    shape[axis] = k.size
    return np.fft.irfft(F * (1j * k.reshape(shape)), n=n, axis=axis) * stretch


class Oracle3D:
    """State a[6][nz][ny][nx] (xi_x, xi_y, xi_z, v_x, v_y, v_z), non-dimensional."""

    def __init__(self, p: Params3D, model: np.ndarray, c3d=None, flow=None, adjoint=False):
        self.p = p
        nx, ny, nz = p.nx, p.ny, p.nz
        self.adjoint = bool(adjoint)
        dimc, dimrho, diml = 1e6, 1e-3, RSUN
        m = np.asarray(model, dtype=np.float64)[:nz]
        self.z = m[:, 0].copy()
        self.x = np.arange(nx) / nx
        self.y = np.arange(ny) / ny
        self.stretchx = RSUN / p.xlength
        self.stretchy = RSUN / p.ylength
        self.stretch = 1.0 / _dbyd2(self.z, 1)[:, 0]                      # initialize.f90:466-470
        self.deltat = p.timestep * dimc / diml                           # initialize.f90:474
        c = (m[:, 1] / dimc)[:, None, None] * np.ones((1, ny, nx))
        if c3d is not None:
            c = np.asarray(c3d, dtype=np.float64).reshape(nz, ny, nx) / dimc
        self.c2 = c * c
        rho0 = m[:, 2] / dimrho
        p0 = m[:, 3] / (dimrho * dimc ** 2)
        self.rho0 = rho0[:, None, None] * np.ones((1, ny, nx))
        g = m[:, 4] * diml / dimc ** 2
        # e_rad, o_rad: the last row with the smallest distance (physics.f90:44-60)
        t2 = 1.0 + p.excitdep / RSUN
        t, self.e_rad = 1.0, 0
        for k in range(nz):
            if abs(self.z[k] - t2) <= t:
                t, self.e_rad = abs(self.z[k] - t2), k
        tr = 1.0 + p.obsheight / RSUN
        t, self.o_rad = 1.0, 0
        for k in range(nz):
            if abs(self.z[k] - tr) <= t:
                t, self.o_rad = abs(self.z[k] - tr), k
        # source depth profile (physics.f90:105), -10.0**(7.0)*4.0 in REAL
        self.source_dep = np.exp(-f32(4.0e7) * (self.z - self.z[self.e_rad]) ** 2)
        self.rhoinv = 1.0 / self.rho0
        self.c2rho0 = self.c2 * self.rho0
        gr0 = self.ddz(self.rho0, 1)                                      # :207
        gp0 = self.ddz(p0[:, None, None] * np.ones((1, ny, nx)), 1)       # :208
        # z-PML (physics.f90:250-296)
        nb, nt = p.npmlbot, p.npmltop
        cs = np.sqrt(self.c2)
        a0 = PI * f32(0.005) * diml / dimc
        az = np.zeros((nz, ny, nx))
        bz = np.zeros((nz, ny, nx))
        z = self.z
        for k in range(nz):
            alphaz, dampz, inpml = a0, np.zeros((ny, nx)), False
            if k < nb:
                alphaz = a0 * (z[k] - z[0]) / (z[nb - 1] - z[0])
                rr = (z[nb - 1] - z[k]) / (z[nb - 1] - z[0])
                dampz = 18.0 * cs[0] / (z[nb - 1] - z[0]) * (rr * rr)
                inpml = True
            if k > nz - nt - 1:
                alphaz = a0 * (z[nz - 1] - z[k]) / (z[nz - 1] - z[nz - nt])
                rr = (z[nz - nt] - z[k]) / (z[nz - 1] - z[nz - nt])
                dampz = 9.0 * cs[nz - 1] / (z[nz - 1] - z[nz - nt]) * (rr * rr)
                inpml = True
            if inpml:
                fd = alphaz / (dampz + alphaz)
                gp0[k] *= fd
                gr0[k] *= fd
                g[k] *= fd[0, 0]
            bz[k] = -alphaz - dampz
            az[k] = -dampz
        self.gp0, self.gr0, self.g = gp0, gr0, g
        self.pml_rows = [k for k in range(nz) if k < nb or k >= nz - nt]
        self.az = az[self.pml_rows]
        self.bz = bz[self.pml_rows]
        # flow (cm/s -> non-dimensional), reversed for the adjoint (driver.f90:144-147)
        sgn = -1.0 if self.adjoint else 1.0
        if flow is None:
            self.v0 = np.zeros((3, nz, ny, nx))
        else:
            self.v0 = sgn * np.asarray(flow, dtype=np.float64).reshape(3, nz, ny, nx) / dimc
        # RK coefficients (step.f90:24-41)
        o2, o3 = 0.5, f32(0.1665579725791151184)
        o4, o5 = f32(0.03950410250586729987), f32(0.007810706393337838236)
        b = [0.0] * 5
        b[4] = o2
        b[3] = o3 / b[4]
        b[2] = o4 / (b[4] * b[3])
        b[1] = o5 / (b[4] * b[3] * b[2])
        self.betas = [v * self.deltat for v in b]
        self.a = np.zeros((6, nz, ny, nx))
        self.scr = np.zeros((6, nz, ny, nx))
        npml = len(self.pml_rows)
        self.pmlvars = np.zeros((2, npml, ny, nx))
        self.scrpml = np.zeros((2, npml, ny, nx))
        self.vr = np.zeros((ny, nx))
        self.fwd = np.zeros(0)

    # ---- derivatives (ddxyz) ----
    def ddx(self, f):
        return _deriv_fft(f, -1, self.stretchx)

    def ddy(self, f):
        return _deriv_fft(f, -2, self.stretchy)

    def ddz(self, f, ibc):
        nz = f.shape[0]
        d = _dbyd2(f.reshape(nz, -1), ibc).reshape(f.shape)
        return d * self.stretch.reshape((nz,) + (1,) * (f.ndim - 1))

    # ---- sources ----
    def set_forward_source(self, xsrc_mm: float, ysrc_mm, maxtime: int):
        """ysrc_mm None: the reference's vr, the x profile repeated along y
        (DETERMINE_STATUS, driver.f90:536-553, `vr(:,j,1)` for every j: a line
        source); a number: the same profile in the horizontal distance to
        (xsrc, ysrc) (a point source)."""
        p = self.p
        e8 = f32(1e-8)
        xd = (self.x - 0.5) * p.xlength * e8 - xsrc_mm
        dx = (self.x[1] - self.x[0]) * p.xlength * e8
        if ysrc_mm is None:
            rd = np.broadcast_to(xd[None, :], (p.ny, p.nx))
            imin = None
        else:
            yd = (self.y - 0.5) * p.ylength * e8 - ysrc_mm
            rd = np.sqrt(yd[:, None] ** 2 + xd[None, :] ** 2)
        k43 = f32(0.43) * PI
        with np.errstate(invalid="ignore", divide="ignore"):
            sinc = np.sin(k43 * rd) / (k43 * rd)
        if ysrc_mm is None:
            sinc[:, int(np.argmin(np.abs(xd)))] = 1.0
        else:
            sinc.flat[int(np.argmin(rd))] = 1.0
        self.vr = np.exp(-(rd * rd) / (400.0 * (dx * dx))) * sinc
        self.fwd = forward_source_series(p.timestep, p.nupeak, p.nuwidth, maxtime + 8)

    def set_adjoint_source(self, vr: np.ndarray, maxtime: int):
        """vr [nt][ny][nx] slices at the forcing cadence, zero-padded to
        forcing_length + 8 (driver.f90:584-627)."""
        p = self.p
        cfs = max(int(math.floor(p.cadforcing / p.timestep)), 1)
        fl = maxtime // cfs + 1
        ns = int(math.floor(maxtime * p.timestep / p.cadforcing)) + 2
        self.num_steps = 2 * (ns // 2) + 2
        self.limit = fl * cfs
        vr = np.asarray(vr, dtype=np.float64)
        self.vra = np.zeros((fl + 8, p.ny, p.nx))
        n = min(vr.shape[0], fl)
        self.vra[:n] = vr[:n]

    def lagrange(self, time: int) -> np.ndarray:
        """LAGRANGE_INTERP (physics.f90:678-737): 7-point interpolation in time
        of the adjoint source, nodes (time1-4 .. time1+2) * cadforcing."""
        cad = self.p.cadforcing
        time1 = int(math.floor(time * self.p.timestep / cad)) + 1
        xt = time * self.p.timestep
        xs = [(time1 - 4 + q) * cad for q in range(7)]
        f = None
        for q in range(7):
            ok = True
            if q <= 3:
                ok = xs[q] >= 0
            if q == 5:
                ok = time1 < self.num_steps
            if q == 6:
                ok = time1 < self.num_steps - 1
            lc = np.zeros((self.p.ny, self.p.nx))
            if ok:
                den = None
                for r in range(7):
                    if r != q:
                        den = (xs[q] - xs[r]) if den is None else den * (xs[q] - xs[r])
                sl = time1 - 3 + q
                if 1 <= sl <= self.vra.shape[0]:
                    lc = self.vra[sl - 1] / den
            t = lc
            for r in range(7):
                if r != q:
                    t = t * (xt - xs[r])
            f = t if f is None else f + t
        return f

    def forcing(self, time: int) -> np.ndarray:
        if self.adjoint:
            if time > self.limit:
                return np.zeros((self.p.ny, self.p.nx))
            return self.lagrange(time)
        fs = self.fwd[time] if 0 <= time < self.fwd.size else 0.0
        return self.vr * fs

    # ---- right-hand side (displacement.f90:395-475) ----
    def rhs(self, time: int, T: np.ndarray, psi: np.ndarray):
        nz = self.p.nz
        rows = self.pml_rows
        dxixdx = self.ddx(T[0])
        dxiydy = self.ddy(T[1])
        dxizdz = self.ddz(T[2], 1)
        scrpml = np.zeros_like(psi)
        scrpml[0] = self.az * dxizdz[rows] + self.bz * psi[0]            # :403-409
        dxizdz[rows] += psi[0]
        div = dxixdx + dxiydy + dxizdz
        p = -self.c2rho0 * div - T[2] * self.gp0                          # :414
        rho = -self.rho0 * div - T[2] * self.gr0                          # :415
        gp = [self.ddx(p), self.ddy(p), self.ddz(p, 0)]                   # :420
        scrpml[1] = self.az * gp[2][rows] + self.bz * psi[1]              # :425-431
        gp[2][rows] += psi[1]
        scr = np.zeros_like(T)
        v0x, v0y, v0z = self.v0
        for i in range(3):                                               # :436-452
            v = T[3 + i]
            tx, ty, tz = self.ddx(v), self.ddy(v), self.ddz(v, 1)
            scr[3 + i] = -2.0 * (v0x * tx + v0y * ty + v0z * tz) - self.rhoinv * gp[i]
        scr[5] -= self.rhoinv * rho * self.g[:, None, None]               # :458
        f = self.forcing(time)
        for k in range(max(self.e_rad - 4, 0), min(self.e_rad + 4, nz - 1) + 1):
            scr[5, k] += f * self.source_dep[k]                           # :460-462
        scr[0:3] = T[3:6]                                                 # :466-468
        scr[:, 0] = 0.0                                                   # :470-473
        scr[:, nz - 1] = 0.0
        return scr, scrpml

    def filter(self, a: np.ndarray) -> None:
        """FILTER_VARS_3D_HORIZ_PML (physics.f90:836-882), in place."""
        p = self.p
        c = [0.75390625, 0.41015625, -0.234375, 0.087890625, -0.01953125, 0.001953125]
        k0, k1 = p.npmlbot, p.nz - p.npmltop
        t = c[0] * a[:, k0:k1]
        for j in range(1, 6):
            t = t + 0.5 * c[j] * (a[:, k0 - j:k1 - j] + a[:, k0 + j:k1 + j])
        a[:, k0:k1] = t

    def step(self, time: int) -> None:
        """STEP (step.f90:43-119)."""
        for s in range(5):
            if s == 0:
                T, psi = self.a, self.pmlvars
            else:
                T = self.a + self.betas[s] * self.scr
                psi = self.pmlvars + self.betas[s] * self.scrpml
            self.scr, self.scrpml = self.rhs(time, T, psi)
        self.a = self.a + self.deltat * self.scr
        self.pmlvars = self.pmlvars + self.deltat * self.scrpml
        if (time + 1) % 2 == 0:
            self.filter(self.a)

    def run(self, maxtime: int) -> np.ndarray:
        """Time indices 0..maxtime; returns vz at o_rad every `steps` steps."""
        out = []
        for t in range(maxtime + 1):
            self.step(t)
            if t % self.p.steps == 0:
                out.append(self.a[5, self.o_rad].copy())
        return np.array(out)


def kernels3d(p: Params3D, model: np.ndarray, fwd: np.ndarray, adj: np.ndarray, c3d=None) -> dict:
    """PRODUCE_KERNELS with its y terms (kernels.f90:13-612 with test_in_2d
    false, the 3D branches of :289-355, :498-522, :563, :579) on in-memory
    snapshots.  fwd [nt][9][nz][ny][nx]: xi_x, xi_y, xi_z, v_x, v_y, v_z,
    acc_x, acc_y, acc_z; adj [nt][3][nz][ny][nx]: v_x, v_y, v_z.  Forward
    snapshot k meets adjoint snapshot nt-1-k.  The background is kernel_back's
    (initialize.f90:937-979): rho0 = rho / 10.0**(-3.0), c2 = (c / 10.0**6.)**2.0,
    with c from c3d when given (model_c, kernels.f90:35-42).  ddzkern is dbyd2
    (ibc 1) times stretch followed by FILTER_VAR_Z (:648-661).  The psi kernel
    is the y component of CURL_KERN (:815-836) of (K_vx, 0, K_vz) before their
    unit scaling, ddzkern with bcz = 0 minus ddxkern, in the units of
    :511-513 (UNKNOWN_FACTOR 1)."""
    nx, ny, nz = p.nx, p.ny, p.nz
    m = np.asarray(model, dtype=np.float64)[:nz]
    dimc, dimrho, diml = f32(1e6), f32(1e-3), RSUN
    z = m[:, 0]
    stretch = 1.0 / _dbyd2(z, 1)[:, 0]
    c = (m[:, 1][:, None, None] * np.ones((1, ny, nx))) if c3d is None else np.asarray(c3d).reshape(nz, ny, nx)
    c2 = (c / dimc) ** 2.0
    rho0 = (m[:, 2] / dimrho)[:, None, None]
    sx, sy = RSUN / p.xlength, RSUN / p.ylength

    def ddzk(f, ibc=1):
        d = _dbyd2(f.reshape(nz, -1), ibc) * stretch[:, None]
        return _filter_var_z(d).reshape(f.shape)

    nt = fwd.shape[0]
    H = np.zeros((nz, ny, nx))
    KC = np.zeros_like(H)
    KV = np.zeros((3, nz, ny, nx))
    for k in range(nt):
        f = fwd[k]
        a = adj[nt - 1 - k]
        H = H + f[8] * f[8] + f[6] * f[6] + f[7] * f[7]                  # :293-294
        divf = _deriv_fft(f[0], -1, sx)                                    # compute_div :843-861
        if ny > 4:
            divf = divf + _deriv_fft(f[1], -2, sy)
        divf = divf + ddzk(f[2])
        diva = _deriv_fft(a[0], -1, sx)
        if ny > 4:
            diva = diva + _deriv_fft(a[1], -2, sy)
        diva = diva + ddzk(a[2])
        KC = KC + divf * diva                                              # :303-306
        d = [[_deriv_fft(f[3 + j], -1, sx) for j in range(3)],            # compute_del :866-887
             [_deriv_fft(f[3 + j], -2, sy) for j in range(3)],
             [ddzk(f[3 + j]) for j in range(3)]]
        for i in range(3):                                                 # :338-355
            KV[i] = d[i][0] * a[0] + d[i][1] * a[1] + d[i][2] * a[2] + KV[i]
    steps = int(math.floor(p.outputcad / p.timestep))
    e8, e5 = f32(1e-8), f32(1e-5)
    dimen = (diml * e8) ** (-3.0) * (dimc * e5) ** (-1.0)                  # :503
    out = {"hess": H}
    curl_y = ddzk(KV[0], 0) - _deriv_fft(KV[2], -1, sx)                    # CURL_KERN :832-834
    out["kpsi"] = 2.0 * curl_y * dimen * 1.0 * (rho0 * c2 ** 0.5) * steps * p.timestep * dimen   # :511-513
    for i, name in enumerate(("kvx", "kvy", "kvz")):                       # :520-522
        out[name] = -2.0 * KV[i] * rho0 * steps * p.timestep * dimen
    dimen_c = (diml * e8) ** (-3.0)                                        # :563
    out["kc"] = -KC * rho0 * c2 * steps * p.timestep * dimen_c * 2.0       # :579
    return out
