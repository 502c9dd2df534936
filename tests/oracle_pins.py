"""Checks that pin the oracle's RHS (orc_rhs) and PRODUCE_KERNELS
(orc_kernels) to what the mathematics fixes, independently of the oracle's
own formulas.  Each check raises AssertionError when it fails, so that
tests/test_oracle_physics.py runs them on the committed oracle and
tests/test_oracle_mutants.py runs them on copies with one planted mistake.

RHS (MP_QUIET_PML_2D_DISPL, physics2d.f90:943-1094):
* force_free: a uniform xi_z in a hydrostatic background with constant g
  (the polytrope) is a rigid translation: div xi = 0, p' = -xi_z dp0/dz,
  rho' = -xi_z drho0/dz, and RHS v_z = (xi_z/rho0) (d2p0/dz2 + g drho0/dz)
  vanishes because dp0/dz = -rho0 g.
* horizontal_mode: xi_x = sin(2 pi m x) alone: RHS v_x = c^2 d2xi_x/dx2
  (c^2 rho0 does not depend on x) and RHS v_z = (1 - Gamma) g div xi, since
  d(c^2 rho0)/dz = Gamma dp0/dz = -Gamma rho0 g for constant Gamma; on the
  top row the closure d xi_z/dz = -d xi_x/dx (:954) makes div xi = 0, so
  RHS v_x = 0 there.
* flow_increment: RHS(u; v0) - RHS(u; 0) = -2 (v0x d/dx + v0z d/dz) v for
  each velocity component (:1065-1079), against closed-form derivatives of
  separable velocity fields.
* rho0_symmetry: with v = 0 and no flow, rho0 times the wave operator
  xi -> RHS v is symmetric: <rho0 L xi, eta> = <xi, rho0 L eta> for fields
  away from the boundaries and layers (the self-adjointness of the adiabatic
  wave operator in hydrostatic balance; spectral d/dx is antisymmetric and
  the compact d/dz is antisymmetric away from its boundary closures).
* pml_auxiliaries: the auxiliary equations of the PMLs, RHS psi =
  a * D + b * psi, take for D the x-derivative after its own auxiliary
  correction (:970-986) and the z-derivative before it (:957-967).

Kernels (PRODUCE_KERNELS, kernels.f90:13-612):
* kernels_numpy: a numpy restatement on numpy's FFT and the reference's own
  dbyd2.f / mtridag.f compiled from source (the oracle's dbyd2 where that
  build is absent; the two are bit-identical), on random snapshots.
* kernels_chain_rule: K_psi is the gradient with respect to psi of what K_v
  is the gradient of with respect to v0 = curl(rho0 c psi y)/rho0:
  sum K_psi dpsi = -(diml/1e8) dimen sum (K_vx dv0x + K_vz dv0z) for a smooth
  dpsi (summation by parts), with dv0 from the oracle's flow set-up (pinned
  by a closed form in test_oracle.py).
* kernels_finite_differences: the reference's adjoint recipe (adjoint source
  = time-reversed dL/dvz as a force density, adjoint run with reversed flow,
  kernels) against central differences of a linear functional
  L = sum W vz of the traces, for a sound-speed direction and a flow
  direction: the continuous-adjoint identities fix the scalings
  (dL/d(rho0 c^2) = -dx dz dT sum div xi div v_a,
   dL/dv0x = -2 rho0 dx dz dT sum d_x v_j v_a,j; dT = steps * deltat).
"""
from __future__ import annotations

import functools

import numpy as np

import oracle
from sparc_flows_amd import background as B
from sparc_flows_amd.params import SparcParams, f32

RSUN = 6.95989467700e10


def _model(nx, nz, npml, nph, **kw):
    bg = B.polytrope(nz)
    p = SparcParams(nx=nx, nz=nz, xlength=200e8, timestep=B.stable_timestep(bg), npmlbot=npml,
                    npmltop=npml, npmlhor=nph, **kw)
    return p, bg


# ------------------------------------------------------------------ RHS ----
def force_free(tol=1e-2):
    nx, nz, npml = 32, 150, 12
    p, bg = _model(nx, nz, npml, 5)
    o = oracle.Oracle(p, bg)
    rho = o.array("rho0").reshape(nz, nx)
    g = o.array("g")
    gr = o.array("gradrho0_z").reshape(nz, nx)
    u = np.zeros((4, nz, nx))
    u[1] = 1.0
    s, _, _ = o.rhs(3, u)
    inner = slice(npml + 6, nz - npml - 6)
    # what a sign error in rho' g or in xi_z dp0/dz leaves: 2 g |drho0/dz| / rho0
    scale = np.abs(2.0 * g[:, None] * gr / rho)[inner].max(axis=1)
    ratio = np.abs(s[3][inner]).max(axis=1) / scale
    assert ratio.max() < tol, f"uniform xi_z is not force free: {ratio.max():.3e}"
    assert np.abs(s[2]).max() == 0.0 and np.abs(s[0]).max() == 0.0 and np.abs(s[1]).max() == 0.0


def horizontal_mode(tol_x=1e-6, tol_z=2e-5):
    nx, nz, npml = 32, 150, 12
    p, bg = _model(nx, nz, npml, 5)
    o = oracle.Oracle(p, bg)
    m = 3
    x = np.arange(nx) / nx
    k = 2 * np.pi * m * o.info("stretchx")
    u = np.zeros((4, nz, nx))
    u[0] = np.sin(2 * np.pi * m * x)[None, :]
    s, _, _ = o.rhs(3, u)
    c2 = o.array("c2").reshape(nz, nx)
    g = o.array("g")
    gam = bg[:, 5][:, None]
    ex_vx = -c2 * k * k * np.sin(2 * np.pi * m * x)[None, :]
    # rows 1 .. below the top layer (the top closure reaches ~10 rows down);
    # columns 0 and nx-1 are zeroed by the horizontal PML (:1089-1092)
    rows = slice(1, nz - npml - 6)
    err_x = np.abs(s[2] - ex_vx)[rows, 1:-1].max() / np.abs(ex_vx).max()
    assert err_x < tol_x, f"RHS v_x != c^2 d2 xi_x/dx2: {err_x:.3e}"
    inner = slice(npml + 6, nz - npml - 6)
    ex_vz = (1.0 - gam) * g[:, None] * k * np.cos(2 * np.pi * m * x)[None, :]
    err_z = np.abs(s[3] - ex_vz)[inner, 1:-1].max() / np.abs(ex_vz[inner]).max()
    assert err_z < tol_z, f"RHS v_z != (1 - Gamma) g div xi: {err_z:.3e}"
    top = np.abs(s[2][nz - 1]).max() / np.abs(ex_vx).max()
    assert top < 1e-10, f"top row: div xi != 0 under the closure (:954): {top:.3e}"


def flow_increment(tol=1e-5):
    nx, nz, npml = 32, 150, 12
    p, bg = _model(nx, nz, npml, 5)
    o = oracle.Oracle(p, bg)
    z = o.array("z")
    x = np.arange(nx) / nx
    sx = o.info("stretchx")
    zc, w = 0.5 * (z[0] + z[-1]), 0.08 * (z[-1] - z[0])
    G1 = np.exp(-((z - zc) / w) ** 2)
    G2 = np.exp(-((z - zc - 0.3 * w) / (1.3 * w)) ** 2)
    dG1 = -2.0 * (z - zc) / w ** 2 * G1
    dG2 = -2.0 * (z - zc - 0.3 * w) / (1.3 * w) ** 2 * G2
    m1, m2 = 2, 3
    vx = np.sin(2 * np.pi * m1 * x)[None, :] * G1[:, None]
    vz = np.cos(2 * np.pi * m2 * x)[None, :] * G2[:, None]
    dxvx = 2 * np.pi * m1 * sx * np.cos(2 * np.pi * m1 * x)[None, :] * G1[:, None]
    dzvx = np.sin(2 * np.pi * m1 * x)[None, :] * dG1[:, None]
    dxvz = -2 * np.pi * m2 * sx * np.sin(2 * np.pi * m2 * x)[None, :] * G2[:, None]
    dzvz = np.cos(2 * np.pi * m2 * x)[None, :] * dG2[:, None]
    U = 150.0 * (1.0 + 0.5 * np.cos(2 * np.pi * x)[None, :]) * np.ones((nz, 1))     # m/s
    W = -90.0 * (1.0 + 0.3 * np.sin(4 * np.pi * x)[None, :]) * np.ones((nz, 1))
    u = np.zeros((4, nz, nx))
    u[2], u[3] = vx, vz
    s0, sp0, _ = o.rhs(3, u)
    o.set_flow(U, W)
    s1, sp1, _ = o.rhs(3, u)
    Un, Wn = U * 1e2 / 1e6, W * 1e2 / 1e6
    ex_x = -2.0 * (Un * dxvx + Wn * dzvx)
    ex_z = -2.0 * (Un * dxvz + Wn * dzvz)
    for ex in (ex_x, ex_z):
        ex[0] = 0.0
        ex[:, 0] = ex[:, -1] = 0.0
    ex_scale = max(np.abs(ex_x).max(), np.abs(ex_z).max())
    ex_x_err = np.abs((s1[2] - s0[2]) - ex_x).max() / ex_scale
    ex_z_err = np.abs((s1[3] - s0[3]) - ex_z).max() / ex_scale
    assert ex_x_err < tol and ex_z_err < tol, f"flow increment: {ex_x_err:.3e} {ex_z_err:.3e}"
    assert np.array_equal(s1[:2], s0[:2]) and np.array_equal(sp1, sp0)


def rho0_symmetry(tol=1e-8):
    nx, nz = 64, 150
    p, bg = _model(nx, nz, 12, 5)
    o = oracle.Oracle(p, bg)
    rho = o.array("rho0").reshape(nz, nx)
    kk, ii = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")

    def bump(k0, i0, w):
        return np.exp(-((kk - k0) / w) ** 2 - ((ii - i0) / w) ** 2)

    xi = np.zeros((4, nz, nx))
    xi[0], xi[1] = bump(70, 30, 5), bump(80, 34, 6)
    eta = np.zeros((4, nz, nx))
    eta[0], eta[1] = 0.7 * bump(75, 36, 6), -bump(72, 28, 5)

    def L(u):
        return o.rhs(3, u)[0][2:4]

    a = float((rho * L(xi) * eta[:2]).sum())
    b = float((rho * L(eta) * xi[:2]).sum())
    assert abs(a - b) < tol * max(abs(a), abs(b)), f"rho0 L not symmetric: {a:.9e} {b:.9e}"


def pml_auxiliaries(tol=1e-12):
    nx, nz, npml, nph = 32, 60, 8, 6
    p, bg = _model(nx, nz, npml, nph)
    o = oracle.Oracle(p, bg)
    rng = np.random.default_rng(4)
    x = np.arange(nx) / nx
    z = o.array("z")
    u = np.zeros((4, nz, nx))
    u[0] = np.sin(2 * np.pi * 2 * x)[None, :] * (1.0 + 0.2 * z[:, None])
    u[1] = np.cos(2 * np.pi * x)[None, :] * (z[:, None] - z.mean()) * 50.0
    nzp = 2 * npml
    psipml = rng.standard_normal((2, nzp, nx)) * 5.0
    psipmlx = rng.standard_normal((2, nz, 2 * nph)) * 5.0
    s, sp, spx = o.rhs(3, u, psipml, psipmlx)
    ax = o.array("axpml").reshape(nz, 2 * nph)
    bx = o.array("bxpml").reshape(nz, 2 * nph)
    az = o.array("az").reshape(nz, nx)
    bz = o.array("bzpml").reshape(nz, nx)
    # pinned operators (test_oracle.py): spectral d/dx, dbyd2 d/dz
    dxix = o.ddx(u[0])
    cols = np.r_[0:nph, nx - nph:nx]
    D = dxix[:, cols] + psipmlx[0]                     # corrected before its auxiliary
    ex = ax * D + bx * psipmlx[0]
    ex_err = np.abs(spx[0] - ex).max() / np.abs(ex).max()
    assert ex_err < tol, f"x-PML auxiliary: {ex_err:.3e}"
    preset = np.zeros((nz, nx))
    preset[nz - 1] = -dxix[nz - 1] * o.array("unstretch")[nz - 1]
    dz = o.ddz(u[1], 3, preset)
    rows = np.r_[0:npml, nz - npml:nz]
    ez = az[rows] * dz[rows] + bz[rows] * psipml[0]     # raw derivative, then corrected
    ez_err = np.abs(sp[0] - ez).max() / np.abs(ez).max()
    assert ez_err < tol, f"z-PML auxiliary: {ez_err:.3e}"


# -------------------------------------------------------------- kernels ----
def _ddx_np(f, stretchx):
    nx = f.shape[-1]
    k = np.arange(nx // 2 + 1, dtype=np.float64)
    mul = 1j * 2 * np.pi * k
    mul[-1] = 0.0                     # the c2r transform drops the imaginary Nyquist part
    return np.fft.irfft(np.fft.rfft(f, axis=-1) * mul, n=nx, axis=-1) * stretchx


def _dbyd2_ref(f, ibc):
    return oracle.ref_dbyd2(f, ibc) if oracle.ref_available() else oracle.dbyd2(f, ibc)


def _filter_var_z_np(v):
    """FILTER_VAR_Z (kernels.f90:891-928) by explicit convolution."""
    nz = v.shape[0]
    c = np.array([0.001953125, -0.01953125, 0.087890625, -0.234375, 0.41015625])
    w = np.concatenate([0.5 * c, [0.75390625], 0.5 * c[::-1]])
    out = v.copy()
    for k in range(1, nz - 1):
        if 5 <= k <= nz - 6:
            out[k] = sum(w[j] * v[k - 5 + j] for j in range(11))
        else:
            out[k] = 0.5 * (v[k - 1] + v[k + 1])
    return out


def kernels_numpy(p, bg, fwd, adj, c2d=None):
    """PRODUCE_KERNELS restated in numpy (kernels.f90:293-355, 498-579)."""
    nt = fwd.shape[0]
    dimc, dimrho = f32(1e6), f32(1e-3)
    rho0 = np.repeat((bg[:, 2] / dimrho)[:, None], p.nx, 1)
    c = (np.repeat(bg[:, 1][:, None], p.nx, 1) if c2d is None else c2d) / dimc
    c2 = c ** 2
    stretch = 1.0 / _dbyd2_ref(bg[:, 0].copy(), 1)[:, 0]
    sx = RSUN / p.xlength
    ddx = functools.partial(_ddx_np, stretchx=sx)

    def ddz(f, ibc):
        return _filter_var_z_np(_dbyd2_ref(f, ibc) * stretch[:, None])

    H = np.zeros((p.nz, p.nx))
    Kc, Kx, Kz = np.zeros_like(H), np.zeros_like(H), np.zeros_like(H)
    for k in range(nt):
        f, a = fwd[k], adj[nt - 1 - k]
        H += f[5] ** 2 + f[4] ** 2
        Kc += (ddx(f[0]) + ddz(f[1], 1)) * (ddx(a[0]) + ddz(a[1], 1))
        Kx += ddx(f[2]) * a[0] + ddx(f[3]) * a[1]
        Kz += ddz(f[2], 1) * a[0] + ddz(f[3], 1) * a[1]
    e8, e5 = f32(1e-8), f32(1e-5)
    dimen = 1.0 / (RSUN * e8 * p.xlength * e8 * dimc * e5)
    dimen_c = 1.0 / (RSUN * e8 * p.xlength * e8)
    T = p.steps * p.timestep
    return {
        "kpsi": 2.0 * (ddz(Kx, 0) - ddx(Kz)) * dimen * rho0 * c * T * dimen,
        "kvx": -2.0 * Kx * rho0 * T * dimen,
        "kvz": -2.0 * Kz * rho0 * T * dimen,
        "kc": -Kc * rho0 * c2 * T * dimen_c * 2.0,
        "hess": H,
    }


def kernels_restatement(tol=1e-12):
    nx, nz, nt = 16, 24, 3
    p, bg = _model(nx, nz, 4, 3)
    rng = np.random.default_rng(21)
    fwd = rng.standard_normal((nt, 6, nz, nx))
    adj = rng.standard_normal((nt, 4, nz, nx))
    c2d = B.sound_speed_bump(nx, p.xlength, bg, amp=0.2, sigx_mm=40.0, sigz_mm=5.0, z0_mm=-8.0)
    got = oracle.kernels(p, bg, fwd, adj, c2d=c2d)
    ref = kernels_numpy(p, bg, fwd, adj, c2d=c2d)
    for key in ("kc", "kpsi", "kvx", "kvz", "hess"):
        err = np.abs(got[key] - ref[key]).max() / np.abs(ref[key]).max()
        assert err < tol, f"{key}: oracle vs numpy restatement {err:.3e}"


def kernels_chain_rule(tol=1e-3):
    nx, nz, nt = 64, 80, 3
    p, bg = _model(nx, nz, 6, 4)
    rng = np.random.default_rng(3)
    fwd = rng.standard_normal((nt, 6, nz, nx))
    adj = rng.standard_normal((nt, 4, nz, nx))
    c2d = B.sound_speed_bump(nx, p.xlength, bg)
    K = oracle.kernels(p, bg, fwd, adj, c2d=c2d)
    e8 = f32(1e-8)
    dimen = 1.0 / (RSUN * e8 * p.xlength * e8 * f32(1e6) * f32(1e-5))
    kk, ii = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    for sh in range(3):
        d = np.exp(-((kk - nz / 2 - 3 * sh) / 6.0) ** 2 - ((ii - nx / 2 + 5 * sh) / 6.0) ** 2)
        o = oracle.Oracle(p, bg, c2d=c2d, psi=d)
        dvx = o.array("v0x").reshape(nz, nx)
        dvz = o.array("v0z").reshape(nz, nx)
        lhs = float((K["kpsi"] * d).sum())
        rhs = -(RSUN / 1e8) * dimen * float((K["kvx"] * dvx + K["kvz"] * dvz).sum())
        assert abs(lhs - rhs) < tol * abs(lhs), f"chain rule K_psi <-> K_v: {lhs:.6e} {rhs:.6e}"


@functools.lru_cache(maxsize=1)
def born_data():
    """Forward/adjoint snapshots of one source and central differences of
    L = sum W vz along a sound-speed and a flow direction (committed oracle).
    64 x 80 polytrope, 10 s cadence (dt divides it), 800 steps, receivers at -3 Mm."""
    nx, nz, maxtime = 64, 80, 800
    bg = B.polytrope(nz)
    dt = B.cadence_aligned_timestep(B.stable_timestep(bg), 10.0)
    p = SparcParams(nx=nx, nz=nz, xlength=200e8, timestep=dt, npmlbot=10, npmltop=10, npmlhor=5,
                    outputcad=10.0, cadforcing=10.0, obsheight=-3e8)
    x = B.grid_x_mm(nx, p.xlength)
    c0 = np.repeat(bg[:, 1][:, None], nx, 1)
    zero = np.zeros((nz, nx))

    def traces(c2d, flow=None, snaps=False):
        o = oracle.Oracle(p, bg, c2d=c2d)
        if flow is not None:
            o.set_flow(*flow)
        o.set_forward_source(0.0, maxtime)
        return o.run(maxtime, snapshots=snaps)

    vz, fsn = traces(c0, snaps=True)
    nout = vz.shape[0]
    t = np.arange(nout)[:, None]
    W = (np.exp(-((x[None, :] - 20.0) / 6.0) ** 2) * np.sin(2 * np.pi * t / 37.0)
         * np.exp(-((t - 0.65 * nout) / (0.15 * nout)) ** 2))
    dT = p.steps * dt * 1e6 / RSUN
    dxn = p.xlength / RSUN / nx
    oa = oracle.Oracle(p, bg, c2d=c0, adjoint=True)
    oa.set_adjoint_source(W[::-1].copy() / (dT * dxn), maxtime)
    _, asn = oa.run(maxtime, snapshots=True)
    kk, ii = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    dirc = np.exp(-((kk - 40) / 7.0) ** 2 - ((ii - 38) / 6.0) ** 2) \
        - 0.7 * np.exp(-((kk - 50) / 5.0) ** 2 - ((ii - 30) / 5.0) ** 2)
    U = 100.0 * np.exp(-((kk - 45) / 6.0) ** 2 - ((ii - 36) / 7.0) ** 2)          # m/s
    Wz = -60.0 * np.exp(-((kk - 38) / 5.0) ** 2 - ((ii - 40) / 6.0) ** 2)
    h, hv = 1e-4, 1e-2
    L = lambda v: float((W * v).sum())  # noqa: E731
    fd_c = (L(traces(c0 * (1 + h * dirc))) - L(traces(c0 * (1 - h * dirc)))) / (2 * h)
    fd_v = (L(traces(c0, (hv * U, hv * Wz))) - L(traces(c0, (-hv * U, -hv * Wz)))) / (2 * hv)
    return dict(p=p, bg=bg, c0=c0, fsn=fsn, asn=asn, dirc=dirc, U=U, Wz=Wz, fd_c=fd_c, fd_v=fd_v,
                dT=dT, dxn=dxn, zero=zero)


def kernels_finite_differences(tol=0.1):
    d = born_data()
    p, bg, c0 = d["p"], d["bg"], d["c0"]
    K = oracle.kernels(p, bg, d["fsn"], d["asn"], c2d=c0)
    e8 = f32(1e-8)
    dimen_c = 1.0 / (RSUN * e8 * p.xlength * e8)
    dimen = 1.0 / (RSUN * e8 * p.xlength * e8 * f32(1e6) * f32(1e-5))
    dzn = np.gradient(bg[:, 0])[:, None]
    cell = d["dxn"] * dzn
    # dL/dc (cm/s) = 2 rho0 c dL/d(rho0 c^2) / dimc, and dL/dv0 (m/s) = 1e2/dimc dL/dv0_nd
    gc = K["kc"] * cell * (1e6 / RSUN) / (dimen_c * c0 / 1e6) / 1e6
    T = p.steps * p.timestep
    gvx = K["kvx"] * cell * d["dT"] / (T * dimen) * 1e-4
    gvz = K["kvz"] * cell * d["dT"] / (T * dimen) * 1e-4
    ad_c = float((gc * c0 * d["dirc"]).sum())
    ad_v = float((gvx * d["U"] + gvz * d["Wz"]).sum())
    ec = abs(ad_c - d["fd_c"]) / abs(d["fd_c"])
    ev = abs(ad_v - d["fd_v"]) / abs(d["fd_v"])
    assert ec < tol and ev < tol, f"kernels vs finite differences: c {ec:.3e}, flow {ev:.3e}"
    return ec, ev
