"""Pins of the 3D oracle (oracle/oracle3d.py).  The reference cannot produce
output for this path (its kernel-mode forward source is never set and its flow
setup is 2D only), so the restatement is pinned by what the mathematics
fixes: derivative closed forms, the x <-> y exchange symmetry of the
equations, linearity, the z filter's moments, and the zeroed edge rows."""
import numpy as np
import pytest

import oracle
from oracle.oracle3d import Oracle3D, Params3D
from sparc_flows_amd import background as B


def make(nx=16, ny=16, nz=40, flow=True, dc=True, adjoint=False, seed=1):
    bg = B.polytrope(nz)
    p = Params3D(nx=nx, ny=ny, nz=nz, timestep=B.stable_timestep(bg), npmlbot=6, npmltop=6)
    rng = np.random.default_rng(seed)
    c3d = None
    if dc:
        x = (np.arange(nx) / nx - 0.5)[None, None, :]
        y = (np.arange(ny) / ny - 0.5)[None, :, None]
        zz = np.arange(nz)[:, None, None] / nz
        c3d = bg[:, 1][:, None, None] * (1 + 0.05 * np.exp(-(x ** 2 + y ** 2) / 0.02 - (zz - 0.6) ** 2 / 0.02))
    fl = None
    if flow:
        fl = 2e4 * rng.standard_normal((3, nz, ny, nx))
    return p, bg, c3d, fl


def swap_xy(f):
    return np.swapaxes(f, -1, -2).copy()


def test_fft_derivatives_closed_forms():
    p, bg, _, _ = make(nx=32, ny=16, flow=False, dc=False)
    o = Oracle3D(p, bg)
    x = np.arange(p.nx) / p.nx
    y = np.arange(p.ny) / p.ny
    X, Y = np.meshgrid(x, y)
    f = np.sin(2 * np.pi * 3 * X) * np.cos(2 * np.pi * 2 * Y)
    assert np.allclose(o.ddx(f), 2 * np.pi * 3 * np.cos(2 * np.pi * 3 * X) * np.cos(2 * np.pi * 2 * Y) * o.stretchx,
                       rtol=0, atol=1e-12 * o.stretchx * 20)
    assert np.allclose(o.ddy(f), -2 * np.pi * 2 * np.sin(2 * np.pi * 3 * X) * np.sin(2 * np.pi * 2 * Y) * o.stretchy,
                       rtol=0, atol=1e-12 * o.stretchy * 20)
    # Nyquist modes and constants differentiate to zero
    assert np.abs(o.ddx(np.cos(np.pi * p.nx * X))).max() < 1e-10
    assert np.abs(o.ddy(np.cos(np.pi * p.ny * Y) + 3.0)).max() < 1e-10


def test_rhs_exchange_symmetry():
    """Swapping x and y (fields, flow components, sound speed) swaps the RHS:
    catches an x/y mix-up in the derivatives, the flow terms or the gradient."""
    p, bg, c3d, fl = make()
    rng = np.random.default_rng(5)
    T = 1e-8 * rng.standard_normal((6, p.nz, p.ny, p.nx))
    psi = 1e-9 * rng.standard_normal((2, 12, p.ny, p.nx))
    c3s = swap_xy(c3d)
    c3d_sym = 0.5 * (c3d + c3s)
    o1 = Oracle3D(p, bg, c3d=c3d_sym, flow=fl)
    flow2 = np.stack([swap_xy(fl[1]), swap_xy(fl[0]), swap_xy(fl[2])])
    o2 = Oracle3D(p, bg, c3d=swap_xy(c3d_sym), flow=flow2)
    o1.set_forward_source(0.0, 0.0, 50)
    o2.set_forward_source(0.0, 0.0, 50)
    r1, q1 = o1.rhs(20, T, psi)
    T2 = np.stack([swap_xy(T[1]), swap_xy(T[0]), swap_xy(T[2]), swap_xy(T[4]), swap_xy(T[3]), swap_xy(T[5])])
    r2, q2 = o2.rhs(20, T2, np.stack([swap_xy(psi[0]), swap_xy(psi[1])]))
    perm = [1, 0, 2, 4, 3, 5]
    for i in range(6):
        assert np.allclose(swap_xy(r2[perm[i]]), r1[i], rtol=1e-12, atol=1e-12 * np.abs(r1).max()), i
    assert np.allclose(swap_xy(q2[0]), q1[0], rtol=1e-12, atol=1e-20)


def test_rhs_linear_without_source_and_zero_edges():
    p, bg, c3d, fl = make()
    o = Oracle3D(p, bg, c3d=c3d, flow=fl)
    rng = np.random.default_rng(7)
    T1, T2 = 1e-8 * rng.standard_normal((2, 6, p.nz, p.ny, p.nx))
    P1, P2 = 1e-9 * rng.standard_normal((2, 2, 12, p.ny, p.nx))
    r1, _ = o.rhs(0, T1, P1)
    r2, _ = o.rhs(0, T2, P2)
    r3, _ = o.rhs(0, 2.0 * T1 - 3.0 * T2, 2.0 * P1 - 3.0 * P2)
    assert np.allclose(r3, 2.0 * r1 - 3.0 * r2, rtol=1e-10, atol=1e-12 * np.abs(r3).max())
    assert np.all(r1[:, 0] == 0) and np.all(r1[:, -1] == 0)
    z, _ = o.rhs(0, np.zeros_like(T1), np.zeros_like(P1))
    assert np.all(z == 0)


def test_y_invariant_input_gives_y_invariant_rhs():
    p, bg, _, _ = make(nx=16, ny=16, flow=False, dc=False)
    rng = np.random.default_rng(3)
    nz, ny, nx = p.nz, p.ny, p.nx
    T = np.repeat(1e-8 * rng.standard_normal((6, nz, 1, nx)), ny, axis=2)
    T[1] = 0.0
    T[4] = 0.0
    fl = np.repeat(1e4 * rng.standard_normal((3, nz, 1, nx)), ny, axis=2)
    fl[1] = 0.0
    o = Oracle3D(p, bg, flow=fl)
    r, _ = o.rhs(0, T, np.zeros((2, 12, ny, nx)))
    assert np.abs(r - r[:, :, :1]).max() <= 1e-14 * np.abs(r).max()
    assert np.abs(r[1]).max() == 0 and np.abs(r[4]).max() < 1e-14 * np.abs(r).max()


def test_filter_moments():
    """FILTER_VARS_3D_HORIZ_PML keeps constants and linear profiles and removes
    the z-Nyquist mode on the filtered rows (coefficients sum to 1, alternate to 0)."""
    p, bg, _, _ = make(flow=False, dc=False)
    o = Oracle3D(p, bg)
    nz = p.nz
    k = np.arange(nz)[None, :, None, None]
    a = np.broadcast_to(3.0 + 0.25 * k + np.cos(np.pi * k), (6, nz, p.ny, p.nx)).copy()
    o.filter(a)
    rows = slice(p.npmlbot, nz - p.npmltop)
    kk = np.arange(nz)[rows][None, :, None, None]
    assert np.allclose(a[:, rows], 3.0 + 0.25 * kk, rtol=0, atol=1e-12)
    assert np.allclose(a[:, :p.npmlbot], (3.0 + 0.25 * k + np.cos(np.pi * k))[:, :p.npmlbot], atol=0)


def test_forward_source_is_radial_and_centred():
    p, bg, _, _ = make(flow=False, dc=False)
    o = Oracle3D(p, bg)
    o.set_forward_source(0.0, 0.0, 40)
    assert np.allclose(o.vr, o.vr.T)
    j, i = np.unravel_index(np.argmax(o.vr), o.vr.shape)
    assert (j, i) == (p.ny // 2, p.nx // 2) and o.vr[j, i] == 1.0
    assert np.allclose(o.forcing(10), o.vr * oracle.forward_source_series(p.timestep, p.nupeak, p.nuwidth, 48)[10])


def test_lagrange_interpolates_cubic_exactly():
    p, bg, _, _ = make(flow=False, dc=False, adjoint=True)
    o = Oracle3D(p, bg, adjoint=True)
    maxtime = 200
    cfs = int(p.cadforcing // p.timestep)
    nt = maxtime // cfs + 1
    t = np.arange(nt) * p.cadforcing
    poly = lambda s: 1e-3 * (1 + 0.1 * (s / 100) - 0.03 * (s / 100) ** 2 + 0.002 * (s / 100) ** 3)
    vr = poly(t)[:, None, None] * np.ones((1, p.ny, p.nx))
    o.set_adjoint_source(vr, maxtime)
    for time in (60, 75, 120):   # interior nodes: all seven present and inside the data
        f = o.lagrange(time)
        assert np.allclose(f, poly(time * p.timestep), rtol=1e-11)


@pytest.mark.parametrize("adjoint", [False, True])
def test_short_run_bounded(adjoint):
    p, bg, c3d, _ = make(flow=False)
    o = Oracle3D(p, bg, c3d=c3d, adjoint=adjoint)
    if adjoint:
        rng = np.random.default_rng(0)
        o.set_adjoint_source(1e-6 * rng.standard_normal((10, p.ny, p.nx)), 40)
    else:
        o.set_forward_source(0.0, 0.0, 40)
    vz = o.run(40)
    assert np.isfinite(vz).all() and np.abs(vz).max() > 0


def test_kernels3d_reduce_to_the_2d_kernels():
    """y-invariant snapshots with no y components: the 3D kernels are the 2D
    oracle's (orc_kernels, pinned against the reference's formulas) up to the
    3D units, dimen3 / dimen2 for the flow kernels and dimen3 / dimen_c for c."""
    from sparc_flows_amd.params import SparcParams
    from oracle.oracle3d import kernels3d
    nx, ny, nz, nt = 16, 8, 40, 4
    bg = B.polytrope(nz)
    dt = B.stable_timestep(bg)
    p3 = Params3D(nx=nx, ny=ny, nz=nz, timestep=dt, xlength=150e8, ylength=80e8)
    p2 = SparcParams(nx=nx, nz=nz, xlength=150e8, timestep=dt, npmlbot=6, npmltop=6, npmlhor=3)
    rng = np.random.default_rng(11)
    f2 = rng.standard_normal((nt, 6, nz, nx))     # xi_x, xi_z, v_x, v_z, acc_x, acc_z
    a2 = rng.standard_normal((nt, 4, nz, nx))     # v_x, v_z, acc_x, acc_z
    rep = lambda g: np.repeat(g[:, None, :], ny, axis=1)
    zero = np.zeros((nz, ny, nx))
    f3 = np.array([[rep(f[0]), zero, rep(f[1]), rep(f[2]), zero, rep(f[3]), rep(f[4]), zero, rep(f[5])] for f in f2])
    a3 = np.array([[rep(a[0]), zero, rep(a[1])] for a in a2])
    k2 = oracle.kernels(p2, bg, f2, a2)
    k3 = kernels3d(p3, bg, f3, a3)
    e8, e5 = oracle.oracle3d.f32(1e-8), oracle.oracle3d.f32(1e-5)
    diml, dimc = oracle.oracle3d.RSUN, oracle.oracle3d.f32(1e6)
    d3 = (diml * e8) ** -3.0 * (dimc * e5) ** -1.0
    d2 = (diml * e8 * p2.xlength * e8 * dimc * e5) ** -1.0
    dc3 = (diml * e8) ** -3.0
    dc2 = (diml * e8 * p2.xlength * e8) ** -1.0
    for j in range(ny):
        assert np.allclose(k3["hess"][:, j], k2["hess"], rtol=1e-13, atol=0)
        assert np.allclose(k3["kvx"][:, j] * d2 / d3, k2["kvx"], rtol=1e-11, atol=1e-13 * np.abs(k2["kvx"]).max())
        assert np.allclose(k3["kvz"][:, j] * d2 / d3, k2["kvz"], rtol=1e-11, atol=1e-13 * np.abs(k2["kvz"]).max())
        assert np.allclose(k3["kc"][:, j] * dc2 / dc3, k2["kc"], rtol=1e-11, atol=1e-13 * np.abs(k2["kc"]).max())
        # K_psi carries dimen twice (kernels.f90:511-513)
        assert np.allclose(k3["kpsi"][:, j] * (d2 / d3) ** 2, k2["kpsi"], rtol=1e-11,
                           atol=1e-13 * np.abs(k2["kpsi"]).max())
    assert np.all(k3["kvy"] == 0)


def test_line_source_is_the_reference_x_profile():
    """ysrc None: vr is the 2D x profile (orc_set_forward_source's) in every y row."""
    from sparc_flows_amd.params import SparcParams
    p, bg, _, _ = make(nx=32, ny=8, flow=False, dc=False)
    o = Oracle3D(p, bg)
    o.set_forward_source(7.5, None, 40)
    p2 = SparcParams(nx=32, nz=p.nz, xlength=p.xlength, timestep=p.timestep, npmlbot=6, npmltop=6, npmlhor=3)
    o2 = oracle.Oracle(p2, bg)
    o2.set_forward_source(7.5, 40)
    vr2 = o2.array("vr")[:32]
    for j in range(p.ny):
        assert np.allclose(o.vr[j], vr2, rtol=1e-14, atol=1e-16)   # numpy vs libm sin, exp
