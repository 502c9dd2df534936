"""GPU parity of the 3D solver (option 16) against the 3D oracle on the same
seeded inputs, through the C ABI (include/sparc_hip3d.h, ctypes).

Tolerances: the FFT derivatives 1e-12 relative, dbyd2 1e-11, one step 1e-11, 200-step vz
traces and states 1e-10 relative L2 (FP64; the reference's 1e-9 bar)."""
import dataclasses

import numpy as np
import pytest

from tests.cases import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from oracle.oracle3d import Oracle3D, Params3D as OParams  # noqa: E402
from sparc_flows_amd import background as B  # noqa: E402
from sparc_flows_amd.solver3d import Params3D, Sparc3D  # noqa: E402


def case(nx=32, ny=16, nz=40, flow=True, dc=True, npml=6, seed=0):
    bg = B.polytrope(nz)
    p = Params3D(nx=nx, ny=ny, nz=nz, timestep=B.stable_timestep(bg), npmlbot=npml, npmltop=npml,
                 xlength=100e8, ylength=100e8 * ny / nx)
    x = (np.arange(nx) / nx - 0.5)[None, None, :]
    y = (np.arange(ny) / ny - 0.5)[None, :, None]
    zz = (np.arange(nz) / nz)[:, None, None]
    c3d = None
    if dc:
        c3d = bg[:, 1][:, None, None] * (1 + 0.05 * np.exp(-(x ** 2 + y ** 2) / 0.03 - (zz - 0.7) ** 2 / 0.01))
    fl = None
    if flow:
        rng = np.random.default_rng(seed)
        amp = rng.uniform(0.5, 1.5, 3) * 2.4e4       # 240 m/s scale
        env = np.exp(-((x - 0.1) ** 2 + (y + 0.05) ** 2) / 0.02 - (zz - 0.75) ** 2 / 0.005)
        fl = np.stack([amp[0] * env * np.cos(2 * np.pi * y), amp[1] * env * np.sin(2 * np.pi * x),
                       amp[2] * env * 0.3]).astype(np.float64)
        fl = np.broadcast_to(fl, (3, nz, ny, nx)).copy()
    return p, bg, c3d, fl


def pair(p, bg, c3d, fl, adjoint=False):
    o = Oracle3D(OParams(**dataclasses.asdict(p)), bg, c3d=c3d, flow=fl, adjoint=adjoint)
    s = Sparc3D(p, bg, c3d=c3d, adjoint=adjoint)
    if fl is not None:
        s.set_flow(*fl)
    return o, s


# (48, 21), (20, 75), (24, 300), (18, 600): the circulant forms for an nx or
# ny that is not a power of two (every slot count of the y form; nx even)
@pytest.mark.parametrize("nx,ny", [(32, 16), (48, 21), (20, 75), (24, 300), (18, 600)])
def test_derivatives_match_oracle(nx, ny):
    p, bg, c3d, fl = case(nx=nx, ny=ny)
    o, s = pair(p, bg, c3d, fl)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((2, p.nz, p.ny, p.nx))
    ft = torch.as_tensor(f, device="cuda")
    assert rel_l2(s.ddx(ft).cpu().numpy(), o.ddx(f)) < 1e-12
    assert rel_l2(s.ddy(ft).cpu().numpy(), o.ddy(f)) < 1e-12
    for ibc in (0, 1):
        got = s.ddz(ft, ibc).cpu().numpy()
        # white noise through the Thomas sweeps: 1e-11, as the 2D operator tests
        assert rel_l2(got[0], o.ddz(f[0], ibc)) < 1e-11 and rel_l2(got[1], o.ddz(f[1], ibc)) < 1e-11


def test_background_facts_match_oracle():
    p, bg, c3d, fl = case()
    o, s = pair(p, bg, c3d, fl)
    assert s.query("e_rad") == o.e_rad and s.query("o_rad") == o.o_rad
    assert s.query("deltat") == o.deltat
    assert s.query("npml") == len(o.pml_rows)


@pytest.mark.parametrize("t", [0, 1])
def test_one_step_matches_oracle(t):
    """One STEP from a seeded state (odd t: the z filter runs)."""
    p, bg, c3d, fl = case()
    o, s = pair(p, bg, c3d, fl)
    o.set_forward_source(3.0, -2.0, 50)
    s.set_forward_source(3.0, -2.0, 50)
    rng = np.random.default_rng(2)
    a0 = 1e-8 * rng.standard_normal((6, p.nz, p.ny, p.nx))
    o.a = a0.copy()
    s.set_state(torch.as_tensor(a0, device="cuda"))
    o.step(t)
    s.step(t)
    got = s.state().cpu().numpy()
    for f in range(6):
        assert rel_l2(got[f], o.a[f]) < 1e-11, f
    assert rel_l2(s.pml().cpu().numpy(), o.pmlvars) < 1e-11


@pytest.mark.parametrize("flow,dc,nx,ny", [(False, False, 32, 16), (True, True, 32, 16), (True, True, 48, 21)])
def test_forward_run_matches_oracle(flow, dc, nx, ny):
    p, bg, c3d, fl = case(flow=flow, dc=dc, nx=nx, ny=ny)
    o, s = pair(p, bg, c3d, fl)
    maxtime = 200
    o.set_forward_source(5.0, 0.0, maxtime)
    ovz = o.run(maxtime)
    vz = s.forward(5.0, 0.0, maxtime).cpu().numpy()
    e = rel_l2(vz, ovz)
    es = rel_l2(s.state().cpu().numpy(), o.a)
    print("3D forward vz", e, "state", es)
    assert np.abs(ovz).max() > 0
    assert e < 1e-10 and es < 1e-10


def test_adjoint_run_matches_oracle():
    p, bg, c3d, fl = case()
    o, s = pair(p, bg, c3d, fl, adjoint=True)
    maxtime = 120
    nt = maxtime // int(p.cadforcing // p.timestep) + 1
    rng = np.random.default_rng(4)
    vr = 1e-6 * rng.standard_normal((nt, p.ny, p.nx))
    o.set_adjoint_source(vr, maxtime)
    ovz = o.run(maxtime)
    vz = s.adjoint_run(torch.as_tensor(vr, device="cuda"), maxtime).cpu().numpy()
    e = rel_l2(vz, ovz)
    es = rel_l2(s.state().cpu().numpy(), o.a)
    print("3D adjoint vz", e, "state", es)
    assert e < 1e-10 and es < 1e-10


def test_full_size_exchange_symmetry():
    """256 x 256 x 300: a centred source in an x <-> y symmetric medium and flow
    keeps xi_x(x, y) = xi_y(y, x), v_x(x, y) = v_y(y, x) and xi_z, v_z symmetric."""
    nx = ny = 256
    nz = 300
    bg = B.polytrope(nz)
    p = Params3D(nx=nx, ny=ny, nz=nz, timestep=B.stable_timestep(bg), npmlbot=12, npmltop=12)
    x = (np.arange(nx) / nx - 0.5)[None, None, :]
    y = (np.arange(ny) / ny - 0.5)[None, :, None]
    zz = (np.arange(nz) / nz)[:, None, None]
    env = np.exp(-(x ** 2 + y ** 2) / 0.01 - (zz - 0.85) ** 2 / 0.002)
    fl = np.stack([2e4 * env * x, 2e4 * env * y, 5e3 * env + 0 * x * y])
    fl = np.broadcast_to(fl, (3, nz, ny, nx))
    s = Sparc3D(p, bg)
    s.set_flow(*fl)
    vz = s.forward(0.0, 0.0, 12)
    a = s.state()
    assert torch.isfinite(a).all() and torch.isfinite(vz).all()
    sw = lambda t: t.transpose(-1, -2)
    scale = a.abs().max().item()
    assert scale > 0
    for i, j in ((0, 1), (3, 4), (2, 2), (5, 5)):
        d = (a[i] - sw(a[j])).abs().max().item()
        assert d <= 1e-11 * scale, (i, j, d, scale)


def test_kernels_match_oracle():
    from oracle.oracle3d import kernels3d
    p, bg, c3d, fl = case()
    s = Sparc3D(p, bg, c3d=c3d, adjoint=True)
    rng = np.random.default_rng(9)
    nt = 3
    f = rng.standard_normal((nt, 9, p.nz, p.ny, p.nx))
    a = rng.standard_normal((nt, 3, p.nz, p.ny, p.nx))
    K = s.kernels(torch.as_tensor(f, device="cuda"), torch.as_tensor(a, device="cuda"))
    ref = kernels3d(OParams(**dataclasses.asdict(p)), bg, f, a, c3d=c3d)
    for k in ("kc", "kpsi", "kvx", "kvy", "kvz", "hess"):
        e = rel_l2(K[k].cpu().numpy(), ref[k])
        print("3D kernel", k, e)
        assert e < 1e-11, k


def _oracle_gradient(p, bg, c3d, fl, xs, ys, maxtime, window_s):
    from oracle.oracle3d import kernels3d
    from sparc_flows_amd.gradient import waveform_adjoint_source
    from sparc_flows_amd.solver3d import receiver_mask_3d
    op = OParams(**dataclasses.asdict(p))
    o = Oracle3D(op, bg, c3d=c3d, flow=fl)
    o.set_forward_source(xs, ys, maxtime)
    fs, vz = [], []
    for t in range(maxtime + 1):
        o.step(t)
        if t % p.steps == 0:
            fs.append(np.concatenate([o.a, o.scr[3:6]]))
            vz.append(o.a[5, o.o_rad].copy())
    vz = np.array(vz)
    nt = vz.shape[0]
    mask = receiver_mask_3d(p, xs, ys).reshape(1, -1)
    _, vr = waveform_adjoint_source(torch.as_tensor(vz.reshape(1, nt, -1)), None, mask, window_s,
                                    p.steps * p.timestep)
    vr = vr.numpy().reshape(nt, p.ny, p.nx)
    oa = Oracle3D(op, bg, c3d=c3d, flow=fl, adjoint=True)
    oa.set_adjoint_source(vr, maxtime)
    asn = []
    for t in range(maxtime + 1):
        oa.step(t)
        if t % p.steps == 0:
            asn.append(oa.a[3:6].copy())
    return kernels3d(op, bg, np.array(fs), np.array(asn), c3d=c3d), vz


def test_checkpointed_gradient_matches_oracle():
    """Forward, waveform adjoint source, adjoint and kernels with checkpoints
    every 40 steps against the oracle's run with every snapshot stored."""
    from sparc_flows_amd.solver3d import Gradient3D
    p, bg, c3d, fl = case()
    p.outputcad = p.cadforcing = 10 * p.timestep
    maxtime, window = 120, 1e9
    g = Gradient3D(p, bg, c3d=c3d, flow=fl, interval=40, window_s=window)
    K = g.run(5.0, -3.0, maxtime)
    ref, ovz = _oracle_gradient(p, bg, c3d, fl, 5.0, -3.0, maxtime, window)
    assert rel_l2(K["vz"].cpu().numpy(), ovz) < 1e-10
    for k in ("kc", "kpsi", "kvx", "kvy", "kvz", "hess"):
        e = rel_l2(K[k].cpu().numpy(), ref[k])
        print("3D gradient", k, e)
        assert np.abs(ref[k]).max() > 0 and e < 1e-9, k
    # the same run with one segment (every snapshot held), with segments that
    # do not fall on output times (35 steps; a remainder segment at the start)
    # and with the HBM-sized interval (one segment here) is bit for bit equal
    for interval in (maxtime + 1, 35, "auto"):
        K1 = Gradient3D(p, bg, c3d=c3d, flow=fl, interval=interval, window_s=window).run(5.0, -3.0, maxtime)
        if interval == "auto":
            assert K1["interval"] == maxtime + 1
        for k in ("kc", "kpsi", "kvx", "kvy", "kvz", "hess"):
            assert torch.equal(K1[k], K[k]), (interval, k)


def test_line_source_matches_oracle_and_is_y_invariant():
    """The reference's vr (the x profile in every y row): with no flow and no
    sound-speed bump the whole run stays uniform in y."""
    p, bg, c3d, fl = case(flow=False, dc=False)
    o, s = pair(p, bg, c3d, fl)
    maxtime = 100
    o.set_forward_source(5.0, None, maxtime)
    assert np.allclose(o.vr, o.vr[:1])
    ovz = o.run(maxtime)
    vz = s.forward(5.0, None, maxtime).cpu().numpy()
    assert rel_l2(vz, ovz) < 1e-10
    a = s.state().cpu().numpy()
    assert np.abs(a - a[:, :, :1]).max() <= 1e-14 * np.abs(a).max()
    assert np.abs(a[1]).max() <= 1e-14 * np.abs(a).max() and np.abs(a[4]).max() <= 1e-14 * np.abs(a).max()


def test_profile_times_every_pass():
    p, bg, c3d, fl = case()
    s = Sparc3D(p, bg, c3d=c3d)
    s.set_forward_source(0.0, 0.0, 10)
    ms = s.profile(2)
    for k in ("ddx", "ddy", "z_p", "z_v", "filter"):
        assert ms[k] > 0, k
    # the profile starts from a zeroed state: two steps from rest match two plain steps
    a = s.state()
    s.reset()
    s.step(0)
    s.step(1)
    assert torch.equal(a, s.state())


def test_odd_nx_is_rejected():
    """The y-derivative pairs columns (16-byte loads), so 3D nx must be even."""
    p, bg, c3d, fl = case(nx=33, ny=16, flow=False, dc=False)
    with pytest.raises(ValueError):
        Sparc3D(p, bg)
