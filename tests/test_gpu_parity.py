"""GPU parity: the HIP path (through the C ABI via the torch binding) against
the CPU oracle on the same seeded inputs.

Tolerances: the reference's bar is 1e-9 relative L2 on vz traces after 1000
steps (floating point, FP64).  Single operators are held to 1e-11 or tighter,
multi-step runs to 1e-10.
"""
import dataclasses

import numpy as np
import pytest

from tests.cases import adjoint_source, rel_l2, small_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import oracle  # noqa: E402
from sparc_flows_amd import background as B  # noqa: E402
from sparc_flows_amd.solver import Sparc  # noqa: E402


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("nx", [16, 64, 128, 256, 512])
def test_ddx_matches_oracle(nx):
    p, bg, _, _ = small_case(nx=nx, nz=40, flow=False, dc=False)
    s = Sparc(p, bg)
    o = oracle.Oracle(p, bg)
    rng = np.random.default_rng(nx)
    f = rng.standard_normal((3, p.nz, nx))
    got = s.ddx(_dev(f)).cpu().numpy()
    ref = np.stack([o.ddx(f[i]) for i in range(3)])
    print("ddx", nx, rel_l2(got, ref))
    assert rel_l2(got, ref) < 1e-13


@pytest.mark.parametrize("nz", [20, 60, 150, 300])
@pytest.mark.parametrize("ibc", [0, 1, 2, 3])
def test_ddz_matches_oracle(nz, ibc):
    p, bg, _, _ = small_case(nx=64, nz=nz, npml=4, flow=False, dc=False)
    s = Sparc(p, bg)
    o = oracle.Oracle(p, bg)
    rng = np.random.default_rng(nz * 10 + ibc)
    f = rng.standard_normal((2, nz, 64))
    pre = rng.standard_normal((2, nz, 64))
    got = s.ddz(_dev(f), ibc, _dev(pre)).cpu().numpy()
    ref = np.stack([o.ddz(f[i], ibc, preset=pre[i]) for i in range(2)])
    print("ddz", nz, ibc, rel_l2(got, ref))
    # white-noise input is the worst case for rounding growth through the
    # non-diagonally-dominant boundary closure (uppr(1) = 4)
    assert rel_l2(got, ref) < 1e-10


@pytest.mark.parametrize("nz", [60, 300])
@pytest.mark.parametrize("ibc", [0, 1, 2, 3])
def test_ddz_smooth_fields(nz, ibc):
    p, bg, _, _ = small_case(nx=64, nz=nz, npml=4, flow=False, dc=False)
    s = Sparc(p, bg)
    o = oracle.Oracle(p, bg)
    zz = np.linspace(0, 1, nz)[:, None]
    xx = np.arange(64)[None, :] / 64
    f = np.stack([np.sin(3 * zz + 2 * np.pi * xx) * np.exp(zz), np.cos(5 * zz) + xx])
    pre = np.zeros_like(f)
    pre[:, 0] = 0.3
    pre[:, -1] = -0.2
    got = s.ddz(_dev(f), ibc, _dev(pre)).cpu().numpy()
    ref = np.stack([o.ddz(f[i], ibc, preset=pre[i]) for i in range(2)])
    print("ddz smooth", nz, ibc, rel_l2(got, ref))
    assert rel_l2(got, ref) < 1e-11


def test_background_matches_oracle():
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    for name in ("stretch", "unstretch", "g", "rhoinv", "c2rho0", "gradp0_z", "gradrho0_z",
                 "v0x", "v0z", "az", "bzpml", "axpml", "bxpml"):
        a, b = s.background(name), o.array(name)
        assert a.shape == b.shape, name
        # stretch = 1/dbyd2(z): z ~ 1 while dz ~ 1e-4, and dbyd2's float32
        # boundary coefficients cancel to ~1e-4 of their size, so FMA versus
        # separate multiply-add rounding is amplified to ~1e-12 there
        assert rel_l2(a, b) < 1e-11, (name, rel_l2(a, b))
    for name in ("e_rad", "o_rad", "steps"):
        assert s.query(name) == o.info(name)
    for name in ("deltat", "delta_width"):
        assert abs(s.query(name) - o.info(name)) <= 1e-15 * abs(o.info(name))


@pytest.mark.parametrize("flow,dc", [(False, False), (True, True)])
def test_steps_match_oracle(flow, dc):
    """Full STEPs (RK + PML + filter + source) against the oracle."""
    p, bg, c2d, psi = small_case(nx=64, nz=60, flow=flow, dc=dc)
    maxtime = 60
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    s.set_forward_sources([3.0], maxtime)
    o.set_forward_source(3.0, maxtime)
    s.reset()
    for t in range(40):
        s.step(t)
        o.step(t)
    got = s.state()[:, 0].cpu().numpy()
    ref = o.state()
    err = rel_l2(got, ref)
    print("40 steps", flow, dc, err, [rel_l2(got[i], ref[i]) for i in range(4)])
    assert err < 1e-11
    pv, px = s.pml()
    opv, opx = o.pml()
    assert rel_l2(pv[:, 0].cpu().numpy(), opv) < 1e-10
    assert rel_l2(px[:, 0].cpu().numpy(), opx) < 1e-10


@pytest.mark.parametrize("wide", ["0", "1", "2"])
def test_steps_match_oracle_256_columns(wide, monkeypatch):
    """40 full steps at nx = 256 with each row-pass team width: 16 lanes per
    row, 32 lanes (three-stage transform), 64 lanes (four radix-4 stages)."""
    monkeypatch.setenv("SPARC_ROW_WIDE", wide)
    p, bg, c2d, psi = small_case(nx=256, nz=60, flow=True, dc=True)
    maxtime = 60
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    assert s.query("row_wide") == int(wide)
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    s.set_forward_sources([3.0], maxtime)
    o.set_forward_source(3.0, maxtime)
    s.reset()
    for t in range(40):
        s.step(t)
        o.step(t)
    err = rel_l2(s.state()[:, 0].cpu().numpy(), o.state())
    print("40 steps nx=256 wide", wide, err)
    assert err < 1e-11
    pv, px = s.pml()
    opv, opx = o.pml()
    assert rel_l2(px[:, 0].cpu().numpy(), opx) < 1e-10


def test_forward_trace_1000_steps():
    """vz trace and snapshots of a 1000-step forward run with flow and dc."""
    p, bg, c2d, psi = small_case(nx=64, nz=60, flow=True, dc=True)
    maxtime = 999
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    vz, sn = s.forward([-4.0], maxtime, snapshots=True)
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    o.set_forward_source(-4.0, maxtime)
    ovz, osn = o.run(maxtime, snapshots=True)
    e = rel_l2(vz[0].cpu().numpy(), ovz)
    es = rel_l2(sn[0].cpu().numpy(), osn)
    print("1000-step vz", e, "snapshots", es)
    assert np.abs(ovz).max() > 0
    assert e < 1e-10 and es < 1e-10


def test_batched_sources_equal_single_runs(monkeypatch):
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    xs = [-10.0, 0.0, 7.5]
    maxtime = 200
    # with the same column segmentation and form for every batch size, a
    # batch is bit for bit the single runs
    monkeypatch.setenv("SPARC_COLSEG_MS", "6")
    monkeypatch.setenv("SPARC_COLSEG_TC", "2")
    sb = Sparc(p, bg, nsrc=3, c2d=c2d, psi=psi)
    vzb, _ = sb.forward(xs, maxtime)
    for i, x in enumerate(xs):
        s1 = Sparc(p, bg, c2d=c2d, psi=psi)
        vz1, _ = s1.forward([x], maxtime)
        assert torch.equal(vzb[i], vz1[0]), i
    # the autotuner may segment a batch differently: equal to rounding
    monkeypatch.delenv("SPARC_COLSEG_MS")
    monkeypatch.delenv("SPARC_COLSEG_TC")
    sb = Sparc(p, bg, nsrc=3, c2d=c2d, psi=psi)
    vzb, _ = sb.forward(xs, maxtime)
    for i, x in enumerate(xs):
        s1 = Sparc(p, bg, c2d=c2d, psi=psi)
        vz1, _ = s1.forward([x], maxtime)
        assert rel_l2(vzb[i].cpu().numpy(), vz1[0].cpu().numpy()) < 1e-13, i


def test_adjoint_and_kernels_match_oracle():
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    maxtime = 4 * p.steps * 5  # multiple of the cadence
    nt = maxtime // p.cadforcing_step + 1
    # forward
    sf = Sparc(p, bg, c2d=c2d, psi=psi)
    vzf, fsn = sf.forward([2.0], maxtime, snapshots=True)
    of = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    of.set_forward_source(2.0, maxtime)
    ovz, ofsn = of.run(maxtime, snapshots=True)
    assert rel_l2(fsn[0].cpu().numpy(), ofsn) < 1e-10
    # adjoint
    vr = adjoint_source(p, nt, p.nx)
    sa = Sparc(p, bg, adjoint=True, c2d=c2d, psi=psi)
    _, asn = sa.adjoint(_dev(vr[None]), maxtime, snapshots=True)
    oa = oracle.Oracle(p, bg, c2d=c2d, psi=psi, adjoint=True)
    oa.set_adjoint_source(vr, maxtime)
    _, oasn = oa.run(maxtime, snapshots=True)
    ea = rel_l2(asn[0].cpu().numpy(), oasn)
    print("adjoint snapshots", ea)
    assert np.abs(oasn).max() > 0 and ea < 1e-10
    # kernels
    K = sa.kernels(fsn, asn)
    oK = oracle.kernels(p, bg, ofsn, oasn, c2d=c2d)
    for name in ("kc", "kpsi", "kvx", "kvz", "hess"):
        e = rel_l2(K[name][0].cpu().numpy(), oK[name])
        print("kernel", name, e)
        assert e < 1e-10, name


def test_explicit_flow_equals_psi_flow():
    """sparc_set_flow (flow in m/s, the compute_data branch) given the flow the
    psi path derives reproduces that background and that run; an adjoint
    handle reverses it."""
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    v0x, v0z = s.background("v0x"), s.background("v0z")
    to_ms = 1e6 / 1e2                       # units of dimc -> cm/s -> m/s
    se = Sparc(p, bg, c2d=c2d)
    se.set_flow((v0x * to_ms).reshape(p.nz, p.nx), (v0z * to_ms).reshape(p.nz, p.nx))
    assert rel_l2(se.background("v0x"), v0x) < 1e-15
    assert rel_l2(se.background("v0z"), v0z) < 1e-15
    vz1, _ = s.forward([1.0], 120)
    vz2, _ = se.forward([1.0], 120)
    assert rel_l2(vz2.cpu().numpy(), vz1.cpu().numpy()) < 1e-13
    sa = Sparc(p, bg, adjoint=True, c2d=c2d)
    sa.set_flow((v0x * to_ms).reshape(p.nz, p.nx), (v0z * to_ms).reshape(p.nz, p.nx))
    assert rel_l2(sa.background("v0x"), -v0x) < 1e-15
    se.set_flow(None, None)                 # back to the psi flow of construction (none here)
    assert np.abs(se.background("v0x")).max() == 0.0


def test_unsupported_sizes_are_rejected():
    """nx must be in [16, 1024] (the parameter check says so before any
    device work); nz needs a segmentation of the column pass into at most 64
    segments (1600 has none), which the library rejects at creation."""
    for nx, nz in ((2048, 60), (64, 1600)):
        p, bg, _, _ = small_case(nx=64, nz=min(nz, 300), flow=False, dc=False)
        bg = B.polytrope(nz)
        p = dataclasses.replace(p, nx=nx, nz=nz, timestep=B.stable_timestep(bg))
        with pytest.raises((ValueError, RuntimeError)):
            Sparc(p, bg)


def _shapes(nz):
    """The autotuner's candidates (sparc_capi.hip: tune_lengths): usable
    segment lengths up to 12, or the shortest usable one."""
    sizes = (4, 6, 8, 10, 12, 16, 19, 20, 24)
    usable = [ms for ms in sizes if -(-nz // ms) <= 64 and nz - (-(-nz // ms) - 1) * ms >= 3]
    tuned = [ms for ms in usable if ms <= 12] or usable[:1]
    for ms in tuned:
        P = -(-nz // ms)
        for tc in (1, 2, 4, 8, 16):
            if 3 * (-(-tc * P // 64) * 64) <= 384:
                yield ms, tc


@pytest.mark.parametrize("nz", [150, 300])
def test_every_column_pass_shape_matches_oracle(nz, monkeypatch):
    """Every segment length and block width the autotuner can choose at
    256 x nz with 8 sources in one batch (flow and sound-speed bump, 60
    steps): every source against its own single-source run with the same
    shape, and two sources against the oracle, so that no timing outcome can
    select a shape whose results differ."""
    p, bg, c2d, psi = small_case(nx=256, nz=nz, npml=12, nph=9)
    maxtime, nsteps = 80, 60
    xs = [-60.0, -45.0, -30.0, -10.0, 0.0, 12.0, 33.0, 55.0]
    refs = {}
    for i in (0, 6):
        o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
        o.set_forward_source(xs[i], maxtime)
        for t in range(nsteps):
            o.step(t)
        refs[i] = o.state()
    bad = []
    for ms, tc in _shapes(p.nz):
        monkeypatch.setenv("SPARC_COLSEG_MS", str(ms))
        monkeypatch.setenv("SPARC_COLSEG_TC", str(tc))
        s = Sparc(p, bg, nsrc=8, c2d=c2d, psi=psi)
        s.set_forward_sources(xs, maxtime)
        s.reset()
        s.run_range(0, nsteps - 1)               # applies the forced shape (autotuner entry)
        assert (s.query("colseg_ms"), s.query("colseg_tc")) == (ms, tc)
        got = s.state().cpu().numpy()
        e_or = max(rel_l2(got[:, i], refs[i]) for i in refs)
        e_b = 0.0
        for i in range(8):
            s1 = Sparc(p, bg, nsrc=1, c2d=c2d, psi=psi)
            s1.set_forward_sources([xs[i]], maxtime)
            s1.reset()
            s1.run_range(0, nsteps - 1)
            e_b = max(e_b, rel_l2(got[:, i], s1.state()[:, 0].cpu().numpy()))
        print("nz", nz, "shape MS", ms, "TC", tc, "vs oracle", e_or, "batch vs single", e_b, flush=True)
        if not (np.isfinite(e_or) and e_or < 1e-11 and e_b < 1e-13):
            bad.append((ms, tc, e_or, e_b))
    assert not bad, bad


@pytest.mark.parametrize("nz", [700, 1000])
def test_steps_match_oracle_deep_columns(nz):
    """Columns deeper than 384 rows, where the column pass takes segments of
    16-32 rows (the only lengths with at most 32 segments) and its LDS holds
    the segment tables of three segments only: 20 full steps with flow and a
    sound-speed bump against the oracle."""
    p, bg, c2d, psi = small_case(nx=32, nz=nz, npml=10, nph=4)
    maxtime = 30
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    s.set_forward_sources([3.0], maxtime)
    o.set_forward_source(3.0, maxtime)
    s.reset()
    for t in range(20):
        s.step(t)
        o.step(t)
    print("deep column nz", nz, "MS", s.query("colseg_ms"), "TC", s.query("colseg_tc"), flush=True)
    err = rel_l2(s.state()[:, 0].cpu().numpy(), o.state())
    print("20 steps, nz", nz, err)
    assert err < 1e-11
