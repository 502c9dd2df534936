"""Grids whose nx is not a power of two (the reference takes any nx through
FFTW, derivatives.f90:28-58): the row pass's and the x-derivative's
circulant form against the oracle, whose ddx takes a direct DFT for such nx
(pinned by the Fourier-mode closed forms of tests/test_oracle.py, odd nx
included).  The slot counts 64 N2 >= nx cover every instantiated form:
N2 = 2 (nx <= 128), 3, 4, 5, 6 (<= 384), 8 (<= 512), 10, 12, 16 (<= 1024).

Tolerances as for the power-of-two grids: single operators 1e-12 (the
circulant sums round differently from the FFT), multi-step runs 1e-10.
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
from sparc_flows_amd.solver import Sparc  # noqa: E402


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("nx", [48, 75, 100, 129, 200, 300, 350, 511, 600, 700, 1000])
def test_ddx_any_nx_matches_oracle(nx):
    p, bg, _, _ = small_case(nx=nx, nz=20, npml=4, flow=False, dc=False)
    s = Sparc(p, bg)
    o = oracle.Oracle(p, bg)
    rng = np.random.default_rng(nx)
    f = rng.standard_normal((3, p.nz, nx))
    got = s.ddx(_dev(f)).cpu().numpy()
    ref = np.stack([o.ddx(f[i]) for i in range(3)])
    print("ddx", nx, rel_l2(got, ref), flush=True)
    assert rel_l2(got, ref) < 1e-12


@pytest.mark.parametrize("nx,maxtime", [(75, 399), (100, 399), (300, 119)])
def test_forward_trace_any_nx(nx, maxtime):
    """vz trace and snapshots of a forward run with flow and sound-speed bump."""
    p, bg, c2d, psi = small_case(nx=nx, nz=60, flow=True, dc=True)
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    vz, sn = s.forward([-4.0], maxtime, snapshots=True)
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    o.set_forward_source(-4.0, maxtime)
    ovz, osn = o.run(maxtime, snapshots=True)
    e = rel_l2(vz[0].cpu().numpy(), ovz)
    es = rel_l2(sn[0].cpu().numpy(), osn)
    print("nx", nx, "vz", e, "snapshots", es, flush=True)
    assert np.abs(ovz).max() > 0
    assert e < 1e-10 and es < 1e-10


def test_steps_any_nx_wide_rows():
    """20 full steps at nx = 600 (the 10-slot form) against the oracle."""
    p, bg, c2d, psi = small_case(nx=600, nz=40, npml=4, flow=True, dc=True)
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    s.set_forward_sources([3.0], 60)
    o.set_forward_source(3.0, 60)
    s.reset()
    for t in range(20):
        s.step(t)
        o.step(t)
    err = rel_l2(s.state()[:, 0].cpu().numpy(), o.state())
    print("nx 600, 20 steps", err, flush=True)
    assert err < 1e-11


def test_adjoint_and_kernels_any_nx():
    """Adjoint snapshots and all five kernels at nx = 96."""
    p, bg, c2d, psi = small_case(nx=96, nz=60)
    maxtime = 4 * p.steps * 5
    nt = maxtime // p.cadforcing_step + 1
    sf = Sparc(p, bg, c2d=c2d, psi=psi)
    _, fsn = sf.forward([2.0], maxtime, snapshots=True)
    of = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    of.set_forward_source(2.0, maxtime)
    _, ofsn = of.run(maxtime, snapshots=True)
    assert rel_l2(fsn[0].cpu().numpy(), ofsn) < 1e-10
    vr = adjoint_source(p, nt, p.nx)
    sa = Sparc(p, bg, adjoint=True, c2d=c2d, psi=psi)
    _, asn = sa.adjoint(_dev(vr[None]), maxtime, snapshots=True)
    oa = oracle.Oracle(p, bg, c2d=c2d, psi=psi, adjoint=True)
    oa.set_adjoint_source(vr, maxtime)
    _, oasn = oa.run(maxtime, snapshots=True)
    ea = rel_l2(asn[0].cpu().numpy(), oasn)
    print("adjoint snapshots", ea, flush=True)
    assert np.abs(oasn).max() > 0 and ea < 1e-10
    K = sa.kernels(fsn, asn)
    oK = oracle.kernels(p, bg, ofsn, oasn, c2d=c2d)
    for name in ("kc", "kpsi", "kvx", "kvz", "hess"):
        e = rel_l2(K[name][0].cpu().numpy(), oK[name])
        print("kernel", name, e, flush=True)
        assert e < 1e-10, name


def test_discrete_gradient_any_nx():
    """The exact discrete adjoint at nx = 75 (odd: no Nyquist bin) against the
    oracle's transposes."""
    p, bg, c2d, psi = small_case(nx=75, nz=60)
    maxtime = 200
    nt = p.nt_out(maxtime)
    gvz = np.random.default_rng(7).standard_normal((1, nt, p.nx))
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    s.set_forward_sources([5.0], maxtime)
    _, gc, gcoef = s.discrete_gradient(_dev(gvz), interval=0, psi=True)
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    o.set_forward_source(5.0, maxtime)
    g2, gx, gz, _ = o.discrete_gradient(maxtime, gvz[0])
    for name, got, ref in (("c2rho0", gcoef[0, 0], g2), ("v0x", gcoef[1, 0], gx), ("v0z", gcoef[2, 0], gz)):
        e = rel_l2(got.cpu().numpy(), ref)
        print(name, e, flush=True)
        assert e < 1e-9


def test_batch_any_nx_equals_single_runs(monkeypatch):
    """Two sources in one batch at nx = 75 (only 1-column blocks divide it)
    equal their single-source runs bit for bit at a fixed segmentation."""
    p, bg, c2d, psi = small_case(nx=75, nz=60)
    monkeypatch.setenv("SPARC_COLSEG_MS", "6")
    monkeypatch.setenv("SPARC_COLSEG_TC", "1")
    xs = [-10.0, 7.5]
    sb = Sparc(p, bg, nsrc=2, c2d=c2d, psi=psi)
    vzb, _ = sb.forward(xs, 200)
    for i, x in enumerate(xs):
        vz1, _ = Sparc(p, bg, c2d=c2d, psi=psi).forward([x], 200)
        assert torch.equal(vzb[i], vz1[0]), i


def test_out_of_range_nx_is_rejected():
    for nx in (8, 2048):
        p, bg, _, _ = small_case(nx=64, nz=40, flow=False, dc=False)
        p = dataclasses.replace(p, nx=nx, npmlhor=2)
        with pytest.raises((ValueError, RuntimeError)):
            Sparc(p, bg)
