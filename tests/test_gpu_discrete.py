"""The exact discrete adjoint on the GPU (sparc_discrete_gradient,
sparc_ddz_T) against the oracle's transposes (orc_discrete_gradient,
orc_ddz_T), which tests/test_oracle_adjoint.py pins with dot-product tests,
the whole-run adjoint identity and finite differences."""
import numpy as np
import pytest

from tests.cases import rel_l2, small_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import oracle  # noqa: E402
from sparc_flows_amd.solver import Sparc  # noqa: E402


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def test_ddz_transpose_matches_oracle():
    """White-noise columns, 1e-10 as for the forward ddz: the boundary
    closure is not diagonally dominant and amplifies rounding."""
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    rng = np.random.default_rng(0)
    f = rng.standard_normal((2, p.nz, p.nx))
    for ibc in (1, 2, 3):
        got = s.ddz_T(_dev(f), ibc)
        if ibc == 3:
            got, pb = got
        for b in range(2):
            ref = o.ddz_T(f[b], ibc)
            if ibc == 3:
                ref, pref = ref
                assert rel_l2(pb[b].cpu().numpy(), pref) < 1e-10
            e = rel_l2(got[b].cpu().numpy(), ref)
            print("ddz_T ibc", ibc, "source", b, e)
            assert e < 1e-10


@pytest.mark.parametrize("flow,dc", [(False, False), (True, True)])
def test_discrete_gradient_matches_oracle(flow, dc):
    """dJ/d(c^2 rho0), dJ/dv0x, dJ/dv0z and their map to psi and c for
    J = sum gvz vz over 400 steps of two sources at 64 x 60, against the
    oracle source by source, to 1e-9."""
    p, bg, c2d, psi = small_case(nx=64, nz=60, flow=flow, dc=dc)
    maxtime = 400
    xs = [-30.0, 12.0]
    nt = p.nt_out(maxtime)
    rng = np.random.default_rng(5)
    gvz = rng.standard_normal((2, nt, p.nx))
    s = Sparc(p, bg, nsrc=2, c2d=c2d, psi=psi)
    s.set_forward_sources(xs, maxtime)
    gpsi, gc, gcoef = s.discrete_gradient(_dev(gvz), interval=0, psi=psi is not None)
    c = c2d if c2d is not None else np.repeat(bg[:p.nz, 1][:, None], p.nx, axis=1)
    for b, x in enumerate(xs):
        o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
        o.set_forward_source(x, maxtime)
        g2, gx, gz, _ = o.discrete_gradient(maxtime, gvz[b])
        for name, got, ref in (("c2rho0", gcoef[0, b], g2), ("v0x", gcoef[1, b], gx), ("v0z", gcoef[2, b], gz)):
            if not flow and name != "c2rho0":
                continue
            e = rel_l2(got.cpu().numpy(), ref)
            print(name, "source", b, e, flush=True)
            assert e < 1e-9
        rpsi, rc = o.model_gradient(maxtime, gvz[b], c, psi)
        e = rel_l2(gc[b].cpu().numpy(), rc)
        print("c source", b, e)
        assert e < 1e-9
        if psi is not None:
            e = rel_l2(gpsi[b].cpu().numpy(), rpsi)
            print("psi source", b, e)
            assert e < 1e-9


@pytest.mark.parametrize("interval", [1, 37, 120])
def test_checkpointed_trajectory_equals_kept_states(interval, monkeypatch):
    """Recomputing the trajectory from checkpoints every `interval` steps
    gives the gradient of the run that keeps every state, bit for bit (the
    same kernels in the same order), and so does recomputing each step's
    stage quantities in its transposed step instead of keeping them from the
    segment's recompute."""
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    maxtime = 300
    s = Sparc(p, bg, nsrc=1, c2d=c2d, psi=psi)
    s.set_forward_sources([5.0], maxtime)
    gvz = _dev(np.random.default_rng(7).standard_normal((1, p.nt_out(maxtime), p.nx)))
    a = s.discrete_gradient(gvz, interval=0)
    b = s.discrete_gradient(gvz, interval=interval)
    monkeypatch.setenv("SPARC_DISCRETE_KEEP_STAGES", "0")
    c = s.discrete_gradient(gvz, interval=interval)
    for x, y, z in zip(a[:2], b[:2], c[:2]):
        assert torch.equal(x, y) and torch.equal(x, z)
    assert torch.equal(a[2], b[2]) and torch.equal(a[2], c[2])


def test_batch_equals_single_sources():
    """A batch of three sources gives each source's gradient of its own run."""
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    maxtime = 200
    xs = [-40.0, 0.0, 25.0]
    nt = p.nt_out(maxtime)
    g = np.random.default_rng(8).standard_normal((3, nt, p.nx))
    s = Sparc(p, bg, nsrc=3, c2d=c2d, psi=psi)
    s.set_forward_sources(xs, maxtime)
    _, gc, _ = s.discrete_gradient(_dev(g))
    for b, x in enumerate(xs):
        s1 = Sparc(p, bg, nsrc=1, c2d=c2d, psi=psi)
        s1.set_forward_sources([x], maxtime)
        _, gc1, _ = s1.discrete_gradient(_dev(g[b:b + 1]))
        assert rel_l2(gc[b].cpu().numpy(), gc1[0].cpu().numpy()) < 1e-13
