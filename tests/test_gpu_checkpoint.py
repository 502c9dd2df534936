"""Checkpoint-and-recompute: a gradient from fixed-interval checkpoints
equals the one from stored snapshots (vz bit for bit; kernels up to the
order of the snapshot-pair sums), and sparc_run_range / checkpoints restart
a run exactly."""
import numpy as np
import pytest

from tests.cases import rel_l2, small_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from sparc_flows_amd.gradient import CheckpointedGradient, GradientPipeline  # noqa: E402
from sparc_flows_amd.solver import Sparc  # noqa: E402


def test_run_range_and_checkpoint_restart_exactly():
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    maxtime = 10 * p.steps + 3
    nt = maxtime // p.steps + 1
    s = Sparc(p, bg, c2d=c2d, psi=psi)
    vz_ref, sn_ref = s.forward([4.0], maxtime, snapshots=True)
    # same run in three ranges, with a checkpoint taken and restored in between
    s.set_forward_sources([4.0], maxtime)
    s.reset()
    vz = torch.zeros_like(vz_ref)
    sn = torch.zeros_like(sn_ref)
    ck = s.checkpoint_buffer(1)
    cut1, cut2 = 3 * p.steps + 5, 7 * p.steps
    s.run_range(0, cut1 - 1, vz, sn, 0)
    s.save_checkpoint(ck[0])
    s.run_range(cut1, cut2 - 1, vz, sn, 0)
    s.reset()                                   # wipe the state, restore the checkpoint
    s.load_checkpoint(ck[0])
    s.run_range(cut1, maxtime, vz, sn, 0)
    assert torch.equal(vz, vz_ref)
    assert torch.equal(sn, sn_ref)
    # a window of slots: outputs 4..6 only
    s.reset()
    part = torch.zeros(1, 3, 6, p.nz, p.nx, dtype=torch.float64, device="cuda")
    s.run_range(0, maxtime, None, part, 4)
    assert torch.equal(part, sn_ref[:, 4:7])
    assert nt > 7


@pytest.mark.parametrize("interval", [37, 100, 10000])
def test_checkpointed_gradient_equals_stored_snapshots(interval):
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    maxtime = 12 * p.steps + 7
    xs = [-15.0, 10.0]
    ref = GradientPipeline(p, bg, nsrc=2, c2d=c2d, psi=psi).run(xs, maxtime)
    got = CheckpointedGradient(p, bg, nsrc=2, c2d=c2d, psi=psi, interval=interval).run(xs, maxtime)
    # two solvers, each with its own timed column-pass shape: equal to rounding
    assert rel_l2(got["vz"].cpu().numpy(), ref["vz"].cpu().numpy()) < 1e-12
    for k in ("kc", "kpsi", "kvx", "kvz", "hess"):
        e = rel_l2(got[k].cpu().numpy(), ref[k].cpu().numpy())
        print(interval, k, e)
        assert np.abs(ref[k].cpu().numpy()).max() > 0
        assert e < 1e-12, k
