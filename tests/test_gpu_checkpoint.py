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
    # the second pair of solvers takes the column-pass shape the first pair
    # timed for this grid, so the forward runs are the same bit for bit; the
    # kernel sums are grouped by segment, so they agree to rounding
    assert torch.equal(got["vz"], ref["vz"])
    for k in ("kc", "kpsi", "kvx", "kvz", "hess"):
        e = rel_l2(got[k].cpu().numpy(), ref[k].cpu().numpy())
        print(interval, k, e)
        assert np.abs(ref[k].cpu().numpy()).max() > 0
        assert e < 1e-12, k


def test_one_segment_checkpointing_allocates_no_recompute_buffer():
    """With the interval longer than the run nothing is recomputed: the
    checkpointed gradient's peak device memory is that of the stored-snapshot
    gradient (plus the one checkpoint row), not twice it."""
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    maxtime = 20 * p.steps
    xs = [-15.0, 10.0]
    pipe = GradientPipeline(p, bg, nsrc=2, c2d=c2d, psi=psi)
    ck = CheckpointedGradient(p, bg, nsrc=2, c2d=c2d, psi=psi, interval=maxtime + 1)
    pipe.run(xs, maxtime)
    ck.run(xs, maxtime)                         # warm both (tuning, graphs, kernel buffers)
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    pipe.run(xs, maxtime)
    torch.cuda.synchronize()
    peak_ref = torch.cuda.max_memory_allocated() - base
    torch.cuda.reset_peak_memory_stats()
    ck.run(xs, maxtime)
    torch.cuda.synchronize()
    peak_ck = torch.cuda.max_memory_allocated() - base
    snap = 2 * (maxtime // p.steps + 1) * 6 * p.nz * p.nx * 8
    print("peak bytes: stored", peak_ref, "one-segment checkpointed", peak_ck, "one snapshot set", snap)
    assert peak_ref > snap
    assert peak_ck <= peak_ref + 2 * 1024 * 1024


@pytest.mark.parametrize("slots", [0, 2, 5])
def test_binomial_checkpointing_equals_stored_snapshots(slots):
    """Binomial checkpointing (at most `slots` checkpoints, one snapshot per
    run; 0 slots recomputes from the start every time) gives the stored-
    snapshot gradient: vz bit for bit, kernels to rounding (one pair per
    accumulation instead of all pairs at once)."""
    from sparc_flows_amd.gradient import BinomialCheckpointedGradient
    p, bg, c2d, psi = small_case(nx=64, nz=60)
    maxtime = 9 * p.steps + 5
    xs = [-15.0, 10.0]
    ref = GradientPipeline(p, bg, nsrc=2, c2d=c2d, psi=psi).run(xs, maxtime)
    got = BinomialCheckpointedGradient(p, bg, nsrc=2, c2d=c2d, psi=psi, slots=slots).run(xs, maxtime)
    assert torch.equal(got["vz"], ref["vz"])
    for k in ("kc", "kpsi", "kvx", "kvz", "hess"):
        e = rel_l2(got[k].cpu().numpy(), ref[k].cpu().numpy())
        print("binomial", slots, k, e)
        assert np.abs(ref[k].cpu().numpy()).max() > 0 and e < 1e-12, k
