"""Checkpoint segments of the checkpoint-and-recompute gradients
(gradient.checkpoint_segments): a partition of 0..maxtime into runs of at most
`interval` steps whose last run is whole, so that the forward run can keep the
last run's snapshots instead of recomputing them."""
import pytest

from sparc_flows_amd.gradient import checkpoint_segments


@pytest.mark.parametrize("maxtime,interval", [(0, 5), (10, 1), (120, 40), (449, 150), (99, 1000),
                                              (3999, 100), (3999, 37), (2999, 1470)])
def test_segments_partition_the_run(maxtime, interval):
    segs = checkpoint_segments(maxtime, interval)
    assert segs[0][0] == 0 and segs[-1][1] == maxtime
    assert all(a[1] + 1 == b[0] for a, b in zip(segs, segs[1:]))
    assert all(1 <= t1 - t0 + 1 <= interval for t0, t1 in segs)
    assert segs[-1][1] - segs[-1][0] + 1 == min(interval, maxtime + 1)
    assert len(segs) == -(-(maxtime + 1) // interval)


def test_remainder_goes_first():
    assert checkpoint_segments(120, 40) == [(0, 0), (1, 40), (41, 80), (81, 120)]
    assert checkpoint_segments(2999, 1470) == [(0, 59), (60, 1529), (1530, 2999)]
