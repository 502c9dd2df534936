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


def test_3d_interval_counts_the_checkpoints():
    """solver3d.choose_interval: one segment of snapshots plus the checkpoints
    of the other segments fit the budget (checked with a mocked budget)."""
    from sparc_flows_amd.solver3d import choose_interval
    snap, ck, nt, steps, maxtime = 9 * 100, 5 * 100, 30, 106, 2999
    assert choose_interval(nt * snap, snap, ck, nt, steps, maxtime) == maxtime + 1
    for budget in (snap, 3 * snap, 10 * snap, 29 * snap):
        I = choose_interval(budget, snap, ck, nt, steps, maxtime)
        n = I // steps
        assert n >= 1 and I == n * steps
        nseg = -(-nt // n)
        if n > 1:
            assert n * snap + (nseg - 1) * ck <= budget
        # one more output per segment would not fit
        if n + 1 < nt:
            assert (n + 1) * snap + (-(-nt // (n + 1)) - 1) * ck > budget or n + 1 > nt
    # snapshots alone would fit 29 outputs, the checkpoints of a second segment do not
    I = choose_interval(29 * snap, snap, ck, nt, steps, maxtime)
    assert I // steps < 29


@pytest.mark.parametrize("nt,slots", [(1, 0), (1, 3), (2, 1), (5, 2), (10, 0), (30, 3), (38, 4),
                                      (134, 8), (134, 3), (101, 1), (300, 10)])
def test_binomial_schedule_reverses_the_run(nt, slots):
    """gradient.binomial_schedule against a mock solver: every forward output
    k = nt-1, ..., 0 is taken from the state at position k, loads only touch
    checkpoints that exist (the forward pass's descent, or saved since), no
    more than `slots` are alive at once, and no output interval is advanced
    more often than the binomial bound allows (r + 1 times, r the least
    repetition number with C(slots + r, slots) >= nt; slots = 0: nt times)."""
    import math
    from sparc_flows_amd.gradient import binomial_descent, binomial_schedule
    have = set(binomial_descent(nt, slots))
    assert len(have) <= slots and all(0 < p < nt for p in have)
    cur, snaps, runs = None, [], [0] * nt
    for op in binomial_schedule(nt, slots):
        if op[0] == "load":
            assert op[1] == 0 or op[1] in have
            cur = op[1]
        elif op[0] == "advance":
            assert cur == op[1] and op[2] > op[1]
            for j in range(op[1], op[2]):
                runs[j] += 1
            cur = op[2]
        elif op[0] == "snap":
            assert cur == op[1]
            runs[op[1]] += 1
            snaps.append(op[1])
            cur += 1
        elif op[0] == "save":
            assert cur == op[1] and op[1] not in have
            have.add(op[1])
            assert len(have) <= slots
        elif op[0] == "free":
            have.discard(op[1])
    assert snaps == list(range(nt - 1, -1, -1))
    if slots == 0:
        assert max(runs) <= nt
    else:
        r = 1
        while math.comb(slots + r, slots) < nt:
            r += 1
        assert max(runs) <= r + 1, (max(runs), r)
