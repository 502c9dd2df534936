"""Forward + adjoint + kernel pipeline: the in-HBM counterpart of full.py.

full.py:24-61 runs three `sparc` processes per source -- forward ('Spectral'
instruction), adjoint ('Adjoint') and kernel (no instruction) -- exchanging
vz traces, adjoint sources and ~1 GB of snapshot FITS files on disk.  Here one
:class:`GradientPipeline` keeps a forward and an adjoint batch on the GPU and
passes snapshots between them in HBM.

The adjoint source of the reference is ADJOINT_SOURCE_FILT (driver.f90:
1148-1754: ridge filters, travel-time windows, time reversal; its host
restatement is adjsrc.py).  The pipeline's misfit is either the windowed
cross-correlation travel-time misfit against reference traces
(misfit="travel_time", misfit.py: Gizon-Birch travel times over the receivers
5-60 Mm from the source, params.0 of the reference, and a 30-minute window)
or the windowed vz residual of a waveform misfit over the same receivers and
window (misfit="waveform", :func:`waveform_adjoint_source`).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .params import RSUN, SparcParams
from .solver import Sparc


def receiver_mask(p: SparcParams, xsrc, dmin_mm: float = 5.0, dmax_mm: float = 60.0) -> np.ndarray:
    """[nsrc][nx] 1 where the horizontal distance to the source is within [dmin, dmax] Mm."""
    x = (np.arange(p.nx) / p.nx - 0.5) * p.xlength * 1e-8
    xs = np.atleast_1d(np.asarray(xsrc, dtype=np.float64))
    d = np.abs(x[None, :] - xs[:, None])
    return ((d >= dmin_mm) & (d <= dmax_mm)).astype(np.float64)


def waveform_adjoint_source(vz: torch.Tensor, vz_obs: torch.Tensor | None, mask: np.ndarray,
                            window_s: float, cadence_s: float):
    """Misfit 0.5*sum(w (vz - obs)^2) and its time-reversed adjoint source.

    vz: [nsrc][nt][nx].  The window keeps the first window_s seconds of the
    record, tapered with a cosine over its last 10 %.
    Returns (misfit [nsrc], vr [nsrc][nt][nx]).
    """
    nsrc, nt, nx = vz.shape
    t = torch.arange(nt, dtype=torch.float64, device=vz.device) * cadence_s
    w = torch.ones(nt, dtype=torch.float64, device=vz.device)
    taper = 0.1 * window_s
    w = torch.where(t > window_s, torch.zeros_like(w), w)
    ramp = (t > window_s - taper) & (t <= window_s)
    w = torch.where(ramp, 0.5 * (1 + torch.cos(np.pi * (t - (window_s - taper)) / taper)), w)
    m = torch.as_tensor(mask, dtype=torch.float64, device=vz.device)[:, None, :]
    res = vz if vz_obs is None else vz - vz_obs
    wres = res * w[None, :, None] * m
    misfit = 0.5 * (wres * res).sum(dim=(1, 2))
    vr = torch.flip(wres, dims=[1]).contiguous()
    return misfit, vr


class GradientPipeline:
    """Forward + misfit + adjoint + kernels for a batch of sources.

    misfit "waveform": 0.5 sum w (vz - vz_obs)^2 (vz_obs None: 0);
    misfit "travel_time": 0.5 sum (tau - tau_obs)^2 with tau the linearised
    travel times of vz against vz_obs (required) and tau_obs 0 unless given."""

    def __init__(self, params: SparcParams, model: np.ndarray, nsrc: int = 1, c2d=None, psi=None,
                 window_s: float = 1800.0, misfit: str = "waveform"):
        if misfit not in ("waveform", "travel_time"):
            raise ValueError(f"unknown misfit {misfit!r}")
        self.p = params
        self.nsrc = nsrc
        self.window_s = window_s
        self.misfit = misfit
        self.fwd = Sparc(params, model, nsrc=nsrc, adjoint=False, c2d=c2d, psi=psi)
        self.adj = Sparc(params, model, nsrc=nsrc, adjoint=True, c2d=c2d, psi=psi)

    def adjoint_source(self, vz: torch.Tensor, xsrc, vz_obs: torch.Tensor | None, tau_obs=None):
        """(misfit, vr) for the traces vz [nsrc][nt][nx]."""
        mask = receiver_mask(self.p, xsrc)
        cad = self.fwd.steps * self.p.timestep
        if self.misfit == "waveform":
            return waveform_adjoint_source(vz, vz_obs, mask, self.window_s, cad)
        from .misfit import travel_time_adjoint_source, window_weights
        if vz_obs is None:
            raise ValueError("the travel-time misfit needs reference traces vz_obs")
        nt = vz.shape[1]
        m = torch.as_tensor(mask, dtype=torch.float64, device=vz.device)
        w = window_weights(nt, cad, self.window_s, device=vz.device)
        tau0 = torch.zeros(vz.shape[0], vz.shape[2], dtype=torch.float64, device=vz.device) \
            if tau_obs is None else torch.as_tensor(tau_obs, dtype=torch.float64, device=vz.device)
        dT = self.fwd.steps * self.p.timestep * 1e6 / RSUN
        dx = self.p.xlength / RSUN / self.p.nx
        return travel_time_adjoint_source(vz, vz_obs, tau0, m, w, cad, dT, dx)

    def run(self, xsrc, maxtime: int, vz_obs: torch.Tensor | None = None, tau_obs=None) -> dict:
        vz, fsn = self.fwd.forward(xsrc, maxtime, snapshots=True)
        misfit, vr = self.adjoint_source(vz, xsrc, vz_obs, tau_obs)
        _, asn = self.adj.adjoint(vr, maxtime, snapshots=True)
        K = self.adj.kernels(fsn, asn)
        K["vz"] = vz
        K["misfit"] = misfit
        return K


def checkpoint_segments(maxtime: int, interval: int) -> list:
    """Segments (t0, t1) of time indices 0..maxtime, `interval` steps each
    counted back from maxtime, so that the last one is whole and the first
    one takes the remainder."""
    last = max(0, maxtime + 1 - interval)
    starts = sorted({0} | set(range(last, 0, -interval)))
    return [(t0, t1 - 1) for t0, t1 in zip(starts, starts[1:] + [maxtime + 1])]


def forward_with_checkpoints(fwd: Sparc, xsrc, maxtime: int, interval: int, keep_last: bool = False):
    """Forward run keeping the vz traces and a full-state checkpoint at the
    start of each segment of checkpoint_segments (checkpoint j holds the state
    at the start time of segment j).
    Returns (vz [nsrc][nt][nx], checkpoints [nseg][numel]); with keep_last,
    (vz, checkpoints, last) where `last` holds the snapshots of the last
    segment's outputs [nsrc][n][6][nz][nx] (None if it has none); that
    segment then needs no checkpoint and only the first nseg - 1 are
    returned."""
    steps = fwd.steps
    p = fwd.params
    nt = maxtime // steps + 1
    segs = checkpoint_segments(maxtime, interval)
    nseg = len(segs)
    fwd.set_forward_sources(xsrc, maxtime)
    fwd.reset()
    ck = fwd.checkpoint_buffer(nseg)
    vz = torch.zeros(fwd.nsrc, nt, p.nx, dtype=torch.float64, device=fwd.device)
    last = None
    for j, (t0, t1) in enumerate(segs):
        ka, kb = math.ceil(t0 / steps), t1 // steps + 1
        if keep_last and j == nseg - 1 and kb > ka:
            last = torch.empty(fwd.nsrc, kb - ka, 6, p.nz, p.nx, dtype=torch.float64, device=fwd.device)
            vzl = torch.zeros(fwd.nsrc, kb - ka, p.nx, dtype=torch.float64, device=fwd.device)
            fwd.run_range(t0, t1, vz=vzl, snaps=last, k0=ka)
            vz[:, ka:kb] = vzl
            ck = ck[: nseg - 1]
            break
        fwd.save_checkpoint(ck[j])
        fwd.run_range(t0, t1, vz=vz, k0=0)
    return (vz, ck, last) if keep_last else (vz, ck)


def kernels_from_checkpoints(fwd: Sparc, adj: Sparc, ck: torch.Tensor, vr: torch.Tensor, maxtime: int,
                             interval: int, last: torch.Tensor | None = None) -> dict:
    """Adjoint run driven by vr with the forward snapshots recomputed segment
    by segment, in reverse, from the checkpoints; forward output k is
    correlated with adjoint output nt-1-k (kernels.f90:200-360) as soon as
    both exist.  `last`: the last segment's snapshots kept by
    forward_with_checkpoints(keep_last=True), used instead of a recompute
    (then ck holds the checkpoints of the other segments only).

    Memory: the last segment is correlated first, and this function drops its
    reference to `last` before it allocates the recompute buffer, sized to the
    longest recomputed segment; a caller that passes `last` without keeping a
    reference of its own therefore peaks at one segment of forward snapshots.
    With a single segment nothing is recomputed and no buffer is allocated.
    Returns the kernels of sparc_kernels_end."""
    p, C, steps = fwd.params, interval, fwd.steps
    nt = maxtime // steps + 1
    dev = fwd.device
    segs = checkpoint_segments(maxtime, C)
    nseg = len(segs)
    nck = ck.shape[0]
    if nck != nseg and not (nck == nseg - 1 and last is not None):
        raise ValueError("checkpoints and interval do not match")
    spans = [(math.ceil(t0 / steps), t1 // steps + 1) for t0, t1 in segs]   # forward outputs ka..kb-1
    kept = nseg - 1 if last is not None else None
    if kept is not None and last.shape[1] != spans[kept][1] - spans[kept][0]:
        raise ValueError("`last` does not hold the last segment's snapshots")
    nrec = max([kb - ka for j, (ka, kb) in enumerate(spans) if j != kept] + [0])
    nall = max([kb - ka for ka, kb in spans] + [0])
    adj.set_adjoint_sources(vr, maxtime)
    adj.reset()
    adj.kernels_begin()
    ta = 0
    abuf = torch.empty(fwd.nsrc, nall, 4, p.nz, p.nx, dtype=torch.float64, device=dev)
    fbuf = None
    for j in reversed(range(nseg)):
        t0, t1 = segs[j]
        ka, kb = spans[j]
        n = kb - ka
        if n <= 0:
            continue
        if j == kept:
            fs = last
            last = None
        else:
            if fbuf is None:
                fbuf = torch.empty(fwd.nsrc, nrec, 6, p.nz, p.nx, dtype=torch.float64, device=dev)
            fs = fbuf[:, :n].contiguous() if n < nrec else fbuf
            fwd.load_checkpoint(ck[j])
            fwd.run_range(t0, t1, snaps=fs, k0=ka)
        tend = (nt - 1 - ka) * steps                        # adjoint outputs nt-kb .. nt-1-ka
        as_ = abuf[:, :n].contiguous() if n < nall else abuf
        adj.run_range(ta, tend, snaps=as_, k0=nt - kb)
        ta = tend + 1
        adj.kernels_add(fs, as_)
        fs = None
    return adj.kernels_end()


class CheckpointedGradient(GradientPipeline):
    """The same gradient with fixed-interval checkpoint-and-recompute instead
    of keeping every forward snapshot.

    The reference dumps the forward snapshots to disk and re-reads them in the
    kernel step (WRITE_OUT_PARTIAL_STATE, all_modules.f90:750-784;
    kernels.f90:200-280).  Here the forward run keeps only a full-state
    checkpoint every `interval` steps (plus the vz traces); the adjoint then
    advances in its own time while the forward segments are recomputed from
    their checkpoints in reverse order (all but the last, whose snapshots the
    forward run keeps), and each segment's
    snapshots are correlated with the matching adjoint snapshots
    (forward output k with adjoint output nt-1-k) as soon as both exist.
    Forward snapshot memory drops from nt to one segment's worth (the kept
    last segment is released before the recompute buffer is allocated).
    """

    def __init__(self, params: SparcParams, model: np.ndarray, nsrc: int = 1, c2d=None, psi=None,
                 window_s: float = 1800.0, interval: int = 100, misfit: str = "waveform"):
        super().__init__(params, model, nsrc, c2d, psi, window_s, misfit)
        if interval < 1:
            raise ValueError("checkpoint interval must be positive")
        self.interval = int(interval)

    def run(self, xsrc, maxtime: int, vz_obs: torch.Tensor | None = None, tau_obs=None) -> dict:
        box = list(forward_with_checkpoints(self.fwd, xsrc, maxtime, self.interval, keep_last=True))
        vz, ck = box[0], box[1]
        misfit, vr = self.adjoint_source(vz, xsrc, vz_obs, tau_obs)
        # hand `last` over without keeping a reference here (see kernels_from_checkpoints)
        K = kernels_from_checkpoints(self.fwd, self.adj, ck, vr, maxtime, self.interval, box.pop())
        K["vz"] = vz
        K["misfit"] = misfit
        return K


# ---------------------------------------------------------------------------
# Binomial checkpointing: the backward pass with a fixed number of checkpoint
# slots, recomputing forward states by the binomial (revolve) schedule of
# Griewank & Walther (2000, ACM TOMS 26:19-45) instead of keeping one
# checkpoint per fixed interval.
#
# Positions: pos 0 is the state at time index 0 (zero state, before step 0);
# advancing pos j -> j+1 runs the steps T(j)..T(j+1)-1 with T(0) = 0 and
# T(j) = (j-1)*steps + 1, so it ends with the step of output j (time index
# j*steps): forward output j is the snapshot taken on the way from pos j to
# pos j+1.  The backward pass needs outputs nt-1, ..., 0 in that order.

def _beta(c: int, r: int) -> int:
    return math.comb(c + r, c)


def binomial_split(a: int, b: int, c: int) -> int:
    """Checkpoint position in (a, b) for reversing outputs a..b-1 from a
    stored state at a with c free slots: with r the least repetition number
    such that beta(c, r) >= b - a, the right part b - m <= beta(c-1, r) and the
    left part m - a <= beta(c, r-1) (beta(c, r) = C(c+r, c))."""
    n = b - a
    r = 1
    while _beta(c, r) < n:
        r += 1
    lo = max(1, n - _beta(c - 1, r))
    hi = min(n - 1, _beta(c, r - 1))
    return a + max(lo, min(hi, n - 1))


def binomial_descent(nt: int, slots: int) -> list:
    """Positions the forward pass checkpoints on its way (the first split of
    every level of the schedule), so the backward pass starts without
    recomputing them."""
    out, a, c = [], 0, slots
    while c > 0 and nt - a > 1:
        a = binomial_split(a, nt, c)
        out.append(a)
        c -= 1
    return out


def binomial_schedule(nt: int, slots: int):
    """Operations of the backward pass over outputs nt-1..0 with `slots`
    checkpoint slots, given the descent checkpoints: ("load", pos) restores a
    checkpoint (pos 0: reset), ("advance", a, b) runs pos a -> b without
    snapshots, ("snap", k) runs pos k -> k+1 keeping output k and hands it to
    the adjoint, ("save", pos) / ("free", pos) take and release a slot."""
    have = set(binomial_descent(nt, slots))
    state = {"cur": None}

    def goto(a):
        if state["cur"] != a:
            state["cur"] = a
            yield ("load", a)

    def rev(a, b, c):
        if b - a == 1:
            yield from goto(a)
            yield ("snap", a)
            state["cur"] = a + 1
            return
        if c == 0:
            for k in range(b - 1, a - 1, -1):
                yield from goto(a)
                if k > a:
                    yield ("advance", a, k)
                    state["cur"] = k
                yield ("snap", k)
                state["cur"] = k + 1
            return
        m = binomial_split(a, b, c)
        if m not in have:
            yield from goto(a)
            yield ("advance", a, m)
            yield ("save", m)
            have.add(m)
            state["cur"] = m
        yield from rev(m, b, c - 1)
        yield ("free", m)
        have.discard(m)
        if state["cur"] == m:
            state["cur"] = None
        yield from rev(a, m, c)

    yield from rev(0, nt, slots)


def _pos_time(j: int, steps: int) -> int:
    return 0 if j == 0 else (j - 1) * steps + 1


def forward_binomial(fwd: Sparc, xsrc, maxtime: int, slots: int):
    """Forward run keeping the vz traces and the descent checkpoints of the
    binomial schedule.  Returns (vz [nsrc][nt][nx], {pos: checkpoint})."""
    steps = fwd.steps
    nt = maxtime // steps + 1
    fwd.set_forward_sources(xsrc, maxtime)
    fwd.reset()
    vz = torch.zeros(fwd.nsrc, nt, fwd.params.nx, dtype=torch.float64, device=fwd.device)
    ck, t = {}, 0
    for pos in binomial_descent(nt, slots):
        tp = _pos_time(pos, steps)
        fwd.run_range(t, tp - 1, vz=vz, k0=0)
        ck[pos] = fwd.checkpoint_buffer(1)[0]
        fwd.save_checkpoint(ck[pos])
        t = tp
    fwd.run_range(t, maxtime, vz=vz, k0=0)
    return vz, ck


def kernels_binomial(fwd: Sparc, adj: Sparc, ck: dict, vr: torch.Tensor, maxtime: int, slots: int) -> dict:
    """Adjoint run driven by vr with the forward snapshots recomputed by the
    binomial schedule from the descent checkpoints `ck` (forward_binomial);
    forward output k is correlated with adjoint output nt-1-k as it appears.
    Holds at most `slots` checkpoints and one snapshot of each run."""
    steps, p = fwd.steps, fwd.params
    nt = maxtime // steps + 1
    dev = fwd.device
    ck = dict(ck)
    adj.set_adjoint_sources(vr, maxtime)
    adj.reset()
    adj.kernels_begin()
    fs = torch.empty(fwd.nsrc, 1, 6, p.nz, p.nx, dtype=torch.float64, device=dev)
    as_ = torch.empty(fwd.nsrc, 1, 4, p.nz, p.nx, dtype=torch.float64, device=dev)
    ta = 0
    for op in binomial_schedule(nt, slots):
        kind = op[0]
        if kind == "load":
            if op[1] == 0:
                fwd.reset()
            else:
                fwd.load_checkpoint(ck[op[1]])
        elif kind == "advance":
            fwd.run_range(_pos_time(op[1], steps), _pos_time(op[2], steps) - 1)
        elif kind == "save":
            ck[op[1]] = fwd.checkpoint_buffer(1)[0]
            fwd.save_checkpoint(ck[op[1]])
        elif kind == "free":
            ck.pop(op[1], None)
        else:                                   # snap k: forward output k, adjoint output nt-1-k
            k = op[1]
            fwd.run_range(_pos_time(k, steps), _pos_time(k + 1, steps) - 1, snaps=fs, k0=k)
            tend = (nt - 1 - k) * steps
            adj.run_range(ta, tend, snaps=as_, k0=nt - 1 - k)
            ta = tend + 1
            adj.kernels_add(fs, as_)
    return adj.kernels_end()


class BinomialCheckpointedGradient(GradientPipeline):
    """The gradient with binomial checkpointing: `slots` full-state
    checkpoints (2D: 4 fields and the PML variables per source) and one
    snapshot per run in memory, whatever the run length; the forward states
    are recomputed by the binomial schedule (binomial_schedule)."""

    def __init__(self, params: SparcParams, model: np.ndarray, nsrc: int = 1, c2d=None, psi=None,
                 window_s: float = 1800.0, slots: int = 8, misfit: str = "waveform"):
        super().__init__(params, model, nsrc, c2d, psi, window_s, misfit)
        if slots < 0:
            raise ValueError("checkpoint slots must be >= 0")
        self.slots = int(slots)

    def run(self, xsrc, maxtime: int, vz_obs: torch.Tensor | None = None, tau_obs=None) -> dict:
        vz, ck = forward_binomial(self.fwd, xsrc, maxtime, self.slots)
        misfit, vr = self.adjoint_source(vz, xsrc, vz_obs, tau_obs)
        K = kernels_binomial(self.fwd, self.adj, ck, vr, maxtime, self.slots)
        K["vz"] = vz
        K["misfit"] = misfit
        return K
