"""Forward + adjoint + kernel pipeline: the in-HBM counterpart of full.py.

full.py:24-61 runs three `sparc` processes per source -- forward ('Spectral'
instruction), adjoint ('Adjoint') and kernel (no instruction) -- exchanging
vz traces, adjoint sources and ~1 GB of snapshot FITS files on disk.  Here one
:class:`GradientPipeline` keeps a forward and an adjoint batch on the GPU and
passes snapshots between them in HBM.

The adjoint source of the reference is ADJOINT_SOURCE_FILT (driver.f90:
1148-1754: ridge filters, travel-time windows, time reversal).  Until that is
built, :func:`waveform_adjoint_source` supplies the windowed, time-reversed vz
residual of a waveform misfit over the same receiver band (5-60 Mm from the
source, params.0 of the reference) and time window.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .params import SparcParams
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
    """Forward + misfit + adjoint + kernels for a batch of sources."""

    def __init__(self, params: SparcParams, model: np.ndarray, nsrc: int = 1, c2d=None, psi=None,
                 window_s: float = 1800.0):
        self.p = params
        self.nsrc = nsrc
        self.window_s = window_s
        self.fwd = Sparc(params, model, nsrc=nsrc, adjoint=False, c2d=c2d, psi=psi)
        self.adj = Sparc(params, model, nsrc=nsrc, adjoint=True, c2d=c2d, psi=psi)

    def run(self, xsrc, maxtime: int, vz_obs: torch.Tensor | None = None) -> dict:
        vz, fsn = self.fwd.forward(xsrc, maxtime, snapshots=True)
        mask = receiver_mask(self.p, xsrc)
        cad = self.fwd.steps * self.p.timestep
        misfit, vr = waveform_adjoint_source(vz, vz_obs, mask, self.window_s, cad)
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
    segment's outputs [nsrc][n][6][nz][nx] (None if it has none) and that
    segment's checkpoint is not written: it need never be recomputed."""
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
    forward_with_checkpoints(keep_last=True), used instead of a recompute.
    Returns the kernels of sparc_kernels_end."""
    p, C, steps = fwd.params, interval, fwd.steps
    nt = maxtime // steps + 1
    nseg = ck.shape[0]
    dev = fwd.device
    adj.set_adjoint_sources(vr, maxtime)
    adj.reset()
    adj.kernels_begin()
    ta = 0
    kmax = (C + steps - 1) // steps + 1
    fbuf = torch.empty(fwd.nsrc, kmax, 6, p.nz, p.nx, dtype=torch.float64, device=dev)
    abuf = torch.empty(fwd.nsrc, kmax, 4, p.nz, p.nx, dtype=torch.float64, device=dev)
    segs = checkpoint_segments(maxtime, C)
    if len(segs) != nseg:
        raise ValueError("checkpoints and interval do not match")
    for j in reversed(range(nseg)):
        t0, t1 = segs[j]
        ka, kb = math.ceil(t0 / steps), t1 // steps + 1      # forward outputs in this segment
        n = kb - ka
        if n <= 0:
            continue
        if j == nseg - 1 and last is not None:
            fs = last
        else:
            fs = fbuf[:, :n].contiguous() if n < kmax else fbuf
            fwd.load_checkpoint(ck[j])
            fwd.run_range(t0, t1, snaps=fs, k0=ka)
        tend = (nt - 1 - ka) * steps                        # adjoint outputs nt-kb .. nt-1-ka
        as_ = abuf[:, :n].contiguous() if n < kmax else abuf
        adj.run_range(ta, tend, snaps=as_, k0=nt - kb)
        ta = tend + 1
        adj.kernels_add(fs, as_)
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
    Forward snapshot memory drops from nt to one segment's worth.
    """

    def __init__(self, params: SparcParams, model: np.ndarray, nsrc: int = 1, c2d=None, psi=None,
                 window_s: float = 1800.0, interval: int = 100):
        super().__init__(params, model, nsrc, c2d, psi, window_s)
        if interval < 1:
            raise ValueError("checkpoint interval must be positive")
        self.interval = int(interval)

    def run(self, xsrc, maxtime: int, vz_obs: torch.Tensor | None = None) -> dict:
        vz, ck, last = forward_with_checkpoints(self.fwd, xsrc, maxtime, self.interval, keep_last=True)
        mask = receiver_mask(self.p, xsrc)
        misfit, vr = waveform_adjoint_source(vz, vz_obs, mask, self.window_s, self.fwd.steps * self.p.timestep)
        K = kernels_from_checkpoints(self.fwd, self.adj, ck, vr, maxtime, self.interval, last)
        K["vz"] = vz
        K["misfit"] = misfit
        return K
