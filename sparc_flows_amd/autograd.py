"""Differentiable 2D solve: (psi, c) -> vz traces as a torch.autograd.Function.

Forward: the background is rebuilt from the stream function psi (model_psi,
driver.f90:94-126) and the 2D sound speed c (model_c, cm/s), and every source
of the batch is run to maxtime with its snapshots kept in HBM
(TIMESTEPPING, driver.f90:337-419; WRITE_OUT_PARTIAL_STATE,
all_modules.f90:750-784).

Backward: the reference's adjoint recipe.  The gradient of the loss with
respect to the vz traces, time reversed, becomes the adjoint source at the
observation row (as ADJOINT_SOURCE writes adj(i, nt-j+1) from sample j and
scales it by nx, driver.f90:805-812); the adjoint batch runs with the flow
reversed (driver.f90:144-147) and PRODUCE_KERNELS (kernels.f90:13-612)
correlates the two.  The kernels come in the reference's physical units
(kernels.f90:495-580); :func:`kernel_to_gradient` turns them into the
gradient with respect to the grid values of psi and c.

Why the recipe gives the gradient.  Write the solver as u' = M u + f with
u = (xi, v) and observations d_k = v_z(x, z_o, t_k).  In the inner product
<a, b> = sum rho0 a.b dx dz the wave operator is symmetric and the flow
operator -2 v0.grad is antisymmetric, so the Lagrange multiplier of the v
equation at time t is -v_a(T - t), where v_a solves the same equations with
v0 -> -v0, forced at row z_o by g(x, T - t) / (rho0 dx dz) and g is the time
density of dL/dd.  The injection term vr * delta_width / rho0
(physics2d.f90:1052-1054) has delta_width = 1/dz, so the adjoint source is
vr = g / dx, with g = (dL/dd_k) / (steps * deltat) between samples.  Then

    dL/dv0x = -2 rho0 dx dz dT sum_k sum_j d_x v_j v_a,j      (Kx of kernels.f90:340-355)
    dL/d(rho0 c^2) = -dx dz dT sum_k div xi . div v_a        (Kc of kernels.f90:303-306)

with dT = steps * deltat, and psi enters through rho0 v0 = curl(rho0 c psi / (diml/1e8) y).
These are the continuous-adjoint identities; the discrete solver matches them
up to its boundary closures, the z filter and the sampling of the time
integral at the output cadence, so finite differences agree to about a
percent, not to rounding.

With ``adjoint="discrete"`` the backward pass is instead the exact transpose
of the discrete solver (Sparc.discrete_gradient, sparc_discrete.hip): every
RK stage, the PMLs, the z filter and the compact-FD and FFT derivatives are
transposed, and dL/dvz enters at the output steps as it is, so the gradient
is the derivative of the computed traces themselves and finite differences
reproduce it to their own accuracy.
"""
from __future__ import annotations


import numpy as np
import torch

from .gradient import (forward_binomial, forward_with_checkpoints, kernels_binomial,
                       kernels_from_checkpoints)
from .misfit import travel_time_misfit, travel_times, window_weights  # noqa: F401  (public names)
from .params import SparcParams
from .solver import Sparc

RSUN = 6.95989467700e10          # diml, initialize.f90:37
DIMC = 1e6                        # initialize.f90:753-776


def _f32(v: float) -> float:
    return float(np.float32(v))


def grid_cells(p: SparcParams, model: np.ndarray):
    """Non-dimensional cell sizes: dx (scalar) and dz per row ([nz])."""
    z = np.asarray(model, dtype=np.float64)[: p.nz, 0]
    dx = p.xlength / RSUN / p.nx
    dz = np.gradient(z)
    return dx, dz


def kernel_to_gradient(p: SparcParams, model: np.ndarray, K: dict, psi: np.ndarray | torch.Tensor,
                       c: np.ndarray | torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Gradient of the loss with respect to psi (model_psi units) and c (cm/s)
    on the grid, from the reference-scaled kernels of one adjoint run.

    K["kpsi"], K["kc"]: [nsrc][nz][nx] as PRODUCE_KERNELS writes them
    (kernels.f90:510-513, 579); sources are summed.
    """
    dev = K["kpsi"].device
    dx, dz = grid_cells(p, model)
    dz = torch.as_tensor(dz, dtype=torch.float64, device=dev)[:, None]
    e8, e5 = _f32(1e-8), _f32(1e-5)
    dimc32 = _f32(1e6)
    dimen = 1.0 / (RSUN * e8 * p.xlength * e8 * dimc32 * e5)        # kernels.f90:503
    dimen_c = 1.0 / (RSUN * e8 * p.xlength * e8)                     # kernels.f90:563
    t_scale = DIMC / RSUN                                            # deltat / timestep
    s = 1e8 / RSUN                                                   # psi' = rho0 c psi s
    psi = torch.as_tensor(psi, dtype=torch.float64, device=dev)
    c = torch.as_tensor(c, dtype=torch.float64, device=dev)
    c_nd = c / DIMC
    # Kpsi = 2 dimen^2 rho0 c steps timestep curl(K); dL/dpsi = -2 dx dz dT rho0 c s curl(K)
    gpsi = -K["kpsi"].sum(0) * (dx * dz * t_scale * s / dimen ** 2)
    # Kc = -2 rho0 c^2 steps timestep dimen_c Kc_raw; dL/dc = -2 rho0 c dx dz dT Kc_raw / dimc
    gc = K["kc"].sum(0) * (dx * dz * t_scale / (dimen_c * c_nd * DIMC))
    # c also scales the flow through psi' = rho0 c psi s: dL/dc += dL/dpsi * psi / c
    gc = gc + gpsi * psi / c
    return gpsi, gc


def _take(obj, name):
    """obj.name, leaving None in its place."""
    v = getattr(obj, name)
    setattr(obj, name, None)
    return v


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi, c, pipe, xsrc, maxtime):
        pipe.set_model(c, psi)
        ctx.binomial = pipe.slots is not None
        ctx.pipe, ctx.maxtime, ctx.model_id, ctx.xsrc = pipe, maxtime, pipe.model_id, xsrc
        if pipe.adjoint == "discrete":
            # the backward pass recomputes the trajectory from its own
            # checkpoints: nothing but the traces is kept here
            vz, _ = pipe.fwd.forward(xsrc, maxtime)
            ctx.fsn = ctx.ck = ctx.last = None
            ctx.save_for_backward(psi, c)
            return vz.clone()
        if ctx.binomial:
            vz, ctx.ck = forward_binomial(pipe.fwd, xsrc, maxtime, pipe.slots)
            ctx.fsn = ctx.last = None
        elif pipe.interval:
            # the last segment's snapshots stay from the forward run: the
            # backward pass recomputes only the segments before it
            vz, ctx.ck, ctx.last = forward_with_checkpoints(pipe.fwd, xsrc, maxtime, pipe.interval,
                                                            keep_last=True)
            ctx.fsn = None
        else:
            vz, ctx.fsn = pipe.fwd.forward(xsrc, maxtime, snapshots=True)
            ctx.ck = ctx.last = None
        ctx.pipe, ctx.maxtime, ctx.model_id = pipe, maxtime, pipe.model_id
        ctx.save_for_backward(psi, c)
        return vz

    @staticmethod
    def backward(ctx, gvz):
        psi, c = ctx.saved_tensors
        pipe = ctx.pipe
        if pipe.model_id != ctx.model_id:        # another solve changed the model since
            pipe.set_model(c, psi)
        if pipe.adjoint == "discrete":
            pipe.fwd.set_forward_sources(ctx.xsrc, ctx.maxtime)
            gpsi, gc, _ = pipe.fwd.discrete_gradient(gvz.to(torch.float64).contiguous(),
                                                     interval=pipe.discrete_interval())
            return gpsi.sum(0).to(psi.device), gc.sum(0).to(c.device), None, None, None
        vr = pipe.adjoint_source_from_grad(gvz)
        if ctx.binomial:
            K = kernels_binomial(pipe.fwd, pipe.adj, ctx.ck, vr, ctx.maxtime, pipe.slots)
        elif ctx.ck is not None:
            # the forward model of this solve: checkpoints were taken with it
            # ctx gives up `last` here, so the kept segment is freed before the
            # recompute buffer is allocated (kernels_from_checkpoints)
            K = kernels_from_checkpoints(pipe.fwd, pipe.adj, ctx.ck, vr, ctx.maxtime, pipe.interval,
                                         _take(ctx, "last"))
        else:
            _, asn = pipe.adj.adjoint(vr, ctx.maxtime, snapshots=True)
            K = pipe.adj.kernels(ctx.fsn, asn)
        ctx.fsn = ctx.ck = ctx.last = None
        gpsi, gc = kernel_to_gradient(pipe.p, pipe.model, K, psi, c)
        return gpsi.to(psi.device), gc.to(c.device), None, None, None


class DifferentiableSolver:
    """Forward and adjoint batches sharing one background, for autograd.

    ``solve(psi, c)`` returns vz [nsrc][nout][nx] (non-dimensional, the
    vz_cc records) and is differentiable with respect to psi [nz][nx]
    (model_psi) and c [nz][nx] (cm/s).  With ``checkpoint_interval`` n > 0
    the forward pass keeps a full-state checkpoint every n steps and the
    snapshots of the last n steps only (that segment is never recomputed);
    the backward pass correlates that segment first, releases it, and
    recomputes the others segment by segment, so forward snapshots peak at
    one segment from the forward pass until the end of the backward pass.
    With ``checkpoint_slots`` c (not None) the backward pass uses binomial
    checkpointing instead (gradient.binomial_schedule): at most c full-state
    checkpoints and one snapshot per run, whatever the run length.
    ``adjoint``: "sparc" (default) runs the reference's recipe above;
    "discrete" differentiates the discrete solver exactly
    (Sparc.discrete_gradient), recomputing the forward trajectory from a
    checkpoint every ``checkpoint_interval`` steps (100 when 0).
    """

    def __init__(self, params: SparcParams, model: np.ndarray, xsrc, maxtime: int,
                 checkpoint_interval: int = 0, checkpoint_slots: int | None = None,
                 adjoint: str = "sparc"):
        if adjoint not in ("sparc", "discrete"):
            raise ValueError(f"unknown adjoint {adjoint!r}: 'sparc' (the reference's recipe) or 'discrete'")
        self.adjoint = adjoint
        if params.steps != params.cadforcing_step:
            raise ValueError("the adjoint source is sampled at the output cadence: "
                             "outputcad and cadforcing must give the same step count")
        self.p = params
        self.model = np.asarray(model, dtype=np.float64)
        self.xsrc = list(np.atleast_1d(np.asarray(xsrc, dtype=np.float64)))
        self.maxtime = int(maxtime)
        n = len(self.xsrc)
        self.fwd = Sparc(params, model, nsrc=n, adjoint=False)
        self.adj = Sparc(params, model, nsrc=n, adjoint=True) if adjoint == "sparc" else None
        self.dx, _ = grid_cells(params, model)
        self.dT = params.steps * params.timestep * DIMC / RSUN
        # 0: keep every forward snapshot; n > 0: checkpoint every n steps and
        # recompute the snapshots during the backward pass
        self.interval = int(checkpoint_interval)
        self.slots = None if checkpoint_slots is None else int(checkpoint_slots)
        if self.slots is not None and self.slots < 0:
            raise ValueError("checkpoint_slots must be >= 0")
        self.model_id = 0

    def discrete_interval(self) -> int:
        """Checkpoint interval of the discrete adjoint's trajectory recompute:
        checkpoint_interval when set, else 100 steps."""
        return self.interval if self.interval > 0 else min(100, self.maxtime + 1)

    def set_model(self, c, psi) -> None:
        cn = c.detach().cpu().numpy() if torch.is_tensor(c) else np.asarray(c)
        pn = psi.detach().cpu().numpy() if torch.is_tensor(psi) else np.asarray(psi)
        self.fwd.set_model(cn, pn)
        if self.adj is not None:
            self.adj.set_model(cn, pn)
        self.model_id += 1

    def adjoint_source_from_grad(self, gvz: torch.Tensor) -> torch.Tensor:
        """vr [nsrc][nt][nx]: time-reversed dL/dvz as a force density."""
        g = torch.flip(gvz.to(torch.float64), dims=[1])
        return (g / (self.dT * self.dx)).contiguous()

    def solve(self, psi: torch.Tensor, c: torch.Tensor) -> torch.Tensor:
        return _Solve.apply(psi, c, self, self.xsrc, self.maxtime)
