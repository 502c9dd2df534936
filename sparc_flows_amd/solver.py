"""Host-side API of the gfx950 solver: the counterpart of one `sparc SS LL` run.

The reference drives its solver as a program: ``full.py`` copies an
instruction file (``Spectral`` or ``Adjoint``) and runs ``sparc`` once per
source and phase (full.py:24-61); the program reads params.i, the model files
and master.pixels, and writes vz traces and snapshots to disk.  Here a
:class:`Sparc` object is one batch of sources sharing a background model, with
its state and snapshots resident in HBM:

* ``forward(xsrc, maxtime)``   -- COMPUTE_FORWARD (driver.f90:242-277)
* ``adjoint(vr, maxtime)``     -- COMPUTE_ADJOINT (driver.f90:279-298)
* ``kernels(fwd, adj)``        -- CONSTRUCT_KERNELS (kernels.f90:13-612)

All compute runs in libsparc_hip.so through the torch binding
(torch.classes.sparc.Solver); nothing here falls back to another path.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from .params import SparcParams

_LIB_DIR = Path(__file__).resolve().parent / "lib"
TORCH_SO = _LIB_DIR / "libsparc_torch.so"
HIP_SO = _LIB_DIR / "libsparc_hip.so"
_loaded = False


def load_native():
    """Load the torch binding (and with it libsparc_hip.so); fail loudly if absent."""
    global _loaded
    if not _loaded:
        if not TORCH_SO.exists() or not HIP_SO.exists():
            raise RuntimeError(
                f"sparc native libraries missing in {_LIB_DIR}; build them with "
                "`python -m sparc_flows_amd.build`")
        torch.ops.load_library(str(TORCH_SO))
        _loaded = True
    return torch.classes.sparc.Solver


def _as_tensor(a, n=None):
    if a is None:
        return torch.empty(0, dtype=torch.float64)
    t = torch.as_tensor(np.ascontiguousarray(a) if isinstance(a, np.ndarray) else a,
                        dtype=torch.float64).contiguous()
    return t


class Sparc:
    """A batch of ``nsrc`` sources on one GPU sharing one background model."""

    def __init__(self, params: SparcParams, model: np.ndarray, nsrc: int = 1, adjoint: bool = False,
                 c2d: np.ndarray | None = None, psi: np.ndarray | None = None, device=None):
        params.validate()
        if not torch.cuda.is_available():
            raise RuntimeError("the sparc solver needs a HIP GPU")
        self.params = params
        self.nsrc = int(nsrc)
        self.is_adjoint = bool(adjoint)
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                   else torch.device(device).index or 0)
        Solver = load_native()
        p = params
        scalars = [float(p.xlength), float(p.timestep), float(p.cadforcing), float(p.outputcad),
                   float(p.obsheight), float(p.excitdep), float(p.nupeak), float(p.nuwidth)]
        with torch.cuda.device(self.device):
            self._s = Solver(p.nx, p.nz, p.npmlbot, p.npmltop, p.npmlhor, self.nsrc, int(adjoint), scalars,
                             _as_tensor(model), _as_tensor(c2d), _as_tensor(psi), self.device.index)
        self.maxtime = None
        self._own_stream = True

    # ---- plumbing ----
    def _stream(self):
        """The binding, queuing on torch's current stream.  On torch's default
        stream the handle queues on a stream of its own instead (a HIP stream
        that orders with the default stream but not with other handles'), so
        then every call that queues device work waits for it (_done): a
        handle's outputs are complete, and its inputs free, when the call
        returns, whichever handle or torch op uses them next."""
        cs = torch.cuda.current_stream(self.device).cuda_stream
        self._own_stream = cs == 0
        self._s.set_stream(cs)
        return self._s

    def _done(self, out=None):
        if self._own_stream:
            torch.cuda.synchronize(self.device)
        return out

    def query(self, name: str):
        try:
            return self._s.query_int(name)
        except RuntimeError:
            return self._s.query_double(name)

    def background(self, name: str) -> np.ndarray:
        return self._s.background(name).numpy()

    @property
    def steps(self) -> int:
        return self._s.query_int("steps")

    def set_model(self, c2d=None, psi=None) -> None:
        torch.cuda.synchronize(self.device)
        self._s.set_model(_as_tensor(c2d), _as_tensor(psi))

    def set_flow(self, vx_ms=None, vz_ms=None) -> None:
        """Background flow given directly in m/s, [nz][nx] (the compute_data
        branch, driver.f90:65-73); None restores the flow from psi."""
        torch.cuda.synchronize(self.device)
        self._s.set_flow(_as_tensor(vx_ms), _as_tensor(vz_ms))

    # ---- runs ----
    def set_forward_sources(self, xsrc, maxtime: int) -> None:
        x = np.atleast_1d(np.asarray(xsrc, dtype=np.float64))
        if x.size != self.nsrc:
            raise ValueError(f"{x.size} source positions for a batch of {self.nsrc}")
        self._stream().set_forward_sources(torch.as_tensor(x), int(maxtime))
        self._done()
        self.maxtime = int(maxtime)

    def set_forward_profile(self, vr, zprof, series, maxtime: int) -> None:
        """Separable forward source RHS v_z += vr[s][x] * series[t] * zprof[z]
        (sources.gaussian_force builds one) in place of the reference's
        one-row form."""
        vr = np.asarray(vr, dtype=np.float64).reshape(self.nsrc, self.params.nx)
        self._stream().set_forward_profile(torch.as_tensor(vr), _as_tensor(zprof), _as_tensor(series),
                                           int(maxtime))
        self._done()
        self.maxtime = int(maxtime)

    def forward_profile(self, vr, zprof, series, maxtime: int, snapshots: bool = False):
        if self.is_adjoint:
            raise RuntimeError("forward_profile() on an adjoint solver")
        self.set_forward_profile(vr, zprof, series, maxtime)
        return self.run(snapshots)

    def set_adjoint_sources(self, vr: torch.Tensor, maxtime: int) -> None:
        vr = torch.as_tensor(vr, dtype=torch.float64, device=self.device).contiguous()
        if vr.dim() == 2:
            vr = vr.unsqueeze(0)
        self._stream().set_adjoint_sources(vr, int(maxtime))
        self._done()
        self.maxtime = int(maxtime)

    def run(self, snapshots: bool = False):
        """TIMESTEPPING for time indices 0..maxtime; returns (vz, snaps)."""
        if self.maxtime is None:
            raise RuntimeError("set sources before run()")
        vz, sn = self._done(self._stream().run(self.maxtime, bool(snapshots)))
        return vz, (sn if snapshots else None)

    def forward(self, xsrc, maxtime: int, snapshots: bool = False):
        if self.is_adjoint:
            raise RuntimeError("forward() on an adjoint solver")
        self.set_forward_sources(xsrc, maxtime)
        return self.run(snapshots)

    def adjoint(self, vr, maxtime: int, snapshots: bool = True):
        if not self.is_adjoint:
            raise RuntimeError("adjoint() on a forward solver")
        self.set_adjoint_sources(vr, maxtime)
        return self.run(snapshots)

    def run_range(self, t0: int, t1: int, vz: torch.Tensor | None = None, snaps: torch.Tensor | None = None,
                  k0: int = 0) -> None:
        """Steps t0..t1 from the current state (which must be at t0) into
        caller-allocated slot buffers; slot j holds output index k0 + j."""
        empty = torch.empty(0, dtype=torch.float64, device=self.device)
        self._stream().run_range(int(t0), int(t1), empty if vz is None else vz,
                                 empty if snaps is None else snaps, int(k0))
        self._done()

    def checkpoint_buffer(self, n: int = 1) -> torch.Tensor:
        """Device storage for n full-state checkpoints, [n][numel]."""
        return torch.empty(n, self._stream().checkpoint_numel(), dtype=torch.float64, device=self.device)

    def save_checkpoint(self, buf: torch.Tensor) -> None:
        self._stream().save_checkpoint(buf)
        self._done()

    def load_checkpoint(self, buf: torch.Tensor) -> None:
        self._stream().load_checkpoint(buf)
        self._done()

    def kernels_begin(self) -> None:
        self._stream().kernels_begin()
        self._done()

    def kernels_add(self, fwd_snaps: torch.Tensor, adj_snaps: torch.Tensor) -> None:
        self._stream().kernels_add(fwd_snaps.contiguous(), adj_snaps.contiguous())
        self._done()

    def kernels_end(self) -> dict:
        return dict(zip(("kc", "kpsi", "kvx", "kvz", "hess"), self._done(self._stream().kernels_end())))

    def reset(self) -> None:
        self._stream().reset()
        self._done()

    def step(self, t: int) -> None:
        self._stream().step(int(t))
        self._done()

    def state(self) -> torch.Tensor:
        """[4][nsrc][nz][nx]: xi_x, xi_z, v_x, v_z."""
        return self._done(self._stream().get_state())

    def set_state(self, a: torch.Tensor) -> None:
        self._stream().set_state(torch.as_tensor(a, dtype=torch.float64, device=self.device).contiguous())
        self._done()

    def pml(self):
        return self._done(self._stream().get_pml())

    # ---- operators ----
    def ddx(self, f: torch.Tensor) -> torch.Tensor:
        return self._done(self._stream().ddx(torch.as_tensor(f, dtype=torch.float64, device=self.device).contiguous()))

    def ddz(self, f: torch.Tensor, ibc: int, preset: torch.Tensor | None = None) -> torch.Tensor:
        f = torch.as_tensor(f, dtype=torch.float64, device=self.device).contiguous()
        pre = torch.empty(0, dtype=torch.float64, device=self.device) if preset is None else \
            torch.as_tensor(preset, dtype=torch.float64, device=self.device).contiguous()
        return self._done(self._stream().ddz(f, int(ibc), pre))

    def record_norms(self, n: int | None, t0: int = 0):
        """From the next run on, record sum(v_z^2) over the grid after every
        step t0 .. t0+n-1 (TIMESTEPPING's nor1 squared, driver.f90:364) into
        self.norm2 [nsrc][n]; None switches it off."""
        if n is None:
            self._s.set_norm_output(torch.empty(0, dtype=torch.float64, device=self.device), 0)
            self.norm2 = None
            return None
        self.norm2 = torch.zeros(self.nsrc, int(n), dtype=torch.float64, device=self.device)
        self._s.set_norm_output(self.norm2, int(t0))
        return self.norm2

    def ddz_T(self, f: torch.Tensor, ibc: int):
        """Transpose of ddz (ibc 1..3, kept rows 0); for ibc 3 also the
        adjoint of the kept last row [nbatch][nx]."""
        f = torch.as_tensor(f, dtype=torch.float64, device=self.device).contiguous()
        out, pb = self._done(self._stream().ddz_T(f, int(ibc)))
        return (out, pb) if ibc == 3 else out

    def discrete_gradient(self, gvz: torch.Tensor, interval: int = 0, psi: bool = True):
        """Exact gradient of J = sum(gvz * vz) over the forward run of the
        sources set for self.maxtime, by the transpose of the discrete step
        (sparc_discrete_gradient): dJ/dpsi, dJ/dc and [dJ/d(c^2 rho0),
        dJ/dv0x, dJ/dv0z], each [nsrc][nz][nx].  interval: checkpoint
        interval of the trajectory recompute (0: keep every state)."""
        if self.is_adjoint:
            raise ValueError("the discrete adjoint differentiates a forward solver")
        if self.maxtime is None:
            raise ValueError("set the forward sources first")
        g = torch.as_tensor(gvz, dtype=torch.float64, device=self.device).contiguous()
        gpsi, gc, gcoef = self._done(self._stream().discrete_gradient(int(self.maxtime), g, int(interval),
                                                                      bool(psi)))
        return (gpsi if psi else None), gc, gcoef

    def ddz_kern(self, f: torch.Tensor, ibc: int) -> torch.Tensor:
        f = torch.as_tensor(f, dtype=torch.float64, device=self.device).contiguous()
        return self._done(self._stream().ddz_kern(f, int(ibc)))

    def profile_kernels(self, reps: int = 200) -> dict:
        """Average ms per launch of each pass over `reps` real time steps (HIP
        events around every launch on this stream); resets the state."""
        ms = self._stream().profile_kernels(int(reps))
        return dict(zip(("col", "col_last", "row", "row_first", "row_last"), ms))

    def kernels(self, fwd_snaps: torch.Tensor, adj_snaps: torch.Tensor) -> dict:
        """PRODUCE_KERNELS from in-HBM snapshots; arrays [nsrc][nz][nx]."""
        o = self._done(self._stream().kernels(fwd_snaps.contiguous(), adj_snaps.contiguous()))
        return dict(zip(("kc", "kpsi", "kvx", "kvz", "hess"), o))
