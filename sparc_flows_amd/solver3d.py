"""Host-side API of the 3D solver with flows (option 16, MP_FLOWS_PML_3D_DISPL,
displacement.f90:377-477): ctypes over the C ABI of include/sparc_hip3d.h in
libsparc_hip.so, with torch tensors as the device buffers.

* ``Sparc3D(params, model, c3d=None, adjoint=False)``  -- sparc3d_create
* ``set_flow(vx, vy, vz)``      -- the background flow, cm/s [nz][ny][nx]
* ``forward(xsrc, ysrc, maxtime)`` -- forward source + TIMESTEPPING; vz traces
* ``adjoint(vr, maxtime)``      -- adjoint source (LAGRANGE_INTERP) + TIMESTEPPING
* ``step(t)``, ``state()``, ``ddx/ddy/ddz`` -- single steps and operators

All compute runs in libsparc_hip.so; nothing here falls back to another path.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np
import torch

from .params import f32
from .solver import HIP_SO


@dataclasses.dataclass
class Params3D:
    """params.i values of a 3D run (runtime here)."""
    nx: int = 256
    ny: int = 256
    nz: int = 300
    xlength: float = 200.0e8          # cm
    ylength: float = 200.0e8          # cm
    timestep: float = 2.0             # s
    cadforcing: float = 30.0          # s
    outputcad: float = 30.0           # s
    obsheight: float = 2.0e7          # cm
    excitdep: float = -1.5e7          # cm
    npmlbot: int = 12
    npmltop: int = 12
    nupeak: float = f32(0.0017)
    nuwidth: float = f32(0.0055)

    @property
    def steps(self) -> int:
        return int(math.floor(self.outputcad / self.timestep))

    def nt_out(self, maxtime: int) -> int:
        return maxtime // self.steps + 1

    def validate(self) -> None:
        for n, v in (("nx", self.nx), ("ny", self.ny)):
            if v < 16 or v > 1024:
                raise ValueError(f"{n}={v}: the horizontal derivatives take [16, 1024]")
        if self.nx % 2:
            raise ValueError(f"nx={self.nx}: the y-derivative pairs columns, nx must be even")
        if self.npmlbot < 5 or self.npmltop < 5 or self.npmlbot + self.npmltop + 1 > self.nz:
            raise ValueError("z-PML rows of at least 5 each inside nz are needed")


class _Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("npmlbot", ctypes.c_int), ("npmltop", ctypes.c_int), ("adjoint", ctypes.c_int),
                ("xlength", ctypes.c_double), ("ylength", ctypes.c_double),
                ("timestep", ctypes.c_double), ("cadforcing", ctypes.c_double),
                ("outputcad", ctypes.c_double), ("obsheight", ctypes.c_double),
                ("excitdep", ctypes.c_double), ("nupeak", ctypes.c_double),
                ("nuwidth", ctypes.c_double), ("model", ctypes.c_void_p), ("c3d", ctypes.c_void_p)]


_lib = None


def lib() -> ctypes.CDLL:
    """libsparc_hip.so with the 3D entry points typed; raises if it is not built."""
    global _lib
    if _lib is None:
        if not HIP_SO.exists():
            raise RuntimeError(f"{HIP_SO} missing; build it with `python -m sparc_flows_amd.build`")
        L = ctypes.CDLL(str(HIP_SO))
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        sig = {
            "sparc3d_create": [ctypes.POINTER(_Config), ctypes.POINTER(vp)],
            "sparc3d_destroy": [vp],
            "sparc3d_set_stream": [vp, vp],
            "sparc3d_set_flow": [vp, vp, vp, vp],
            "sparc3d_set_forward_source": [vp, dp, dp, ip],
            "sparc3d_set_adjoint_source": [vp, vp, ip, ip],
            "sparc3d_reset": [vp],
            "sparc3d_step": [vp, ip],
            "sparc3d_run": [vp, ip, vp],
            "sparc3d_run_range": [vp, ip, ip, vp, ip, ip],
            "sparc3d_checkpoint_size": [vp, ctypes.POINTER(ctypes.c_size_t)],
            "sparc3d_save_checkpoint": [vp, vp],
            "sparc3d_load_checkpoint": [vp, vp],
            "sparc3d_get_snapshot": [vp, vp],
            "sparc3d_kernels_begin": [vp],
            "sparc3d_kernels_add": [vp, vp, vp],
            "sparc3d_kernels_end": [vp, vp, vp, vp, vp, vp, vp],
            "sparc3d_get_state": [vp, vp],
            "sparc3d_set_state": [vp, vp],
            "sparc3d_get_pml": [vp, vp],
            "sparc3d_get_scr": [vp, vp],
            "sparc3d_ddx": [vp, vp, vp, ip],
            "sparc3d_ddy": [vp, vp, vp, ip],
            "sparc3d_ddz": [vp, vp, vp, ip, ip],
            "sparc3d_query_int": [vp, ctypes.c_char_p, ctypes.POINTER(ip)],
            "sparc3d_query_double": [vp, ctypes.c_char_p, ctypes.POINTER(dp)],
            "sparc3d_profile": [vp, ip, ctypes.POINTER(dp)],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.sparc3d_last_error.argtypes = [vp]
        L.sparc3d_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Sparc3D:
    """One 3D source on one GPU; state [6][nz][ny][nx] in HBM."""

    def __init__(self, params: Params3D, model: np.ndarray, c3d: np.ndarray | None = None,
                 adjoint: bool = False, device=None):
        params.validate()
        if not torch.cuda.is_available():
            raise RuntimeError("the sparc 3D solver needs a HIP GPU")
        self.p = params
        self.device = torch.device(device or "cuda")
        self.adjoint = bool(adjoint)
        nz = params.nz
        self._model = np.ascontiguousarray(np.asarray(model, dtype=np.float64)[:nz])
        if self._model.shape != (nz, 6):
            raise ValueError("model must be [nz][6]")
        self._c3d = None
        if c3d is not None:
            self._c3d = np.ascontiguousarray(np.asarray(c3d, dtype=np.float64).reshape(nz, params.ny, params.nx))
        cfg = _Config(params.nx, params.ny, nz, params.npmlbot, params.npmltop, int(self.adjoint),
                      params.xlength, params.ylength, params.timestep, params.cadforcing,
                      params.outputcad, params.obsheight, params.excitdep, params.nupeak, params.nuwidth,
                      self._model.ctypes.data, None if self._c3d is None else self._c3d.ctypes.data)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            self._check(lib().sparc3d_create(ctypes.byref(cfg), ctypes.byref(h)), None)
        self.h = h
        # queue on torch's stream, so that torch's own writes to the buffers
        # passed in (zero fills, copies) are ordered with the solver's work
        self._call("sparc3d_set_stream", ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        self.maxtime = -1

    # ---- plumbing ----
    def _check(self, rc: int, h) -> None:
        if rc != 0:
            msg = lib().sparc3d_last_error(h).decode(errors="replace")
            exc = ValueError if rc == -1 else RuntimeError
            raise exc(f"sparc3d error {rc}: {msg}")

    def _call(self, name: str, *args) -> None:
        self._check(getattr(lib(), name)(self.h, *args), self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().sparc3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _zeros(self, *shape) -> torch.Tensor:
        return torch.zeros(shape, dtype=torch.float64, device=self.device)

    def query(self, name: str):
        v = ctypes.c_int()
        if lib().sparc3d_query_int(self.h, name.encode(), ctypes.byref(v)) == 0:
            return v.value
        d = ctypes.c_double()
        self._check(lib().sparc3d_query_double(self.h, name.encode(), ctypes.byref(d)), self.h)
        return d.value

    # ---- set-up ----
    def set_flow(self, vx=None, vy=None, vz=None) -> None:
        """Background flow in cm/s, [nz][ny][nx] each (None: no flow)."""
        if vx is None:
            self._call("sparc3d_set_flow", None, None, None)
            return
        arrs = [np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(self.p.nz, self.p.ny, self.p.nx))
                for v in (vx, vy, vz)]
        self._call("sparc3d_set_flow", *[a.ctypes.data for a in arrs])

    def set_forward_source(self, xsrc_mm: float, ysrc_mm, maxtime: int) -> None:
        """ysrc_mm None: the reference's line source at x = xsrc (uniform in y);
        a number: a point source at (xsrc, ysrc)."""
        y = float("nan") if ysrc_mm is None else float(ysrc_mm)
        self._call("sparc3d_set_forward_source", float(xsrc_mm), y, int(maxtime))
        self.maxtime = int(maxtime)

    def set_adjoint_source(self, vr: torch.Tensor, maxtime: int) -> None:
        """vr: [nt][ny][nx] adjoint source slices at the forcing cadence."""
        vr = torch.as_tensor(vr, dtype=torch.float64, device=self.device).contiguous()
        self._vr = vr
        self._call("sparc3d_set_adjoint_source", _ptr(vr), int(vr.shape[0]), int(maxtime))
        self.maxtime = int(maxtime)

    # ---- time stepping ----
    def reset(self) -> None:
        self._call("sparc3d_reset")

    def step(self, t: int) -> None:
        self._call("sparc3d_step", int(t))

    def run(self, maxtime: int | None = None) -> torch.Tensor:
        """Steps 0..maxtime from the current state; vz at o_rad [nout][ny][nx]."""
        maxtime = self.maxtime if maxtime is None else int(maxtime)
        vz = self._zeros(self.p.nt_out(maxtime), self.p.ny, self.p.nx)
        self._call("sparc3d_run", maxtime, _ptr(vz))
        return vz

    def forward(self, xsrc_mm: float, ysrc_mm, maxtime: int) -> torch.Tensor:
        self.set_forward_source(xsrc_mm, ysrc_mm, maxtime)
        self.reset()
        return self.run(maxtime)

    def adjoint_run(self, vr, maxtime: int) -> torch.Tensor:
        self.set_adjoint_source(vr, maxtime)
        self.reset()
        return self.run(maxtime)

    def run_range(self, t0: int, t1: int, vz_out: torch.Tensor | None = None, k0: int = 0) -> None:
        """Steps t0..t1 from the current state; vz of time t goes to slot t/steps - k0."""
        n = 0 if vz_out is None else int(vz_out.shape[0])
        self._call("sparc3d_run_range", int(t0), int(t1), _ptr(vz_out), int(k0), n)

    # ---- checkpoints and snapshots ----
    def checkpoint_size(self) -> int:
        b = ctypes.c_size_t()
        self._call("sparc3d_checkpoint_size", ctypes.byref(b))
        return b.value

    def save_checkpoint(self, buf: torch.Tensor | None = None) -> torch.Tensor:
        if buf is None:
            buf = torch.empty(self.checkpoint_size() // 8, dtype=torch.float64, device=self.device)
        self._call("sparc3d_save_checkpoint", _ptr(buf))
        return buf

    def load_checkpoint(self, buf: torch.Tensor) -> None:
        self._call("sparc3d_load_checkpoint", _ptr(buf))

    def snapshot(self, out: torch.Tensor | None = None) -> torch.Tensor:
        """Forward: [9][nz][ny][nx] (xi, v, acc); adjoint: [3][nz][ny][nx] (v)."""
        nf = 3 if self.adjoint else 9
        if out is None:
            out = torch.empty((nf, self.p.nz, self.p.ny, self.p.nx), dtype=torch.float64, device=self.device)
        self._call("sparc3d_get_snapshot", _ptr(out))
        return out

    # ---- kernels (PRODUCE_KERNELS with the y terms) ----
    def kernels_begin(self) -> None:
        self._call("sparc3d_kernels_begin")

    def kernels_add(self, fwd_snap: torch.Tensor, adj_snap: torch.Tensor) -> None:
        self._call("sparc3d_kernels_add", _ptr(fwd_snap.contiguous()), _ptr(adj_snap.contiguous()))

    def kernels_end(self) -> dict:
        keys = ("kc", "kpsi", "kvx", "kvy", "kvz", "hess")
        out = {k: self._zeros(self.p.nz, self.p.ny, self.p.nx) for k in keys}
        self._call("sparc3d_kernels_end", *[_ptr(out[k]) for k in keys])
        return out

    def kernels(self, fwd_snaps: torch.Tensor, adj_snaps: torch.Tensor) -> dict:
        """All pairs at once: forward snapshot k with adjoint snapshot nt-1-k."""
        nt = fwd_snaps.shape[0]
        self.kernels_begin()
        for k in range(nt):
            self.kernels_add(fwd_snaps[k], adj_snaps[nt - 1 - k])
        return self.kernels_end()

    # ---- state and operators ----
    def state(self) -> torch.Tensor:
        a = self._zeros(6, self.p.nz, self.p.ny, self.p.nx)
        self._call("sparc3d_get_state", _ptr(a))
        return a

    def set_state(self, a: torch.Tensor) -> None:
        a = torch.as_tensor(a, dtype=torch.float64, device=self.device).contiguous()
        self._call("sparc3d_set_state", _ptr(a))

    def pml(self) -> torch.Tensor:
        npml = self.query("npml")
        t = self._zeros(2, npml, self.p.ny, self.p.nx)
        self._call("sparc3d_get_pml", _ptr(t))
        return t

    def scr(self) -> torch.Tensor:
        t = self._zeros(6, self.p.nz, self.p.ny, self.p.nx)
        self._call("sparc3d_get_scr", _ptr(t))
        return t

    def ddx(self, f: torch.Tensor) -> torch.Tensor:
        f = torch.as_tensor(f, dtype=torch.float64, device=self.device).contiguous()
        out = torch.empty_like(f)
        self._call("sparc3d_ddx", _ptr(f), _ptr(out), f.numel() // (self.p.ny * self.p.nx))
        return out

    def ddy(self, f: torch.Tensor) -> torch.Tensor:
        f = torch.as_tensor(f, dtype=torch.float64, device=self.device).contiguous()
        out = torch.empty_like(f)
        self._call("sparc3d_ddy", _ptr(f), _ptr(out), f.numel() // (self.p.ny * self.p.nx))
        return out

    def ddz(self, f: torch.Tensor, ibc: int) -> torch.Tensor:
        f = torch.as_tensor(f, dtype=torch.float64, device=self.device).contiguous()
        out = torch.zeros_like(f)
        n = f.numel() // (self.p.nz * self.p.ny * self.p.nx)
        self._call("sparc3d_ddz", _ptr(f), _ptr(out), n, int(ibc))
        return out

    def profile(self, reps: int = 2) -> dict:
        """Mean ms per launch of each pass kind over `reps` steps (HIP events)."""
        ms = (ctypes.c_double * 8)()
        self._check(lib().sparc3d_profile(self.h, int(reps), ms), self.h)
        names = ["unused0", "ddx", "ddy", "z_p", "z_v", "filter", "unused6", "unused7"]
        return {n: ms[i] for i, n in enumerate(names)}


def receiver_mask_3d(p: Params3D, xsrc: float, ysrc, dmin_mm: float = 5.0, dmax_mm: float = 60.0):
    """[ny][nx] 1 where the horizontal distance to the source (to the line
    x = xsrc for a line source, ysrc None) is within [dmin, dmax] Mm."""
    x = (np.arange(p.nx) / p.nx - 0.5) * p.xlength * 1e-8 - xsrc
    y = (np.arange(p.ny) / p.ny - 0.5) * p.ylength * 1e-8 - (0.0 if ysrc is None else ysrc)
    d = np.sqrt(x[None, :] ** 2 + (0.0 if ysrc is None else 1.0) * y[:, None] ** 2)
    return ((d >= dmin_mm) & (d <= dmax_mm)).astype(np.float64)


def choose_interval(budget: float, snap_bytes: int, ckpt_bytes: int, nt: int, steps: int, maxtime: int) -> int:
    """Checkpoint interval (time steps) of a gradient over outputs 0..nt-1:
    the whole run (maxtime + 1, no recompute) when all nt snapshots fit the
    budget, else the largest n outputs per segment with
    n * snap_bytes + (ceil(nt / n) - 1) * ckpt_bytes <= budget (one segment of
    snapshots plus the checkpoints of the others), at least one output."""
    if nt * snap_bytes <= budget:
        return maxtime + 1
    best = 1
    for n in range(1, nt):
        if n * snap_bytes + (-(-nt // n) - 1) * ckpt_bytes <= budget:
            best = n
    return best * steps


class Gradient3D:
    """Forward + misfit + adjoint + kernels of one 3D source with fixed-interval
    checkpoint-and-recompute: the run is cut into segments of `interval` steps
    counted back from its end (the first segment takes the remainder); the
    forward run keeps its state at each segment start in HBM and the last
    segment's snapshots themselves.  The adjoint run then meets each forward
    snapshot k (time k*steps) at its own output nt-1-k (kernels.f90:215, 256),
    taking the kept snapshots first and then recomputing the earlier segments
    last to first, holding one segment's snapshots at a time.  Cost: one
    forward run, a second run of all but the last `interval` steps, and one
    adjoint run; memory: the checkpoints plus one segment of snapshots.
    interval "auto" (auto_interval) sizes the segment to the free HBM, so that
    a run whose snapshots all fit is never recomputed.

    The adjoint source is the windowed, time-reversed vz residual over the
    receivers 5-60 Mm from the source (gradient.waveform_adjoint_source), which
    needs outputcad == cadforcing."""

    def __init__(self, params: Params3D, model, c3d=None, flow=None, interval: int | str = "auto",
                 window_s: float = 1800.0, hbm_fraction: float = 0.5):
        if abs(params.outputcad - params.cadforcing) > 1e-12:
            raise ValueError("the adjoint source from vz needs outputcad == cadforcing")
        self.p = params
        self.interval = interval if interval == "auto" else int(interval)
        if self.interval != "auto" and self.interval < 1:
            raise ValueError("checkpoint interval must be positive")
        self.hbm_fraction = float(hbm_fraction)
        self.window_s = window_s
        self.fwd = Sparc3D(params, model, c3d=c3d)
        self.adj = Sparc3D(params, model, c3d=c3d, adjoint=True)
        if flow is not None:
            self.fwd.set_flow(*flow)
            self.adj.set_flow(*flow)

    def auto_interval(self, maxtime: int) -> int:
        """The longest segment whose forward snapshots (9 fields each), with the
        checkpoints of the other segments, fit in hbm_fraction of the free
        device memory (what the device reports free plus the blocks torch's
        allocator holds unused); see choose_interval."""
        dev = self.fwd.device
        free, _ = torch.cuda.mem_get_info(dev)
        free += torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev)
        p = self.p
        return choose_interval(self.hbm_fraction * free, 9 * p.nz * p.ny * p.nx * 8,
                               self.fwd.checkpoint_size(), p.nt_out(maxtime), p.steps, maxtime)

    def run(self, xsrc: float, ysrc, maxtime: int, vz_obs: torch.Tensor | None = None) -> dict:
        from .gradient import checkpoint_segments, waveform_adjoint_source

        p = self.p
        steps = p.steps
        I = self.interval if self.interval != "auto" else self.auto_interval(maxtime)
        nt = maxtime // steps + 1
        fwd, adj = self.fwd, self.adj
        fwd.set_forward_source(xsrc, ysrc, maxtime)
        fwd.reset()
        vz = fwd._zeros(nt, p.ny, p.nx)
        # segments counted back from the end: the last one, [last, maxtime], is
        # a whole interval and keeps its snapshots
        segs = checkpoint_segments(maxtime, I)
        starts, ends = [a for a, _ in segs], dict(segs)
        last = starts[-1]
        ckpt = {}
        kept = {}
        for t0 in starts[:-1]:
            ckpt[t0] = fwd.save_checkpoint()
            fwd.run_range(t0, ends[t0], vz)
        for t in range(last, maxtime + 1):                # the last segment, snapshots kept
            fwd.run_range(t, t, vz)
            if t % steps == 0:
                kept[t // steps] = fwd.snapshot()
        ckpt[last] = None
        # adjoint source
        mask = receiver_mask_3d(p, xsrc, ysrc)[None]
        misfit, vr = waveform_adjoint_source(vz.reshape(1, nt, -1), None if vz_obs is None else
                                             vz_obs.reshape(1, nt, -1), mask.reshape(1, -1),
                                             self.window_s, steps * p.timestep)
        vr = vr.reshape(nt, p.ny, p.nx)
        adj.set_adjoint_source(vr, maxtime)
        adj.reset()
        adj.kernels_begin()
        asnap = None
        t_adj = 0
        for t0 in sorted(ckpt, reverse=True):
            if ckpt[t0] is None:
                snaps, kept = kept, None
            else:
                fwd.load_checkpoint(ckpt[t0])
                snaps = {}
                for t in range(t0, ends[t0] + 1):
                    fwd.step(t)
                    if t % steps == 0:
                        snaps[t // steps] = fwd.snapshot()
            for k in sorted(snaps, reverse=True):
                ta = (nt - 1 - k) * steps
                if ta >= t_adj:
                    adj.run_range(t_adj, ta)
                    t_adj = ta + 1
                asnap = adj.snapshot(asnap)
                adj.kernels_add(snaps[k], asnap)
            del snaps
        K = adj.kernels_end()
        K["vz"] = vz
        K["misfit"] = misfit
        K["interval"] = I
        return K
