"""ctypes access to the CPU oracle (sparc_oracle.c) and to the reference's own
compact-FD solver compiled from its sources (oracle/_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker or the CPU baseline.
The product (sparc_flows_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_dbyd2.so"

_dp = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("nz", ctypes.c_int), ("npmlbot", ctypes.c_int),
                ("npmltop", ctypes.c_int), ("npmlhor", ctypes.c_int), ("adjoint", ctypes.c_int),
                ("xlength", ctypes.c_double), ("timestep", ctypes.c_double),
                ("cadforcing", ctypes.c_double), ("outputcad", ctypes.c_double),
                ("obsheight", ctypes.c_double), ("excitdep", ctypes.c_double),
                ("nupeak", ctypes.c_double), ("nuwidth", ctypes.c_double)]


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    if force or not LIB.exists() or LIB.stat().st_mtime < (HERE / "sparc_oracle.c").stat().st_mtime:
        subprocess.check_call(["make", "-s", "-C", str(HERE), "liboracle.so"])
    if os.path.isdir("/root/reference") and (force or not REF_LIB.exists()):
        subprocess.check_call(["make", "-s", "-C", str(HERE), "ref"])


_lib = None


def _bind(path) -> ctypes.CDLL:
    L = ctypes.CDLL(str(path))
    L.orc_create.restype = ctypes.c_void_p
    L.orc_create.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp, _dp]
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    L.orc_set_flow.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc_set_forward_source.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    L.orc_set_forward_profile.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
    L.orc_set_adjoint_source.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int]
    L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.orc_get_state.argtypes = [ctypes.c_void_p, _dp]
    L.orc_set_state.argtypes = [ctypes.c_void_p, _dp]
    L.orc_get_pml.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc_get_scr.argtypes = [ctypes.c_void_p, _dp]
    L.orc_info_int.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    L.orc_info_double.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    L.orc_info_double.restype = ctypes.c_double
    L.orc_get_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp, ctypes.c_int]
    L.orc_dbyd2.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_ddx.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
    L.orc_ddz.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
    L.orc_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_filter_vars_2d.argtypes = [ctypes.c_void_p, _dp]
    L.orc_filter_var_z.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
    L.orc_lagrange.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.orc_forward_source_series.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int, _dp]
    L.orc_fft.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int]
    L.orc_kernels.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp, ctypes.c_int, _dp, _dp,
                              _dp, _dp, _dp, _dp, _dp]
    L.orc_ddz_T.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp]
    L.orc_ddx_T.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
    L.orc_rhs_T.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_filter_vars_2d_T.argtypes = [ctypes.c_void_p, _dp]
    L.orc_discrete_gradient.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    return L


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = _bind(LIB)
    return _lib


class using:
    """Context manager: route every oracle call to another build of
    sparc_oracle.c (the mutant copies of tests/test_oracle_mutants.py)."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        self.prev = lib()
        _lib = _bind(self.path)
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self.prev
        return False


def build_variant(mutant: int, outdir) -> Path:
    """Compile sparc_oracle.c with ORC_MUTANT=mutant into outdir."""
    out = Path(outdir) / f"liboracle_m{mutant}.so"
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared",
                           f"-DORC_MUTANT={int(mutant)}", "-o", str(out), str(HERE / "sparc_oracle.c"), "-lm"])
    return out


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def make_params(p, adjoint: bool = False) -> OrcParams:
    return OrcParams(p.nx, p.nz, p.npmlbot, p.npmltop, p.npmlhor, int(adjoint), p.xlength,
                     p.timestep, p.cadforcing, p.outputcad, p.obsheight, p.excitdep,
                     p.nupeak, p.nuwidth)


class Oracle:
    """One source's CPU restatement of the sparc option-17 solver."""

    def __init__(self, params, bg: np.ndarray, c2d=None, psi=None, adjoint: bool = False):
        self.p = params
        self.nx, self.nz = params.nx, params.nz
        self._op = make_params(params, adjoint)
        self._bg = np.ascontiguousarray(bg, dtype=np.float64)
        self._c2d = None if c2d is None else np.ascontiguousarray(c2d, dtype=np.float64)
        self._psi = None if psi is None else np.ascontiguousarray(psi, dtype=np.float64)
        self.adjoint = adjoint
        self._L = lib()
        self.h = self._L.orc_create(ctypes.byref(self._op), _ptr(self._bg), _ptr(self._c2d),
                                  _ptr(self._psi))
        if not self.h:
            raise ValueError("orc_create rejected the configuration")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self._L.orc_destroy(h)
            self.h = None

    # ---- info ----
    def info(self, name: str):
        v = self._L.orc_info_int(self.h, name.encode())
        if v != -1:
            return v
        return self._L.orc_info_double(self.h, name.encode())

    def array(self, name: str) -> np.ndarray:
        cap = 4 * self.nx * self.nz + 1_000_000
        out = np.zeros(cap)
        n = self._L.orc_get_array(self.h, name.encode(), _ptr(out), cap)
        if n < 0:
            raise KeyError(name)
        return out[:n].copy()

    def set_flow(self, vx_ms: np.ndarray, vz_ms: np.ndarray) -> None:
        """compute_data flow (true_vx / true_vz in m/s), replacing the psi flow."""
        self._vx = np.ascontiguousarray(vx_ms, dtype=np.float64)
        self._vz = np.ascontiguousarray(vz_ms, dtype=np.float64)
        assert self._L.orc_set_flow(self.h, _ptr(self._vx), _ptr(self._vz)) == 0

    # ---- sources ----
    def set_forward_source(self, xloc: float, maxtime: int) -> None:
        assert self._L.orc_set_forward_source(self.h, float(xloc), int(maxtime)) == 0

    def set_forward_profile(self, vr, zprof, series, maxtime: int) -> None:
        """Separable forward source vr[nx] * series[t] * zprof[nz]."""
        self._pv = [np.ascontiguousarray(a, dtype=np.float64) for a in (vr, zprof, series)]
        assert self._L.orc_set_forward_profile(self.h, *[_ptr(a) for a in self._pv],
                                               self._pv[2].size, int(maxtime)) == 0

    def set_adjoint_source(self, vr: np.ndarray, maxtime: int) -> None:
        vr = np.ascontiguousarray(vr, dtype=np.float64)
        assert self._L.orc_set_adjoint_source(self.h, _ptr(vr), vr.shape[0], int(maxtime)) == 0

    # ---- stepping ----
    def step(self, t: int) -> None:
        self._L.orc_step(self.h, int(t))

    def run(self, maxtime: int, snapshots: bool = False):
        nout = maxtime // self.p.steps + 1
        vz = np.zeros((nout, self.nx))
        nf = 4 if self.adjoint else 6
        snaps = np.zeros((nout, nf, self.nz, self.nx)) if snapshots else None
        n = self._L.orc_run(self.h, int(maxtime), _ptr(vz), _ptr(snaps))
        assert n == nout
        return (vz, snaps) if snapshots else vz

    def state(self) -> np.ndarray:
        a = np.zeros((4, self.nz, self.nx))
        self._L.orc_get_state(self.h, _ptr(a))
        return a

    def set_state(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._L.orc_set_state(self.h, _ptr(a))

    def pml(self):
        nzp = self.p.npmlbot + self.p.npmltop
        pv = np.zeros((2, nzp, self.nx))
        px = np.zeros((2, self.nz, 2 * self.p.npmlhor))
        self._L.orc_get_pml(self.h, _ptr(pv), _ptr(px))
        return pv, px

    def scr(self) -> np.ndarray:
        s = np.zeros((4, self.nz, self.nx))
        self._L.orc_get_scr(self.h, _ptr(s))
        return s

    # ---- operators ----
    def ddx(self, f: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros_like(f)
        self._L.orc_ddx(self.h, _ptr(f), _ptr(out), f.size // self.nx)
        return out

    def ddz(self, f: np.ndarray, ibc: int, preset: np.ndarray | None = None) -> np.ndarray:
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros_like(f) if preset is None else np.ascontiguousarray(preset, dtype=np.float64).copy()
        self._L.orc_ddz(self.h, _ptr(f), _ptr(out), int(ibc))
        return out

    def rhs(self, t: int, temp4, psipml=None, psipmlx=None):
        nzp = self.p.npmlbot + self.p.npmltop
        temp4 = np.ascontiguousarray(temp4, dtype=np.float64)
        psipml = np.zeros((2, nzp, self.nx)) if psipml is None else np.ascontiguousarray(psipml)
        psipmlx = (np.zeros((2, self.nz, 2 * self.p.npmlhor)) if psipmlx is None
                   else np.ascontiguousarray(psipmlx))
        scr = np.zeros_like(temp4)
        scrpml = np.zeros_like(psipml)
        scrpmlx = np.zeros_like(psipmlx)
        self._L.orc_rhs(self.h, int(t), _ptr(temp4), _ptr(psipml), _ptr(psipmlx), _ptr(scr),
                      _ptr(scrpml), _ptr(scrpmlx))
        return scr, scrpml, scrpmlx

    def filter_vars_2d(self, a4: np.ndarray) -> np.ndarray:
        a4 = np.ascontiguousarray(a4, dtype=np.float64).copy()
        self._L.orc_filter_vars_2d(self.h, _ptr(a4))
        return a4

    def lagrange(self, t: int) -> np.ndarray:
        out = np.zeros(self.nx)
        self._L.orc_lagrange(self.h, int(t), _ptr(out))
        return out

    # ---- exact discrete adjoint (transposes of the operators above) ----
    def ddz_T(self, f: np.ndarray, ibc: int):
        """Dz_ibc^T f; for ibc 3 also the adjoint of the kept last row [nx]."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros_like(f)
        pb = np.zeros(self.nx)
        self._L.orc_ddz_T(self.h, _ptr(f), _ptr(out), int(ibc), _ptr(pb))
        return (out, pb) if ibc == 3 else out

    def ddx_T(self, f: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros_like(f)
        self._L.orc_ddx_T(self.h, _ptr(f), _ptr(out), f.size // self.nx)
        return out

    def rhs_T(self, Rb4, Rbpz, Rbpx):
        """Transpose of rhs() without its source: (Yb4, Ybpz, Ybpx, pbar)."""
        Rb4, Rbpz, Rbpx = (np.ascontiguousarray(a, dtype=np.float64) for a in (Rb4, Rbpz, Rbpx))
        Yb4, Ybpz, Ybpx = np.zeros_like(Rb4), np.zeros_like(Rbpz), np.zeros_like(Rbpx)
        pb = np.zeros((self.nz, self.nx))
        self._L.orc_rhs_T(self.h, _ptr(Rb4), _ptr(Rbpz), _ptr(Rbpx), _ptr(Yb4), _ptr(Ybpz), _ptr(Ybpx),
                          _ptr(pb))
        return Yb4, Ybpz, Ybpx, pb

    def filter_vars_2d_T(self, a4: np.ndarray) -> np.ndarray:
        a4 = np.ascontiguousarray(a4, dtype=np.float64).copy()
        self._L.orc_filter_vars_2d_T(self.h, _ptr(a4))
        return a4

    def discrete_gradient(self, maxtime: int, gvz: np.ndarray):
        """dJ/d(c^2 rho0), dJ/dv0x, dJ/dv0z [nz][nx] (code units) and the
        adjoint of the initial state, for J = sum_k <gvz[k], vz_out[k]>."""
        gvz = np.ascontiguousarray(gvz, dtype=np.float64)
        assert gvz.shape == (maxtime // self.p.steps + 1, self.nx)
        out = [np.zeros((self.nz, self.nx)) for _ in range(3)]
        lam0 = np.zeros((4, self.nz, self.nx))
        r = self._L.orc_discrete_gradient(self.h, int(maxtime), _ptr(gvz), *[_ptr(a) for a in out], _ptr(lam0))
        assert r == 0
        return out[0], out[1], out[2], lam0

    def model_gradient(self, maxtime: int, gvz: np.ndarray, c: np.ndarray, psi: np.ndarray | None):
        """dJ/dpsi and dJ/dc [nz][nx] for psi in the model_psi units and c in
        cm/s, through c^2 rho0 = (c/dimc)^2 rho0 and the flow map
        rho0 v0 = curl(rho0 c psi / (Rsun/1e8) y) (driver.f90:94-126).  The
        PML coefficients' dependence on c (rows 1 and nz-5, column 1) is not
        included: directions that vanish in the layers see none of it."""
        g2, gx, gz, _ = self.discrete_gradient(maxtime, gvz)
        dimc, diml = self.info("dimc"), self.info("diml")
        rho0 = self.array("rho0").reshape(self.nz, self.nx)
        cc = np.asarray(c, dtype=np.float64) / dimc
        # pv = rho0 c psi / (diml/1e8); v0x = -Dz1(pv)/rho0, v0z = Dx(pv)/rho0
        gpv = -self.ddz_T(gx / rho0, 1) + self.ddx_T(gz / rho0)
        scale = diml / 1e8
        gpsi = gpv * rho0 * cc / scale
        gc = g2 * 2.0 * cc * rho0 / dimc
        if psi is not None:
            gc = gc + gpv * rho0 * np.asarray(psi, dtype=np.float64) / scale / dimc
        return gpsi, gc


def dbyd2(B: np.ndarray, ibc: int, A: np.ndarray | None = None) -> np.ndarray:
    """Oracle dbyd2 on B[N2][N1] (derivative along axis 0, unit interval)."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B[:, None]
    out = np.zeros_like(B) if A is None else np.ascontiguousarray(A, dtype=np.float64).reshape(B.shape).copy()
    lib().orc_dbyd2(_ptr(out), _ptr(B), B.shape[1], B.shape[0], int(ibc))
    return out


def filter_var_z(var: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(var, dtype=np.float64).copy()
    lib().orc_filter_var_z(v.shape[1], v.shape[0], _ptr(v))
    return v


def forward_source_series(timestep, nupeak, nuwidth, nt) -> np.ndarray:
    out = np.zeros(nt)
    lib().orc_forward_source_series(float(timestep), float(nupeak), float(nuwidth), int(nt), _ptr(out))
    return out


def fft(re: np.ndarray, im: np.ndarray, sign: int):
    re = np.ascontiguousarray(re, dtype=np.float64).copy()
    im = np.ascontiguousarray(im, dtype=np.float64).copy()
    lib().orc_fft(_ptr(re), _ptr(im), re.size, int(sign))
    return re, im


def kernels(params, bg, fwd_snaps, adj_snaps, c2d=None):
    """PRODUCE_KERNELS on in-memory snapshots; returns dict of [nz][nx] arrays."""
    op = make_params(params)
    bg = np.ascontiguousarray(bg, dtype=np.float64)
    fwd = np.ascontiguousarray(fwd_snaps, dtype=np.float64)
    adj = np.ascontiguousarray(adj_snaps, dtype=np.float64)
    c2d = None if c2d is None else np.ascontiguousarray(c2d, dtype=np.float64)
    nt = fwd.shape[0]
    assert adj.shape[0] == nt
    outs = {k: np.zeros((params.nz, params.nx)) for k in ("kc", "kpsi", "kvx", "kvz", "hess")}
    r = lib().orc_kernels(ctypes.byref(op), _ptr(bg), _ptr(c2d), nt, _ptr(fwd), _ptr(adj),
                          _ptr(outs["kc"]), _ptr(outs["kpsi"]), _ptr(outs["kvx"]),
                          _ptr(outs["kvz"]), _ptr(outs["hess"]))
    assert r == 0
    return outs


# ---- the reference's own dbyd2.f/mtridag.f (compiled from its sources) ----
_ref = None


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_dbyd2(B: np.ndarray, ibc: int, A: np.ndarray | None = None) -> np.ndarray:
    """Call the reference DBYD2 (Fortran, column-major A(N1,N2)) on B[N2][N1]."""
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(str(REF_LIB))
        ip = ctypes.POINTER(ctypes.c_int)
        _ref.dbyd2_.argtypes = [_dp, _dp, ip, ip, ip]
    B = np.ascontiguousarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B[:, None]
    out = np.zeros_like(B) if A is None else np.ascontiguousarray(A, dtype=np.float64).reshape(B.shape).copy()
    n1 = ctypes.c_int(B.shape[1])
    n2 = ctypes.c_int(B.shape[0])
    bc = ctypes.c_int(int(ibc))
    _ref.dbyd2_(_ptr(out), _ptr(B), ctypes.byref(n1), ctypes.byref(n2), ctypes.byref(bc))
    return out
