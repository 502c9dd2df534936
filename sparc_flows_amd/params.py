"""Run parameters of the 2D solver: the params.i constants the option-17 path reads.

The reference fixes these at compile time in ``params.i`` (``params.i:25-28``
grid and time step, ``:37`` cadforcing, ``:66`` nupeak/nuwidth, ``:89`` npmlhor,
``:122`` npmlbot/npmltop, ``:139`` obsheight, ``:143`` outputcad, ``:152``
excitdep).  Here they are runtime values so one build serves every grid.

``nupeak`` and ``nuwidth`` are written in ``params.i`` as default-REAL
(float32) literals assigned to real64 parameters, so the solver uses the
float32-widened values; :func:`f32` reproduces that widening.
"""
from __future__ import annotations

import dataclasses
import math
import re
import struct
from pathlib import Path

RSUN = 6.95989467700e10  # cm, initialize.f90:37


def f32(v: float) -> float:
    """Round a Python float to float32 and widen it back (a Fortran REAL literal)."""
    return struct.unpack("f", struct.pack("f", float(v)))[0]


@dataclasses.dataclass
class SparcParams:
    nx: int = 512
    nz: int = 300
    ny: int = 1                       # > 1: a 3D build (option 16, sparc_flows_amd.solver3d)
    xlength: float = 800.0e8          # cm (params.i:28)
    ylength: float = 0.0              # cm; 0: = xlength (params.i:28)
    timestep: float = 2.0             # s
    cadforcing: float = 30.0          # s
    outputcad: float = 30.0           # s
    obsheight: float = 2.0e7          # cm (params.i:139)
    excitdep: float = -1.5e7          # cm (params.i:152)
    npmlbot: int = 10
    npmltop: int = 10
    npmlhor: int = 9
    nupeak: float = f32(0.0017)
    nuwidth: float = f32(0.0055)
    solartime: float = 4.0            # hours (params.i:131)
    file_data: str = "solar_deep"

    # ---- derived integers (driver.f90:214, initialize.f90:991-994) ----
    @property
    def steps(self) -> int:
        return int(math.floor(self.outputcad / self.timestep))

    @property
    def cadforcing_step(self) -> int:
        return int(math.floor(self.cadforcing / self.timestep))

    def reference_maxtime(self) -> int:
        """maxtime as READ_PARAMETERS derives it (initialize.f90:991-993)."""
        cfs = self.cadforcing_step
        mt = int(math.floor((self.solartime * 3600.0 + 2.0 / self.nupeak) / self.timestep)) + 2
        return cfs * (mt // cfs)

    def forcing_length(self, maxtime: int) -> int:
        return maxtime // self.cadforcing_step + 1

    def nt_out(self, maxtime: int) -> int:
        """Number of vz outputs for time indices 0..maxtime."""
        return maxtime // self.steps + 1

    def validate(self) -> None:
        if self.nx < 16 or self.nx > 1024:
            raise ValueError(f"nx={self.nx}: the x-derivative takes nx in [16, 1024]")
        if self.nz < 16 or self.npmlbot + self.npmltop >= self.nz:
            raise ValueError(f"nz={self.nz} too small for npml {self.npmlbot}+{self.npmltop}")
        if 2 * self.npmlhor >= self.nx:
            raise ValueError("horizontal PML wider than the grid")
        if self.steps < 1 or self.cadforcing_step < 1:
            raise ValueError("timestep longer than the output/forcing cadence")


_PARAM_RE = re.compile(r"parameter\s*\((.*)\)", re.IGNORECASE)


def _fortran_number(tok: str) -> float:
    tok = tok.strip().lower().replace("d", "e").replace(" ", "")
    m = re.fullmatch(r"([0-9.e+\-]+)\s*\*\s*10\s*\*\*\s*\(?\s*([+\-]?\d+)\s*\)?", tok)
    if m:
        return float(m.group(1)) * 10.0 ** int(m.group(2))
    return float(tok)


def read_params_i(path: str | Path) -> SparcParams:
    """Parse the numeric PARAMETER statements of a params.i file.

    Mirrors what the reference's read_params.py extracts by text matching.
    REAL literals without a d-exponent are float32 in Fortran; only nupeak and
    nuwidth reach the solver un-rounded-away, so only those are widened.
    """
    text = Path(path).read_text()
    vals: dict[str, str] = {}
    for line in text.splitlines():
        line = line.split("!")[0]
        m = _PARAM_RE.search(line)
        if not m:
            continue
        for part in re.split(r",(?![^(]*\))", m.group(1)):
            if "=" in part:
                k, v = part.split("=", 1)
                vals[k.strip().lower()] = v.strip()
    p = SparcParams()
    ints = {"nx": "nx", "ny": "ny", "nz": "nz", "npmlbot": "npmlbot", "npmltop": "npmltop", "npmlhor": "npmlhor"}
    for key, attr in ints.items():
        if key in vals:
            setattr(p, attr, int(vals[key]))
    floats = {"timestep": "timestep", "cadforcing": "cadforcing", "outputcad": "outputcad",
              "obsheight": "obsheight", "excitdep": "excitdep", "solartime": "solartime"}
    for key, attr in floats.items():
        if key in vals:
            setattr(p, attr, _fortran_number(vals[key]))
    if "xlength" in vals:
        p.xlength = _fortran_number(vals["xlength"].split(",")[0])
    if "ylength" in vals and vals["ylength"].strip().lower() != "xlength":
        p.ylength = _fortran_number(vals["ylength"])
    for key in ("nupeak", "nuwidth"):
        if key in vals:
            setattr(p, key, f32(_fortran_number(vals[key])))
    fd = re.search(r"file_data\s*=\s*'([^']*)'", text)
    if fd:
        p.file_data = fd.group(1)
    return p


def write_params_i(p: SparcParams, path: str | Path, directory: str = "") -> None:
    """Write a minimal params.i with the reference's PARAMETER statement forms
    (params.i:26-152) for the values :func:`read_params_i` reads."""
    def d(v: float) -> str:
        r = repr(float(v))
        return r.replace("e", "d") if "e" in r else r + "d0"
    lines = [
        f"character (LEN = *), PARAMETER :: directory = '{directory}'",
        f"character (LEN = *), PARAMETER :: file_data = '{p.file_data}'",
        f"parameter (nx = {p.nx}, ny = {p.ny}, nz = {p.nz})",
        f"parameter (xlength = {d(p.xlength)}, ylength = "
        f"{'xlength' if not p.ylength or p.ylength == p.xlength else d(p.ylength)}, timestep = {d(p.timestep)})",
        f"parameter( cadforcing = {d(p.cadforcing)})",
        f"parameter(nupeak={repr(float(p.nupeak))}, nuwidth = {repr(float(p.nuwidth))})",
        f"parameter (npmlhor = {p.npmlhor})",
        f"parameter (npmlbot = {p.npmlbot}, npmltop = {p.npmltop})",
        f"parameter (solartime = {d(p.solartime)})",
        f"parameter (obsheight = {d(p.obsheight)})",
        f"parameter (outputcad = {d(p.outputcad)})",
        f"parameter (excitdep = {d(p.excitdep)})",
    ]
    Path(path).write_text("\n".join(lines) + "\n")
