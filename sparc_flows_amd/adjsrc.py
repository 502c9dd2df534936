"""Travel-time adjoint source and misfit: ADJOINT_SOURCE_FILT (driver.f90:1758-2238)
and the per-mode, per-band misfit summary MISFIT_ALL (driver.f90:2294-2651).

After a forward run the reference filters the synthetic vz traces and the
observed ones with ridge filters in (k, nu) (f mode, p1..p7, all p modes;
FMODE_FILTER .. ALL_PMODE_FILTER, driver.f90:1148-1718) and optional
frequency filters (FREQ_FILTER, :1721-1733), picks a window around the
largest synthetic arrival of each receiver, measures the Gizon & Birch travel
time (COMPUTE_TT_GIZONBIRCH, :894-940) and builds the adjoint source

    adj(x, nt+1-j) += Re[ F^-1( filter * FFT_t(ccdot_i * window) * e^{-i k x_i} ) ](x, j) * con,
    con = -tau / (sum_window ccdot_i^2 dt),    adj *= nt / 60,

with ccdot the time derivative of the filtered synthetic.  The misfit is
0.5 sum tau^2 over all measurements (minutes^2).

This module restates that arithmetic with numpy FFTs in FFTW's conventions
(forward e^{-i...} unnormalised, backward unnormalised).  Arrays are
[nt][nx] (time, x), the transpose of the reference's (nx, 1, nt).  REAL
literals of the reference are float32 and are widened the same way here.

Deviations, each where the reference's behaviour is undefined:
* SIMPSON_REGULAR (integrals.f90:107-131) accumulates into an uninitialised
  variable; the sums start at 0 here.
* the search range timest..timefin of a receiver is clipped to the record,
  and a receiver whose range lies wholly outside it is not measured (the
  reference indexes outside the array in both cases); windows are clipped
  to the record the same way.
* a line search whose stored window list is shorter than its receivers
  stops measuring that mode (the reference reuses the last window read).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .params import f32

PI = math.pi


def distmat(n: int) -> np.ndarray:
    """distmat(n, 1, f) (initialize.f90:1181-1198): FFT index order with the
    Nyquist index positive."""
    i = np.arange(n)
    return np.where(i > n - i, -(n - i), i).astype(np.float64)


def _kk(nx: int, xlength_cm: float) -> np.ndarray:
    # k = abs(k) * 2 pi / (xlength * 10.**(-8.) * nx / (nx - 1.))  (Mm^-1)
    return np.abs(distmat(nx)) * 2.0 * PI / (xlength_cm * f32(1e-8) * nx / (nx - 1.0))


def _ff(nt: int, outputcad: float) -> np.ndarray:
    # w = abs(w) * 2 pi / (nt dt); f = w / (2 pi) * 1e3  (mHz)
    w = np.abs(distmat(nt)) * 2.0 * PI / (nt * outputcad)
    return w / (2.0 * PI) * 1e3


def _low_cut(filt: np.ndarray, f: np.ndarray, f_low: float, df: float) -> np.ndarray:
    """Zero below f_low, raised-cosine taper up to f_low + df (f per row)."""
    out = filt.copy()
    out[f < f_low, :] = 0.0
    taper = f < f_low + df
    out[taper, :] *= (0.5 * (1.0 + np.cos(PI * (f[taper] - (f_low + df)) / df)))[:, None]
    return out


def _band(f: np.ndarray, f0: np.ndarray, f1: np.ndarray) -> np.ndarray:
    """Raised cosine over 0 < f - f0 < f1 - f0 for every (f, k)."""
    delta = (f1 - f0)[None, :]
    d = f[:, None] - f0[None, :]
    inside = (d < delta) & (d > 0)
    with np.errstate(invalid="ignore", divide="ignore"):
        v = 0.5 * (1.0 + np.cos(PI * (np.abs(d) - np.abs(delta) * 0.5) / (np.abs(delta) * 0.5)))
    return np.where(inside, v, 0.0)


# (Polylow, Poly) quadratic ridge boundaries in mHz against k in Mm^-1
# (driver.f90:1148-1718); p2 uses f0 = const * sqrt(k).
_RIDGES = {
    "f": ((0.7, 1.7, -0.2), (1.1, 1.9, -0.2), 1.1),
    "p1": ((1.1, 2.4, -0.3), (1.4, 3.0, -0.5), 1.6),
    "p2": (4.1004, (1.6, 3.8, -0.65), 1.6),
    "p3": ((2.0, 3.55, -0.7), (2.0, 4.1, -0.8), 1.6),
    "p4": ((2.15, 4.0, -0.7), (2.3, 4.4, -0.7), 1.6),
    "p5": ((2.2, 4.7, -1.0), (2.35, 5.6, -1.1), 1.6),
    "p6": ((1.80035417224, 7.42939105658, -2.84595764385), (1.98567115899, 8.09108986838, -3.20316331815), 1.6),
    "p7": ((1.98297334673, 7.94931076885, -3.00725356897), (2.18544600032, 8.68183289647, -3.84478880142), 1.6),
}


def ridge_filter(name: str, nt: int, nx: int, xlength_cm: float, outputcad: float) -> np.ndarray:
    """FMODE_FILTER / P1..P7MODE_FILTER / ALL_PMODE_FILTER as [nt][nx]."""
    k = _kk(nx, xlength_cm)
    f = _ff(nt, outputcad)
    df = f32(0.5)
    if name == "allp":
        # ALL_PMODE_FILTER starts from 1 everywhere and sets the band to 1
        filt = np.ones((nt, nx))
        return _low_cut(filt, f, f32(1.6), df)
    low, high, f_low = _RIDGES[name]
    poly = [f32(c) for c in high]
    f1 = poly[0] + poly[1] * k + poly[2] * k ** 2.0
    if name == "p2":
        f0 = f32(low) * np.abs(k) ** 0.5
    else:
        pl = [f32(c) for c in low]
        f0 = pl[0] + pl[1] * k + pl[2] * k ** 2.0
    return _low_cut(_band(f, f0, f1), f, f32(f_low), df)


def freq_filter(f1: float, f2: float, nt: int, outputcad: float) -> np.ndarray:
    """FREQ_FILTER (driver.f90:1721-1733): band pass between f1 and f2 mHz, [nt]."""
    w = np.abs(distmat(nt)) * 1e3 / (nt * outputcad)
    wid = f32(0.4) * (w[1] - w[0])
    return 1.0 / (1.0 + np.exp((f1 - w) / wid)) - 1.0 / (1.0 + np.exp((f2 - w) / wid))


def simpson_regular(fv: np.ndarray, dvar: float) -> float:
    """SIMPSON_REGULAR (integrals.f90:107-131): weights 1, 4, 2, 4, ..., 1 by
    1-based index parity, times dvar/3 (the sum starts at 0)."""
    n = fv.shape[0]
    wts = np.where(np.arange(1, n + 1) % 2 == 0, 4.0, 2.0)
    wts[0] = wts[-1] = 1.0
    return float((wts * fv).sum() * dvar / 3.0)


def tt_gizon_birch(u0: np.ndarray, u: np.ndarray, dt: float, lef: int, rig: int) -> float:
    """COMPUTE_TT_GIZONBIRCH (driver.f90:894-940); lef, rig 0-based inclusive."""
    nt = u0.shape[0]
    window = np.zeros(nt)
    window[lef:rig + 1] = 1.0
    u0w = np.fft.rfft(u0)
    kk = np.arange(nt // 2)
    u0w[: nt // 2] = u0w[: nt // 2] * (1j * 2 * PI * kk / (nt * dt))
    u0w[nt // 2] = 0.0
    u0dot = np.fft.irfft(u0w, n=nt) * nt / nt          # c2r unnormalised, then / nt
    num = simpson_regular(window * u0dot * (u0 - u), dt)
    den = simpson_regular(window * u0dot ** 2, dt)
    return num / den


@dataclass
class Measurement:
    mode: int           # 0 f, 1..7 p1..p7, 8 first-bounce p modes
    pixel: int          # 0-based receiver column
    tau_s: float        # travel time, s
    lef: int            # window, 1-based inclusive (as windows.<mode> holds it)
    rig: int
    loc: int
    timest: int
    timefin: int


@dataclass
class AdjointSource:
    adj: np.ndarray                      # [nt][nx], written as source.fits
    misfit: float                        # 0.5 sum tau^2 (minutes^2)
    nmeas: int
    measurements: list = field(default_factory=list)


_MODE_NAMES = {0: "f", 1: "p1", 2: "p2", 3: "p3", 4: "p4", 5: "p5", 6: "p6", 7: "p7", 8: "allp"}
_VEL = [0.44, 0.60, 0.75, 0.9, 1.2, 1.4, 1.7, 1.9]          # Mm/min (driver.f90:1946-1954)


def _fftw_fwd(a):
    return np.fft.fft2(a)


def _fftw_bwd(a):
    return np.fft.ifft2(a) * a.shape[0] * a.shape[1]


def adjoint_source_filt(acc: np.ndarray, dat: np.ndarray, x: np.ndarray, x00: float, xlength_cm: float,
                        outputcad: float, dt_min: float, mode_params: dict,
                        freq_params: dict | None = None, windows: dict | None = None) -> AdjointSource:
    """ADJOINT_SOURCE_FILT for one source.

    acc, dat: [nt][nx] synthetic and observed vz (vz_cc.fits, data/SS.fits);
    x: grid x in [0, 1) (x = (i-1)/nx); x00: source position (Mm);
    dt_min: sample spacing in minutes (from the timeline file);
    mode_params: {mode: (mindist, maxdist, window_min)} for the params.<mode>
    files present; freq_params: {"allmodes"|0|1|2: (f1, f2)} for the
    filter.params.* files present; windows: {mode: [(lef, rig), ...]} to reuse
    (the linesearch branch reads them from windows.<mode>).
    """
    nt, nx = acc.shape
    freq_params = freq_params or {}
    distances = np.abs((x - 0.5) * xlength_cm * f32(1e-8) - x00)
    eyekh = 1j * distmat(nx) * 2.0 * PI
    freqnu = distmat(nt) / (nt * dt_min * 60.0)
    # frequency filters as the reference applies them (driver.f90:1862-1935):
    # one running filter, replaced by each filter.params.* present, multiplies
    # every ridge filter once and then the f, p1, p2 filters again in turn
    filt = np.ones(nt)
    if "allmodes" in freq_params:
        filt = freq_filter(*freq_params["allmodes"], nt, outputcad)
    filters = {m: ridge_filter(name, nt, nx, xlength_cm, outputcad) * filt[:, None]
               for m, name in _MODE_NAMES.items()}
    for m in (0, 1, 2):
        if m in freq_params:
            filt = freq_filter(*freq_params[m], nt, outputcad)
        filters[m] = filters[m] * filt[:, None]
    tempout = _fftw_fwd(acc.astype(np.complex128))
    tempdat = _fftw_fwd(dat.astype(np.complex128))
    adj = np.zeros((nt, nx))
    misfit = 0.0
    nmeas = 0
    meas = []
    for pord in range(9):
        if pord not in mode_params:
            continue
        mindist, maxdist, window = mode_params[pord]
        halftime = int(math.floor(window / (2.0 * dt_min) + 0.5))      # nint of a positive value
        filt = filters[pord]
        accf = _fftw_bwd(tempout * filt)
        datf = _fftw_bwd(tempdat * filt)
        ccdot = _fftw_bwd(tempout * filt * (1j * freqnu * 2.0 * PI)[:, None])
        wlist = list(windows.get(pord, [])) if windows else None
        for i in range(nx):
            if not (distances[i] > mindist and distances[i] < maxdist):
                continue
            if wlist is None:
                if pord != 8:
                    timest = int(math.floor(distances[i] / f32(_VEL[pord]) * 1.0 / dt_min))
                    timefin = timest + 40
                else:
                    dist = distances[i]
                    # default-REAL literals: float32 constants (driver.f90:2046-2049)
                    timest = int(math.floor((f32(-0.000417147774671) * dist ** 2 + f32(0.313350998096) * dist
                                             + f32(17.9609631186)) * 1.0 / dt_min))
                    timefin = int(math.floor((f32(-0.00028361249034) * dist ** 2 + f32(0.29270337114) * dist
                                              + f32(36.4398734578)) * 1.0 / dt_min))
                # 1-based timest..timefin, clipped to the record; a receiver
                # whose search range lies outside the record is not measured
                t0, t1 = max(timest, 1), min(timefin, nt)
                if t0 > t1:
                    continue
                seg = np.abs(accf[t0 - 1:t1, i].real)
                loc = int(np.argmax(seg)) + t0              # 1-based
                lef, rig = max(loc - halftime, 1), min(loc + halftime, nt)
            else:
                if not wlist:
                    break                                   # the stored list is exhausted
                lef, rig = wlist.pop(0)
                loc, timest, timefin = 0, 0, 0
            tau = tt_gizon_birch(accf[:, i].real, datf[:, i].real, dt_min, lef - 1, rig - 1)
            meas.append(Measurement(pord, i, tau * 60.0, lef, rig, loc, timest, timefin))
            win = np.zeros(nt)
            win[lef - 1:rig] = 1.0
            ccdotone = np.fft.fft(ccdot[:, i] * win)
            spec = filt * ccdotone[:, None] * np.exp(-eyekh * x[i])[None, :]
            filtex = _fftw_bwd(spec)
            misfit += tau ** 2.0
            nmeas += 1
            con = -tau / (np.sum(ccdot[lef - 1:rig, i] ** 2.0) * dt_min)
            adj[::-1, :] += (filtex * con).real
    adj = adj * nt / 60.0
    return AdjointSource(adj=adj, misfit=0.5 * misfit, nmeas=nmeas, measurements=meas)


# MISFIT_ALL (driver.f90:2294-2651): the misfit of every mode in three
# frequency bands, written to kernel/misfit_all_SS_LL.
MISFIT_BANDS = ((2.0, 3.5), (3.5, 4.5), (4.5, 6.0))      # mHz (driver.f90:2400-2411)


@dataclass
class MisfitAll:
    bands: list          # per band: (left_mHz, right_mHz, nmeas[0..10], 0.5 * sum tau^2 [0..10])
    windows: dict        # {(mode, band 1..3): [(lef, rig), ...]}, as windows.all.MM.B holds them

    def text(self, index: int) -> str:
        """The file misfit_all_SS_LL: per band a '#' line with the source
        index and band edges, a '#' line with the 11 measurement counts and a
        line with the 11 half misfits (driver.f90:2627-2632; list-directed
        output, read back with numpy.loadtxt(comments='#'))."""
        out = []
        for f1, f2, n, m in self.bands:
            out.append(f" # {index:d} {f1:.16E} {f2:.16E}")
            out.append(" # " + " ".join(f"{int(v):d}" for v in n))
            out.append(" " + " ".join(f"{v:.16E}" for v in m))
        return "\n".join(out) + "\n"


def misfit_all(acc: np.ndarray, dat: np.ndarray, x: np.ndarray, x00: float, xlength_cm: float,
               outputcad: float, dt_min: float, mode_params: dict, windows: dict | None = None) -> MisfitAll:
    """MISFIT_ALL for one source: for each band of MISFIT_BANDS, every ridge
    filter with a params.<mode> file times the band's FREQ_FILTER (no
    chaining between filters here), windows picked as in
    ADJOINT_SOURCE_FILT except the all-p-mode polynomials
    (driver.f90:2503-2512), the Gizon-Birch travel time per receiver, and
    sums of tau^2 per mode.  `windows` ({(mode, band): [(lef, rig)]}) replaces
    the picking, as the linesearch branch reads windows.all.MM.B."""
    nt, nx = acc.shape
    distances = np.abs((x - 0.5) * xlength_cm * f32(1e-8) - x00)
    tempout = _fftw_fwd(acc.astype(np.complex128))
    tempdat = _fftw_fwd(dat.astype(np.complex128))
    ridges = {m: ridge_filter(name, nt, nx, xlength_cm, outputcad) for m, name in _MODE_NAMES.items()}
    bands, wins = [], {}
    for band, (f1, f2) in enumerate(MISFIT_BANDS, start=1):
        filt = freq_filter(f1, f2, nt, outputcad)
        mis = np.zeros(11)
        nmeas = np.zeros(11, dtype=np.int64)
        for pord in range(9):
            if pord not in mode_params:
                continue
            mindist, maxdist, window = mode_params[pord]
            halftime = int(math.floor(window / (2.0 * dt_min) + 0.5))
            F = ridges[pord] * filt[:, None]
            accf = _fftw_bwd(tempout * F).real
            datf = _fftw_bwd(tempdat * F).real
            given = list(windows.get((pord, band), [])) if windows else None
            picked = []
            for i in range(nx):
                if not (distances[i] > mindist and distances[i] < maxdist):
                    continue
                if given is None:
                    if pord != 8:
                        timest = int(math.floor(distances[i] / f32(_VEL[pord]) * 1.0 / dt_min))
                        timefin = timest + 40
                    else:
                        d_i = distances[i]
                        timest = int(math.floor((f32(-1.45511095e-04) * d_i ** 2 + f32(2.72140632e-01) * d_i
                                                 + f32(2.16331969e+01)) * 1.0 / dt_min))
                        timefin = int(math.floor((f32(3.85746516e-08) * d_i ** 2 + f32(2.21324786e-01) * d_i
                                                  + f32(4.36702155e+01)) * 1.0 / dt_min))
                    t0, t1 = max(timest, 1), min(timefin, nt)
                    if t0 > t1:
                        continue
                    loc = int(np.argmax(np.abs(accf[t0 - 1:t1, i]))) + t0
                    lef, rig = max(loc - halftime, 1), min(loc + halftime, nt)
                else:
                    if not given:
                        break
                    lef, rig = given.pop(0)
                picked.append((lef, rig))
                tau = tt_gizon_birch(accf[:, i], datf[:, i], dt_min, lef - 1, rig - 1)
                mis[pord] += tau ** 2
                nmeas[pord] += 1
            wins[(pord, band)] = picked
        bands.append((f1, f2, nmeas, 0.5 * mis))
    return MisfitAll(bands=bands, windows=wins)


def read_mode_params(directory: Path) -> tuple[dict, dict]:
    """params.0 .. params.8 (mindist, maxdist, window) and filter.params.*
    (f1, f2 in mHz) present in the data directory."""
    directory = Path(directory)
    modes, freqs = {}, {}
    for m in range(9):
        f = directory / f"params.{m}"
        if f.exists():
            v = [float(t) for t in f.read_text().split()[:3]]
            modes[m] = tuple(v)
    for key, name in (("allmodes", "filter.params.allmodes"), (0, "filter.params.0"), (1, "filter.params.1"),
                      (2, "filter.params.2")):
        f = directory / name
        if f.exists():
            freqs[key] = tuple(float(t) for t in f.read_text().split()[:2])
    return modes, freqs


def adjoint_source_filt_torch(acc, dat, x: np.ndarray, x00: float, xlength_cm: float, outputcad: float,
                              dt_min: float, mode_params: dict, freq_params: dict | None = None,
                              windows: dict | None = None, device="cuda") -> AdjointSource:
    """adjoint_source_filt on a torch device (the GPU by default), every
    receiver of a mode at once.

    The same measurements as the loop above: filtered traces by 2D FFTs,
    the window of each receiver around the largest filtered synthetic sample
    in its search range (first maximum), Gizon & Birch travel times with the
    time derivative by a batched real FFT and SIMPSON_REGULAR weights, and
    the adjoint source.  The one change of order: the per-receiver terms
    Re[F^-1(filter * c_i(nu) e^{-i k x_i})] * con_i of a mode are summed
    before the inverse transform, as filter * (C E) with C[nu][i] = con_i c_i(nu)
    and E[i][k] = e^{-i k x_i} (one matrix product and one inverse FFT per
    mode instead of one inverse FFT per receiver); that changes rounding
    only.  acc, dat: [nt][nx] arrays or tensors; the filters and the
    per-receiver search ranges are built on the host as above."""
    import torch

    dev = torch.device(device)
    acc_t = torch.as_tensor(np.asarray(acc, dtype=np.float64) if not torch.is_tensor(acc) else acc,
                            dtype=torch.float64, device=dev)
    dat_t = torch.as_tensor(np.asarray(dat, dtype=np.float64) if not torch.is_tensor(dat) else dat,
                            dtype=torch.float64, device=dev)
    nt, nx = acc_t.shape
    freq_params = freq_params or {}
    distances = np.abs((x - 0.5) * xlength_cm * f32(1e-8) - x00)
    eyekh = torch.as_tensor(1j * distmat(nx) * 2.0 * PI, device=dev)
    freqnu = torch.as_tensor(distmat(nt) / (nt * dt_min * 60.0), device=dev)
    filt = np.ones(nt)
    if "allmodes" in freq_params:
        filt = freq_filter(*freq_params["allmodes"], nt, outputcad)
    filters = {m: ridge_filter(name, nt, nx, xlength_cm, outputcad) * filt[:, None]
               for m, name in _MODE_NAMES.items()}
    for m in (0, 1, 2):
        if m in freq_params:
            filt = freq_filter(*freq_params[m], nt, outputcad)
        filters[m] = filters[m] * filt[:, None]
    scale = nt * nx                                                   # FFTW's unnormalised backward
    tempout = torch.fft.fft2(acc_t.to(torch.complex128))
    tempdat = torch.fft.fft2(dat_t.to(torch.complex128))
    # SIMPSON_REGULAR weights (1, 4, 2, 4, ..., 1 by 1-based parity) and the
    # time-derivative multiplier of COMPUTE_TT_GIZONBIRCH
    wts = torch.where(torch.arange(1, nt + 1, device=dev) % 2 == 0, 4.0, 2.0).to(torch.float64)
    wts[0] = 1.0
    wts[-1] = 1.0
    nr = nt // 2 + 1
    kk = torch.arange(nr, device=dev, dtype=torch.float64)
    dmul = torch.where(kk < nt // 2, 1j * 2 * PI * kk / (nt * dt_min), torch.zeros_like(kk) * 1j)
    tt = torch.arange(nt, device=dev)
    adj = torch.zeros((nt, nx), dtype=torch.float64, device=dev)
    misfit = 0.0
    nmeas = 0
    meas = []
    for pord in range(9):
        if pord not in mode_params:
            continue
        mindist, maxdist, window = mode_params[pord]
        halftime = int(math.floor(window / (2.0 * dt_min) + 0.5))
        fm = torch.as_tensor(filters[pord], device=dev)
        accf = torch.fft.ifft2(tempout * fm) * scale
        datf = torch.fft.ifft2(tempdat * fm) * scale
        ccdot = torch.fft.ifft2(tempout * fm * (1j * freqnu * 2.0 * PI)[:, None]) * scale
        # receivers, search ranges (1-based inclusive) and windows on the host
        rec, lefs, rigs, locs, tsts, tfins = [], [], [], [], [], []
        wlist = list(windows.get(pord, [])) if windows else None
        cand = [i for i in range(nx) if distances[i] > mindist and distances[i] < maxdist]
        if wlist is None:
            t0s, t1s, ts_, tf_ = [], [], [], []
            for i in cand:
                if pord != 8:
                    timest = int(math.floor(distances[i] / f32(_VEL[pord]) * 1.0 / dt_min))
                    timefin = timest + 40
                else:
                    dist = distances[i]
                    timest = int(math.floor((f32(-0.000417147774671) * dist ** 2 + f32(0.313350998096) * dist
                                             + f32(17.9609631186)) * 1.0 / dt_min))
                    timefin = int(math.floor((f32(-0.00028361249034) * dist ** 2 + f32(0.29270337114) * dist
                                              + f32(36.4398734578)) * 1.0 / dt_min))
                t0, t1 = max(timest, 1), min(timefin, nt)
                if t0 > t1:
                    continue
                rec.append(i)
                t0s.append(t0)
                t1s.append(t1)
                ts_.append(timest)
                tf_.append(timefin)
            if rec:
                ri = torch.as_tensor(rec, device=dev)
                a0 = torch.as_tensor(t0s, device=dev)[None, :]
                a1 = torch.as_tensor(t1s, device=dev)[None, :]
                inrange = (tt[:, None] + 1 >= a0) & (tt[:, None] + 1 <= a1)
                mag = torch.where(inrange, accf[:, ri].real.abs(), torch.full_like(accf[:, ri].real, -1.0))
                loc = (torch.argmax(mag, dim=0) + 1).cpu().tolist()       # first maximum, 1-based
                for j, i in enumerate(rec):
                    lefs.append(max(loc[j] - halftime, 1))
                    rigs.append(min(loc[j] + halftime, nt))
                    locs.append(loc[j])
                    tsts.append(ts_[j])
                    tfins.append(tf_[j])
        else:
            for i in cand:
                if not wlist:
                    break
                lef, rig = wlist.pop(0)
                rec.append(i)
                lefs.append(lef)
                rigs.append(rig)
                locs.append(0)
                tsts.append(0)
                tfins.append(0)
        if not rec:
            continue
        ri = torch.as_tensor(rec, device=dev)
        lt = torch.as_tensor(lefs, device=dev)[None, :]
        rt = torch.as_tensor(rigs, device=dev)[None, :]
        win = ((tt[:, None] + 1 >= lt) & (tt[:, None] + 1 <= rt)).to(torch.float64)     # [nt][nrec]
        u0 = accf[:, ri].real
        u = datf[:, ri].real
        u0dot = torch.fft.irfft(torch.fft.rfft(u0, dim=0) * dmul[:, None], n=nt, dim=0)
        num = (wts[:, None] * (win * u0dot * (u0 - u))).sum(0) * dt_min / 3.0
        den = (wts[:, None] * (win * u0dot ** 2)).sum(0) * dt_min / 3.0
        tau = num / den                                                   # [nrec], minutes
        cc = ccdot[:, ri]
        ccdotone = torch.fft.fft(cc * win, dim=0)                          # [nt][nrec]
        con = -tau / ((cc * cc * win).sum(0) * dt_min)                     # complex, as the reference's sum
        xr = torch.as_tensor(x[rec], device=dev)
        E = torch.exp(-eyekh[None, :] * xr[:, None])                      # [nrec][nx]
        spec = fm * ((ccdotone * con[None, :]) @ E)
        adj = adj + torch.flip((torch.fft.ifft2(spec) * scale).real, dims=[0])
        tau_h = tau.cpu().tolist()
        for j, i in enumerate(rec):
            meas.append(Measurement(pord, i, tau_h[j] * 60.0, lefs[j], rigs[j], locs[j], tsts[j], tfins[j]))
            misfit += tau_h[j] ** 2.0
            nmeas += 1
    adj = adj * nt / 60.0
    return AdjointSource(adj=adj.cpu().numpy(), misfit=0.5 * misfit, nmeas=nmeas, measurements=meas)
