"""The `sparc SS LL` drop-in (sparc_flows_amd.cli): the reference's command
line, instruction and flag files, FITS outputs and exit code.

The CPU tests cover the parameter file and the argument/instruction
handling; the GPU tests run the three phases of full.py:17-61 (forward,
adjoint, kernel) as child processes in a scratch directory, for a 2D and a
3D params.i, and compare the files they write with the same computation
through the Python API (whose kernel call pairs forward output k with
adjoint output nt-1-k, as kernels.f90:215, 256 read the partial files).
"""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from sparc_flows_amd import background as B
from sparc_flows_amd.cli import main
from sparc_flows_amd.fitsio import read_fits, write_fits
from sparc_flows_amd.params import SparcParams, read_params_i, write_params_i

ROOT = Path(__file__).resolve().parents[1]


def _setup(tmp: Path, nx=64, nz=60):
    bg = B.polytrope(nz)
    p = SparcParams(nx=nx, nz=nz, xlength=200e8, timestep=B.stable_timestep(bg), npmlbot=6, npmltop=6,
                    npmlhor=5, file_data="poly.dat")
    data = tmp / "data"
    data.mkdir()
    B.write_model_file(tmp / "poly.dat", bg)
    write_params_i(p, tmp / "params.i", str(data) + "/")
    (data / "master.pixels").write_text("-12.0\n7.5\n")
    psi = B.supergranule_psi(nx, p.xlength, bg)
    write_fits(data / "model_psi_ls00.fits", psi[:, None, :])
    return p, bg, psi, data


def test_params_file_round_trip(tmp_path):
    p, _, _, _ = _setup(tmp_path)
    assert read_params_i(tmp_path / "params.i") == p


def test_params_file_round_trip_3d(tmp_path):
    p = SparcParams(nx=32, ny=16, nz=40, xlength=100e8, ylength=50e8, npmlbot=6, npmltop=6)
    write_params_i(p, tmp_path / "params.i", str(tmp_path) + "/")
    assert read_params_i(tmp_path / "params.i") == p
    q = SparcParams(nx=32, ny=32, nz=40)                 # ylength = xlength, written symbolically
    write_params_i(q, tmp_path / "params.i", str(tmp_path) + "/")
    assert "ylength = xlength" in (tmp_path / "params.i").read_text()
    assert read_params_i(tmp_path / "params.i") == q


def test_usage_and_unknown_instruction(tmp_path, monkeypatch, capsys):
    assert main([]) == 1
    _setup(tmp_path)
    monkeypatch.chdir(tmp_path)
    (tmp_path / "Instruction_src01_ls00").write_text("bogus\n")
    assert main(["01", "00"]) == 1
    assert not (tmp_path / "Instruction_src01_ls00").exists()   # read and removed (driver.f90:243)


def _cli(tmp: Path, data: Path, maxtime: int, *args):
    env = dict(os.environ, SPARC_DIRECTORY=str(data) + "/", SPARC_MAXTIME=str(maxtime),
               PYTHONPATH=str(ROOT))
    r = subprocess.run([sys.executable, "-m", "sparc_flows_amd.cli", *args], cwd=tmp, env=env,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    return r.returncode


def _check_text_outputs(p, bg, psi, d, maxtime, adjoint, vr=None):
    """timeline, kernel_info and rms_hist of a forward or adjoint phase:
    * timeline: time and solar time of every output step (driver.f90:386-390),
      from -solartime (forward) or from final_time - (forcing_length - 1)
      cadforcing_step timestep (adjoint);
    * kernel_info: init and maxtime (forward, :249-252), the number of adjoint
      source slices and the last forcing time (adjoint, :283-292);
    * rms_hist: nor1 = sqrt(sum v_z^2) after every step (:364, :384), against
      the oracle stepping the same run."""
    import oracle
    steps, dt, cfs = p.steps, p.timestep, p.cadforcing_step
    tl = np.loadtxt(d / "timeline", ndmin=2)
    outs = np.arange(0, maxtime + 1, steps)
    assert np.array_equal(tl[:, 0].astype(int), outs)
    start = (p.solartime * 3600.0 - (p.forcing_length(maxtime) - 1.0) * cfs * dt) if adjoint \
        else -p.solartime * 3600.0
    assert np.allclose(tl[:, 1], start + (outs + 1) * dt, rtol=0, atol=1e-9 * p.solartime * 3600.0)
    ki = [int(v) for v in (d / "kernel_info").read_text().split()]
    assert ki == ([maxtime // cfs + 1, (maxtime // cfs) * cfs] if adjoint else [0, maxtime])
    rms = np.loadtxt(d / "rms_hist")
    assert rms.shape == (maxtime + 1,)
    o = oracle.Oracle(p, bg, psi=psi, adjoint=adjoint)
    if adjoint:
        o.set_adjoint_source(vr[:p.forcing_length(maxtime)], maxtime)      # the slices the phase reads
    else:
        o.set_forward_source(-12.0, maxtime)
    ref = np.empty(maxtime + 1)
    for t in range(maxtime + 1):
        o.step(t)
        ref[t] = np.sqrt((o.state()[3] ** 2).sum())
    assert np.abs(ref).max() > 0
    assert np.abs(rms - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.gpu
def test_forward_adjoint_kernel_phases(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sparc_flows_amd.solver import Sparc
    p, bg, psi, data = _setup(tmp_path)
    maxtime = 60 * p.steps
    # observed traces: the same source without the flow; the all-p-mode
    # measurement (params.8) between 3 and 40 Mm with a 10 min window
    s0 = Sparc(p, bg)
    vz_obs, _ = s0.forward([-12.0], maxtime)
    (data / "data").mkdir(exist_ok=True)
    write_fits(data / "data" / "01.fits", vz_obs[0].cpu().numpy()[:, None, :])
    (data / "params.8").write_text("3.0\n40.0\n10.0\n")
    # forward (instruction 'spectral'): traces, snapshots, adjoint source
    (tmp_path / "Instruction_src01_ls00").write_text("spectral\n")
    assert _cli(tmp_path, data, maxtime, "01", "00") == 0
    fdir = data / "forward_src01_ls00"
    vz = read_fits(fdir / "vz_cc.fits")[:, 0, :]
    s = Sparc(p, bg, psi=psi)
    vz_ref, _ = s.forward([-12.0], maxtime)
    # another process's solver, which times its own column-pass shape: equal to rounding
    vr0 = vz_ref[0].cpu().numpy()
    assert np.linalg.norm(vz - vr0) <= 1e-12 * np.linalg.norm(vr0)
    assert np.array_equal(np.fromfile(fdir / "vz_cc.bin", dtype="<f8").reshape(vz.shape), vz)
    assert (fdir / "kernel_info").exists() and (fdir / f"xix_{0:06d}_partial.fits").exists()
    _check_text_outputs(p, bg, psi, fdir, maxtime, adjoint=False)
    from sparc_flows_amd.adjsrc import adjoint_source_filt, adjoint_source_filt_torch
    args = (vz, vz_obs[0].cpu().numpy(), np.arange(p.nx) / p.nx, -12.0, p.xlength, p.outputcad,
            p.steps * p.timestep / 60.0, {8: (3.0, 40.0, 10.0)})
    r = adjoint_source_filt_torch(*args, device="cuda")
    assert r.nmeas > 0
    adir = data / "adjoint_src01"
    vr = read_fits(adir / "source.fits")[:, 0, :]
    assert np.array_equal(vr, r.adj)                   # the GPU routine the program ran
    rn = adjoint_source_filt(*args)                    # the numpy restatement
    assert rn.nmeas == r.nmeas and np.abs(vr - rn.adj).max() <= 1e-11 * np.abs(rn.adj).max()
    assert (fdir / "ttdiff.8").read_text().count("\n") == r.nmeas
    assert (data / "kernel" / "misfit_01_00").exists() and (adir / "information").exists()
    # MISFIT_ALL: three bands of 11 half misfits, and the band windows it picked
    from sparc_flows_amd.adjsrc import misfit_all
    ma = misfit_all(vz, vz_obs[0].cpu().numpy(), np.arange(p.nx) / p.nx, -12.0, p.xlength, p.outputcad,
                    p.steps * p.timestep / 60.0, {8: (3.0, 40.0, 10.0)})
    v = np.loadtxt(data / "kernel" / "misfit_all_01_00", comments="#")
    assert v.shape == (3, 11) and np.allclose(v, [b[3] for b in ma.bands], rtol=1e-15, atol=0)
    assert (fdir / "windows.all.08.1").read_text().count("\n") == len(ma.windows[(8, 1)])
    # adjoint (instruction 'adjoint') from that source.fits
    (tmp_path / "Instruction_src01_ls00").write_text("adjoint\n")
    assert _cli(tmp_path, data, maxtime, "01", "00") == 0
    assert (adir / "kernel_info").exists()
    _check_text_outputs(p, bg, psi, adir, maxtime, adjoint=True, vr=vr)
    # kernels (no instruction file)
    assert _cli(tmp_path, data, maxtime, "01", "00") == 0
    kc = read_fits(data / "kernel" / "kernel_c_01.fits")[:, 0, :]
    kpsi = read_fits(data / "kernel" / "kernel_psi_01.fits")[:, 0, :]
    assert not list(fdir.glob("*partial*"))                        # kernels.f90:591-597
    # the same three phases through the Python API
    _, fsn = s.forward([-12.0], maxtime, snapshots=True)
    sa = Sparc(p, bg, adjoint=True, psi=psi)
    _, asn = sa.adjoint(torch.as_tensor(vr[None], device="cuda"), maxtime, snapshots=True)
    K = Sparc(p, bg).kernels(fsn, asn)
    for got, ref in ((kc, K["kc"]), (kpsi, K["kpsi"])):
        ref = ref[0].cpu().numpy()
        assert np.abs(ref).max() > 0
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_flag_file_branches(tmp_path):
    """The working-directory flag files of a forward run (READ_PARAMETERS,
    initialize.f90:1036-1043):
    * compute_data: the flow is v0_x, v0_z in m/s from true_vx.fits /
      true_vz.fits (driver.f90:65-73), no partial snapshots (:412-414) and no
      adjoint source or misfit (:274);
    * linesearch: no snapshots; travel times in the windows a plain run wrote
      to forward_srcSS_ls00 (:2023-2029); source.fits and information are
      left alone (:2194, :2210); the misfit file is written;
    * compute_synth: no snapshots and no source.fits (:2194); the windows and
      information are written as in a plain run.
    Traces are compared with the Python API (another process's solver:
    rounding), files by presence and content."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sparc_flows_amd.solver import Sparc
    p, bg, psi, data = _setup(tmp_path)
    maxtime = 60 * p.steps
    s0 = Sparc(p, bg)
    vz_obs, _ = s0.forward([-12.0], maxtime)
    (data / "data").mkdir(exist_ok=True)
    write_fits(data / "data" / "01.fits", vz_obs[0].cpu().numpy()[:, None, :])
    (data / "params.8").write_text("3.0\n40.0\n10.0\n")
    for ls in ("01", "02", "03"):
        write_fits(data / f"model_psi_ls{ls}.fits", (0.8 * psi)[:, None, :])
    def forward(ls):
        (tmp_path / f"Instruction_src01_ls{ls}").write_text("spectral\n")
        assert _cli(tmp_path, data, maxtime, "01", ls) == 0
        return data / f"forward_src01_ls{ls}"

    def partials(d):
        return sorted(x.name for x in d.glob("*_partial.fits"))

    def api_vz(**kw):
        s = Sparc(p, bg, psi=kw.get("psi"))
        if "flow" in kw:
            s.set_flow(*kw["flow"])
        v, _ = s.forward([-12.0], maxtime)
        return v[0].cpu().numpy()

    def close(a, b):
        return np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)

    # a plain forward run: snapshots, windows in ls00, source.fits, information
    f0 = forward("00")
    adir = data / "adjoint_src01"
    assert partials(f0) and (f0 / "windows.8").exists() and (adir / "source.fits").exists()
    src0, info0 = (adir / "source.fits").read_bytes(), (adir / "information").read_text()
    win0 = (f0 / "windows.8").read_text()
    # compute_data: the true flow from true_vx / true_vz (m/s)
    sp = Sparc(p, bg, psi=psi)
    to_ms = 1e6 / 1e2
    vx = 1.5 * sp.background("v0x").reshape(p.nz, p.nx) * to_ms
    vz_t = 1.5 * sp.background("v0z").reshape(p.nz, p.nx) * to_ms
    write_fits(tmp_path / "true_vx.fits", vx[:, None, :])
    write_fits(tmp_path / "true_vz.fits", vz_t[:, None, :])
    (tmp_path / "compute_data").touch()
    fd = forward("01")
    (tmp_path / "compute_data").unlink()
    vz_d = read_fits(fd / "vz_cc.fits")[:, 0, :]
    assert close(vz_d, api_vz(flow=(vx, vz_t)))
    assert not close(vz_d, api_vz(psi=0.8 * psi))            # not the model_psi flow of ls01
    assert not partials(fd) and not list(fd.glob("ttdiff.*"))
    assert not (data / "kernel" / "misfit_01_01").exists()
    assert (adir / "source.fits").read_bytes() == src0
    # linesearch: ls00's windows, no source.fits / information, misfit written
    (tmp_path / "linesearch").touch()
    fl = forward("02")
    (tmp_path / "linesearch").unlink()
    assert close(read_fits(fl / "vz_cc.fits")[:, 0, :], api_vz(psi=0.8 * psi))
    assert not partials(fl)
    nwin = win0.count("\n")
    assert (fl / "ttdiff.8").read_text().count("\n") == nwin
    assert (f0 / "windows.8").read_text() == win0
    assert (adir / "source.fits").read_bytes() == src0 and (adir / "information").read_text() == info0
    assert (data / "kernel" / "misfit_01_02").exists()
    # compute_synth: no snapshots, no source.fits; windows and information written
    (tmp_path / "compute_synth").touch()
    fs = forward("03")
    (tmp_path / "compute_synth").unlink()
    assert not partials(fs) and (fs / "ttdiff.8").exists()
    assert (adir / "source.fits").read_bytes() == src0
    assert (f0 / "windows.8").exists() and (adir / "information").exists()
    assert (data / "kernel" / "misfit_01_03").exists()
    assert not any(tmp_path.glob("Instruction_*"))


@pytest.mark.gpu
def test_forward_adjoint_kernel_phases_3d(tmp_path):
    """A params.i with ny > 1: the three phases on the 3D solver (line source
    at master.pixels' x, a 3D model_c, a given adjoint source.fits) against
    the same runs through sparc_flows_amd.solver3d."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from sparc_flows_amd.solver3d import Params3D, Sparc3D
    nx, ny, nz = 32, 16, 40
    bg = B.polytrope(nz)
    p = SparcParams(nx=nx, ny=ny, nz=nz, xlength=100e8, ylength=50e8, timestep=B.stable_timestep(bg),
                    npmlbot=6, npmltop=6, file_data="poly.dat")
    data = tmp_path / "data"
    data.mkdir()
    B.write_model_file(tmp_path / "poly.dat", bg)
    write_params_i(p, tmp_path / "params.i", str(data) + "/")
    (data / "master.pixels").write_text("-12.0\n")
    x = (np.arange(nx) / nx - 0.5)[None, None, :]
    y = (np.arange(ny) / ny - 0.5)[None, :, None]
    zz = (np.arange(nz) / nz)[:, None, None]
    c3d = bg[:, 1][:, None, None] * (1 + 0.05 * np.exp(-(x ** 2 + y ** 2) / 0.03 - (zz - 0.7) ** 2 / 0.01))
    write_fits(data / "model_c_ls00.fits", c3d)
    q = Params3D(nx=nx, ny=ny, nz=nz, xlength=100e8, ylength=50e8, timestep=p.timestep, npmlbot=6, npmltop=6,
                 cadforcing=p.cadforcing, outputcad=p.outputcad, obsheight=p.obsheight, excitdep=p.excitdep,
                 nupeak=p.nupeak, nuwidth=p.nuwidth)
    steps = q.steps
    maxtime = 12 * steps
    # forward
    (tmp_path / "Instruction_src01_ls00").write_text("spectral\n")
    assert _cli(tmp_path, data, maxtime, "01", "00") == 0
    fdir = data / "forward_src01_ls00"
    vz = read_fits(fdir / "vz_cc.fits")
    s = Sparc3D(q, bg, c3d=c3d)
    vz_ref = s.forward(-12.0, None, maxtime).cpu().numpy()
    assert vz.shape == (maxtime // steps + 1, ny, nx) and np.abs(vz_ref).max() > 0
    assert np.array_equal(vz, vz_ref)
    assert np.array_equal(np.fromfile(fdir / "vz_cc.bin", dtype="<f8").reshape(vz.shape), vz)
    # adjoint from a given source.fits, (nx, ny, nt) at the forcing cadence
    cfs = int(q.cadforcing // q.timestep)
    fl = maxtime // cfs + 1
    rng = np.random.default_rng(3)
    vr = rng.standard_normal((fl, ny, nx)) * np.exp(-(x ** 2) / 0.02)[0]
    adir = data / "adjoint_src01"
    adir.mkdir()
    write_fits(adir / "source.fits", vr)
    (tmp_path / "Instruction_src01_ls00").write_text("adjoint\n")
    assert _cli(tmp_path, data, maxtime, "01", "00") == 0
    assert (adir / "kernel_info").exists() and (adir / f"vy_{0:06d}_partial.fits").exists()
    # kernels
    assert _cli(tmp_path, data, maxtime, "01", "00") == 0
    assert not list(fdir.glob("*partial*")) and not list(adir.glob("*partial*"))
    # the same through the Python API: snapshots every `steps` steps
    fs, asn = [], []
    s.set_forward_source(-12.0, None, maxtime)
    s.reset()
    sa = Sparc3D(q, bg, c3d=c3d, adjoint=True)
    sa.set_adjoint_source(torch.as_tensor(vr, device="cuda"), maxtime)
    sa.reset()
    for t in range(maxtime + 1):
        s.run_range(t, t)
        sa.run_range(t, t)
        if t % steps == 0:
            fs.append(s.snapshot())
            asn.append(sa.snapshot())
    K = sa.kernels(torch.stack(fs), torch.stack(asn))
    for key, name in (("kc", "kernel_c"), ("kpsi", "kernel_psi"), ("kvx", "kernel_vx"), ("kvy", "kernel_vy"),
                      ("kvz", "kernel_vz"), ("hess", "hessian")):
        got = read_fits(data / "kernel" / f"{name}_01.fits")
        ref = K[key].cpu().numpy()
        assert np.abs(ref).max() > 0, key
        assert np.array_equal(got, ref), key
