"""The `sparc SS LL` drop-in (sparc_flows_amd.cli): the reference's command
line, instruction and flag files, FITS outputs and exit code.

The CPU tests cover the parameter file and the argument/instruction
handling; the GPU test runs the three phases of full.py:17-61 (forward,
adjoint, kernel) as child processes in a scratch directory and compares the
files they write with the same computation through the Python API.
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
    assert np.array_equal(vz, vz_ref[0].cpu().numpy())
    assert np.array_equal(np.fromfile(fdir / "vz_cc.bin", dtype="<f8").reshape(vz.shape), vz)
    assert (fdir / "kernel_info").exists() and (fdir / f"xix_{0:06d}_partial.fits").exists()
    from sparc_flows_amd.adjsrc import adjoint_source_filt
    r = adjoint_source_filt(vz, vz_obs[0].cpu().numpy(), np.arange(p.nx) / p.nx, -12.0, p.xlength,
                            p.outputcad, p.steps * p.timestep / 60.0, {8: (3.0, 40.0, 10.0)})
    assert r.nmeas > 0
    adir = data / "adjoint_src01"
    vr = read_fits(adir / "source.fits")[:, 0, :]
    assert np.array_equal(vr, r.adj)
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
