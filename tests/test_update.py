"""Gradient post-processing and model update (sparc_flows_amd.update), on the
CPU.  grad.py needs astropy and matplotlib, absent here, so it is not run;
its pieces are pinned by what fixes them:

* the z smoothing equals scipy.ndimage.gaussian_filter1d (the routine
  grad.py calls) for several widths;
* the x low-pass multiplies a Fourier mode of wavenumber k by
  exp(-k^2 / (2*30^2));
* (anti)symmetrisation: odd/even parts, idempotent, each other's kernel;
* the 11-point z filter equals the stencil evaluated by brute force, with the
  reference's zero rows;
* Polak-Ribiere CG with exact line searches minimises a 2-D quadratic in two
  steps; the L-BFGS two-loop operator maps the newest gradient change to the
  newest model change (the secant equation) and minimises an n-D quadratic
  in n steps;
* the line-search model is the closed form, and the directory driver writes
  the reference's files and eps bookkeeping.
"""
import numpy as np
import pytest

from sparc_flows_amd import update as U
from sparc_flows_amd.fitsio import read_fits, write_fits


def test_gaussian_z_equals_scipy():
    from scipy.ndimage import gaussian_filter1d
    rng = np.random.default_rng(0)
    a = rng.standard_normal((40, 12))
    for sigma in (0.3, 1.0, 2.0, 3.7):
        ref = gaussian_filter1d(a, sigma=sigma, axis=0)
        assert np.allclose(U.gaussian_filter_z(a, sigma), ref, rtol=0, atol=1e-14), sigma


def test_filterx_on_fourier_modes():
    nx, nz = 64, 5
    x = np.arange(nx)
    for k in (0, 3, 17, 32):
        f = np.cos(2 * np.pi * k * x / nx)[None, :].repeat(nz, 0)
        out = U.filterx(f, 30.0)
        assert np.allclose(out, np.exp(-k ** 2 / (2 * 30.0 ** 2)) * f, atol=1e-13)


def test_symmetry_parts():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((6, 16))
    odd, even = U.antisymmetrize(a), U.symmetrize(a)
    assert np.allclose(odd + even, a)
    assert np.allclose(odd[:, ::-1], -odd) and np.allclose(even[:, ::-1], even)
    assert np.allclose(U.antisymmetrize(odd), odd) and np.allclose(U.antisymmetrize(even), 0)


def test_smooth_z_brute_force():
    rng = np.random.default_rng(2)
    nz = 30
    a = rng.standard_normal((nz, 3))
    out = U.smooth_z(a)
    c = U._SMOOTH
    for k in range(nz):
        if 5 <= k <= nz - 6:
            ref = c[0] * a[k] + sum(0.5 * c[m] * (a[k - m] + a[k + m]) for m in range(1, 6))
        elif 1 <= k <= 3 or nz - 5 <= k <= nz - 2:
            ref = 0.5 * (a[k - 1] + a[k + 1])
        elif k == 0:
            ref = a[0]
        else:                       # rows 4 and nz-1: never set in the reference
            ref = np.zeros(3)
        assert np.allclose(out[k], ref, atol=1e-14), k


def test_preconditioner_and_gradient_shape():
    rng = np.random.default_rng(3)
    H = [rng.standard_normal((20, 16)) for _ in range(3)]
    rho = np.linspace(1.0, 0.01, 20)
    h = U.preconditioner(H, rho)
    assert abs(h.max() - 1.0) < 1e-15 and h.min() >= 5e-3
    tot = np.abs(H[0]) + np.abs(H[1]) + np.abs(H[2])
    ref = tot * rho[:, None]
    ref = np.maximum(ref / ref.max(), 5e-3)
    assert np.allclose(h, ref)
    K = [rng.standard_normal((20, 16)) for _ in range(3)]
    g = U.gradient_psi(K, H, rho)
    assert g.shape == (20, 16) and np.allclose(g[:, ::-1], -g)


def _exact_step(A, x, p):
    g = A @ x
    return x - (g @ p) / (p @ A @ p) * p


def test_polak_ribiere_cg_two_steps_on_quadratic():
    A = np.array([[3.0, 1.0], [1.0, 2.0]])
    x0 = np.array([1.0, -2.0])
    g0 = A @ x0
    p0 = U.sd_update(g0)
    x1 = _exact_step(A, x0, p0)
    g1 = A @ x1
    p1 = U.cg_update(g1, g0, p0)
    x2 = _exact_step(A, x1, p1)
    assert np.linalg.norm(x2) < 1e-12 * np.linalg.norm(x0)


def test_lbfgs_secant_and_finite_termination():
    rng = np.random.default_rng(4)
    n = 4
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    xs = [rng.standard_normal(n)]
    gs = [A @ xs[0]]
    for _ in range(n):
        d = U.lbfgs_direction(gs[-1], xs, gs, m=n)
        xs.append(_exact_step(A, xs[-1], d))
        gs.append(A @ xs[-1])
        if len(xs) >= 2:
            s, y = xs[-1] - xs[-2], gs[-1] - gs[-2]
            # secant equation of the two-loop operator for the newest pair
            assert np.allclose(-U.lbfgs_direction(y, xs, gs, m=n), s, rtol=1e-10, atol=1e-12)
    assert np.linalg.norm(xs[-1]) < 1e-9 * np.linalg.norm(xs[0])


def test_linesearch_model_closed_form():
    m = np.array([[1.0, -2.0], [3.0, 0.5]])
    u = np.array([[0.5, -4.0], [2.0, 1.0]])
    t = U.linesearch_model(m, u, 0.03)
    assert np.allclose(t, m + 0.03 * (u / 2.0) * np.sqrt(np.mean(m ** 2)))
    assert np.allclose(U.linesearch_model(np.zeros((2, 2)), u, 0.1), 0.1 * u / 2.0)


def test_directory_driver(tmp_path):
    rng = np.random.default_rng(5)
    nz, nx, nsrc = 24, 32, 2
    (tmp_path / "kernel").mkdir()
    (tmp_path / "update").mkdir()
    (tmp_path / "master.pixels").write_text("-10.0\n15.0\n")
    for s in (1, 2):
        write_fits(tmp_path / "kernel" / f"kernel_psi_{s:02d}.fits", rng.standard_normal((nz, 1, nx)))
        write_fits(tmp_path / "kernel" / f"hessian_{s:02d}.fits", rng.standard_normal((nz, 1, nx)))
    model = rng.standard_normal((nz, nx))
    write_fits(tmp_path / "update" / "model_psi_00.fits", model[:, None, :])
    (tmp_path / "update" / "misfit_00").write_text("0\n")
    rho = np.linspace(1.0, 0.1, nz)
    r = U.run(tmp_path, rho, algo="cg", eps=[0.01, 0.02])
    assert r["iteration"] == 0
    g = read_fits(tmp_path / "update" / "gradient_psi_00.fits")[:, 0, :]
    u = read_fits(tmp_path / "update" / "update_psi_00.fits")[:, 0, :]
    assert np.allclose(u, -g)                       # iteration 0 is steepest descent
    t2 = read_fits(tmp_path / "update" / "test_psi_2.fits")[:, 0, :]
    assert np.allclose(t2, U.linesearch_model(model, u, 0.02))
    with np.load(tmp_path / "epslist.npz") as f:
        assert np.allclose(f["0"][0], [0.01, 0.02]) and np.allclose(f["0"][1:], 0)
    # second pass of the same iteration shifts the eps rows (grad.py:405-410)
    U.run(tmp_path, rho, algo="cg", eps=[0.03, 0.04])
    with np.load(tmp_path / "epslist.npz") as f:
        assert np.allclose(f["0"][0], [0.03, 0.04]) and np.allclose(f["0"][2], [0.01, 0.02])
    # iteration 1 with CG uses the previous gradient and update
    (tmp_path / "update" / "misfit_01").write_text("0\n")
    write_fits(tmp_path / "update" / "model_psi_01.fits", (model + 0.1)[:, None, :])
    r1 = U.run(tmp_path, rho, algo="cg", eps=[0.01])
    g1 = r1["gradient"]
    beta = np.sum(g1 * (g1 - g)) / np.sum(g * g)
    assert np.allclose(r1["update"], -g1 + beta * u)
    with pytest.raises(ValueError):
        U.run(tmp_path, rho, algo="newton", eps=[0.01])
