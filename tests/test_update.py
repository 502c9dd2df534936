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


def _compare_device(device, tmp_path):
    """The torch versions (*_t) and the driver on a torch device against the
    numpy functions: same results to 1e-13 (order of sums only), for the psi
    gradient, both z filters, CG, L-BFGS, the line-search models and two
    iterations of the directory driver with each algorithm."""
    import torch
    rng = np.random.default_rng(11)
    nz, nx = 40, 64
    K = [rng.standard_normal((nz, nx)) for _ in range(3)]
    H = [rng.standard_normal((nz, nx)) for _ in range(3)]
    rho = np.linspace(2.0, 0.2, nz)
    g = U.gradient_psi(K, H, rho)
    gt = U.gradient_psi_t(K, H, rho, device)
    assert gt.device.type == torch.device(device).type
    close = lambda a, b: np.abs(a.cpu().numpy() - b).max() <= 1e-13 * np.abs(b).max()
    assert close(gt, g)
    a = rng.standard_normal((nz, nx))
    at = torch.as_tensor(a, device=device)
    assert close(U.smooth_z_t(at), U.smooth_z(a))
    assert close(U.filter_and_symmetrize_t(at, at.abs() + 1.0, "sym", "smooth"),
                 U.filter_and_symmetrize(a, np.abs(a) + 1.0, "sym", "smooth"))
    g0, u0 = rng.standard_normal((nz, nx)), rng.standard_normal((nz, nx))
    T = lambda v: torch.as_tensor(v, device=device)
    assert close(U.cg_update_t(T(g), T(g0), T(u0)), U.cg_update(g, g0, u0))
    models = [rng.standard_normal((nz, nx)) for _ in range(6)]
    grads = [rng.standard_normal((nz, nx)) for _ in range(6)]
    assert close(U.lbfgs_direction_t(T(g), [T(v) for v in models], [T(v) for v in grads], 4),
                 U.lbfgs_direction(g, models, grads, 4))
    assert close(U.linesearch_model_t(T(a), T(g), 0.02), U.linesearch_model(a, g, 0.02))
    # the directory driver, numpy and device runs in two copies of one layout
    for algo in ("cg", "LBFGS"):
        outs = []
        for dev in (None, device):
            d = tmp_path / f"{algo}_{dev}"
            (d / "kernel").mkdir(parents=True)
            (d / "update").mkdir()
            (d / "master.pixels").write_text("-10.0\n15.0\n")
            r2 = np.random.default_rng(5)
            for s in (1, 2):
                write_fits(d / "kernel" / f"kernel_psi_{s:02d}.fits", r2.standard_normal((nz, 1, nx)))
                write_fits(d / "kernel" / f"hessian_{s:02d}.fits", r2.standard_normal((nz, 1, nx)))
            model = r2.standard_normal((nz, nx))
            write_fits(d / "update" / "model_psi_00.fits", model[:, None, :])
            (d / "update" / "misfit_00").write_text("0\n")
            U.run(d, rho, algo=algo, eps=[0.01, 0.02], device=dev)
            for s in (1, 2):                         # the next iteration's kernels
                write_fits(d / "kernel" / f"kernel_psi_{s:02d}.fits", r2.standard_normal((nz, 1, nx)))
            (d / "update" / "misfit_01").write_text("0\n")
            # a model step that is not orthogonal to the gradient change (a
            # constant step makes s.y ~ 0 and the L-BFGS scaling ill-conditioned)
            write_fits(d / "update" / "model_psi_01.fits",
                       (model + 0.1 * r2.standard_normal((nz, nx)))[:, None, :])
            outs.append(U.run(d, rho, algo=algo, eps=[0.01, 0.02], device=dev))
        ref, got = outs
        for key in ("gradient", "update"):
            assert np.abs(got[key] - ref[key]).max() <= 1e-12 * np.abs(ref[key]).max(), (algo, key)
        for t0, t1 in zip(ref["tests"], got["tests"]):
            assert np.abs(t1 - t0).max() <= 1e-12 * np.abs(t0).max()


def test_torch_update_matches_numpy_on_cpu(tmp_path):
    _compare_device("cpu", tmp_path)


@pytest.mark.gpu
def test_torch_update_matches_numpy_on_gpu(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _compare_device("cuda", tmp_path)
