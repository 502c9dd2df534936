"""Gradient post-processing and model update of the inversion loop
(grad.py:1-433): the step between the kernels of all sources and the
line-search models of the next iteration.

    python -m sparc_flows_amd.update [algo=sd|cg|LBFGS] [device=cuda] [eps1 eps2 ...]

run in the directory that holds ``params.i`` (or with SPARC_DIRECTORY set),
as ``python grad.py`` is.  Arrays here are numpy (nz, nx) -- the reference
transposes its FITS arrays to (nx, 1, nz) (grad.py:12-27); every function
states its axis.  The work is a few passes over nx*nz doubles per iteration;
it runs on the host with numpy, or on a torch device (device=cuda: the *_t
functions, the same steps on GPU tensors, for an inversion loop whose
kernels are already on the GPU).

What it does, in the reference's order:

* sum K_psi over the sources, and |H| over the sources (grad.py:184-189);
* preconditioner: hess * rho0(z), normalised to max 1, floored at 5e-3
  (grad.py:191-193);
* gradient = filter_and_symmetrize(sum K_psi / hess): Gaussian low-pass in
  x wavenumber (sigma 30 wavenumbers, grad.py:43-53), antisymmetrised in x
  (grad.py:100), Gaussian smoothing in z of 2 pixels (scipy's
  gaussian_filter1d, reflect boundary, truncate 4; grad.py:93-94)
  (grad.py:109-118, 197);
* update direction: steepest descent, or Polak-Ribiere conjugate gradient
  against the previous iteration (grad.py:203-250), or L-BFGS (below);
* line-search models model + eps * (update / max(update)) * rms(model)
  (grad.py:343-357), one per eps;
* the eps list bookkeeping in ``epslist.npz`` (grad.py:400-426).

L-BFGS: the reference's branch (grad.py:252-328) cannot run -- it calls
``np.dot`` with one argument, reads models through ``read_grad`` and stores
rho in place of alpha -- so this module implements the standard two-loop
recursion (Nocedal & Wright, Algorithm 7.4) with m = 4 pairs from the stored
models and gradients, H0 = s.y / y.y of the newest pair.
"""
from __future__ import annotations

import fnmatch
import os
import sys
from pathlib import Path

import numpy as np

from .fitsio import read_fits, write_fits

RSUN_MM = 6.95989467700e2


# --------------------------------------------------------------------------
# Smoothing and symmetry (axis 0 = z, axis 1 = x)


def filterx(kern: np.ndarray, nk: float = 30.0) -> np.ndarray:
    """Gaussian low-pass in x wavenumber, exp(-k^2 / (2 nk^2)) with k the
    integer wavenumber 0..nx/2 (grad.py:43-53)."""
    nx = kern.shape[-1]
    k = np.fft.rfftfreq(nx, 1.0 / nx)
    filt = np.exp(-k ** 2 / (2.0 * nk ** 2))
    return np.fft.irfft(np.fft.rfft(kern, axis=-1) * filt, n=nx, axis=-1)


def antisymmetrize(a: np.ndarray) -> np.ndarray:
    """0.5 (a - a reversed in x), index i <-> nx-1-i (grad.py:100)."""
    return 0.5 * (a - a[..., ::-1])


def symmetrize(a: np.ndarray) -> np.ndarray:
    """0.5 (a + a reversed in x) (grad.py:101)."""
    return 0.5 * (a + a[..., ::-1])


def gaussian_filter_z(a: np.ndarray, sigma: float, truncate: float = 4.0) -> np.ndarray:
    """Gaussian smoothing along z (axis 0) with the conventions of
    scipy.ndimage.gaussian_filter1d: weights exp(-t^2 / (2 sigma^2)) over
    |t| <= int(truncate * sigma + 0.5), normalised to sum 1, and the 'reflect'
    boundary (the row before the first is the first, d c b a | a b c d)."""
    r = int(truncate * float(sigma) + 0.5)
    t = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * t ** 2 / float(sigma) ** 2)
    w /= w.sum()
    nz = a.shape[0]
    idx = np.arange(-r, nz + r)
    period = 2 * nz
    m = np.mod(idx, period)
    src = np.where(m < nz, m, period - 1 - m)
    pad = a[src]
    out = np.zeros_like(a, dtype=np.float64)
    for q in range(2 * r + 1):
        out += w[q] * pad[q:q + nz]
    return out


_SMOOTH = (0.75390625, 0.41015625, -0.234375, 0.087890625, -0.01953125, 0.001953125)


def smooth_z(a: np.ndarray) -> np.ndarray:
    """The 11-point z filter of filterz(algo='smooth') (grad.py:58-83): the
    stencil on rows 5..nz-6, two-point means on rows 1..3 and nz-5..nz-2,
    row 0 kept.  The reference copies rows 1..nz-1 of its scratch array
    back, and rows 4 and nz-1 of that array are never set, so they become
    0; this keeps that."""
    nz = a.shape[0]
    c = _SMOOTH
    temp = np.zeros_like(a)
    for k in range(5, nz - 5):
        temp[k] = (c[0] * a[k] + 0.5 * c[1] * (a[k - 1] + a[k + 1]) + 0.5 * c[2] * (a[k - 2] + a[k + 2])
                   + 0.5 * c[3] * (a[k - 3] + a[k + 3]) + 0.5 * c[4] * (a[k - 4] + a[k + 4])
                   + 0.5 * c[5] * (a[k - 5] + a[k + 5]))
    temp[nz - 5:nz - 1] = 0.5 * (a[nz - 4:] + a[nz - 6:nz - 2])
    temp[1:4] = 0.5 * (a[:3] + a[2:5])
    out = a.copy()
    out[1:nz] = temp[1:nz]
    return out


def filter_and_symmetrize(totkern: np.ndarray, hess: np.ndarray, sym: str | None = None,
                          z_filt_algo: str = "gaussian", z_filt_pix: float = 0.3) -> np.ndarray:
    """kern = totkern / hess, x low-pass (30 wavenumbers), optional
    (anti)symmetrisation in x, z smoothing (grad.py:109-118)."""
    kern = filterx(totkern / hess, 30.0)
    if sym == "sym":
        kern = symmetrize(kern)
    elif sym == "asym":
        kern = antisymmetrize(kern)
    if z_filt_algo == "gaussian":
        kern = gaussian_filter_z(kern, z_filt_pix)
    elif z_filt_algo == "smooth":
        kern = smooth_z(kern)
    else:
        raise ValueError(f"z filter {z_filt_algo!r} is not available (gaussian, smooth)")
    return kern


def preconditioner(hessians, rho0: np.ndarray, floor: float = 5e-3) -> np.ndarray:
    """sum |H_src| * rho0(z), normalised to max 1 and floored (grad.py:189-193)."""
    h = np.zeros_like(np.asarray(hessians[0], dtype=np.float64))
    for H in hessians:
        h += np.abs(H)
    h = h * np.asarray(rho0, dtype=np.float64)[:, None]
    h = h / np.abs(h).max()
    h[h < floor] = floor
    return h


def gradient_psi(kernels_psi, hessians, rho0: np.ndarray) -> np.ndarray:
    """The psi gradient of one iteration (grad.py:184-199)."""
    tot = np.zeros_like(np.asarray(kernels_psi[0], dtype=np.float64))
    for K in kernels_psi:
        tot += K
    return filter_and_symmetrize(tot, preconditioner(hessians, rho0), sym="asym",
                                 z_filt_algo="gaussian", z_filt_pix=2.0)


# --------------------------------------------------------------------------
# Update directions


def sd_update(grad: np.ndarray) -> np.ndarray:
    """Steepest descent (grad.py:206-210)."""
    return -grad


def polak_ribiere_beta(grad: np.ndarray, lastgrad: np.ndarray) -> float:
    """beta = g.(g - g_prev) / g_prev.g_prev (grad.py:222-235)."""
    return float(np.sum(grad * (grad - lastgrad)) / np.sum(lastgrad ** 2))


def cg_update(grad: np.ndarray, lastgrad: np.ndarray, lastupdate: np.ndarray) -> np.ndarray:
    """-g + beta p_prev, Polak-Ribiere (grad.py:237-245)."""
    return -grad + polak_ribiere_beta(grad, lastgrad) * lastupdate


def lbfgs_direction(grad: np.ndarray, models, grads, m: int = 4) -> np.ndarray:
    """Two-loop L-BFGS: returns -H grad from the last m (model, gradient)
    pairs; models/grads are lists ordered oldest to newest, the last entry
    being the current iterate."""
    s_list, y_list = [], []
    for i in range(max(len(models) - 1 - m, 0), len(models) - 1):
        s_list.append(models[i + 1] - models[i])
        y_list.append(grads[i + 1] - grads[i])
    q = grad.astype(np.float64).copy()
    alphas, rhos = [], []
    for s, y in zip(reversed(s_list), reversed(y_list)):
        rho = 1.0 / float(np.sum(y * s))
        a = rho * float(np.sum(s * q))
        q -= a * y
        alphas.append(a)
        rhos.append(rho)
    if s_list:
        s, y = s_list[-1], y_list[-1]
        q *= float(np.sum(s * y)) / float(np.sum(y * y))
    for (s, y), a, rho in zip(zip(s_list, y_list), reversed(alphas), reversed(rhos)):
        b = rho * float(np.sum(y * q))
        q += s * (a - b)
    return -q


def rms(a: np.ndarray) -> float:
    return float(np.sqrt(np.sum(a ** 2) / a.size))


def linesearch_model(model: np.ndarray, update: np.ndarray, eps: float, kind: str = "linear") -> np.ndarray:
    """model + eps * (update / max(update)) * rms(model) (grad.py:343-357);
    the update is divided by its maximum (not its absolute maximum) unless
    that is 0, and rms(model) = 0 counts as 1."""
    u = np.asarray(update, dtype=np.float64)
    umax = u.max()
    if umax != 0:
        u = u / umax
    if kind != "linear":
        raise ValueError("only the linear line-search model is defined (grad.py:350-355)")
    scale = rms(model)
    if scale == 0:
        scale = 1.0
    return model + eps * u * scale


# --------------------------------------------------------------------------
# The same steps on a torch device (the GPU in the inversion loop, where the
# kernels already are): every function above with a torch tensor in and out,
# the same arithmetic in the same order of operations except for the order
# of the sums (rounding only).


def _t(a, device):
    import torch
    return torch.as_tensor(np.asarray(a, dtype=np.float64) if not torch.is_tensor(a) else a,
                           dtype=torch.float64, device=device)


def filterx_t(kern, nk: float = 30.0):
    import torch
    nx = kern.shape[-1]
    k = torch.arange(nx // 2 + 1, dtype=torch.float64, device=kern.device)
    filt = torch.exp(-k ** 2 / (2.0 * nk ** 2))
    return torch.fft.irfft(torch.fft.rfft(kern, dim=-1) * filt, n=nx, dim=-1)


def antisymmetrize_t(a):
    return 0.5 * (a - a.flip(-1))


def symmetrize_t(a):
    return 0.5 * (a + a.flip(-1))


def gaussian_filter_z_t(a, sigma: float, truncate: float = 4.0):
    import torch
    r = int(truncate * float(sigma) + 0.5)
    t = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * t ** 2 / float(sigma) ** 2)
    w /= w.sum()
    nz = a.shape[0]
    m = np.mod(np.arange(-r, nz + r), 2 * nz)
    src = torch.as_tensor(np.where(m < nz, m, 2 * nz - 1 - m), device=a.device)
    pad = a[src]
    out = torch.zeros_like(a)
    for q in range(2 * r + 1):
        out = out + float(w[q]) * pad[q:q + nz]
    return out


def smooth_z_t(a):
    c = _SMOOTH
    nz = a.shape[0]
    temp = a.new_zeros(a.shape)
    k = slice(5, nz - 5)
    sh = lambda d: a[5 + d:nz - 5 + d]
    temp[k] = (c[0] * sh(0) + 0.5 * c[1] * (sh(-1) + sh(1)) + 0.5 * c[2] * (sh(-2) + sh(2))
               + 0.5 * c[3] * (sh(-3) + sh(3)) + 0.5 * c[4] * (sh(-4) + sh(4))
               + 0.5 * c[5] * (sh(-5) + sh(5)))
    temp[nz - 5:nz - 1] = 0.5 * (a[nz - 4:] + a[nz - 6:nz - 2])
    temp[1:4] = 0.5 * (a[:3] + a[2:5])
    out = a.clone()
    out[1:nz] = temp[1:nz]
    return out


def filter_and_symmetrize_t(totkern, hess, sym: str | None = None, z_filt_algo: str = "gaussian",
                            z_filt_pix: float = 0.3):
    kern = filterx_t(totkern / hess, 30.0)
    if sym == "sym":
        kern = symmetrize_t(kern)
    elif sym == "asym":
        kern = antisymmetrize_t(kern)
    if z_filt_algo == "gaussian":
        return gaussian_filter_z_t(kern, z_filt_pix)
    if z_filt_algo == "smooth":
        return smooth_z_t(kern)
    raise ValueError(f"z filter {z_filt_algo!r} is not available (gaussian, smooth)")


def preconditioner_t(hessians, rho0, floor: float = 5e-3):
    import torch
    h = torch.zeros_like(hessians[0])
    for H in hessians:
        h = h + H.abs()
    h = h * rho0[:, None]
    h = h / h.abs().max()
    return torch.where(h < floor, torch.full_like(h, floor), h)


def gradient_psi_t(kernels_psi, hessians, rho0, device="cuda"):
    """gradient_psi on a torch device; inputs arrays or tensors."""
    import torch
    K = [_t(k, device) for k in kernels_psi]
    tot = torch.zeros_like(K[0])
    for k in K:
        tot = tot + k
    return filter_and_symmetrize_t(tot, preconditioner_t([_t(h, device) for h in hessians], _t(rho0, device)),
                                   sym="asym", z_filt_algo="gaussian", z_filt_pix=2.0)


def cg_update_t(grad, lastgrad, lastupdate):
    beta = (grad * (grad - lastgrad)).sum() / (lastgrad ** 2).sum()
    return -grad + beta * lastupdate


def lbfgs_direction_t(grad, models, grads, m: int = 4):
    s_list, y_list = [], []
    for i in range(max(len(models) - 1 - m, 0), len(models) - 1):
        s_list.append(models[i + 1] - models[i])
        y_list.append(grads[i + 1] - grads[i])
    q = grad.clone()
    alphas, rhos = [], []
    for s, y in zip(reversed(s_list), reversed(y_list)):
        rho = 1.0 / (y * s).sum()
        a = rho * (s * q).sum()
        q = q - a * y
        alphas.append(a)
        rhos.append(rho)
    if s_list:
        s, y = s_list[-1], y_list[-1]
        q = q * ((s * y).sum() / (y * y).sum())
    for (s, y), a, rho in zip(zip(s_list, y_list), reversed(alphas), reversed(rhos)):
        b = rho * (y * q).sum()
        q = q + s * (a - b)
    return -q


def linesearch_model_t(model, update, eps: float, kind: str = "linear"):
    import torch
    if kind != "linear":
        raise ValueError("only the linear line-search model is defined (grad.py:350-355)")
    umax = update.max()
    u = torch.where(umax != 0, update / torch.where(umax != 0, umax, torch.ones_like(umax)), update)
    scale = torch.sqrt((model ** 2).sum() / model.numel())
    scale = torch.where(scale == 0, torch.ones_like(scale), scale)
    return model + eps * u * scale


# --------------------------------------------------------------------------
# The grad.py driver over the reference's directory layout


def _rd(path) -> np.ndarray:
    a = read_fits(path)
    return a[:, 0, :] if a.ndim == 3 else a


def _wr(path, a: np.ndarray) -> None:
    write_fits(path, np.asarray(a, dtype=np.float64)[:, None, :])


def iteration_number(datadir: Path) -> int:
    """Number of update/misfit_NN files minus one (grad.py:29-33)."""
    return len(fnmatch.filter(os.listdir(datadir / "update"), "misfit_[0-9][0-9]")) - 1


def run(datadir, rho0: np.ndarray, algo: str = "conjugate gradient", eps=None, m: int = 4,
        device=None) -> dict:
    """One pass of grad.py over `datadir`: reads kernel/kernel_psi_SS.fits,
    kernel/hessian_SS.fits, master.pixels and update/model_psi_NN.fits (and
    the previous gradient/update for CG, all previous models/gradients for
    L-BFGS); writes update/gradient_psi_NN.fits, update/update_psi_NN.fits,
    update/test_psi_N.fits and epslist.npz.  Returns what it wrote (numpy
    arrays).  device: a torch device to do the array work on (the *_t
    functions), or None for numpy."""
    if device is None:
        grad_fn, cg_fn, lb_fn, ls_fn = gradient_psi, cg_update, lbfgs_direction, linesearch_model
        conv = back = (lambda a: a)
    else:
        def grad_fn(K, H, r):
            return gradient_psi_t(K, H, r, device)
        cg_fn, lb_fn, ls_fn = cg_update_t, lbfgs_direction_t, linesearch_model_t
        conv = (lambda a: _t(a, device))
        back = (lambda a: a.cpu().numpy())
    datadir = Path(datadir)
    upd = datadir / "update"
    it = iteration_number(datadir)
    nsrc = len(np.loadtxt(datadir / "master.pixels", ndmin=1))
    if eps is None or len(eps) == 0:
        nls = len(fnmatch.filter(os.listdir(datadir), "forward_src01_ls[0-9][1-9]"))
        eps = [0.01 * i for i in range(1, nls + 1)]
    eps = [float(e) for e in eps]
    kern = [_rd(datadir / "kernel" / f"kernel_psi_{s:02d}.fits") for s in range(1, nsrc + 1)]
    hess = [_rd(datadir / "kernel" / f"hessian_{s:02d}.fits") for s in range(1, nsrc + 1)]
    g = grad_fn(kern, hess, rho0)
    _wr(upd / f"gradient_psi_{it:02d}.fits", back(g))
    sd = algo in ("sd", "steepest descent")
    cg = algo in ("cg", "conjugate gradient")
    if it == 0 or sd:
        u = -g
    elif cg:
        u = cg_fn(g, conv(_rd(upd / f"gradient_psi_{it - 1:02d}.fits")),
                  conv(_rd(upd / f"update_psi_{it - 1:02d}.fits")))
    elif algo == "LBFGS":
        lo = max(it - m, 0)
        models = [conv(_rd(upd / f"model_psi_{i:02d}.fits")) for i in range(lo, it + 1)]
        grads = [conv(_rd(upd / f"gradient_psi_{i:02d}.fits")) for i in range(lo, it + 1)]
        u = lb_fn(g, models, grads, m)
    else:
        raise ValueError(f"unknown algorithm {algo!r}")
    _wr(upd / f"update_psi_{it:02d}.fits", back(u))
    model = conv(_rd(upd / f"model_psi_{it:02d}.fits"))
    tests = []
    for i, e in enumerate(eps):
        t = back(ls_fn(model, u, e))
        _wr(upd / f"test_psi_{i + 1}.fits", t)
        tests.append(t)
    # epslist.npz: per iteration a 4 x nls array, row 0 the eps of the newest
    # line search, rows 2-3 the previous rows 0-1 (grad.py:400-426)
    path = datadir / "epslist.npz"
    lists = {}
    if path.exists():
        with np.load(path) as f:
            lists = {k: f[k] for k in f.files}
    key = str(it)
    if key in lists:
        arr = lists[key].copy()
        arr[2:4] = arr[0:2]
        arr[0] = eps
        arr[1] = 0
    else:
        arr = np.zeros((4, len(eps)))
        arr[0] = eps
    lists[key] = arr
    np.savez(path, **lists)
    return {"iteration": it, "gradient": back(g), "update": back(u), "tests": tests, "eps": eps}


def main(argv=None) -> int:
    from .background import read_model_file
    from .params import read_params_i

    argv = sys.argv[1:] if argv is None else argv
    algo = "conjugate gradient"
    eps = []
    device = None
    for a in argv:
        if a.startswith("algo="):
            algo = a.split("=", 1)[1]
        elif a.startswith("device="):
            device = a.split("=", 1)[1]          # e.g. device=cuda: the array work on the GPU
        else:
            try:
                eps.append(float(a))
            except ValueError:
                pass
    p = read_params_i("params.i")
    datadir = Path(os.environ.get("SPARC_DIRECTORY") or p.directory or ".")
    bg = read_model_file(p.file_data, p.nz)
    r = run(datadir, bg[:, 2], algo=algo, eps=eps, device=device)
    print(f"iteration {r['iteration']}: gradient, update and {len(r['tests'])} line-search models written")
    return 0


if __name__ == "__main__":
    sys.exit(main())
