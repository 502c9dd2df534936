"""ADJOINT_SOURCE_FILT restatement (sparc_flows_amd.adjsrc), on the CPU.

The reference routine cannot be run here (it needs FFTW3 and the full
program), so its parity is unpinned at the level of whole outputs; these
tests pin the pieces by what the mathematics fixes:

* distmat is the FFT index order with a positive Nyquist index;
* SIMPSON_REGULAR integrates cubics exactly on an odd number of points;
* the Gizon-Birch travel time of a delayed copy is the delay;
* the ridge filters are 1 at their band centre, 0 outside the band and
  below the low-frequency cut, even in k and nu;
* the adjoint source is the negative gradient of the misfit with respect to
  the synthetic traces (time reversed, scaled by nt/60 and by 60/dt as the
  reference writes it), checked against central differences along smooth
  directions.
"""
import numpy as np
import pytest

from sparc_flows_amd import adjsrc as A


def test_distmat_index_order():
    for n in (8, 9, 16):
        d = A.distmat(n)
        f = np.fft.fftfreq(n) * n
        if n % 2 == 0:
            f[n // 2] = n // 2
        assert np.array_equal(d, f)


def test_simpson_exact_for_cubics():
    x = np.linspace(0.0, 2.0, 21)
    fv = 1 + x - 3 * x ** 2 + 0.5 * x ** 3
    exact = 2 + 2 - 8 + 0.125 * 16
    assert abs(A.simpson_regular(fv, x[1] - x[0]) - exact) < 1e-12


def _wave(t, t0=40.0, f=0.2, w=8.0):
    return np.cos(2 * np.pi * f * (t - t0)) * np.exp(-((t - t0) / w) ** 2)


def test_gizon_birch_travel_time_of_a_delayed_copy():
    dt = 0.5
    t = np.arange(256) * dt
    for delay in (0.01, -0.02, 0.05):
        u0 = _wave(t)
        u = _wave(t, t0=40.0 + delay)
        tau = A.tt_gizon_birch(u0, u, dt, 40, 120)
        # u(t) = u0(t - delay): data later than the synthetic by `delay`
        assert abs(tau - delay) < 0.02 * abs(delay), (delay, tau)


def test_ridge_filters_shape():
    nt, nx, xl, cad = 128, 64, 200e8, 30.0
    k = A._kk(nx, xl)
    f = A._ff(nt, cad)
    for name in ("f", "p1", "p2", "p3", "p4", "p5", "p6", "p7"):
        F = A.ridge_filter(name, nt, nx, xl, cad)
        assert F.shape == (nt, nx)
        assert F.min() >= 0.0 and F.max() <= 1.0 + 1e-12
        # even in k and nu (index i <-> n - i)
        assert np.array_equal(F[1:, :], F[1:, :][::-1, :])
        assert np.array_equal(F[:, 1:], F[:, 1:][:, ::-1])
        f_low = A.f32(1.1 if name == "f" else 1.6)
        assert np.all(F[f < f_low, :] == 0.0)
        # inside the band and above the taper the filter is the raised cosine
        low, high, _ = A._RIDGES[name]
        f1 = A.f32(high[0]) + A.f32(high[1]) * k + A.f32(high[2]) * k ** 2
        f0 = (A.f32(low) * np.sqrt(k) if name == "p2"
              else A.f32(low[0]) + A.f32(low[1]) * k + A.f32(low[2]) * k ** 2)
        for j in range(nt):
            if f[j] < f_low + 0.5:
                continue
            for i in range(nx):
                d, delta = f[j] - f0[i], f1[i] - f0[i]
                if 0 < d < delta:
                    assert F[j, i] > 0.0
                    if abs(d - delta / 2) < 1e-9:
                        assert abs(F[j, i] - 1.0) < 1e-12
                else:
                    assert F[j, i] == 0.0
    allp = A.ridge_filter("allp", nt, nx, xl, cad)
    assert np.all(allp[f >= A.f32(1.6) + 0.5, :] == 1.0)


def test_freq_filter_band():
    nt, cad = 256, 30.0
    filt = A.freq_filter(2.0, 4.0, nt, cad)
    w = np.abs(A.distmat(nt)) * 1e3 / (nt * cad)
    assert np.all(filt[(w > 2.5) & (w < 3.5)] > 0.99)
    assert np.all(np.abs(filt[(w < 1.0) | (w > 5.0)]) < 1e-3)


def _traces(nt=240, nx=64, dt_min=0.5, seed=0, x00=0.0, xl=200e8, delay=0.02):
    """Synthetic traces: a dispersive-free wave leaving x00 at 1 Mm/min."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx) / nx
    xmm = (x - 0.5) * xl * A.f32(1e-8)
    t = np.arange(nt) * dt_min
    dist = np.abs(xmm - x00)
    acc = np.zeros((nt, nx))
    dat = np.zeros((nt, nx))
    for i in range(nx):
        arr = 10.0 + dist[i] / 1.0
        acc[:, i] = _wave(t, t0=arr, f=0.18, w=6.0)
        dat[:, i] = _wave(t, t0=arr + delay * (1 + 0.3 * rng.standard_normal()), f=0.18, w=6.0)
    return acc, dat, x, t


def test_adjoint_source_is_the_misfit_gradient():
    dt_min = 0.5
    acc, dat, x, t = _traces(dt_min=dt_min)
    nt, nx = acc.shape
    # every frequency passes: mode 8 (all p modes) above its 1.6 mHz cut, with a
    # window placed by the reference's rule, then frozen for the differences
    modes = {8: (5.0, 60.0, 30.0)}
    base = A.adjoint_source_filt(acc, dat, x, 0.0, 200e8, dt_min * 60.0, dt_min, modes)
    assert base.nmeas > 0 and base.misfit > 0
    wins = {8: [(m.lef, m.rig) for m in base.measurements]}
    # the reference's source, undone: time reversal and the nt/60 factor
    g = -base.adj[::-1, :] * 60.0 / nt * dt_min / 60.0
    rng = np.random.default_rng(3)
    from scipy.ndimage import gaussian_filter
    ok = 0
    for _ in range(4):
        d = gaussian_filter(rng.standard_normal((nt, nx)), sigma=(6.0, 2.0), mode="wrap")
        d /= np.abs(d).max()
        eps = 1e-4
        mp = A.adjoint_source_filt(acc + eps * d, dat, x, 0.0, 200e8, dt_min * 60.0, dt_min, modes,
                                   windows={8: list(wins[8])}).misfit
        mm = A.adjoint_source_filt(acc - eps * d, dat, x, 0.0, 200e8, dt_min * 60.0, dt_min, modes,
                                   windows={8: list(wins[8])}).misfit
        fd = (mp - mm) / (2 * eps)
        ad = float((g * d).sum())
        print("fd", fd, "adjoint", ad, "ratio", ad / fd)
        if abs(ad - fd) < 0.1 * abs(fd):
            ok += 1
    assert ok >= 3


def test_misfit_all_matches_adjoint_source_measurements():
    """MISFIT_ALL's band-1 all-p-mode misfit equals ADJOINT_SOURCE_FILT's
    misfit with the same band as its all-modes frequency filter and the same
    windows (both are 0.5 sum tau^2 of the same filtered traces)."""
    dt_min = 0.5
    acc, dat, x, t = _traces(dt_min=dt_min)
    modes = {8: (5.0, 60.0, 30.0)}
    ma = A.misfit_all(acc, dat, x, 0.0, 200e8, dt_min * 60.0, dt_min, modes)
    wins = ma.windows[(8, 1)]
    assert len(wins) > 0
    band = A.MISFIT_BANDS[0]
    r = A.adjoint_source_filt(acc, dat, x, 0.0, 200e8, dt_min * 60.0, dt_min, modes,
                              freq_params={"allmodes": band}, windows={8: list(wins)})
    f1, f2, nmeas, half = ma.bands[0]
    assert (f1, f2) == band and nmeas[8] == r.nmeas == len(wins)
    assert abs(half[8] - r.misfit) <= 1e-12 * abs(r.misfit)
    # reusing the picked windows (the linesearch branch) reproduces every band
    again = A.misfit_all(acc, dat, x, 0.0, 200e8, dt_min * 60.0, dt_min, modes, windows=ma.windows)
    for (_, _, n0, m0), (_, _, n1, m1) in zip(ma.bands, again.bands):
        assert np.array_equal(n0, n1) and np.allclose(m0, m1, rtol=1e-14, atol=0)


def test_misfit_all_file_reads_with_loadtxt(tmp_path):
    acc, dat, x, t = _traces()
    ma = A.misfit_all(acc, dat, x, 0.0, 200e8, 30.0, 0.5, {8: (5.0, 60.0, 30.0), 0: (3.0, 40.0, 20.0)})
    f = tmp_path / "misfit_all_01_00"
    f.write_text(ma.text(1))
    v = np.loadtxt(f, comments="#")
    assert v.shape == (3, 11)
    for b in range(3):
        assert np.allclose(v[b], ma.bands[b][3], rtol=1e-15)
    assert np.all(v[:, 1:8] == 0) and np.all(v[:, 9:] == 0)


def _compare_torch(device):
    """adjoint_source_filt_torch against adjoint_source_filt: the same
    measurements (receivers, windows, search ranges exactly; travel times to
    1e-12) and the same adjoint source to 1e-11, for the f mode, p1 and all
    p modes with frequency filters, with picked and with stored windows
    (including a list that runs out)."""
    dt_min = 0.5
    acc, dat, x, t = _traces(nt=240, nx=64, dt_min=dt_min, delay=0.05)
    modes = {0: (3.0, 40.0, 20.0), 1: (5.0, 50.0, 25.0), 8: (5.0, 60.0, 30.0)}
    fp = {"allmodes": (1.0, 7.0), 1: (2.0, 6.0)}
    cases = [dict(), dict(freq_params=fp)]
    ref0 = A.adjoint_source_filt(acc, dat, x, 1.5, 200e8, dt_min * 60.0, dt_min, modes)
    wins = {m: [(q.lef, q.rig) for q in ref0.measurements if q.mode == m] for m in modes}
    wins[8] = wins[8][:-3]                                # a stored list shorter than the receivers
    cases.append(dict(windows=wins))
    for kw in cases:
        kwn = {k: ({m: list(v) for m, v in w.items()} if k == "windows" else w) for k, w in kw.items()}
        kwt = {k: ({m: list(v) for m, v in w.items()} if k == "windows" else w) for k, w in kw.items()}
        ref = A.adjoint_source_filt(acc, dat, x, 1.5, 200e8, dt_min * 60.0, dt_min, modes, **kwn)
        got = A.adjoint_source_filt_torch(acc, dat, x, 1.5, 200e8, dt_min * 60.0, dt_min, modes, device=device,
                                          **kwt)
        assert ref.nmeas > 10 and got.nmeas == ref.nmeas
        for a, b in zip(ref.measurements, got.measurements):
            assert (a.mode, a.pixel, a.lef, a.rig, a.loc, a.timest, a.timefin) == \
                (b.mode, b.pixel, b.lef, b.rig, b.loc, b.timest, b.timefin)
            assert abs(a.tau_s - b.tau_s) <= 1e-12 * max(abs(a.tau_s), 1e-3)
        assert abs(got.misfit - ref.misfit) <= 1e-12 * ref.misfit
        err = np.abs(got.adj - ref.adj).max() / np.abs(ref.adj).max()
        print(device, sorted(kw), "adjoint source max rel", err)
        assert err < 1e-11


def test_torch_adjoint_source_matches_numpy_on_cpu():
    _compare_torch("cpu")


@pytest.mark.gpu
def test_torch_adjoint_source_matches_numpy_on_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _compare_torch("cuda")
