"""Pins of the oracle's exact discrete adjoint (orc_rhs_T, orc_ddz_T,
orc_ddx_T, orc_filter_vars_2d_T, orc_discrete_gradient): each transposed
operator satisfies <A u, w> = <u, A^T w> for random u, w; the whole backward
sweep gives the adjoint of the initial state, <gvz, vz(u0)> = <lam0, u0> for
a source-free run from a random u0; and the gradient with respect to psi and
c equals central finite differences of the oracle's own traces."""
import numpy as np
import pytest

import oracle
from tests.cases import small_case


@pytest.fixture(scope="module")
def case():
    oracle.build()
    return small_case()


def _dot(*pairs):
    return sum(float((a * b).sum()) for a, b in pairs)


def test_rhs_transpose(case):
    p, bg, c2d, psi = case
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    rng = np.random.default_rng(1)
    nzp, nph = p.npmlbot + p.npmltop, 2 * p.npmlhor
    u = [rng.standard_normal(s) for s in ((4, p.nz, p.nx), (2, nzp, p.nx), (2, p.nz, nph))]
    w = [rng.standard_normal(s) for s in ((4, p.nz, p.nx), (2, nzp, p.nx), (2, p.nz, nph))]
    R = o.rhs(10 ** 9, *u)                         # a time past the source: the linear part
    Y = o.rhs_T(*w)[:3]
    a, b = _dot(*zip(R, w)), _dot(*zip(u, Y))
    assert abs(a - b) <= 1e-12 * abs(a)


def test_derivative_and_filter_transposes(case):
    p, bg, c2d, psi = case
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    rng = np.random.default_rng(2)
    f, g = rng.standard_normal((p.nz, p.nx)), rng.standard_normal((p.nz, p.nx))
    for ibc in (1, 2, 3):
        kept = np.zeros((p.nz, p.nx))
        if ibc == 3:
            kept[-1] = rng.standard_normal(p.nx)
        r = o.ddz_T(g, ibc)
        gt, pb = r if ibc == 3 else (r, np.zeros(p.nx))
        a = _dot((o.ddz(f, ibc, kept), g))
        b = _dot((f, gt), (kept[-1], pb))
        assert abs(a - b) <= 1e-13 * abs(a), ibc
    a, b = _dot((o.ddx(f), g)), _dot((f, o.ddx_T(g)))
    assert abs(a - b) <= 1e-13 * abs(a)
    a4, w4 = rng.standard_normal((4, p.nz, p.nx)), rng.standard_normal((4, p.nz, p.nx))
    a, b = _dot((o.filter_vars_2d(a4), w4)), _dot((a4, o.filter_vars_2d_T(w4)))
    assert abs(a - b) <= 1e-13 * abs(a)


def test_backward_sweep_is_the_adjoint_of_the_run(case):
    """lam0 from the backward sweep against a source-free forward run from a
    random initial state (PML auxiliaries 0): every step, stage, filter and
    output injection of the sweep is checked at once."""
    p, bg, c2d, psi = case
    maxtime = 150
    nt = p.nt_out(maxtime)
    rng = np.random.default_rng(3)
    gvz = rng.standard_normal((nt, p.nx))
    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    *_, lam0 = o.discrete_gradient(maxtime, gvz)
    u0 = rng.standard_normal((4, p.nz, p.nx))
    o2 = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    o2.set_state(u0)
    vz = o2.run(maxtime)
    a, b = _dot((gvz, vz)), _dot((lam0, u0))
    assert abs(a - b) <= 1e-11 * abs(a)


def test_gradient_matches_finite_differences(case):
    """dJ/dpsi and dJ/dc of J = sum gvz vz (400 steps, flow and sound-speed
    bump) against central differences of the traces, along two directions
    per parameter that vanish in the PML rows and columns: the tapered
    gradient itself (the largest directional derivative) and a Gaussian bump.
    psi directions are |psi|max in size with a step of 0.03, c directions 10 %
    of c with a step of 0.003: there the differences' rounding and their
    O(h^2) error are both near 1e-7 (the flow changes the traces by ~1 %).
    Checked: |ad - fd| / |fd| of each parameter's pair of derivatives."""
    p, bg, c2d, psi = case
    maxtime = 400
    nt = p.nt_out(maxtime)
    rng = np.random.default_rng(4)
    gvz = rng.standard_normal((nt, p.nx))

    def J(ps, cc):
        o = oracle.Oracle(p, bg, c2d=cc, psi=ps)
        o.set_forward_source(5.0, maxtime)
        return _dot((o.run(maxtime), gvz))

    o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
    o.set_forward_source(5.0, maxtime)
    gpsi, gc = o.model_gradient(maxtime, gvz, c2d, psi)
    z, x = np.arange(p.nz)[:, None], np.arange(p.nx)[None, :]
    taper = np.zeros((p.nz, p.nx))
    taper[p.npmlbot + 1: p.nz - p.npmltop - 6, p.npmlhor + 1: p.nx - p.npmlhor - 1] = 1.0
    bump = np.exp(-((x - 30.0) / 6.0) ** 2 - ((z - 28.0) / 5.0) ** 2) * taper
    for g, base, scale, h, shift in ((gpsi, psi, np.abs(psi).max(), 3e-2, "psi"),
                                     (gc, c2d, 0.1 * c2d.max(), 3e-3, "c")):
        ads, fds = [], []
        for d in (g * taper / np.abs(g * taper).max(), bump):
            dd = scale * d
            if shift == "psi":
                fd = (J(base + h * dd, c2d) - J(base - h * dd, c2d)) / (2 * h)
            else:
                fd = (J(psi, base + h * dd) - J(psi, base - h * dd)) / (2 * h)
            ads.append(_dot((g, dd)))
            fds.append(fd)
        ads, fds = np.array(ads), np.array(fds)
        err = float(np.linalg.norm(ads - fds) / np.linalg.norm(fds))
        print(shift, "adjoint", ads, "differences", fds, "relative error", err)
        assert err < 1e-6, shift
