"""CPU emulation of the index arithmetic of the column-pass kernel
(sparc_flows_amd/csrc/sparc_colseg.hip: k_colsplit) for every launch shape
the autotuner can choose: segment lengths MS of SPARC_COLSEG_SIZES with
P = ceil(nz/MS) <= 64 and a last segment of >= 3 real rows, block widths TC in
{1, 2, 4, 8, 16} with 3*roundup(TC*P, 64) <= 384, both segment-count extremes
of every MS, and nz from 16 to 1024.  Every global index (loads and the
stores of real rows) must fall inside its allocation and every LDS index inside the
dynamic LDS the launcher requests (colsplit_lds_bytes).

The expressions below are the kernels' own, evaluated for every thread of a
block and every unrolled row; blocks are checked at the grid's extremes
(first and last column group and source), where the bounds are tightest."""
import numpy as np
import pytest

SIZES = (4, 6, 8, 10, 12, 16, 19, 20, 24)
KSEGMAX, KZT, KZS, KZF, KCB, KPL = 64, 21, 15, 6, 12, 7
Z_PROFILE_ROWS = [0, 1, 17, 18, 19, 20]   # ztab rows of the six epilogue profiles


def stage_prow(r):
    """ztab row of staged profile row r (stage_prow in sparc_device.h)."""
    return 0 if r == 0 else (1 if r == 1 else 15 + r)


def seg_lengths(nz):
    out = []
    for v in SIZES:
        p = -(-nz // v)
        if p <= KSEGMAX and nz - (p - 1) * v >= 3:
            out.append(v)
    return out


def nz_cases():
    """(nz, MS) pairs: nz 16..40 in full with every usable MS, then the two
    ends of every (MS, P) pair up to nz = 1024 with that MS (where the
    indices of that segmentation are tightest)."""
    cases = {(nz, ms) for nz in range(16, 41) for ms in seg_lengths(nz)}
    for ms in SIZES:
        for p in range(1, KSEGMAX + 1):
            for nz in ((p - 1) * ms + 3, p * ms):
                if 16 <= nz <= 1024 and ms in seg_lengths(nz):
                    cases.add((nz, ms))
    return sorted(cases)


class Check:
    def __init__(self, what):
        self.what = what

    def idx(self, name, v, size, mask=None):
        v = np.asarray(v)
        if mask is not None:
            v = v[np.broadcast_to(mask, v.shape)]
        if v.size:
            assert v.min() >= 0 and v.max() < size, f"{self.what}: {name} in [{v.min()}, {v.max()}], size {size}"


def emulate(nz, ms, tc, split, nx, S, nb, nt, hw=3):
    """All indices of one launch of k_colsplit (split 1; split 0 is the
    lane layout of one lane per (column, segment) for all fields)."""
    P = -(-nz // ms)
    nzp, n2P = P * ms, 2 * P
    nzpml = nb + nt
    nband = n2P * (4 * hw + 4)
    LY, hw2 = n2P + 4 * hw + 2, 2 * hw
    assert split, "only the field-split kernel exists"
    G = -(-tc * P // 64) * 64
    block = 3 * G
    nstaged = (3 * nband + KZS * 3 * ms + KZF * nzp + KCB + KPL * nzpml * tc + nzp + 1) & ~1
    lds = nstaged + 4 * tc * LY + 3 * tc * nzp
    zl = 3 * nband
    zf = zl + KZS * 3 * ms
    cbl = zf + KZF * nzp
    pl, yb0 = cbl + KCB, nstaged
    wl = yb0 + 4 * tc * LY
    chk = Check(f"nz={nz} MS={ms} TC={tc} split={split} nx={nx} S={S}")
    tid = np.arange(block)[:, None]
    f = tid // G
    l_ = tid - f * G
    c, s = l_ % tc, l_ // tc
    act = s < P
    a = np.where(act, s, 0) * ms
    i = np.arange(ms)[None, :]
    gridx = nx // tc
    n = S * nzp * 0 + nz * nx
    for bx in (0, gridx - 1):
        for b in (0, S - 1):
            x = bx * tc + c
            col = b * nz * nx + x
            rows = np.minimum(a + i, nz - 1)
            qj = col + rows * nx
            valid = act & (a + i < nz)
            chk.idx("x", x, nx)
            chk.idx("row loads (rowq)", qj, S * n)
            chk.idx("x-varying coefficient (gq)", qj - b * nz * nx, n)
            k6 = np.arange(ms + 6)[None, :]
            chk.idx("rhs window", col + np.clip(a - 3 + k6, 0, nz - 1) * nx, S * n)
            chk.idx("last rows (e)", col + (nz - 5 + np.arange(5)[None, :]) * nx, S * n)
            k10 = np.arange(ms + 10)[None, :]
            chk.idx("filter window", col + np.clip(a - 5 + k10, 0, nz - 1) * nx, S * n)
            chk.idx("topbc", b * nx + x, S * nx)
            chk.idx("rho0 at the source row", x, nx)
            # adjoint-source slices of LAGRANGE_INTERP, loaded with clamped slice
            # numbers for every time1 (nvr slices of [S][nvr][nx])
            nvr = 9
            for time1 in range(-6, nvr + 6):
                sl = np.clip(time1 - 3 + np.arange(7), 1, max(nvr, 1))
                chk.idx("adjoint source slices", (b * nvr + sl[None, :] - 1) * nx + x, S * nvr * nx)
            # z tables: zt = ztab + a, rows zrow(slot, k) = 2 + 5 slot + k and 17..20
            for row in [2 + 5 * sl + k for sl in range(3) for k in range(5)] + [0, 1, 17, 18, 19, 20]:
                chk.idx("ztab", row * nzp + a + i, KZT * nzp + KCB)
            chk.idx("boundary coefficients", KZT * nzp + np.arange(KCB), KZT * nzp + KCB)
            # z-PML rows
            j = a + i
            jc = np.minimum(j, nz - 1)
            pj = np.where(jc < nb, jc, np.minimum(nzpml - 1, nb + np.maximum(jc - (nz - nt), 0)))
            chk.idx("pmlrow", pj, nzpml)
            pz0 = b * nzpml * nx + x
            pz1 = (S + b) * nzpml * nx + x
            chk.idx("pz (psivz)", pz0 + pj * nx, 2 * S * nzpml * nx)
            chk.idx("pz (psip)", pz1 + pj * nx, 2 * S * nzpml * nx)
            chk.idx("az/bz/fz", pj * nx + x, nzpml * nx)
            for k in range(KPL):
                chk.idx("staged z-PML data", pl + (k * nzpml + pj) * tc + c, lds)
            zpl = pl + KPL * nzpml * tc
            chk.idx("staged source z profile (read)", zpl + j, nstaged)
            chk.idx("staged source z profile (read, lower end)", j, nzp)
            # stores: only real rows store (padding rows are masked off)
            chk.idx("valid-row stores", qj, S * n, valid)
            # snapshots and vz output at their extreme slots
            for nsnap, nfield in ((3, 6), (5, 4)):
                slot = (b * nsnap + (nsnap - 1)) * nfield
                for fs in range(nfield):
                    chk.idx("snapshot", ((slot + fs) * nz + j) * nx + x, S * nsnap * nfield * n, valid)
            # LDS: interface inverses, end-value exchange, staged tables, wl
            yend = yb0 + 4 * tc * LY
            for fb in range(4):
                ybase = yb0 + fb * tc * LY + c * LY
                chk.idx("yb end values", ybase + hw2 + 2 * s + np.array([[0, 1]]), yend, act)
                # seg_correct reads rows k1, k2 (clamped to 0 / 2P-1 at the ends) for every active lane
                k1 = np.where(s > 0, 2 * s - 1, 0)
                k2 = np.where(s < P - 1, 2 * s + 2, n2P - 1)
                for k, on in ((k1, s > 0), (k2, s < P - 1)):
                    rd = ybase + 2 * (k // 2) + np.arange(4 * hw + 4)[None, :]
                    chk.idx("yb window of a band row", rd, yend, act)
                    chk.idx("yb window of a band row (lower end)", rd - yb0, 4 * tc * LY, act)
                    # the window covers q = 2 (k/2 - hw) .. + 4hw+3, stored at q + 2hw
                    q = rd - ybase - hw2
                    inside = (q >= 0) & (q < n2P)
                    assert np.all(np.where(np.broadcast_to(act & on, q.shape) & inside,
                                           rd == ybase + hw2 + q, True))
            for k in (k1, k2):
                for base in (0, nband, 2 * nband):
                    chk.idx("banded inverse", base + k * (4 * hw + 4) + np.arange(4 * hw + 4)[None, :],
                            base + nband, act)
            padw = LY - n2P
            qq = np.arange(4 * tc * padw)
            r, u = qq // padw, qq % padw
            pad = r * LY + np.where(u < hw2, u, n2P + u)
            chk.idx("yb pads", yb0 + pad, yend)
            assert len(set(pad.tolist())) == pad.size
            assert not set(pad.tolist()) & {f * LY + hw2 + v for f in range(4 * tc) for v in range(n2P)}
            sc = np.where(act, s, 0)
            zt = zl + np.where(sc == 0, 0, np.where(sc == P - 1, 2, 1)) * ms   # seg_kind
            for r in range(KZS):
                chk.idx("staged segment table", zt + r * 3 * ms + i, zf)
                chk.idx("staged segment table (lower end)", zt + r * 3 * ms + i - zl, KZS * 3 * ms)
            for r in range(KZF):
                chk.idx("staged z profile", zf + r * nzp + a + i, cbl)
            chk.idx("staged boundary coefficients", cbl + np.arange(KCB), pl)
            chk.idx("wl group 1", wl + c * nzp + a + i, lds, act)
            chk.idx("wl group 0 (w0)", wl + (tc + c) * nzp + a + i, lds, act)
            chk.idx("p column (store)", wl + (2 * tc + c) * nzp + j, lds, valid)
            k6 = np.arange(ms + 6)[None, :]
            chk.idx("p column (window)", wl + (2 * tc + c) * nzp + np.clip(a - 3 + k6, 0, nz - 1), lds, act)
            chk.idx("p column (last rows)", wl + (2 * tc + c) * nzp + nz - 5 + np.arange(5)[None, :], lds, act)
    # staging, for the last column group and source (every index grows with
    # both): the inverses and z tables are element q < nA of the contiguous
    # LDS image sg.stage, the z-PML data element q < nzpml*TC of kind k
    bx, b = gridx - 1, S - 1
    nA, npl = 3 * nband + KZS * 3 * ms + KZF * nzp + KCB, nzpml * tc
    assert (nA + KPL * npl + nzp + 1) & ~1 == nstaged
    # the interface correction reads band rows and end-value windows in
    # 16-byte pairs: every row, buffer and window starts on an even element
    assert nband % 2 == 0 and (4 * hw + 4) % 2 == 0 and yb0 % 2 == 0 and LY % 2 == 0
    q = np.arange(nA)
    chk.idx("staging: image", q, nA)
    q = np.arange(npl)
    pj2, xc = q // tc, bx * tc + q % tc
    chk.idx("staging: az/bz/fz", pj2 * nx + xc, nzpml * nx)
    chk.idx("staging: pz slot 0", (b * nzpml + pj2) * nx + xc, 2 * S * nzpml * nx)
    chk.idx("staging: pz slot 1", ((S + b) * nzpml + pj2) * nx + xc, 2 * S * nzpml * nx)
    chk.idx("staging: LDS", nA + np.arange(KPL)[:, None] * npl + q[None, :], nstaged)
    # the source z profile: element q < nzp, loaded from row min(q, nz-1) of [nz]
    q = np.arange(nzp)
    chk.idx("staging: source z profile", np.minimum(q, nz - 1), nz)
    chk.idx("staging: source z profile (LDS)", nA + KPL * npl + q, nstaged)
    assert [stage_prow(r) for r in range(KZF)] == Z_PROFILE_ROWS
    return P


def test_every_autotuner_shape_stays_in_bounds():
    split = 1
    count = 0
    for nz, ms in nz_cases():
        P = -(-nz // ms)
        for tc in (1, 2, 4, 8, 16):
            if split and 3 * (-(-tc * P // 64) * 64) > 384:
                continue
            if not split and tc * P > 256:
                continue
            for nx, S in ((16, 1), (1024, 3)):
                if nx % tc:
                    continue
                npml = min(12, (nz - 1) // 2 - 1)
                hw = (0, 3, P)[count % 3]        # interface-inverse half-widths in turn
                emulate(nz, ms, tc, split, nx, S, npml, npml, hw)
                count += 1
    assert count > 1000


def test_segmentation_covers_every_nz():
    """Every nz from 16 to 1024 has a usable segmentation."""
    gaps = [nz for nz in range(16, 1025) if not seg_lengths(nz)]
    assert gaps == []
