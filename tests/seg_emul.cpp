// CPU run of the column pass's segmented z-derivative of one column, through
// the kernel's own seg_rhs / seg_solve / seg_correct (sparc_seg.h) and tables
// (sparc_segtab.cpp): every segment's right-hand side and local solve, the
// end-value buffer with its zero pads, and the interface correction, as the
// lanes of one column of k_colsplit do them.  Built by
// tests/test_seg_emulation.py with g++.
#include <cstring>
#include <vector>

#include "sparc_seg.h"

using namespace sparc;

template <int MS>
static int run(int nz, int ibc, const double *B, double preset0, double presetN, double *out, int *hw_out)
{
    SegTables t;
    if (ibc < 1 || ibc > 3) return -3;
    if (!build_seg_tables(nz, MS, t)) return -2;
    const int P = t.P, nzp = t.nzp, hw = t.hw, n2P = 2 * P, nw = 4 * hw + 4, hw2 = 2 * hw;
    const int slot = ibc - 1;                     // table sets for ibc 1, 2, 3
    double c1, c2, cb[kCB];
    dbyd2_coefs(nz, c1, c2, cb);
    // the tables as the kernel holds them: three segments' worth, by kind
    std::vector<double> kt;
    if (!seg_kind_tables(t, kt)) return -4;
    std::vector<double> yp(yb_len(P, hw), 0.0);
    std::vector<double> Ys((size_t)P * MS);
    const double *band = t.band.data() + (size_t)slot * n2P * nw;
    for (int s = 0; s < P; s++) {
        const int a = s * MS;
        const bool first = s == 0, lastseg = a + MS > nz - 2;
        auto ld = [&](int k) {
            int r = a - 3 + k;
            r = r < 0 ? 0 : (r > nz - 1 ? nz - 1 : r);
            return B[r];
        };
        auto tab = [&](int tt, int i) { return kt[((size_t)(slot * 5 + tt) * 3 + seg_kind(s, P)) * MS + i]; };
        double e[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (lastseg)
            for (int k = 0; k < 5; k++) e[k] = B[nz - 5 + k];
        double Y[MS];
        seg_rhs<MS>(ld, e, Y, a, nz, first, lastseg, ibc, preset0, presetN, c1, c2, cb);
        seg_solve<MS>(Y, tab);
        yp[hw2 + 2 * s] = Y[0];
        yp[hw2 + 2 * s + 1] = Y[MS - 1];
        std::memcpy(&Ys[(size_t)s * MS], Y, sizeof Y);
    }
    for (int s = 0; s < P; s++) {
        const int a = s * MS;
        auto tab = [&](int tt, int i) { return kt[((size_t)(slot * 5 + tt) * 3 + seg_kind(s, P)) * MS + i]; };
        double Y[MS];
        std::memcpy(Y, &Ys[(size_t)s * MS], sizeof Y);
        seg_correct<MS>(Y, s, P, hw, yp.data(), band, tab);
        for (int i = 0; i < MS && a + i < nz; i++) out[a + i] = Y[i];
    }
    *hw_out = hw;
    return 0;
}

// dbyd2 of one column B[nz] with segments of ms rows; preset0 / presetN are
// the kept first / last rows of ibc 2 / 3.  Returns 0, or < 0 on bad input.
extern "C" int seg_ddz_column(int nz, int ms, int ibc, const double *B, double preset0, double presetN,
                              double *out, int *hw)
{
    switch (ms) {
    case 4: return run<4>(nz, ibc, B, preset0, presetN, out, hw);
    case 6: return run<6>(nz, ibc, B, preset0, presetN, out, hw);
    case 8: return run<8>(nz, ibc, B, preset0, presetN, out, hw);
    case 10: return run<10>(nz, ibc, B, preset0, presetN, out, hw);
    case 12: return run<12>(nz, ibc, B, preset0, presetN, out, hw);
    case 16: return run<16>(nz, ibc, B, preset0, presetN, out, hw);
    case 19: return run<19>(nz, ibc, B, preset0, presetN, out, hw);
    case 20: return run<20>(nz, ibc, B, preset0, presetN, out, hw);
    case 24: return run<24>(nz, ibc, B, preset0, presetN, out, hw);
    case 32: return run<32>(nz, ibc, B, preset0, presetN, out, hw);
    }
    return -1;
}
