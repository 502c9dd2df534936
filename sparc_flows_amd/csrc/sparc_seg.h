// sparc_seg.h -- the segmented (SPIKE) form of dbyd2 (dbyd2.f:8-200,
// mtridag.f:95-112) that the column pass runs, one segment per lane: the
// segment's right-hand side, its local Thomas solve and the interface
// correction, plus the host construction of their tables.
//
// The three templates are the kernel's own code (k_colsplit includes this
// header); the same source also compiles as host C++, so that the algorithm
// is checked on the CPU against the oracle's dbyd2 for every segment length.
#pragma once

#include <vector>

#if defined(__HIPCC__)
#define SPARC_HD __host__ __device__ __forceinline__
#else
#define SPARC_HD inline
#endif

namespace sparc {

// The boundary-row coefficients of dbyd2 (dbyd2.f:103-124, 156-162), read
// only by the lanes of the first and last segments: f1..f4, b1..b5, alpha2,
// beta1, cnst.  The column pass keeps them after its z tables.
enum { CB_F = 0, CB_B = 4, CB_ALPHA2 = 9, CB_BETA1 = 10, CB_CNST = 11, kCB = 12 };

// dbyd2 right-hand side (dbyd2.f:91-172) of rows j = a..a+MS-1 of one
// segment; ld(k) returns B(a - 3 + k) for k in [1, MS + 5) (the rows the
// formulas of real rows use; anything finite outside the column).  Padding
// rows (j >= nz) get 0.  The one-sided rows nz-2 and nz-1 exist only in the
// last segment (lastseg: it holds >= 3 real rows, so both); their inputs
// e[k] = B(nz - 5 + k) come from the caller.  preset0 / presetN: the kept
// first / last row of ibc 2 / 3.
template <int MS, typename Loader>
SPARC_HD void seg_rhs(Loader ld, const double (&e)[5], double (&R)[MS], int a, int nz, bool first, bool lastseg,
                      int ibc, double preset0, double presetN, double c1, double c2, const double *cb)
{
#pragma unroll
    for (int i = 0; i < MS; i++) {
        const double r = c1 * (ld(i + 4) - ld(i + 2)) + c2 * (ld(i + 5) - ld(i + 1));
        R[i] = (a + i >= nz) ? 0.0 : r;
    }
    if (first) {
        // row 0 from B(0..3) = ld(3..6), row 1 from B(0..4) = ld(3..7)
        R[0] = (ibc == 1 || ibc == 3) ? cb[CB_F] * ld(3) + cb[CB_F + 1] * ld(4) + cb[CB_F + 2] * ld(5) +
                                            cb[CB_F + 3] * ld(6)
                                      : (ibc == 0 ? 0.0 : preset0);
        R[1] = cb[CB_B] * ld(3) + cb[CB_B + 1] * ld(4) + cb[CB_B + 2] * ld(5) + cb[CB_B + 3] * ld(6) +
               cb[CB_B + 4] * ld(7);
        if (ibc & 1) R[0] = (cb[CB_ALPHA2] * R[0] - cb[CB_BETA1] * R[1]) * cb[CB_CNST];   // dbyd2.f:163-167
    }
    if (lastseg) {
        const double rm2 = -(cb[CB_B] * e[4] + cb[CB_B + 1] * e[3] + cb[CB_B + 2] * e[2] +
                             cb[CB_B + 3] * e[1] + cb[CB_B + 4] * e[0]);
        double rm1 = (ibc == 1 || ibc == 2)
                         ? -(cb[CB_F] * e[4] + cb[CB_F + 1] * e[3] + cb[CB_F + 2] * e[2] + cb[CB_F + 3] * e[1])
                         : (ibc == 0 ? 0.0 : presetN);
        if (ibc == 1 || ibc == 2)                                                           // dbyd2.f:168-172
            rm1 = (cb[CB_ALPHA2] * rm1 - cb[CB_BETA1] * rm2) * cb[CB_CNST];
        const int i2 = nz - 2 - a;
#pragma unroll
        for (int i = 1; i < MS; i++) {
            R[i] = (i == i2) ? rm2 : R[i];
            R[i] = (i == i2 + 1) ? rm1 : R[i];
        }
    }
}

// Local Thomas solve of the segment (factorization restarted at its start);
// tab(t, i): row i of the segment in table t (0 lowr, 1 rbl, 2 gl, 3 V, 4 W).
template <int MS, typename Tab>
SPARC_HD void seg_solve(double (&Y)[MS], Tab tab)
{
    Y[0] = Y[0] * tab(1, 0);
#pragma unroll
    for (int i = 1; i < MS; i++) Y[i] = (Y[i] - tab(0, i) * Y[i - 1]) * tab(1, i);
#pragma unroll
    for (int i = MS - 2; i >= 0; i--) Y[i] = Y[i] - tab(2, i + 1) * Y[i + 1];
}

// Two consecutive doubles read as one 16-byte access (ds_read_b128 from
// LDS); the pointer must be 16-byte aligned.
struct alignas(16) Dbl2 {
    double x, y;
} __attribute__((may_alias));
SPARC_HD Dbl2 ld2(const double *p) { return *reinterpret_cast<const Dbl2 *>(p); }

// y <- y - L_{s-1} V - F_{s+1} W; the interface values are rows of the
// banded inverse (row k holds entries q = 2 (k/2 - hw) + w, w < 4 hw + 4,
// zero where q is outside 0..2P-1) against the segment end values y, stored
// with 2 hw zeros before and 2 hw + 2 after (y[q] at yp[q + 2 hw]).  band
// and yp must be 16-byte aligned (every row and window then starts on an
// even element: 4 hw + 4 is a multiple of 4, the windows start at 2 (k/2)).
template <int MS, typename Tab>
SPARC_HD void seg_correct(double (&Y)[MS], int s, int P, int hw, const double *yp, const double *band, Tab tab)
{
    const int nw = 4 * hw + 4;
    // both rows in one loop, four entries of each per iteration, so that
    // 16 reads are issued before the first product needs one; rows outside
    // 0..2P-1 are clamped (their windows stay inside the buffer) and weighted 0
    const int k1 = s > 0 ? 2 * s - 1 : 0, k2 = s < P - 1 ? 2 * s + 2 : 2 * P - 1;
    const double *r1 = band + k1 * nw, *y1 = yp + 2 * (k1 / 2);   // y[2 (k/2 - hw) + w] = yp[2 (k/2) + w]
    const double *r2 = band + k2 * nw, *y2 = yp + 2 * (k2 / 2);
    double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
#pragma unroll 2
    for (int w = 0; w < nw; w += 4) {
        const Dbl2 pA = ld2(r1 + w), pB = ld2(r1 + w + 2), uA = ld2(y1 + w), uB = ld2(y1 + w + 2);
        const Dbl2 qA = ld2(r2 + w), qB = ld2(r2 + w + 2), vA = ld2(y2 + w), vB = ld2(y2 + w + 2);
        a0 += pA.x * uA.x + pB.x * uB.x;
        a1 += pA.y * uA.y + pB.y * uB.y;
        b0 += qA.x * vA.x + qB.x * vB.x;
        b1 += qA.y * vA.y + qB.y * vB.y;
    }
    const double Lp = s > 0 ? a0 + a1 : 0.0;
    const double Fn = s < P - 1 ? b0 + b1 : 0.0;
#pragma unroll
    for (int i = 0; i < MS; i++) Y[i] = Y[i] - Lp * tab(3, i) - Fn * tab(4, i);
}

// Length of one column's end-value buffer: 2P values between 2hw and 2hw+2 zeros.
inline int yb_len(int P, int hw) { return 2 * P + 4 * hw + 2; }

// ---------------------------------------------------------------------------
// Host side (sparc_segtab.cpp).

// dbyd2's coefficients for N2 rows (dbyd2.f:85-162), float32 scalars widened:
// c1, c2 (interior), then the kCB boundary-row coefficients in CB_ order.
void dbyd2_coefs(int N2, double &c1, double &c2, double cb[kCB]);

// The tridiagonal dbyd2 matrix of one ibc (dbyd2.f:151-195): sub- and
// super-diagonals (diagonal 1).
void dbyd2_matrix(int N2, int ibc, std::vector<double> &lowr, std::vector<double> &uppr);

// Tables of the segmented solve for nz rows cut into segments of m rows
// (P = ceil(nz/m), nzp = P*m, the last segment padded with identity rows).
// slots: [3][5][nzp], for ibc 1, 2, 3: lowr, rbl, gl (Thomas factors restarted
// at each segment start), V, W (the spikes: the segment's response to the row
// above and below it).  band: [3][2P][4hw+4], rows of the exact inverse of the
// 2P x 2P interface system, entries below 1e-24 of their row maximum dropped
// (hw: the largest distance in segments to a kept entry).
struct SegTables {
    int P = 0, m = 0, nzp = 0, hw = 0;
    std::vector<double> slots, band;
};
// false if the interface system is singular.
bool build_seg_tables(int nz, int m, SegTables &t);

// Segment kind of segment s of P: 0 the first, 2 the last, 1 any other.
// Every interior segment has the same tables (the dbyd2 coefficients are
// constant away from the boundary rows, and each segment's factorization
// restarts at its start), so the column pass keeps three segments' worth.
SPARC_HD int seg_kind(int s, int P) { return s == 0 ? 0 : (s == P - 1 ? 2 : 1); }
// The tables of the first, an interior and the last segment, [15][3][m] (row
// r = 5 slot + k of t.slots, then the segment kind); false if some interior
// segment's tables differ from the first interior one's.
bool seg_kind_tables(const SegTables &t, std::vector<double> &out);

}  // namespace sparc
