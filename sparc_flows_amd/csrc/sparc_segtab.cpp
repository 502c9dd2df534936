// sparc_segtab.cpp -- host construction of the dbyd2 coefficients, matrix and
// segmented-solve tables (dbyd2.f:85-195, mtridag.f:95-112) shared by the
// library and the CPU check of the column pass's algorithm.  Plain C++.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sparc_seg.h"

namespace sparc {

void dbyd2_coefs(int N2, double &c1, double &c2, double cb[kCB])
{
    const float d2i = (float)(N2 - 1);
    c1 = (double)(d2i * 7.0f / 9.0f);
    c2 = (double)(d2i / 36.0f);
    cb[CB_B + 0] = (double)(-(43.f / 96.f * d2i));
    cb[CB_B + 1] = (double)(-(5.f / 6.f * d2i));
    cb[CB_B + 2] = (double)(9.f / 8.f * d2i);
    cb[CB_B + 3] = (double)(1.f / 6.f * d2i);
    cb[CB_B + 4] = (double)(-(1.f / 96.f * d2i));
    cb[CB_F + 0] = (double)(-(10.0f / 3.f * d2i));
    cb[CB_F + 1] = (double)(-(3.0f * d2i));
    cb[CB_F + 2] = (double)(6.f * d2i);
    cb[CB_F + 3] = (double)(1.f / 3.f * d2i);
    const float alpha2 = 3.f / 4.f, gamma2 = 1.f / 8.f, beta1 = 3.f;
    cb[CB_ALPHA2] = alpha2;
    cb[CB_BETA1] = beta1;
    cb[CB_CNST] = (double)(1.f / (alpha2 - beta1 * gamma2));
}

void dbyd2_matrix(int N2, int ibc, std::vector<double> &lowr, std::vector<double> &uppr)
{
    const float alpha = 1.f / 3.f, alpha2 = 3.f / 4.f, gamma2 = 1.f / 8.f;
    const float alpha1 = 6.0f, beta1 = 3.f;
    const float cnst = 1.f / (alpha2 - beta1 * gamma2);
    const float up1 = (alpha1 * alpha2 - beta1) * cnst;
    uppr.assign(N2, 0.0);
    lowr.assign(N2, 0.0);
    for (int j = 2; j <= N2 - 3; j++) uppr[j] = alpha;
    uppr[0] = (ibc % 2 == 0) ? 0.0 : (double)up1;
    uppr[1] = alpha2;
    uppr[N2 - 2] = gamma2;
    for (int j = 0; j < N2; j++) lowr[j] = uppr[N2 - 1 - j];
    lowr[N2 - 1] = (ibc != 1 && ibc != 2) ? 0.0 : (double)up1;
    lowr[0] = 0.0;
    uppr[N2 - 1] = 0.0;
}

namespace {

// Invert a dense n x n matrix (Gauss-Jordan, partial pivoting).
bool invert(std::vector<double> A, int n, std::vector<double> &inv)
{
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int pr = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)pr * n + c])) pr = r;
        if (A[(size_t)pr * n + c] == 0.0) return false;
        if (pr != c)
            for (int k = 0; k < n; k++) {
                std::swap(A[(size_t)pr * n + k], A[(size_t)c * n + k]);
                std::swap(inv[(size_t)pr * n + k], inv[(size_t)c * n + k]);
            }
        const double d = 1.0 / A[(size_t)c * n + c];
        for (int k = 0; k < n; k++) { A[(size_t)c * n + k] *= d; inv[(size_t)c * n + k] *= d; }
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = A[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; k++) {
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
                inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
            }
        }
    }
    return true;
}

}  // namespace

bool build_seg_tables(int N, int m, SegTables &t)
{
    const int P = (N + m - 1) / m;
    const int nzp = P * m, n2 = 2 * P;
    t.P = P;
    t.m = m;
    t.nzp = nzp;
    t.slots.assign((size_t)3 * 5 * nzp, 0.0);
    std::vector<double> minv_all;
    int hw = 0;
    const int ibcs[3] = {1, 2, 3};
    for (int slot = 0; slot < 3; slot++) {
        std::vector<double> lo0, up0;
        dbyd2_matrix(N, ibcs[slot], lo0, up0);
        std::vector<double> lowr(nzp, 0.0), uppr(nzp, 0.0);
        for (int j = 0; j < N; j++) { lowr[j] = lo0[j]; uppr[j] = up0[j]; }
        std::vector<double> rbl(nzp, 1.0), gl(nzp, 0.0), V(nzp, 0.0), W(nzp, 0.0), bet(nzp, 1.0);
        for (int s = 0; s < P; s++) {
            const int a0 = s * m;
            for (int j = a0 + 1; j < a0 + m; j++) {
                gl[j] = uppr[j - 1] / bet[j - 1];
                bet[j] = 1.0 - lowr[j] * gl[j];
                rbl[j] = 1.0 / bet[j];
            }
            auto lsolve = [&](std::vector<double> r, double *out) {
                r[0] = r[0] * rbl[a0];
                for (int i = 1; i < m; i++) r[i] = (r[i] - lowr[a0 + i] * r[i - 1]) * rbl[a0 + i];
                for (int i = m - 2; i >= 0; i--) r[i] = r[i] - gl[a0 + i + 1] * r[i + 1];
                for (int i = 0; i < m; i++) out[a0 + i] = r[i];
            };
            std::vector<double> r(m, 0.0);
            if (s > 0) { r[0] = lowr[a0]; lsolve(r, V.data()); }
            std::fill(r.begin(), r.end(), 0.0);
            if (s < P - 1) { r[m - 1] = uppr[a0 + m - 1]; lsolve(r, W.data()); }
        }
        std::vector<double> M((size_t)n2 * n2, 0.0), Minv;
        for (int s = 0; s < P; s++) {
            const int f = s * m, l = s * m + m - 1;
            M[(size_t)(2 * s) * n2 + 2 * s] = 1.0;
            M[(size_t)(2 * s + 1) * n2 + 2 * s + 1] = 1.0;
            if (s > 0) {
                M[(size_t)(2 * s) * n2 + 2 * s - 1] += V[f];
                M[(size_t)(2 * s + 1) * n2 + 2 * s - 1] += V[l];
            }
            if (s < P - 1) {
                M[(size_t)(2 * s) * n2 + 2 * s + 2] += W[f];
                M[(size_t)(2 * s + 1) * n2 + 2 * s + 2] += W[l];
            }
        }
        if (!invert(M, n2, Minv)) return false;
        // window half-width: entries below 1e-24 of their row maximum are dropped
        for (int k = 0; k < n2; k++) {
            double rm = 0.0;
            for (int q = 0; q < n2; q++) rm = std::max(rm, std::fabs(Minv[(size_t)k * n2 + q]));
            for (int q = 0; q < n2; q++)
                if (std::fabs(Minv[(size_t)k * n2 + q]) > 1e-24 * rm) hw = std::max(hw, std::abs(q / 2 - k / 2));
        }
        minv_all.insert(minv_all.end(), Minv.begin(), Minv.end());   // banded below, once hw is known
        const std::vector<double> *tabs[5] = {&lowr, &rbl, &gl, &V, &W};
        for (int k = 0; k < 5; k++)
            std::copy(tabs[k]->begin(), tabs[k]->end(), t.slots.begin() + (size_t)(5 * slot + k) * nzp);
    }
    t.hw = hw;
    // band storage: row k keeps the entries q = 2 (k/2 - hw) + w, w < 4 hw + 4
    // (the window around its own segment), zero outside 0..2P-1
    const int W = 4 * hw + 4;
    t.band.assign((size_t)3 * n2 * W, 0.0);
    for (int slot = 0; slot < 3; slot++)
        for (int k = 0; k < n2; k++)
            for (int w = 0; w < W; w++) {
                const int q = 2 * (k / 2 - hw) + w;
                if (q >= 0 && q < n2)
                    t.band[((size_t)slot * n2 + k) * W + w] = minv_all[((size_t)slot * n2 + k) * n2 + q];
            }
    return true;
}

bool seg_kind_tables(const SegTables &t, std::vector<double> &out)
{
    const int P = t.P, m = t.m, nzp = t.nzp;
    const int sk[3] = {0, P > 1 ? 1 : 0, P - 1};
    out.clear();
    out.reserve((size_t)15 * 3 * m);
    for (int r = 0; r < 15; r++) {
        const double *row = t.slots.data() + (size_t)r * nzp;
        for (int s = 2; s < P - 1; s++)
            if (std::memcmp(row + (size_t)s * m, row + m, (size_t)m * sizeof(double))) return false;
        for (int k = 0; k < 3; k++) out.insert(out.end(), row + (size_t)sk[k] * m, row + (size_t)(sk[k] + 1) * m);
    }
    return true;
}

}  // namespace sparc
