/*
 * sparc_oracle.c -- CPU restatement of SPARC's 2D option-17 path.
 *
 * TEST INFRASTRUCTURE ONLY (see sparc_oracle.h).  Single-threaded, plain C,
 * following the reference Fortran statement by statement, including the
 * float32 literals that the reference widens to real64.  A C float literal
 * written as 0.005f here stands for the Fortran default-REAL literal 0.005.
 *
 * Parity pins (tests/test_oracle.py):
 *   orc_dbyd2 / mtridag  - bit-compared against the reference dbyd2.f +
 *                          mtridag.f compiled from source (oracle/_ref);
 *   orc_ddx / orc_fft    - closed forms (Fourier modes, Nyquist, constants) and
 *                          numpy's FFT, at power-of-two and other lengths;
 *   filters, RK, Lagrange - brute force / closed forms;
 *   orc_forward_source_series - the series' spectrum against the closed-form
 *                          transform of the continuous pulse;
 *   orc_create           - flows against the closed form of a separable
 *                          stream function; PML coefficient structure and
 *                          the absorption of an outgoing pulse;
 *   orc_rhs / orc_step   - the RK polynomial, linearity, mirror symmetry,
 *                          and closed forms (tests/test_oracle_physics.py):
 *                          a uniform xi_z in hydrostatic balance is force
 *                          free, a horizontal xi_x mode gives c^2 d2/dx2 and
 *                          (1 - Gamma) g div, flow increments -2 v0 . grad v,
 *                          the PML auxiliary equations, the top closure;
 *   orc_kernels          - a numpy restatement on the reference-built dbyd2,
 *                          the psi <-> v0 chain rule, and central finite
 *                          differences of a linear functional of the traces.
 * End-to-end traces against the full Fortran build are "parity unpinned":
 * the reference needs FFTW3, which this image does not have.
 *
 * ORC_MUTANT (compile time, default 0) plants one known mistake in a copy of
 * this library so that tests/test_oracle_mutants.py can show that each pin
 * fails on the mistake it is meant to catch.  Every other build has 0.
 */
#include "sparc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RSUN 6.95989467700e10           /* initialize.f90:37 */
#define PI 3.141592653589793            /* initialize.f90:37 */

#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define MUT(n) (ORC_MUTANT == (n))

struct orc_model {
    orc_params p;
    int nx, nz, nzpml, nph;
    double dimc, dimrho, diml, deltat, stretchx, normx;
    double *x, *eyekx, *z, *stretch, *unstretch, *g;
    double *c2, *rho0, *p0, *rhoinv, *c2rho0, *gradp0_z, *gradrho0_z;
    double *v0x, *v0z, *az, *bzpml, *axpml, *bxpml;
    int e_rad, o_rad;                   /* 0-based */
    double delta_width;
    double betas[5];
    /* state: fields 0 xi_x, 1 xi_z, 2 v_x, 3 v_z, each [nz][nx] */
    double *a, *temp, *scr;
    double *pmlvars, *psipml, *scrpml;  /* [2][nzpml][nx]: psivz, psip */
    double *pmlvarsx, *psipmlx, *scrpmlx; /* [2][nz][nph]: psivx, psi_gradp_x */
    /* work arrays (persistent, like the reference module arrays) */
    double *dxixdx, *dxizdz, *pp, *rho, *gradp_x, *gradp_z, *div;
    double *fre, *fim;
    /* sources */
    double *vr;
    int nvr;
    double *fwdsource;
    int nfwd;
    double *forcing;
    double *zprof;                      /* profile source: [nz], NULL for the one-row form */
    int maxtime, steps, cadforcing_step, forcing_length, num_steps;
    /* discrete adjoint: the z-derivative of ibc 1, 2, 3 as dense matrices
     * [3][nz][nz] (output row, input row; stretch included) and the ibc-3
     * response to its kept last row [nz]; built on first use */
    double *dzm, *dzmN;
};

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* ------------------------------------------------------------------------ */
/* MTRIDAG, ispecial = 0 branch: mtridag.f:95-112 (Thomas algorithm over N1
 * independent columns; R is [N2][N1], solution written to U). */
static void orc_mtridag0(const double *A, const double *B, const double *C,
                         const double *R, double *U, int N1, int N2)
{
    double *gam = dalloc(N2), *betad = dalloc(N2), *ud = dalloc(N2);
    betad[0] = B[0];
    for (int j = 1; j < N2; j++) {                 /* mtridag.f:97-100 */
        gam[j] = C[j - 1] / betad[j - 1];
        betad[j] = B[j] - A[j] * gam[j];
    }
    for (int i = 0; i < N1; i++) {                 /* mtridag.f:101-112 */
        ud[0] = R[i] / betad[0];
        for (int j = 1; j < N2; j++)
            ud[j] = (R[(size_t)j * N1 + i] - A[j] * ud[j - 1]) / betad[j];
        for (int j = N2 - 2; j >= 0; j--)
            ud[j] = ud[j] - gam[j + 1] * ud[j + 1];
        for (int j = 0; j < N2; j++)
            U[(size_t)j * N1 + i] = ud[j];
    }
    free(gam); free(betad); free(ud);
}

/* DBYD2: dbyd2.f:8-200, ibc in 0..4 (the non-periodic branch).  All scalars
 * of dbyd2.f are implicit default REAL, so they are float here and are
 * widened to double when they multiply the real64 arrays.  A and B are
 * [N2][N1].  For ibc 2 (3) the caller's A(1) (A(N2)) row is used as is. */
void orc_dbyd2(double *A, const double *B, int N1, int N2, int ibc)
{
    if (ibc < 0 || ibc > 4) return;
    float d2i = (float)(N2 - 1);                   /* dbyd2.f:43-47 */
#define AR(j, i) A[(size_t)(j) * N1 + (i)]
#define BR(j, i) B[(size_t)(j) * N1 + (i)]
    if (N2 < 5) {                                  /* dbyd2.f:50-84 */
        for (int j = 0; j < N2; j++)
            for (int i = 0; i < N1; i++) {
                if (ibc == 0) AR(j, i) = 0.0;
                else if (ibc == 1) AR(j, i) = BR(N2 - 1, i) - BR(0, i);
                else if (ibc == 2) AR(j, i) = AR(0, i);
                else if (ibc == 3) AR(j, i) = AR(N2 - 1, i);
                else AR(j, i) = ((float)j * AR(0, i) + (float)(N2 - 1 - j) * AR(N2 - 1, i)) / d2i;
            }
        return;
    }
    float C1 = d2i * 7.0f / 9.0f;                  /* dbyd2.f:85-86 */
    float C2 = d2i / 36.0f;
    for (int j = 2; j <= N2 - 3; j++)              /* dbyd2.f:91-96 */
        for (int i = 0; i < N1; i++)
            AR(j, i) = (double)C1 * (BR(j + 1, i) - BR(j - 1, i)) +
                       (double)C2 * (BR(j + 2, i) - BR(j - 2, i));
    {
        float c1 = -(43.f / 96.f * d2i);           /* dbyd2.f:103-107 */
        float c2 = -(5.f / 6.f * d2i);
        float c3 = 9.f / 8.f * d2i;
        float c4 = 1.f / 6.f * d2i;
        float c5 = -(1.f / 96.f * d2i);
        for (int i = 0; i < N1; i++)               /* dbyd2.f:109-117 */
            AR(1, i) = (double)c1 * BR(0, i) + (double)c2 * BR(1, i) + (double)c3 * BR(2, i) +
                       (double)c4 * BR(3, i) + (double)c5 * BR(4, i);
        for (int i = 0; i < N1; i++)
            AR(N2 - 2, i) = -((double)c1 * BR(N2 - 1, i) + (double)c2 * BR(N2 - 2, i) +
                              (double)c3 * BR(N2 - 3, i) + (double)c4 * BR(N2 - 4, i) +
                              (double)c5 * BR(N2 - 5, i));
    }
    {
        float f1 = -(10.0f / 3.f * d2i);           /* dbyd2.f:121-124 */
        float f2 = -(3.0f * d2i);
        float f3 = 6.f * d2i;
        float f4 = 1.f / 3.f * d2i;
        if (ibc == 0) {                            /* dbyd2.f:125-134 */
            for (int i = 0; i < N1; i++) AR(0, i) = 0.0;
        } else if (ibc == 1 || ibc == 3) {
            for (int i = 0; i < N1; i++)
                AR(0, i) = (double)f1 * BR(0, i) + (double)f2 * BR(1, i) +
                           (double)f3 * BR(2, i) + (double)f4 * BR(3, i);
        }
        if (ibc == 0) {                            /* dbyd2.f:135-145 */
            for (int i = 0; i < N1; i++) AR(N2 - 1, i) = 0.0;
        } else if (ibc == 1 || ibc == 2) {
            for (int i = 0; i < N1; i++)
                AR(N2 - 1, i) = -((double)f1 * BR(N2 - 1, i) + (double)f2 * BR(N2 - 2, i) +
                                  (double)f3 * BR(N2 - 3, i) + (double)f4 * BR(N2 - 4, i));
        }
    }
    /* Matrix: dbyd2.f:151-195 */
    float alpha = 1.f / 3.f, alpha2 = 3.f / 4.f, gamma2 = 1.f / 8.f;
    float alpha1 = 6.0f, beta1 = 3.f;
    float cnst = 1.f / (alpha2 - beta1 * gamma2);
    float up1 = (alpha1 * alpha2 - beta1) * cnst;
    double *uppr = dalloc(N2), *lowr = dalloc(N2), *diag = dalloc(N2);
    for (int j = 0; j < N2; j++) diag[j] = 1.0;    /* dbyd2.f:30 */
    for (int j = 2; j <= N2 - 3; j++) uppr[j] = alpha;
    if (ibc % 2 != 0)                              /* dbyd2.f:163-167 */
        for (int i = 0; i < N1; i++)
            AR(0, i) = ((double)alpha2 * AR(0, i) - (double)beta1 * AR(1, i)) * (double)cnst;
    if (ibc == 1 || ibc == 2)                      /* dbyd2.f:168-172 */
        for (int i = 0; i < N1; i++)
            AR(N2 - 1, i) = ((double)alpha2 * AR(N2 - 1, i) - (double)beta1 * AR(N2 - 2, i)) *
                            (double)cnst;
    uppr[0] = (ibc % 2 == 0) ? 0.0 : (double)up1;  /* dbyd2.f:174-179 */
    uppr[1] = alpha2;
    uppr[N2 - 2] = gamma2;
    for (int j = 0; j < N2; j++) lowr[j] = uppr[N2 - 1 - j]; /* dbyd2.f:185-187 */
    lowr[N2 - 1] = (ibc != 1 && ibc != 2) ? 0.0 : (double)up1; /* dbyd2.f:189-195 */
    orc_mtridag0(lowr, diag, uppr, A, A, N1, N2);  /* dbyd2.f:200 */
    free(uppr); free(lowr); free(diag);
#undef AR
#undef BR
}

/* ------------------------------------------------------------------------ */
/* Radix-2 complex FFT, sum_j x_j exp(sign*2*pi*i*j*k/n), unnormalized.
 * Stands in for FFTW 3.2's guru r2c/c2r plans (initialize.f90:325-328),
 * which compute the same DFT up to rounding. */
/* DFT for n that is not a power of two (FFTW plans any n; the transform is
 * the same sum): mixed-radix Cooley-Tukey, decimation in time, splitting
 * off the smallest prime factor p of n = p m:
 *   X[k + m q] = sum_{r<p} W_n^{r (k + m q)} Y_r[k],  Y_r = DFT_m(x[r + p j]),
 * with a prime n summed directly.  W_N^e for a sub-transform of length N
 * is the table's W_n0^{e n0/N}; every exponent is reduced exactly. */
static void fft_mixed(const double *xr, const double *xi, int n, int stride, const double *wr,
                      const double *wi, int n0, double *yr, double *yi)
{
    if (n == 1) { yr[0] = xr[0]; yi[0] = xi[0]; return; }
    int p = 2;
    while (p * p <= n && n % p) p++;
    if (n % p) p = n;                               /* n is prime */
    int m = n / p, sc = n0 / n;
    if (p == n) {
        for (int k = 0; k < n; k++) {
            double ar = 0.0, ai = 0.0;
            for (int j = 0, e = 0; j < n; j++) {
                double c = wr[e * sc], sn = wi[e * sc];
                ar += xr[j * stride] * c - xi[j * stride] * sn;
                ai += xr[j * stride] * sn + xi[j * stride] * c;
                e += k; if (e >= n) e -= n;
            }
            yr[k] = ar; yi[k] = ai;
        }
        return;
    }
    double tr[n], ti[n];
    for (int r = 0; r < p; r++)
        fft_mixed(xr + r * stride, xi + r * stride, m, stride * p, wr, wi, n0, tr + r * m, ti + r * m);
    for (int q = 0; q < p; q++)
        for (int k = 0; k < m; k++) {
            int kk = k + m * q;
            double ar = 0.0, ai = 0.0;
            for (int r = 0, e = 0; r < p; r++) {
                double c = wr[e * sc], sn = wi[e * sc];
                ar += tr[r * m + k] * c - ti[r * m + k] * sn;
                ai += tr[r * m + k] * sn + ti[r * m + k] * c;
                e += kk; if (e >= n) e -= n;       /* r kk mod n, kk < n */
            }
            yr[kk] = ar; yi[kk] = ai;
        }
}

static void dft_any(double *re, double *im, int n, int sign)
{
    /* W_n^e = exp(sign 2 pi i e / n), e < n, cached per (n, sign) */
    static int tn = 0, ts = 0;
    static double *wr = NULL, *wi = NULL;
    if (tn != n || ts != sign) {
        free(wr); free(wi);
        wr = dalloc(n); wi = dalloc(n);
        for (int e = 0; e < n; e++) {
            double ang = 2.0 * PI * (double)e / (double)n;
            wr[e] = cos(ang); wi[e] = (sign < 0 ? -1.0 : 1.0) * sin(ang);
        }
        tn = n; ts = sign;
    }
    double yr[n], yi[n];
    fft_mixed(re, im, n, 1, wr, wi, n, yr, yi);
    memcpy(re, yr, (size_t)n * sizeof(double));
    memcpy(im, yi, (size_t)n * sizeof(double));
}

void orc_fft(double *re, double *im, int n, int sign)
{
    if (n & (n - 1)) { dft_any(re, im, n, sign); return; }
    /* twiddle table w_k = exp(-2 pi i k / n), k < n/2, cached per n */
    static int tn = 0;
    static double *twr = NULL, *twi = NULL;
    if (tn != n) {
        free(twr); free(twi);
        twr = dalloc(n / 2 + 1); twi = dalloc(n / 2 + 1);
        for (int k = 0; k < n / 2; k++) {
            double ang = -2.0 * PI * (double)k / (double)n;
            twr[k] = cos(ang); twi[k] = sin(ang);
        }
        tn = n;
    }
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, stride = n / len;
        for (int k = 0; k < half; k++) {
            double wr = twr[k * stride], wi = (sign < 0 ? 1.0 : -1.0) * twi[k * stride];
            for (int s = 0; s < n; s += len) {
                int a = s + k, b = s + k + half;
                double tr = re[b] * wr - im[b] * wi;
                double ti = re[b] * wi + im[b] * wr;
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] += tr; im[a] += ti;
            }
        }
    }
}

/* ddx: derivatives.f90:28-58 (PERIODIC and USE_FFT branch).  r2c, multiply
 * bins 0..nx/2 by eyekx = i*2*pi*k/nx, c2r (imaginary parts of the DC and,
 * for even nx, Nyquist bins ignored, as FFTW's c2r does; for odd nx bin
 * nx/2 = (nx-1)/2 is an ordinary bin), times stretchx. */
static void ddx_rows(int nx, const double *eyekx, double stretchx, double *fre, double *fim,
                     const double *var, double *dvar, int nrows)
{
    int h = nx / 2;
    for (int r = 0; r < nrows; r++) {
        const double *v = var + (size_t)r * nx;
        for (int i = 0; i < nx; i++) { fre[i] = v[i]; fim[i] = 0.0; }
        orc_fft(fre, fim, nx, -1);
        fim[0] = 0.0;                              /* r2c: DC (and Nyquist) are real */
        if (nx % 2 == 0) fim[h] = 0.0;
        for (int k = 0; k <= h; k++) {             /* derivatives.f90:38-44 */
            double ar = fre[k], ai = fim[k], w = eyekx[k];
            fre[k] = ar * 0.0 - ai * w;
            fim[k] = ar * w + ai * 0.0;
        }
        fim[0] = 0.0;                              /* c2r ignores these */
        if (nx % 2 == 0) fim[h] = 0.0;
        for (int k = 1; k < nx - k; k++) { fre[nx - k] = fre[k]; fim[nx - k] = -fim[k]; }
        orc_fft(fre, fim, nx, +1);                 /* derivatives.f90:46 */
        double *d = dvar + (size_t)r * nx;
        for (int i = 0; i < nx; i++) d[i] = fre[i] * stretchx; /* :56 */
    }
}

void orc_ddx(const orc_model *m, const double *var, double *dvar, int nrows)
{
    ddx_rows(m->nx, m->eyekx, m->stretchx, m->fre, m->fim, var, dvar, nrows);
}

/* ddz: derivatives.f90:10-24 (compact_finite_diff branch). */
static void ddz_cols(int nx, int nz, const double *stretch, const double *var, double *dvar,
                     int ibc)
{
    orc_dbyd2(dvar, var, nx, nz, ibc);
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) dvar[(size_t)k * nx + i] *= stretch[k];
}

void orc_ddz(const orc_model *m, const double *var, double *dvar, int ibc)
{
    ddz_cols(m->nx, m->nz, m->stretch, var, dvar, ibc);
}

/* ------------------------------------------------------------------------ */
/* Grid, background, stretch, PML and flow: initialize.f90:194-225,463-474,
 * 731-794; physics2d.f90:23-57,227-231,288-357; driver.f90:94-147. */
orc_model *orc_create(const orc_params *p, const double *bg, const double *c2d,
                      const double *psi)
{
    int nx = p->nx, nz = p->nz;
    if (nx < 8 || nz < 16 || p->npmlbot + p->npmltop >= nz ||
        2 * p->npmlhor >= nx)
        return NULL;
    orc_model *m = (orc_model *)calloc(1, sizeof(orc_model));
    m->p = *p;
    m->nx = nx; m->nz = nz;
    m->nzpml = p->npmlbot + p->npmltop;
    m->nph = 2 * p->npmlhor;
    size_t n2 = (size_t)nx * nz;

    m->x = dalloc(nx);
    m->eyekx = dalloc(nx / 2 + 1);
    m->normx = 1.0 / (double)nx;                   /* initialize.f90:36 */
    for (int i = 0; i < nx; i++) m->x[i] = (double)i / (double)nx;      /* :194-198 */
    for (int k = 0; k <= nx / 2; k++) m->eyekx[k] = (double)k * m->normx * 2.0 * PI; /* :206-208 */
    m->stretchx = RSUN / p->xlength;               /* :225 */

    /* solar_data: initialize.f90:731-794 */
    m->dimc = 1e6; m->dimrho = 1e-3; m->diml = RSUN;
    m->z = dalloc(nz); m->g = dalloc(nz);
    m->c2 = dalloc(n2); m->rho0 = dalloc(n2); m->p0 = dalloc(n2);
    double dimc2 = pow(m->dimc, 2.0);
    for (int k = 0; k < nz; k++) {
        const double *row = bg + 6 * k;
        m->z[k] = row[0];
        double cs = row[1] / m->dimc;
        for (int i = 0; i < nx; i++) {
            double c = c2d ? c2d[(size_t)k * nx + i] / m->dimc : cs;
            m->c2[(size_t)k * nx + i] = c * c;     /* c2 = c_speed**2.0 */
            m->rho0[(size_t)k * nx + i] = row[2] / m->dimrho;
            m->p0[(size_t)k * nx + i] = row[3] / (m->dimrho * dimc2);
        }
        m->g[k] = row[4] * m->diml / dimc2;
    }
    /* stretch: initialize.f90:466-470 */
    m->unstretch = dalloc(nz); m->stretch = dalloc(nz);
    orc_dbyd2(m->unstretch, m->z, 1, nz, 1);
    for (int k = 0; k < nz; k++) m->stretch[k] = 1.0 / m->unstretch[k];
    m->deltat = p->timestep * m->dimc / m->diml;   /* :474 */

    m->fre = dalloc(nx); m->fim = dalloc(nx);

    /* flow from psi: driver.f90:94-126, reversal :144-147 */
    m->v0x = dalloc(n2); m->v0z = dalloc(n2);
    if (psi) {
        double *pv = dalloc(n2);
        for (size_t q = 0; q < n2; q++)
            pv[q] = m->rho0[q] * sqrt(m->c2[q]) * (psi[q] / (m->diml / 1.0e8));
        ddz_cols(nx, nz, m->stretch, pv, m->v0x, 1);
        for (size_t q = 0; q < n2; q++) m->v0x[q] = -m->v0x[q] / m->rho0[q];
        ddx_rows(nx, m->eyekx, m->stretchx, m->fre, m->fim, pv, m->v0z, nz);
        for (size_t q = 0; q < n2; q++) m->v0z[q] = m->v0z[q] / m->rho0[q];
        if (p->adjoint)
            for (size_t q = 0; q < n2; q++) { m->v0x[q] = -m->v0x[q]; m->v0z[q] = -m->v0z[q]; }
        free(pv);
    }

    /* mp_Initialize_RHS_2d: physics2d.f90:23-405 */
    {
        double temp2 = 1.0 + p->excitdep / RSUN, temp = 1.0;      /* :33-41 */
        for (int k = 0; k < nz; k++)
            if (fabs(m->z[k] - temp2) <= temp) { temp = fabs(m->z[k] - temp2); m->e_rad = k; }
        double temp_r = 1.0 + p->obsheight / RSUN;                 /* :44-51 */
        temp = 1.0;
        for (int k = 0; k < nz; k++)
            if (fabs(m->z[k] - temp_r) <= temp) { temp = fabs(m->z[k] - temp_r); m->o_rad = k; }
        int r = p->adjoint ? m->o_rad : m->e_rad;                  /* :56-57 */
        if (r < 1 || r > nz - 2) { orc_destroy(m); return NULL; }
        m->delta_width = 2.0 / (m->z[r + 1] - m->z[r - 1]);
    }
    m->rhoinv = dalloc(n2); m->c2rho0 = dalloc(n2);
    for (size_t q = 0; q < n2; q++) {                              /* :227-228 */
        m->rhoinv[q] = 1.0 / m->rho0[q];
        m->c2rho0[q] = m->c2[q] * m->rho0[q];
    }
    m->gradrho0_z = dalloc(n2); m->gradp0_z = dalloc(n2);
    ddz_cols(nx, nz, m->stretch, m->rho0, m->gradrho0_z, 1);       /* :230 */
    ddz_cols(nx, nz, m->stretch, m->p0, m->gradp0_z, 1);           /* :231 */

    m->az = dalloc(n2); m->bzpml = dalloc(n2);                    /* :288-333 */
    {
        int nb = p->npmlbot, nt = p->npmltop;
        double alpha0 = PI * (double)0.005f * m->diml / m->dimc;
        double *z = m->z;
        for (int k = 0; k < nz; k++) {
            double alphaz = alpha0;
            for (int i = 0; i < nx; i++) {
                double dampz = 0.0;
                size_t q = (size_t)k * nx + i;
                int inpml = 0;
                if (k < nb) {                                      /* :301-310 */
                    alphaz = PI * (double)0.005f * m->diml / m->dimc * (z[k] - z[0]) /
                             (z[nb - 1] - z[0]);
                    double rr = (z[nb - 1] - z[k]) / (z[nb - 1] - z[0]);
                    dampz = 9.0 * sqrt(m->c2[i]) / (z[nb - 1] - z[0]) * (rr * rr);
                    inpml = 1;
                }
                if (k > nz - nt - 1) {                             /* :312-322 */
                    alphaz = PI * (double)0.005f * m->diml / m->dimc * (z[nz - 1] - z[k]) /
                             (z[nz - 1] - z[nz - nt]);
                    double rr = (z[nz - nt] - z[k]) / (z[nz - 1] - z[nz - nt]);
                    dampz = 9.0 * sqrt(m->c2[(size_t)(nz - 6) * nx + i]) /
                            (z[nz - 1] - z[nz - nt]) * (rr * rr);
                    inpml = 1;
                }
                if (inpml) {
                    double fd = alphaz / (dampz + alphaz);
                    m->gradp0_z[q] *= fd;
                    m->gradrho0_z[q] *= fd;
                    if (i == 0) m->g[k] *= fd;             /* g(k)*functiondecay(1,1,k) */
                }
                m->bzpml[q] = -alphaz - dampz;                     /* :325-326 */
                m->az[q] = -dampz;
            }
        }
    }
    m->axpml = dalloc((size_t)m->nph * nz); m->bxpml = dalloc((size_t)m->nph * nz);
    {                                                              /* :339-357 */
        int nh = p->npmlhor;
        double *x = m->x;
        double beta0 = (double)0.005f * m->diml / m->dimc * PI;
        for (int k = 0; k < nz; k++) {
            double c1 = sqrt(m->c2[(size_t)k * nx + 0]);
            for (int i = 0; i < nh; i++) {
                double rr = (x[i] - x[nh - 1]) / (x[nh - 1] - x[0]);
                double ax = -(2.25 * c1 / (x[nh - 1] - x[0]) * (rr * rr));
                m->axpml[(size_t)k * m->nph + i] = ax;
                m->bxpml[(size_t)k * m->nph + i] = ax - beta0 * (x[i] - x[0]) / (x[nh - 1] - x[0]);
            }
            for (int c = 0; c < nh; c++) {
                int i = nx - nh + c;
                double rr = (x[nx - nh] - x[i]) / (x[nx - nh] - x[nx - 1]);
                double ax = -(2.25 * c1 / (x[nx - 1] - x[nx - nh]) * (rr * rr));
                m->axpml[(size_t)k * m->nph + nh + c] = ax;
                m->bxpml[(size_t)k * m->nph + nh + c] =
                    ax - beta0 * (x[nx - 1] - x[i]) / (x[nx - 1] - x[nx - nh]);
            }
        }
    }

    /* INITIALIZE_STEP: step.f90:24-41 */
    {
        double o2 = 0.5, o3 = (double)0.1665579725791151184f;
        double o4 = (double)0.03950410250586729987f, o5 = (double)0.007810706393337838236f;
        double *b = m->betas;
        b[4] = o2;
        b[3] = o3 / b[4];
        b[2] = o4 / (b[4] * b[3]);
        b[1] = o5 / (b[4] * b[3] * b[2]);
        b[0] = 0.0;
        for (int s = 0; s < 5; s++) b[s] *= m->deltat;
    }

    m->a = dalloc(4 * n2); m->temp = dalloc(4 * n2); m->scr = dalloc(4 * n2);
    size_t npz = (size_t)2 * m->nzpml * nx, npx = (size_t)2 * nz * m->nph;
    m->pmlvars = dalloc(npz); m->psipml = dalloc(npz); m->scrpml = dalloc(npz);
    m->pmlvarsx = dalloc(npx); m->psipmlx = dalloc(npx); m->scrpmlx = dalloc(npx);
    m->dxixdx = dalloc(n2); m->dxizdz = dalloc(n2); m->pp = dalloc(n2); m->rho = dalloc(n2);
    m->gradp_x = dalloc(n2); m->gradp_z = dalloc(n2); m->div = dalloc(n2);
    m->forcing = dalloc(nx);

    m->steps = (int)floor(p->outputcad / p->timestep);            /* driver.f90:214 */
    m->cadforcing_step = (int)floor(p->cadforcing / p->timestep); /* initialize.f90:991 */
    return m;
}

void orc_destroy(orc_model *m)
{
    if (!m) return;
    double **ptrs[] = {&m->x, &m->eyekx, &m->z, &m->stretch, &m->unstretch, &m->g, &m->c2,
                       &m->rho0, &m->p0, &m->rhoinv, &m->c2rho0, &m->gradp0_z, &m->gradrho0_z,
                       &m->v0x, &m->v0z, &m->az, &m->bzpml, &m->axpml, &m->bxpml, &m->a,
                       &m->temp, &m->scr, &m->pmlvars, &m->psipml, &m->scrpml, &m->pmlvarsx,
                       &m->psipmlx, &m->scrpmlx, &m->dxixdx, &m->dxizdz, &m->pp, &m->rho,
                       &m->gradp_x, &m->gradp_z, &m->div, &m->fre, &m->fim, &m->vr,
                       &m->fwdsource, &m->forcing, &m->zprof, &m->dzm, &m->dzmN};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(*ptrs[i]);
    free(m);
}

/* Common time bookkeeping: READ_PARAMETERS initialize.f90:991-994 with the
 * step count forced to maxtime, and num_steps from driver.f90:217-218. */
static void set_times(orc_model *m, int maxtime)
{
    m->maxtime = maxtime;
    int cfs = m->cadforcing_step > 0 ? m->cadforcing_step : 1;
    m->forcing_length = maxtime / cfs + 1;
    int ns = (int)floor((double)maxtime * m->p.timestep / m->p.cadforcing) + 2;
    m->num_steps = 2 * (ns / 2) + 2;
}

/* FORWARD_SOURCE: driver.f90:1120-1145. */
void orc_forward_source_series(double timestep, double nupeak, double nuwidth, int nt,
                               double *fwd)
{
    for (int k = 0; k < nt; k++) {
        double ttime = (double)(float)k * timestep;   /* (k-1.)*timestep, k 1-based */
        double d = ttime - (double)0.5f / nupeak;
        fwd[k] = cos(2.0 * PI * nupeak * ttime) * exp(-(2.0 * (d * d) * (nuwidth * nuwidth)));
        if (fabs(fwd[k]) < (double)0.00001f) fwd[k] = 0.0;
    }
    double s = (double)1e-5f;                          /* 10.**(-5.) */
    for (int k = 0; k < nt; k++) fwd[k] = fwd[k] * s;
}

/* DETERMINE_STATUS forward branch: driver.f90:522-571 (x profile of vr). */
int orc_set_forward_source(orc_model *m, double xloc, int maxtime)
{
    if (m->p.adjoint) return -1;
    set_times(m, maxtime);
    int nx = m->nx;
    free(m->fwdsource);
    m->nfwd = maxtime + 8;
    m->fwdsource = dalloc(m->nfwd);
    orc_forward_source_series(m->p.timestep, m->p.nupeak, m->p.nuwidth, m->nfwd, m->fwdsource);
    free(m->zprof);
    m->zprof = NULL;
    free(m->vr);
    m->vr = dalloc(nx);
    m->nvr = 1;
    double e8 = (double)1e-8f;                     /* 10.**(-8) */
    double dx = (m->x[1] - m->x[0]) * m->p.xlength * e8;
    double *xd = dalloc(nx);
    int imin = 0;
    for (int i = 0; i < nx; i++) {
        xd[i] = (m->x[i] - 0.5) * m->p.xlength * e8 - xloc;
        if (fabs(xd[i]) < fabs(xd[imin])) imin = i;  /* minloc: first minimum */
    }
    double k43 = (double)0.43f * PI;
    for (int i = 0; i < nx; i++) {
        double sincx = (i == imin) ? 1.0 : sin(k43 * xd[i]) / (k43 * xd[i]);
        m->vr[i] = exp(-(xd[i] * xd[i]) / (400.0 * (dx * dx))) * sincx;
    }
    free(xd);
    return 0;
}

/* A separable forward source in place of the one-row form of
 * physics2d.f90:1056-1059: RHS v_z(x, z) += vr[x] * series[time] * zprof[z]
 * (zprof holds 1/rho0 and the z normalisation), for time <=
 * forcing_length * cadforcing_step as in :1051. */
int orc_set_forward_profile(orc_model *m, const double *vr, const double *zprof,
                            const double *series, int nseries, int maxtime)
{
    if (m->p.adjoint || nseries < 1) return -1;
    set_times(m, maxtime);
    free(m->fwdsource);
    m->nfwd = nseries;
    m->fwdsource = dalloc(nseries);
    memcpy(m->fwdsource, series, (size_t)nseries * sizeof(double));
    free(m->vr);
    m->vr = dalloc(m->nx);
    m->nvr = 1;
    memcpy(m->vr, vr, (size_t)m->nx * sizeof(double));
    free(m->zprof);
    m->zprof = dalloc(m->nz);
    memcpy(m->zprof, zprof, (size_t)m->nz * sizeof(double));
    return 0;
}

/* DETERMINE_STATUS adjoint branch: driver.f90:584-594 (vr zero-padded). */
int orc_set_adjoint_source(orc_model *m, const double *vr, int nt, int maxtime)
{
    if (!m->p.adjoint) return -1;
    set_times(m, maxtime);
    int nx = m->nx, fl = m->forcing_length;
    free(m->vr);
    m->nvr = fl + 8;
    m->vr = dalloc((size_t)m->nvr * nx);
    int ncopy = nt < fl ? nt : fl;
    memcpy(m->vr, vr, (size_t)ncopy * nx * sizeof(double));
    return 0;
}

/* compute_data branch of the flow set-up: v0_x, v0_z read in m/s from
 * true_vx.fits / true_vz.fits and scaled by 1e2/dimc (driver.f90:65-73),
 * in place of the flow from psi.  An adjoint model reverses it, as the psi
 * flow is reversed (driver.f90:144-147). */
int orc_set_flow(orc_model *m, const double *vx_ms, const double *vz_ms)
{
    size_t n2 = (size_t)m->nx * m->nz;
    double sgn = m->p.adjoint ? -1.0 : 1.0;
    for (size_t q = 0; q < n2; q++) {
        m->v0x[q] = sgn * (vx_ms[q] * 1e2 / m->dimc);
        m->v0z[q] = sgn * (vz_ms[q] * 1e2 / m->dimc);
    }
    return 0;
}

/* LAGRANGE_INTERP: physics.f90:678-737. vr slice j (1-based) is node time
 * (j-1)*cadforcing. */
void orc_lagrange(orc_model *m, int time, double *forcing)
{
    double cad = m->p.cadforcing;
    double xt = (double)time * m->p.timestep / cad;
    int time1 = (int)floor(xt) + 1;
    xt = time * m->p.timestep;
    double xs[7];
    for (int q = 0; q < 7; q++) xs[q] = (double)(time1 - 4 + q) * cad;
    int nx = m->nx;
    for (int i = 0; i < nx; i++) {
        double LC[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 7; q++) {
            int ok = 1;
            if (q <= 3) ok = xs[q] >= 0;
            if (q == 5) ok = time1 < m->num_steps;
            if (q == 6) ok = time1 < m->num_steps - 1;
            if (!ok) continue;
            int slice = time1 - 3 + q;             /* vr(:,:,time1-3+q), 1-based */
            double den = 1.0;
            int first = 1;
            for (int r = 0; r < 7; r++) {
                if (r == q) continue;
                if (first) { den = xs[q] - xs[r]; first = 0; }
                else den = den * (xs[q] - xs[r]);
            }
            double v = (slice >= 1 && slice <= m->nvr) ? m->vr[(size_t)(slice - 1) * nx + i] : 0.0;
            LC[q] = v / den;
        }
        double f = 0.0;
        for (int q = 0; q < 7; q++) {
            double t = LC[q];
            for (int r = 0; r < 7; r++)
                if (r != q) t = t * (xt - xs[r]);
            f = (q == 0) ? t : f + t;
        }
        forcing[i] = f;
    }
}

/* ------------------------------------------------------------------------ */
/* MP_QUIET_PML_2D_DISPL: physics2d.f90:943-1094. */
void orc_rhs(orc_model *m, int time, const double *temp4, const double *psipml,
             const double *psipmlx, double *scr4, double *scrpml, double *scrpmlx)
{
    int nx = m->nx, nz = m->nz, nh = m->p.npmlhor, nph = m->nph;
    size_t n2 = (size_t)nx * nz;
    const double *xi_x = temp4, *xi_z = temp4 + n2, *v_x = temp4 + 2 * n2, *v_z = temp4 + 3 * n2;
    double *RHSxi_x = scr4, *RHSxi_z = scr4 + n2, *RHSv_x = scr4 + 2 * n2, *RHSv_z = scr4 + 3 * n2;
    const double *psivz = psipml, *psip = psipml + (size_t)m->nzpml * nx;
    double *RHSpsivz = scrpml, *RHSpsip = scrpml + (size_t)m->nzpml * nx;
    const double *psivx = psipmlx, *psi_gradp_x = psipmlx + (size_t)nz * nph;
    double *RHSpsivx = scrpmlx, *RHS_psi_gradp_x = scrpmlx + (size_t)nz * nph;
    double *dxixdx = m->dxixdx, *dxizdz = m->dxizdz, *p = m->pp, *rho = m->rho;
    double *gradp_x = m->gradp_x, *gradp_z = m->gradp_z, *div = m->div;

    orc_ddx(m, xi_x, dxixdx, nz);                                  /* :953 */
    for (int i = 0; i < nx; i++)                                   /* :954 */
        dxizdz[(size_t)(nz - 1) * nx + i] = (MUT(14) ? 1.0 : -1.0) *
                                            dxixdx[(size_t)(nz - 1) * nx + i] * m->unstretch[nz - 1];
    orc_ddz(m, xi_z, dxizdz, 3);                                   /* :955 */
    int pmlindex = 0;                                              /* :957-967 */
    for (int k = 0; k < nz; k++) {
        if (k < m->p.npmlbot || k >= nz - m->p.npmltop) {
            for (int i = 0; i < nx; i++) {
                size_t q = (size_t)k * nx + i, qp = (size_t)pmlindex * nx + i;
                if (MUT(6)) dxizdz[q] = dxizdz[q] + psivz[qp];
                RHSpsivz[qp] = m->az[q] * dxizdz[q] + m->bzpml[q] * psivz[qp];
                if (!MUT(6)) dxizdz[q] = dxizdz[q] + psivz[qp];
            }
            pmlindex++;
        }
    }
    for (int k = 0; k < nz; k++)                                   /* :970-986 */
        for (int i = 0; i < nh; i++) {
            size_t q = (size_t)k * nx + i, qx = (size_t)k * nph + i;
            if (MUT(5)) RHSpsivx[qx] = m->axpml[qx] * dxixdx[q] + m->bxpml[qx] * psivx[qx];
            dxixdx[q] = dxixdx[q] + psivx[qx];
            if (!MUT(5)) RHSpsivx[qx] = m->axpml[qx] * dxixdx[q] + m->bxpml[qx] * psivx[qx];
            size_t q2 = (size_t)k * nx + nx - nh + i, qx2 = qx + nh;
            if (MUT(5)) RHSpsivx[qx2] = m->axpml[qx2] * dxixdx[q2] + m->bxpml[qx2] * psivx[qx2];
            dxixdx[q2] = dxixdx[q2] + psivx[qx2];
            if (!MUT(5)) RHSpsivx[qx2] = m->axpml[qx2] * dxixdx[q2] + m->bxpml[qx2] * psivx[qx2];
        }
    for (size_t q = 0; q < n2; q++) div[q] = dxixdx[q] + dxizdz[q];              /* :990 */
    for (size_t q = 0; q < n2; q++)                                               /* :998 */
        p[q] = -(m->c2rho0[q] * div[q]) + (MUT(2) ? 1.0 : -1.0) * m->gradp0_z[q] * xi_z[q];
    orc_ddx(m, p, gradp_x, nz);                                    /* :1000 */
    for (int k = 0; k < nz; k++)                                   /* :1003-1020 */
        for (int i = 0; i < nh; i++) {
            size_t q = (size_t)k * nx + i, qx = (size_t)k * nph + i;
            gradp_x[q] = gradp_x[q] + psi_gradp_x[qx];
            RHS_psi_gradp_x[qx] = m->axpml[qx] * gradp_x[q] + m->bxpml[qx] * psi_gradp_x[qx];
            size_t q2 = (size_t)k * nx + nx - nh + i, qx2 = qx + nh;
            gradp_x[q2] = gradp_x[q2] + psi_gradp_x[qx2];
            RHS_psi_gradp_x[qx2] = m->axpml[qx2] * gradp_x[q2] + m->bxpml[qx2] * psi_gradp_x[qx2];
        }
    /* :1023 with ibc 2 keeps row 1 of gradp_z as it is: the previous call's
     * value, which stays 0 along every run (psi_p of that row starts at 0 and
     * its RHS is then bz psi_p); it is set to 0 here so that one evaluation
     * depends on its arguments only, as in the HIP column pass */
    for (int i = 0; i < nx; i++) gradp_z[i] = 0.0;
    orc_ddz(m, p, gradp_z, 2);
    pmlindex = 0;                                                  /* :1025-1035 */
    for (int k = 0; k < nz; k++) {
        if (k < m->p.npmlbot || k >= nz - m->p.npmltop) {
            for (int i = 0; i < nx; i++) {
                size_t q = (size_t)k * nx + i, qp = (size_t)pmlindex * nx + i;
                RHSpsip[qp] = m->az[q] * gradp_z[q] + m->bzpml[q] * psip[qp];
                gradp_z[q] = gradp_z[q] + psip[qp];
            }
            pmlindex++;
        }
    }
    for (size_t q = 0; q < n2; q++) RHSv_x[q] = -m->rhoinv[q] * gradp_x[q];       /* :1039 */
    for (size_t q = 0; q < n2; q++)                                /* :1043 */
        rho[q] = -(m->rho0[q] * div[q]) + (MUT(7) ? 1.0 : -1.0) * m->gradrho0_z[q] * xi_z[q];
    for (int k = 0; k < nz; k++)                                   /* :1044-1046 */
        for (int i = 0; i < nx; i++) {
            size_t q = (size_t)k * nx + i;
            RHSv_z[q] = -m->rhoinv[q] * (gradp_z[q] + (MUT(1) ? -1.0 : 1.0) * rho[q] * m->g[k]);
        }
    memcpy(RHSxi_x, v_x, n2 * sizeof(double));                     /* :1048-1049 */
    memcpy(RHSxi_z, v_z, n2 * sizeof(double));
    if (time <= m->forcing_length * m->cadforcing_step) {          /* :1051-1061 */
        if (m->p.adjoint) {
            int r = m->o_rad;
            for (int i = 0; i < nx; i++)
                RHSv_z[(size_t)r * nx + i] += m->forcing[i] * m->delta_width / m->rho0[(size_t)r * nx + i];
        } else if (m->fwdsource && m->zprof) {
            double fs = (time >= 0 && time < m->nfwd) ? m->fwdsource[time] : 0.0;
            for (int k = 0; k < nz; k++)
                if (m->zprof[k] != 0.0)
                    for (int i = 0; i < nx; i++)
                        RHSv_z[(size_t)k * nx + i] += (m->vr[i] * fs) * m->zprof[k];
        } else if (m->fwdsource) {
            int r = m->e_rad;
            double fs = (time >= 0 && time < m->nfwd) ? m->fwdsource[time] : 0.0;
            for (int i = 0; i < nx; i++)
                RHSv_z[(size_t)r * nx + i] += m->vr[i] * fs * m->delta_width / m->rho0[(size_t)r * nx + i];
        }
    }
    /* flows, :1065-1079; dxixdx is the reference's scratch here */
    const double two = MUT(3) ? 1.0 : 2.0;
    const double *fx = MUT(4) ? m->v0z : m->v0x, *fz = MUT(4) ? m->v0x : m->v0z;
    orc_ddx(m, v_x, dxixdx, nz);
    for (size_t q = 0; q < n2; q++) RHSv_x[q] = RHSv_x[q] - two * fx[q] * dxixdx[q];
    orc_ddz(m, v_x, dxixdx, 1);
    for (size_t q = 0; q < n2; q++) RHSv_x[q] = RHSv_x[q] - two * fz[q] * dxixdx[q];
    orc_ddx(m, v_z, dxixdx, nz);
    for (size_t q = 0; q < n2; q++) RHSv_z[q] = RHSv_z[q] - two * fx[q] * dxixdx[q];
    orc_ddz(m, v_z, dxixdx, 1);
    for (size_t q = 0; q < n2; q++) RHSv_z[q] = RHSv_z[q] - two * fz[q] * dxixdx[q];
    for (int i = 0; i < nx; i++) {                                 /* :1083-1087 */
        RHSv_x[i] = 0.0; RHSv_z[i] = 0.0; RHSxi_x[i] = 0.0; RHSxi_z[i] = 0.0;
    }
    for (int f = 0; f < 4; f++)                                    /* :1089-1092 */
        for (int k = 0; k < nz; k++) {
            scr4[f * n2 + (size_t)k * nx + 0] = 0.0;
            scr4[f * n2 + (size_t)k * nx + nx - 1] = 0.0;
        }
}

/* FILTER_VARS_2D: physics2d.f90:509-556 (HORIZONTAL_PMLS, USE_PML; the FFT
 * decay branch :561 is skipped under horizontal PMLs). */
static const double FC[6] = {0.75390625, 0.41015625, -0.234375, 0.087890625, -0.01953125,
                             0.001953125};

static double filt11(const double *a, size_t q, size_t s)
{
    return FC[0] * a[q] + 0.5 * FC[1] * (a[q - s] + a[q + s]) +
           0.5 * FC[2] * (a[q - 2 * s] + a[q + 2 * s]) + 0.5 * FC[3] * (a[q - 3 * s] + a[q + 3 * s]) +
           0.5 * FC[4] * (a[q - 4 * s] + a[q + 4 * s]) + 0.5 * FC[5] * (a[q - 5 * s] + a[q + 5 * s]);
}

void orc_filter_vars_2d(const orc_model *m, double *a4)
{
    int nx = m->nx, nz = m->nz, nh = m->p.npmlhor;
    size_t n2 = (size_t)nx * nz;
    double *temp = dalloc(n2);
    for (int f = 0; f < 4; f++) {
        double *a = a4 + f * n2;
        for (int k = 5; k <= nz - 6; k++)
            for (int i = 0; i < nx; i++) temp[(size_t)k * nx + i] = filt11(a, (size_t)k * nx + i, nx);
        for (int k = nz - 5; k <= nz - 2; k++)
            for (int i = 0; i < nx; i++) {
                size_t q = (size_t)k * nx + i;
                temp[q] = 0.5 * (a[q + nx] + a[q - nx]);
            }
        for (int k = 5; k <= nz - 2; k++)                          /* kst=6 .. nz-1 */
            for (int i = nh - 1; i <= nx - nh; i++) a[(size_t)k * nx + i] = temp[(size_t)k * nx + i];
    }
    free(temp);
}

/* FILTER_VAR_Z: kernels.f90:891-928. */
void orc_filter_var_z(int nx, int nz, double *var)
{
    size_t n2 = (size_t)nx * nz;
    double *temp = dalloc(n2);
    for (int k = 1; k <= 4; k++)
        for (int i = 0; i < nx; i++) {
            size_t q = (size_t)k * nx + i;
            temp[q] = 0.5 * (var[q + nx] + var[q - nx]);
        }
    for (int k = 5; k <= nz - 6; k++)
        for (int i = 0; i < nx; i++) temp[(size_t)k * nx + i] = filt11(var, (size_t)k * nx + i, nx);
    for (int k = nz - 5; k <= nz - 2; k++)
        for (int i = 0; i < nx; i++) {
            size_t q = (size_t)k * nx + i;
            temp[q] = 0.5 * (var[q + nx] + var[q - nx]);
        }
    memcpy(var + nx, temp + nx, (size_t)(nz - 2) * nx * sizeof(double));
    free(temp);
}

/* STEP: step.f90:43-119 (option 17, USE_PML, HORIZONTAL_PMLS, TEST_IN_2D). */
void orc_step(orc_model *m, int time)
{
    size_t n4 = (size_t)4 * m->nx * m->nz;
    size_t npz = (size_t)2 * m->nzpml * m->nx, npx = (size_t)2 * m->nz * m->nph;
    if (m->p.adjoint && time <= m->forcing_length * m->cadforcing_step)
        orc_lagrange(m, time, m->forcing);                         /* driver.f90:359-360 */
    for (int s = 0; s < 5; s++) {
        if (s == 0) {
            memcpy(m->temp, m->a, n4 * sizeof(double));
            memcpy(m->psipml, m->pmlvars, npz * sizeof(double));
            memcpy(m->psipmlx, m->pmlvarsx, npx * sizeof(double));
        } else {
            double b = m->betas[s];
            for (size_t q = 0; q < n4; q++) m->temp[q] = m->a[q] + b * m->scr[q];
            for (size_t q = 0; q < npz; q++) m->psipml[q] = m->pmlvars[q] + b * m->scrpml[q];
            for (size_t q = 0; q < npx; q++) m->psipmlx[q] = m->pmlvarsx[q] + b * m->scrpmlx[q];
        }
        orc_rhs(m, time, m->temp, m->psipml, m->psipmlx, m->scr, m->scrpml, m->scrpmlx);
    }
    for (size_t q = 0; q < n4; q++) m->a[q] = m->a[q] + m->deltat * m->scr[q];          /* :107 */
    for (size_t q = 0; q < npz; q++) m->pmlvars[q] = m->pmlvars[q] + m->deltat * m->scrpml[q];
    orc_filter_vars_2d(m, m->a);                                                      /* :110 */
    for (size_t q = 0; q < npx; q++) m->pmlvarsx[q] = m->pmlvarsx[q] + m->deltat * m->scrpmlx[q];
}

/* TIMESTEPPING: driver.f90:337-419 (vz output and partial snapshots). */
int orc_run(orc_model *m, int maxtime, double *vz_out, double *snaps)
{
    int nx = m->nx, nz = m->nz, nout = 0;
    size_t n2 = (size_t)nx * nz;
    if (m->maxtime != maxtime) set_times(m, maxtime);
    for (int i = 0; i <= maxtime; i++) {
        orc_step(m, i);
        if (i % m->steps == 0) {
            if (vz_out)
                memcpy(vz_out + (size_t)nout * nx, m->a + 3 * n2 + (size_t)m->o_rad * nx,
                       nx * sizeof(double));
            if (snaps) {
                if (!m->p.adjoint) {
                    double *s = snaps + (size_t)nout * 6 * n2;
                    memcpy(s, m->a, n2 * sizeof(double));
                    memcpy(s + n2, m->a + n2, n2 * sizeof(double));
                    memcpy(s + 2 * n2, m->a + 2 * n2, n2 * sizeof(double));
                    memcpy(s + 3 * n2, m->a + 3 * n2, n2 * sizeof(double));
                    memcpy(s + 4 * n2, m->scr + 2 * n2, n2 * sizeof(double));
                    memcpy(s + 5 * n2, m->scr + 3 * n2, n2 * sizeof(double));
                } else {
                    double *s = snaps + (size_t)nout * 4 * n2;
                    memcpy(s, m->a + 2 * n2, n2 * sizeof(double));
                    memcpy(s + n2, m->a + 3 * n2, n2 * sizeof(double));
                    memcpy(s + 2 * n2, m->scr + 2 * n2, n2 * sizeof(double));
                    memcpy(s + 3 * n2, m->scr + 3 * n2, n2 * sizeof(double));
                }
            }
            nout++;
        }
    }
    return nout;
}

void orc_get_state(const orc_model *m, double *a4)
{
    memcpy(a4, m->a, (size_t)4 * m->nx * m->nz * sizeof(double));
}

void orc_set_state(orc_model *m, const double *a4)
{
    memcpy(m->a, a4, (size_t)4 * m->nx * m->nz * sizeof(double));
}

void orc_get_pml(const orc_model *m, double *pmlvars, double *pmlvarsx)
{
    if (pmlvars) memcpy(pmlvars, m->pmlvars, (size_t)2 * m->nzpml * m->nx * sizeof(double));
    if (pmlvarsx) memcpy(pmlvarsx, m->pmlvarsx, (size_t)2 * m->nz * m->nph * sizeof(double));
}

void orc_get_scr(const orc_model *m, double *scr4)
{
    memcpy(scr4, m->scr, (size_t)4 * m->nx * m->nz * sizeof(double));
}

int orc_info_int(const orc_model *m, const char *name)
{
    if (!strcmp(name, "e_rad")) return m->e_rad;
    if (!strcmp(name, "o_rad")) return m->o_rad;
    if (!strcmp(name, "steps")) return m->steps;
    if (!strcmp(name, "cadforcing_step")) return m->cadforcing_step;
    if (!strcmp(name, "forcing_length")) return m->forcing_length;
    if (!strcmp(name, "num_steps")) return m->num_steps;
    if (!strcmp(name, "nzpml")) return m->nzpml;
    return -1;
}

double orc_info_double(const orc_model *m, const char *name)
{
    if (!strcmp(name, "deltat")) return m->deltat;
    if (!strcmp(name, "stretchx")) return m->stretchx;
    if (!strcmp(name, "delta_width")) return m->delta_width;
    if (!strcmp(name, "beta1")) return m->betas[1];
    if (!strcmp(name, "beta2")) return m->betas[2];
    if (!strcmp(name, "beta3")) return m->betas[3];
    if (!strcmp(name, "beta4")) return m->betas[4];
    if (!strcmp(name, "dimc")) return m->dimc;
    if (!strcmp(name, "diml")) return m->diml;
    if (!strcmp(name, "dimrho")) return m->dimrho;
    return NAN;
}

int orc_get_array(const orc_model *m, const char *name, double *out, int cap)
{
    size_t n2 = (size_t)m->nx * m->nz;
    struct { const char *n; const double *p; size_t len; } t[] = {
        {"z", m->z, (size_t)m->nz}, {"stretch", m->stretch, (size_t)m->nz},
        {"unstretch", m->unstretch, (size_t)m->nz}, {"g", m->g, (size_t)m->nz},
        {"x", m->x, (size_t)m->nx}, {"eyekx", m->eyekx, (size_t)m->nx / 2 + 1},
        {"rho0", m->rho0, n2}, {"rhoinv", m->rhoinv, n2}, {"c2rho0", m->c2rho0, n2},
        {"c2", m->c2, n2}, {"gradp0_z", m->gradp0_z, n2}, {"gradrho0_z", m->gradrho0_z, n2},
        {"v0x", m->v0x, n2}, {"v0z", m->v0z, n2}, {"az", m->az, n2}, {"bzpml", m->bzpml, n2},
        {"axpml", m->axpml, (size_t)m->nph * m->nz}, {"bxpml", m->bxpml, (size_t)m->nph * m->nz},
        {"fwdsource", m->fwdsource, (size_t)m->nfwd}, {"vr", m->vr, (size_t)m->nvr * m->nx},
        {"forcing", m->forcing, (size_t)m->nx},
    };
    for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); i++)
        if (!strcmp(name, t[i].n)) {
            if (!t[i].p) return 0;
            if ((size_t)cap < t[i].len) return -1;
            memcpy(out, t[i].p, t[i].len * sizeof(double));
            return (int)t[i].len;
        }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* PRODUCE_KERNELS for the 2D flow and sound-speed kernels:
 * kernels.f90:13-612 with INIT_KERNEL/kernel_back initialize.f90:590-724,
 * 937-979; ddzkern :648-661 (dbyd2 * stretchkern then FILTER_VAR_Z),
 * ddxkern :665-691, COMPUTE_DIV :843-861, COMPUTE_DEL :866-887,
 * CURL_KERN :815-838 (bcz = 0: never assigned, static storage). */
int orc_kernels(const orc_params *p, const double *bg, const double *c2d, int nt_kern,
                const double *fwd, const double *adj, double *kc, double *kpsi, double *kvx,
                double *kvz, double *hess)
{
    int nx = p->nx, nz = p->nz;
    size_t n2 = (size_t)nx * nz;
    if (nx < 8 || nz < 16 || nt_kern < 1) return -1;
    double dimc = (double)1e6f;                    /* 10.0**6. */
    double dimrho = (double)1e-3f;                 /* 10.0**(-3.0) */
    double diml = RSUN;
    double *zk = dalloc(nz), *unst = dalloc(nz), *stk = dalloc(nz);
    double *c2 = dalloc(n2), *rho0 = dalloc(n2);
    for (int k = 0; k < nz; k++) {
        zk[k] = bg[6 * k + 0];
        for (int i = 0; i < nx; i++) {
            size_t q = (size_t)k * nx + i;
            double c = bg[6 * k + 1] / dimc;
            c2[q] = pow(c, 2.0);
            rho0[q] = bg[6 * k + 2] / dimrho;
        }
    }
    if (c2d)                                       /* kernels.f90:35-42 */
        for (size_t q = 0; q < n2; q++) { double c = c2d[q] / dimc; c2[q] = pow(c, 2.0); }
    orc_dbyd2(unst, zk, 1, nz, 1);
    for (int k = 0; k < nz; k++) stk[k] = 1.0 / unst[k];
    double *eyekx = dalloc(nx / 2 + 1), *fre = dalloc(nx), *fim = dalloc(nx);
    double normx = 1.0 / (double)nx;
    for (int k = 0; k <= nx / 2; k++) eyekx[k] = (double)k * normx * 2.0 * PI;
    double stretchx = RSUN / p->xlength;
    int stepskern = (int)floor(p->outputcad / p->timestep);

    double *t1 = dalloc(n2), *t2 = dalloc(n2), *t3 = dalloc(n2);
    double *kv1 = dalloc(n2), *kv3 = dalloc(n2), *kcc = dalloc(n2), *hh = dalloc(n2);
    double *d11 = dalloc(n2), *d13 = dalloc(n2), *d31 = dalloc(n2), *d33 = dalloc(n2);
#define DDXK(in, out) ddx_rows(nx, eyekx, stretchx, fre, fim, in, out, nz)
#define DDZK(in, out, bc) do { ddz_cols(nx, nz, stk, in, out, bc); if (!MUT(10)) orc_filter_var_z(nx, nz, out); } while (0)
    for (int k = 0; k < nt_kern; k++) {
        const double *f = fwd + (size_t)k * 6 * n2;
        const double *a = adj + (size_t)(nt_kern - 1 - k) * 4 * n2;
        const double *fxix = f, *fxiz = f + n2, *fvx = f + 2 * n2, *fvz = f + 3 * n2;
        const double *faccx = f + 4 * n2, *faccz = f + 5 * n2;
        const double *avx = a, *avz = a + n2;
        for (size_t q = 0; q < n2; q++)            /* kernels.f90:293-294 */
            hh[q] = hh[q] + faccz[q] * faccz[q] + (MUT(13) ? 0.0 : faccx[q] * faccx[q]) + 0.0 * 0.0;
        DDXK(fxix, t2); DDZK(fxiz, t1, 1);         /* compute_div(f_xi) */
        for (size_t q = 0; q < n2; q++) t2[q] = t2[q] + t1[q];
        DDXK(avx, t3); DDZK(avz, t1, 1);           /* compute_div(a_vel) */
        for (size_t q = 0; q < n2; q++) t3[q] = t3[q] + t1[q];
        for (size_t q = 0; q < n2; q++) kcc[q] = kcc[q] + t2[q] * t3[q]; /* :303-306 */
        DDXK(fvx, d11); DDXK(fvz, d13);            /* compute_del :866-887 */
        DDZK(fvx, d31, 1); DDZK(fvz, d33, 1);
        for (size_t q = 0; q < n2; q++) {          /* :338-355 (y terms are zero) */
            kv1[q] = d11[q] * avx[q] + 0.0 * 0.0 + d13[q] * avz[q] + kv1[q];
            kv3[q] = d31[q] * avx[q] + 0.0 * 0.0 + d33[q] * avz[q] + kv3[q];
        }
    }
    memcpy(hess, hh, n2 * sizeof(double));
    double e8 = (double)1e-8f, e5 = (double)1e-5f;
    double dimen = pow(diml * e8 * p->xlength * e8 * dimc * e5, -1.0) * 1.0; /* :503 */
    DDZK(kv1, t1, MUT(9) ? 1 : 0);                 /* curl_y = ddz(Kx) - ddx(Kz) */
    DDXK(kv3, t2);
    for (size_t q = 0; q < n2; q++) {
        double kp = t1[q] - t2[q];                 /* :510-513 */
        kpsi[q] = 2.0 * kp * dimen * 1.0 * (rho0[q] * sqrt(c2[q])) * (double)stepskern *
                  p->timestep * (MUT(8) ? 1.0 : dimen);
        const double sv = MUT(12) ? 2.0 : -2.0;    /* :520-522 */
        kvx[q] = sv * kv1[q] * rho0[q] * (double)stepskern * p->timestep * dimen;
        kvz[q] = sv * kv3[q] * rho0[q] * (double)stepskern * p->timestep * dimen;
    }
    double dimen_c = pow(diml * e8 * p->xlength * e8, -1.0) * 1.0; /* :563-566 */
    for (size_t q = 0; q < n2; q++)                /* :579 */
        kc[q] = -kcc[q] * rho0[q] * c2[q] * (double)stepskern * p->timestep * dimen_c * (MUT(11) ? 1.0 : 2.0);
#undef DDXK
#undef DDZK
    free(zk); free(unst); free(stk); free(c2); free(rho0); free(eyekx); free(fre); free(fim);
    free(t1); free(t2); free(t3); free(kv1); free(kv3); free(kcc); free(hh);
    free(d11); free(d13); free(d31); free(d33);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Exact discrete adjoint of the option-17 step: the transpose of orc_step's
 * linear map (5-stage low-storage RK of orc_rhs, z filter, PML auxiliaries)
 * and the gradient of a functional of the vz traces with respect to the
 * point coefficients c^2 rho0, v0x, v0z.  Not a routine of the reference,
 * whose adjoint is the continuous recipe (a time-reversed run with -v0,
 * driver.f90:144-147); this is its transpose, statement by statement. */

static void dz_matrices(orc_model *m)
{
    if (m->dzm) return;
    int nz = m->nz;
    m->dzm = dalloc((size_t)3 * nz * nz);
    m->dzmN = dalloc(nz);
    double *e = dalloc(nz), *y = dalloc(nz);
    for (int ibc = 1; ibc <= 3; ibc++) {
        double *M = m->dzm + (size_t)(ibc - 1) * nz * nz;
        for (int j = 0; j < nz; j++) {
            memset(e, 0, nz * sizeof(double));
            memset(y, 0, nz * sizeof(double));       /* kept rows (ibc 2 / 3) at 0 */
            e[j] = 1.0;
            ddz_cols(1, nz, m->stretch, e, y, ibc);
            for (int i = 0; i < nz; i++) M[(size_t)i * nz + j] = y[i];
        }
    }
    memset(e, 0, nz * sizeof(double));
    memset(m->dzmN, 0, nz * sizeof(double));
    m->dzmN[nz - 1] = 1.0;                           /* the kept last row of ibc 3 */
    ddz_cols(1, nz, m->stretch, e, m->dzmN, 3);
    free(e); free(y);
}

/* out = Dz_ibc^T in, column by column; for ibc 3 also presetN_bar[x] =
 * sum_k dzmN[k] in[k][x], the adjoint of the kept last row. */
void orc_ddz_T(orc_model *m, const double *in, double *out, int ibc, double *presetN_bar)
{
    dz_matrices(m);
    int nx = m->nx, nz = m->nz;
    const double *M = m->dzm + (size_t)(ibc - 1) * nz * nz;
    memset(out, 0, (size_t)nx * nz * sizeof(double));
    for (int i = 0; i < nz; i++)
        for (int j = 0; j < nz; j++) {
            double c = M[(size_t)i * nz + j];
            if (c == 0.0) continue;
            for (int x = 0; x < nx; x++) out[(size_t)j * nx + x] += c * in[(size_t)i * nx + x];
        }
    if (presetN_bar) {
        for (int x = 0; x < nx; x++) presetN_bar[x] = 0.0;
        for (int k = 0; k < nz; k++)
            for (int x = 0; x < nx; x++) presetN_bar[x] += m->dzmN[k] * in[(size_t)k * nx + x];
    }
}

/* Dx^T = -Dx: the multiplier i k is odd in k and the Nyquist bin is dropped. */
void orc_ddx_T(const orc_model *m, const double *in, double *out, int nrows)
{
    orc_ddx(m, in, out, nrows);
    for (size_t q = 0; q < (size_t)nrows * m->nx; q++) out[q] = -out[q];
}

/* Transpose of orc_rhs without its source term (physics2d.f90:943-1094 in
 * reverse).  Rb* are the adjoints of the outputs (scr4, scrpml, scrpmlx),
 * Yb* receive the adjoints of the inputs (temp4, psipml, psipmlx); pbar
 * (may be NULL) receives the adjoint of p. */
void orc_rhs_T(orc_model *m, const double *Rb4, const double *Rbpz, const double *Rbpx,
               double *Yb4, double *Ybpz, double *Ybpx, double *pbar)
{
    int nx = m->nx, nz = m->nz, nh = m->p.npmlhor, nph = m->nph, nzpml = m->nzpml;
    size_t n2 = (size_t)nx * nz;
    double *R = dalloc(4 * n2), *t1 = dalloc(n2), *t2 = dalloc(n2);
    double *gpz = dalloc(n2), *gpx = dalloc(n2), *rhob = dalloc(n2), *pb = dalloc(n2);
    double *divb = dalloc(n2), *dxb = dalloc(n2), *dzb = dalloc(n2), *presetb = dalloc(nx);
    memcpy(R, Rb4, 4 * n2 * sizeof(double));
    memset(Yb4, 0, 4 * n2 * sizeof(double));
    memset(Ybpz, 0, (size_t)2 * nzpml * nx * sizeof(double));
    memset(Ybpx, 0, (size_t)2 * nz * nph * sizeof(double));
    double *Rvx = R + 2 * n2, *Rvz = R + 3 * n2;
    double *xxb = Yb4, *xzb = Yb4 + n2, *vxb = Yb4 + 2 * n2, *vzb = Yb4 + 3 * n2;
    const double *Rpsivz = Rbpz, *Rpsip = Rbpz + (size_t)nzpml * nx;
    double *psivzb = Ybpz, *psipb = Ybpz + (size_t)nzpml * nx;
    const double *Rpsivx = Rbpx, *Rpsigpx = Rbpx + (size_t)nz * nph;
    double *psivxb = Ybpx, *psigpxb = Ybpx + (size_t)nz * nph;
    /* :1083-1092 zeroed rows and columns */
    for (int f = 0; f < 4; f++) {
        for (int i = 0; i < nx; i++) R[f * n2 + i] = 0.0;
        for (int k = 0; k < nz; k++) {
            R[f * n2 + (size_t)k * nx] = 0.0;
            R[f * n2 + (size_t)k * nx + nx - 1] = 0.0;
        }
    }
    /* :1065-1079 flows: RHS v -= 2 v0x dv/dx + 2 v0z dv/dz */
    for (int c = 0; c < 2; c++) {
        const double *Rv = c ? Rvz : Rvx;
        double *vb = c ? vzb : vxb;
        for (size_t q = 0; q < n2; q++) t1[q] = -2.0 * m->v0x[q] * Rv[q];
        orc_ddx_T(m, t1, t2, nz);
        for (size_t q = 0; q < n2; q++) vb[q] += t2[q];
        for (size_t q = 0; q < n2; q++) t1[q] = -2.0 * m->v0z[q] * Rv[q];
        orc_ddz_T(m, t1, t2, 1, NULL);
        for (size_t q = 0; q < n2; q++) vb[q] += t2[q];
    }
    /* :1048-1049 RHS xi = v */
    for (size_t q = 0; q < n2; q++) { vxb[q] += R[q]; vzb[q] += R[n2 + q]; }
    /* :1039, :1044-1046 */
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
            size_t q = (size_t)k * nx + i;
            gpz[q] = -m->rhoinv[q] * Rvz[q];
            rhob[q] = -m->rhoinv[q] * m->g[k] * Rvz[q];
            gpx[q] = -m->rhoinv[q] * Rvx[q];
        }
    /* :1025-1035 z-PML of d p/dz */
    int pmlindex = 0;
    for (int k = 0; k < nz; k++) {
        if (k < m->p.npmlbot || k >= nz - m->p.npmltop) {
            for (int i = 0; i < nx; i++) {
                size_t q = (size_t)k * nx + i, qp = (size_t)pmlindex * nx + i;
                psipb[qp] = gpz[q] + m->bzpml[q] * Rpsip[qp];
                gpz[q] += m->az[q] * Rpsip[qp];
            }
            pmlindex++;
        }
    }
    orc_ddz_T(m, gpz, pb, 2, NULL);                                /* :1023 */
    /* :1003-1020 x-PML of d p/dx */
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nph; i++) {
            size_t q = (size_t)k * nx + (i < nh ? i : nx - nh + (i - nh)), qx = (size_t)k * nph + i;
            double tot = gpx[q] + m->axpml[qx] * Rpsigpx[qx];
            psigpxb[qx] = tot + m->bxpml[qx] * Rpsigpx[qx];
            gpx[q] = tot;
        }
    orc_ddx_T(m, gpx, t2, nz);                                     /* :1000 */
    for (size_t q = 0; q < n2; q++) pb[q] += t2[q];
    /* :998, :1043 p and rho' */
    for (size_t q = 0; q < n2; q++) {
        divb[q] = -m->rho0[q] * rhob[q] - m->c2rho0[q] * pb[q];
        xzb[q] += -m->gradrho0_z[q] * rhob[q] - m->gradp0_z[q] * pb[q];
    }
    /* :990 div = dxixdx + dxizdz; :970-986 x-PML of dxi_x/dx */
    memcpy(dxb, divb, n2 * sizeof(double));
    memcpy(dzb, divb, n2 * sizeof(double));
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nph; i++) {
            size_t q = (size_t)k * nx + (i < nh ? i : nx - nh + (i - nh)), qx = (size_t)k * nph + i;
            double tot = dxb[q] + m->axpml[qx] * Rpsivx[qx];
            psivxb[qx] = tot + m->bxpml[qx] * Rpsivx[qx];
            dxb[q] = tot;
        }
    /* :957-967 z-PML of dxi_z/dz */
    pmlindex = 0;
    for (int k = 0; k < nz; k++) {
        if (k < m->p.npmlbot || k >= nz - m->p.npmltop) {
            for (int i = 0; i < nx; i++) {
                size_t q = (size_t)k * nx + i, qp = (size_t)pmlindex * nx + i;
                psivzb[qp] = dzb[q] + m->bzpml[q] * Rpsivz[qp];
                dzb[q] += m->az[q] * Rpsivz[qp];
            }
            pmlindex++;
        }
    }
    /* :955 ibc 3 with the kept last row :954 = -dxi_x/dx unstretch */
    orc_ddz_T(m, dzb, t2, 3, presetb);
    for (size_t q = 0; q < n2; q++) xzb[q] += t2[q];
    for (int i = 0; i < nx; i++) dxb[(size_t)(nz - 1) * nx + i] += -m->unstretch[nz - 1] * presetb[i];
    orc_ddx_T(m, dxb, t2, nz);                                     /* :953 */
    for (size_t q = 0; q < n2; q++) xxb[q] += t2[q];
    if (pbar) memcpy(pbar, pb, n2 * sizeof(double));
    free(R); free(t1); free(t2); free(gpz); free(gpx); free(rhob); free(pb);
    free(divb); free(dxb); free(dzb); free(presetb);
}

/* Transpose of FILTER_VARS_2D (physics2d.f90:509-556) on the 4 fields. */
void orc_filter_vars_2d_T(const orc_model *m, double *a4)
{
    int nx = m->nx, nz = m->nz, nh = m->p.npmlhor;
    size_t n2 = (size_t)nx * nz;
    double *out = dalloc(n2);
    for (int f = 0; f < 4; f++) {
        double *a = a4 + f * n2;
        memcpy(out, a, n2 * sizeof(double));
        for (int i = nh - 1; i <= nx - nh; i++) {
            for (int k = 5; k <= nz - 2; k++) out[(size_t)k * nx + i] = 0.0;   /* replaced rows */
            for (int k = 5; k <= nz - 6; k++) {
                double l = a[(size_t)k * nx + i];
                out[(size_t)k * nx + i] += FC[0] * l;
                for (int d = 1; d <= 5; d++) {
                    out[(size_t)(k - d) * nx + i] += 0.5 * FC[d] * l;
                    out[(size_t)(k + d) * nx + i] += 0.5 * FC[d] * l;
                }
            }
            for (int k = nz - 5; k <= nz - 2; k++) {
                double l = a[(size_t)k * nx + i];
                out[(size_t)(k - 1) * nx + i] += 0.5 * l;
                out[(size_t)(k + 1) * nx + i] += 0.5 * l;
            }
        }
        memcpy(a, out, n2 * sizeof(double));
    }
    free(out);
}

/* Gradient of J = sum_k <gvz[k], vz_out[k]> (vz_out as orc_run records it,
 * after step k*steps) over a forward run of maxtime+1 steps from the zero
 * state, by the exact transpose of every step.  The model must be a forward
 * model (adjoint = 0) with its source set.  Outputs [nz][nx]: dJ/d(c^2 rho0),
 * dJ/dv0x, dJ/dv0z (code units), and (may be NULL) the adjoint of the
 * initial state lam0 [4][nz][nx].  Every forward state is kept in memory.
 * Returns 0, or -1 on an adjoint model. */
int orc_discrete_gradient(orc_model *m, int maxtime, const double *gvz, double *g_c2rho0,
                          double *g_v0x, double *g_v0z, double *lam0)
{
    if (m->p.adjoint) return -1;
    int nx = m->nx, nz = m->nz;
    size_t n2 = (size_t)nx * nz, n4 = 4 * n2;
    size_t npz = (size_t)2 * m->nzpml * nx, npx = (size_t)2 * nz * m->nph;
    size_t nst = n4 + npz + npx;
    if (m->maxtime != maxtime) set_times(m, maxtime);
    /* forward run keeping every state U_t (before step t) */
    memset(m->a, 0, n4 * sizeof(double));
    memset(m->pmlvars, 0, npz * sizeof(double));
    memset(m->pmlvarsx, 0, npx * sizeof(double));
    double *U = dalloc((size_t)(maxtime + 1) * nst);
    for (int t = 0; t <= maxtime; t++) {
        double *u = U + (size_t)t * nst;
        memcpy(u, m->a, n4 * sizeof(double));
        memcpy(u + n4, m->pmlvars, npz * sizeof(double));
        memcpy(u + n4 + npz, m->pmlvarsx, npx * sizeof(double));
        orc_step(m, t);
    }
    memset(g_c2rho0, 0, n2 * sizeof(double));
    memset(g_v0x, 0, n2 * sizeof(double));
    memset(g_v0z, 0, n2 * sizeof(double));
    /* stage inputs and the forward quantities the gradient pairs with */
    double *Y = dalloc(5 * nst), *R = dalloc(nst);
    double *Dd = dalloc(5 * 5 * n2);   /* [stage][div, dvx/dx, dvx/dz, dvz/dx, dvz/dz] */
    double *lam = dalloc(nst), *acc = dalloc(nst), *Rb = dalloc(nst), *Yb = dalloc(nst);
    double *pb = dalloc(n2);
    for (int t = maxtime; t >= 0; t--) {
        if (t % m->steps == 0) {                 /* output k after step t */
            int k = t / m->steps;
            for (int i = 0; i < nx; i++) lam[3 * n2 + (size_t)m->o_rad * nx + i] += gvz[(size_t)k * nx + i];
        }
        const double *u = U + (size_t)t * nst;
        for (int s = 0; s < 5; s++) {            /* orc_step's stages */
            double *y = Y + (size_t)s * nst;
            for (size_t q = 0; q < nst; q++) y[q] = s == 0 ? u[q] : u[q] + m->betas[s] * R[q];
            orc_rhs(m, t, y, y + n4, y + n4 + npz, R, R + n4, R + n4 + npz);
            double *D = Dd + (size_t)s * 5 * n2;
            memcpy(D, m->div, n2 * sizeof(double));
            orc_ddx(m, y + 2 * n2, D + n2, nz);
            orc_ddz(m, y + 2 * n2, D + 2 * n2, 1);
            orc_ddx(m, y + 3 * n2, D + 3 * n2, nz);
            orc_ddz(m, y + 3 * n2, D + 4 * n2, 1);
        }
        /* back through a_new = filter(a + deltat scr), pml += deltat scrpml */
        orc_filter_vars_2d_T(m, lam);
        memcpy(acc, lam, nst * sizeof(double));
        for (size_t q = 0; q < nst; q++) Rb[q] = m->deltat * lam[q];
        for (int s = 4; s >= 0; s--) {
            orc_rhs_T(m, Rb, Rb + n4, Rb + n4 + npz, Yb, Yb + n4, Yb + n4 + npz, pb);
            const double *D = Dd + (size_t)s * 5 * n2;
            for (int k = 0; k < nz; k++)
                for (int i = 0; i < nx; i++) {
                    size_t q = (size_t)k * nx + i;
                    int z = (k == 0 || i == 0 || i == nx - 1);       /* zeroed RHS entries */
                    double rvx = z ? 0.0 : Rb[2 * n2 + q], rvz = z ? 0.0 : Rb[3 * n2 + q];
                    g_c2rho0[q] += -pb[q] * D[q];
                    g_v0x[q] += -2.0 * (rvx * D[n2 + q] + rvz * D[3 * n2 + q]);
                    g_v0z[q] += -2.0 * (rvx * D[2 * n2 + q] + rvz * D[4 * n2 + q]);
                }
            for (size_t q = 0; q < nst; q++) acc[q] += Yb[q];
            if (s > 0)
                for (size_t q = 0; q < nst; q++) Rb[q] = m->betas[s] * Yb[q];
        }
        memcpy(lam, acc, nst * sizeof(double));
    }
    if (lam0) memcpy(lam0, lam, n4 * sizeof(double));
    free(U); free(Y); free(R); free(Dd); free(lam); free(acc); free(Rb); free(Yb); free(pb);
    return 0;
}
