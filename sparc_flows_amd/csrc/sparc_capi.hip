// sparc_capi.hip -- C ABI (include/sparc_hip.h): handle, background set-up,
// time loop and kernel accumulation of the gfx950 solver.
//
// The set-up mirrors the reference's initialisation (initialize.f90,
// physics2d.f90:23-405, driver.f90:94-147, step.f90:24-41) including its
// float32 literals; its z- and x-derivatives run on the GPU with the same
// kernels as the time loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sparc_hip.h"
#include "sparc_device.h"
#include "sparc_launch.h"
#include "sparc_host.h"
#include "sparc_seg.h"

using namespace sparc;

namespace {

constexpr double kRsun = 6.95989467700e10;   // initialize.f90:37
constexpr double kPi = 3.141592653589793;    // initialize.f90:37

}  // namespace

// Work buffer of sparc_ddz_noalloc: the Thomas sweep's per-column scratch
// (k_ddz_cols) in a file-scope device array, so that entry point allocates
// nothing.  Every GPU the library runs on holds its own copy, 8 MiB, until the
// process ends.  Host code never uses the name `scratch` as an address: the
// per-device address comes from hipGetSymbolAddress (scratch_address).
constexpr size_t kScratchDoubles = (size_t)1 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_dev;   // device address of scratch, by device id
}

// This device's address of scratch, resolved once per device with that device
// current (called from sparc_create).
static hipError_t scratch_address(int dev, double **p)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0) return hipErrorInvalidDevice;
    if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize(dev + 1, nullptr);
    if (!g_scratch_dev[dev]) {
        void *q = nullptr;
        const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[dev] = (double *)q;
    }
    *p = g_scratch_dev[dev];
    return hipSuccess;
}

struct sparc_handle {
    // configuration
    int nx = 0, nz = 0, S = 0, nb = 0, nt = 0, nh = 0, adjoint = 0;
    double xlength = 0, timestep = 0, cadforcing = 0, outputcad = 0, obsheight = 0, excitdep = 0;
    double nupeak = 0, nuwidth = 0;
    std::vector<double> model, c2d, psi;
    std::vector<double> flowx, flowz;   // explicit flow (non-dimensional), replaces psi when set
    Dims d{};
    std::string err;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    double *scratch_dev = nullptr;   // this device's scratch (sparc_ddz_noalloc)

    // host background (kept for queries)
    std::vector<double> x, eyekx, z, stretch, unstretch, g, c2, rho0, p0, rhoinv, c2rho0, gp0, gr0,
        v0x, v0z, azf, bzf, az, bz, axpml, bxpml, fwdsource;
    std::vector<double> gp0z, gr0z, fdec, fz;   // unscaled d p0/dz, d rho0/dz; PML factor (full, PML rows)
    int e_rad = 0, o_rad = 0, steps = 1, cfs = 1;
    double delta_width = 0, deltat = 0, stretchx = 0, betas[5] = {0, 0, 0, 0, 0};
    DbyCoef cf{};

    // device memory
    DevArena bg_mem, st_mem, src_mem, kern_mem, kpair_mem;
    double *kpair = nullptr;        // many-pair accumulation: G[8][Kc], Dx[4][Kc], Dz[4][Kc], ddz work[4][Kc]
    int kpair_cap = 0;              // Kc the buffer holds
    Bg bg{};
    Tri tri[4]{};
    State st{};
    double *d_scratch = nullptr;    // [S][nz][nx] general scratch
    double *kbuf = nullptr;         // kernel accumulation: D[8], KC, KX, KZ, H, t1, t2, work
    long kern_pairs = 0;
    double *d_scratch2 = nullptr;
    int *tctr = nullptr;

    // sources
    Src src{};
    int maxtime = -1, forcing_length = 0, num_steps = 0;
    bool have_src = false;
    bool seg_forced = false;             // segment length set by SPARC_COLSEG_MS
    std::vector<SegTab> seg_cands;       // column-pass tables built so far
    // launch timing (sparc_profile_kernels): events recorded around each pass
    struct Mark { hipEvent_t a, b; int kind; };
    std::vector<hipEvent_t> evpool;
    std::vector<Mark> marks;
    bool profiling = false;
    bool vx_pending = false;

    // kernel-mode background (kernel_back, initialize.f90:937-979)
    double *d_krho0 = nullptr, *d_kc2 = nullptr;

    // segmented column pass: tables and launch shape (chosen by autotune)
    SegTab seg{};
    int seg_tc = 16, seg_ms = 20;
    int row_wide = 0;                    // row pass at nx = 256 on 16 (0), 32 (1) or 64 (2) lanes per row
    bool tuned = false;
    // per-step v_z norm output of the runs (sparc_set_norm_output)
    double *norm2 = nullptr;
    int nnorm = 0, tnorm0 = 0;
};

// One launch of the column pass in the handle's chosen shape.
static hipError_t launch_col(const sparc_handle *h, int tc, int ms, const SegTab &sg, const ColArgs &ca,
                             const Clock &clk, const Sink &sk)
{
    return launch_colsplit(tc, ms, h->d, h->bg, h->cf, sg, h->st, ca, h->src, clk, sk, h->stream);
}

static thread_local std::string g_create_err;

#define HCHECK(h, expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
            return SPARC_E_HIP;                                                            \
        }                                                                                  \
    } while (0)

static int fail(sparc_handle *h, int code, const std::string &msg)
{
    h->err = msg;
    return code;
}

// ---------------------------------------------------------------------------
// dbyd2 coefficients and matrix (dbyd2.f:85-195), float32 scalars widened.
DbyCoef make_coef(int N2)
{
    DbyCoef c{};
    double cb[kCB];
    dbyd2_coefs(N2, c.c1, c.c2, cb);
    c.f1 = cb[CB_F]; c.f2 = cb[CB_F + 1]; c.f3 = cb[CB_F + 2]; c.f4 = cb[CB_F + 3];
    c.b1 = cb[CB_B]; c.b2 = cb[CB_B + 1]; c.b3 = cb[CB_B + 2]; c.b4 = cb[CB_B + 3]; c.b5 = cb[CB_B + 4];
    c.alpha2 = cb[CB_ALPHA2];
    c.beta1 = cb[CB_BETA1];
    c.cnst = cb[CB_CNST];
    return c;
}

// Thomas factors of the ibc matrix: lowr, gam = uppr(j-1)/betad(j-1), 1/betad.
void make_tri(int N2, int ibc, std::vector<double> &lowr, std::vector<double> &gam,
                     std::vector<double> &rbet)
{
    const float alpha = 1.f / 3.f, alpha2 = 3.f / 4.f, gamma2 = 1.f / 8.f;
    const float alpha1 = 6.0f, beta1 = 3.f;
    const float cnst = 1.f / (alpha2 - beta1 * gamma2);
    const float up1 = (alpha1 * alpha2 - beta1) * cnst;
    std::vector<double> uppr(N2, 0.0), betad(N2, 0.0);
    lowr.assign(N2, 0.0);
    gam.assign(N2, 0.0);
    rbet.assign(N2, 0.0);
    for (int j = 2; j <= N2 - 3; j++) uppr[j] = alpha;
    uppr[0] = (ibc % 2 == 0) ? 0.0 : (double)up1;
    uppr[1] = alpha2;
    uppr[N2 - 2] = gamma2;
    for (int j = 0; j < N2; j++) lowr[j] = uppr[N2 - 1 - j];
    lowr[N2 - 1] = (ibc != 1 && ibc != 2) ? 0.0 : (double)up1;
    lowr[0] = 0.0;
    betad[0] = 1.0;
    for (int j = 1; j < N2; j++) {                 // mtridag.f:96-100
        gam[j] = uppr[j - 1] / betad[j - 1];
        betad[j] = 1.0 - lowr[j] * gam[j];
    }
    for (int j = 0; j < N2; j++) rbet[j] = 1.0 / betad[j];
}

// Upload a host vector into a new device buffer of the arena.
hipError_t upload(DevArena &ar, double **dst, const std::vector<double> &v, hipStream_t st)
{
    hipError_t e = ar.alloc(dst, v.size());
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, st);
}

// GPU z-derivative of [nb][nz][ncols] host data (used by the set-up).
static int host_ddz(sparc_handle *h, const std::vector<double> &in, std::vector<double> &out,
                    int ncols, int nbatch, int ibc, bool use_stretch, int tri_ibc)
{
    const size_t n = (size_t)ncols * h->nz * nbatch;
    double *di = nullptr, *dout = nullptr, *dw = nullptr;
    HCHECK(h, hipMalloc(&di, n * sizeof(double)));
    HCHECK(h, hipMalloc(&dout, n * sizeof(double)));
    HCHECK(h, hipMalloc(&dw, n * sizeof(double)));
    HCHECK(h, hipMemcpyAsync(di, in.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HCHECK(h, launch_fill(dout, n, 0.0, h->stream));
    HCHECK(h, launch_ddz_cols(di, dout, ncols, h->nz, nbatch, ibc,
                              use_stretch ? h->bg.stretch : nullptr, h->cf, h->tri[tri_ibc], dw, 0,
                              h->stream));
    out.resize(n);
    HCHECK(h, hipMemcpyAsync(out.data(), dout, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HCHECK(h, hipStreamSynchronize(h->stream));
    (void)hipFree(di);
    (void)hipFree(dout);
    (void)hipFree(dw);
    return SPARC_OK;
}

static int host_ddx(sparc_handle *h, const std::vector<double> &in, std::vector<double> &out,
                    int nrows)
{
    const size_t n = (size_t)nrows * h->nx;
    double *di = nullptr, *dout = nullptr;
    HCHECK(h, hipMalloc(&di, n * sizeof(double)));
    HCHECK(h, hipMalloc(&dout, n * sizeof(double)));
    HCHECK(h, hipMemcpyAsync(di, in.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HCHECK(h, launch_ddx_rows(di, dout, nrows, h->nx, h->bg.tw, h->bg.dmul, h->bg.cdx, h->stretchx, h->stream));
    out.resize(n);
    HCHECK(h, hipMemcpyAsync(out.data(), dout, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HCHECK(h, hipStreamSynchronize(h->stream));
    (void)hipFree(di);
    (void)hipFree(dout);
    return SPARC_OK;
}

// Per-bin multiplier of the spectral derivative over the full complex
// spectrum: eyekx(k) below the Nyquist bin, zero at it (FFTW's c2r drops its
// imaginary part, derivatives.f90:46; only an even nx has one), -eyekx(nx-k)
// above.
static std::vector<double> dmul_table(const sparc_handle *h)
{
    const int n = h->nx;
    std::vector<double> m(n, 0.0);
    for (int k = 0; k < n; k++) m[k] = 2 * k < n ? h->eyekx[k] : (2 * k == n ? 0.0 : -h->eyekx[n - k]);
    return m;
}

// The x-derivative's tables: twiddles exp(-2 pi i k / nx) (from the exact
// octant values where they exist) for the FFT form, and for any other nx the
// map's first column c[m] = sum_k i dmul[k] exp(2 pi i k m / nx)
// = -sum_k dmul[k] sin(2 pi (k m mod nx) / nx) (real: dmul is odd in k; the
// inverse transform's 1/nx is inside eyekx, as in ddx), stored twice over
// (circ_deriv_row indexes [0, 2 nx) without a wrap).
static int upload_x_tables(sparc_handle *h)
{
    const int nx = h->nx;
    std::vector<double> tw(2 * (size_t)nx);
    for (int k = 0; k < nx; k++) {
        const double ang = -2.0 * kPi * (double)k / (double)nx;
        double c = std::cos(ang), s = std::sin(ang);
        if (4 * k == nx || 4 * k == 3 * nx) c = 0.0;
        if (2 * k == nx) s = 0.0;
        tw[2 * k] = c;
        tw[2 * k + 1] = s;
    }
    const std::vector<double> dm = dmul_table(h);
    std::vector<double> cc(2 * (size_t)nx, 0.0);
    if (!nx_has_fft(nx)) {
        std::vector<long double> sn(nx);
        for (int q = 0; q < nx; q++) sn[q] = std::sin(2.0L * (long double)kPi * q / nx);
        for (int m = 0; m < nx; m++) {
            long double acc = 0.0L;
            for (int k = 0; k < nx; k++) acc -= (long double)dm[k] * sn[(int)(((long)k * m) % nx)];
            cc[m] = cc[m + nx] = (double)acc;
        }
    }
    double *p = nullptr;
    HCHECK(h, upload(h->bg_mem, &p, tw, h->stream));
    h->bg.tw = (const double2 *)p;
    HCHECK(h, upload(h->bg_mem, &p, h->eyekx, h->stream));
    h->bg.eyekx = p;
    HCHECK(h, upload(h->bg_mem, &p, dm, h->stream));
    h->bg.dmul = p;
    HCHECK(h, upload(h->bg_mem, &p, cc, h->stream));
    h->bg.cdx = p;
    return SPARC_OK;
}

// Grid, derivative tables (initialize.f90:194-225, 466-474).
static int setup_grid(sparc_handle *h)
{
    const int nx = h->nx, nz = h->nz;
    h->x.resize(nx);
    h->eyekx.resize(nx / 2 + 1);
    const double normx = 1.0 / (double)nx;                   // initialize.f90:36
    for (int i = 0; i < nx; i++) h->x[i] = (double)i / (double)nx;
    for (int k = 0; k <= nx / 2; k++) h->eyekx[k] = (double)k * normx * 2.0 * kPi;
    h->stretchx = kRsun / h->xlength;
    if (int r = upload_x_tables(h)) return r;
    double *p = nullptr;
    h->bg.stretchx = h->stretchx;

    h->cf = make_coef(nz);
    for (int ibc = 0; ibc < 4; ibc++) {
        std::vector<double> lowr, gam, rbet;
        make_tri(nz, ibc, lowr, gam, rbet);
        double *a = nullptr, *b = nullptr, *c = nullptr;
        HCHECK(h, upload(h->bg_mem, &a, lowr, h->stream));
        HCHECK(h, upload(h->bg_mem, &b, gam, h->stream));
        HCHECK(h, upload(h->bg_mem, &c, rbet, h->stream));
        h->tri[ibc] = Tri{a, b, c};
    }
    // solar_data z column and stretch = 1/dbyd2(z) (initialize.f90:466-470)
    h->z.resize(nz);
    for (int k = 0; k < nz; k++) h->z[k] = h->model[6 * (size_t)k];
    std::vector<double> us;
    int r = host_ddz(h, h->z, us, 1, 1, 1, false, 1);
    if (r) return r;
    h->unstretch = us;
    h->stretch.resize(nz);
    for (int k = 0; k < nz; k++) h->stretch[k] = 1.0 / h->unstretch[k];
    HCHECK(h, upload(h->bg_mem, &p, h->stretch, h->stream));
    h->bg.stretch = p;
    HCHECK(h, upload(h->bg_mem, &p, h->unstretch, h->stream));
    h->bg.unstretch = p;
    h->deltat = h->timestep * 1e6 / kRsun;                 // initialize.f90:474
    return SPARC_OK;
}

// solar_data (initialize.f90:731-794), mp_Initialize_RHS_2d
// (physics2d.f90:23-357), flow from psi (driver.f90:94-147).
static int setup_background(sparc_handle *h)
{
    const int nx = h->nx, nz = h->nz;
    const size_t n2 = (size_t)nx * nz;
    const double dimc = 1e6, dimrho = 1e-3, diml = kRsun;
    const double dimc2 = std::pow(dimc, 2.0);
    h->c2.assign(n2, 0.0);
    h->rho0.assign(n2, 0.0);
    h->p0.assign(n2, 0.0);
    h->g.assign(nz, 0.0);
    for (int k = 0; k < nz; k++) {
        const double *row = &h->model[6 * (size_t)k];
        const double cs = row[1] / dimc;
        for (int i = 0; i < nx; i++) {
            const size_t q = (size_t)k * nx + i;
            const double c = h->c2d.empty() ? cs : h->c2d[q] / dimc;
            h->c2[q] = c * c;
            h->rho0[q] = row[2] / dimrho;
            h->p0[q] = row[3] / (dimrho * dimc2);
        }
        h->g[k] = row[4] * diml / dimc2;
    }
    // e_rad, o_rad, delta_width (physics2d.f90:33-57)
    {
        const double temp2 = 1.0 + h->excitdep / kRsun;
        double temp = 1.0;
        for (int k = 0; k < nz; k++)
            if (std::fabs(h->z[k] - temp2) <= temp) { temp = std::fabs(h->z[k] - temp2); h->e_rad = k; }
        const double temp_r = 1.0 + h->obsheight / kRsun;
        temp = 1.0;
        for (int k = 0; k < nz; k++)
            if (std::fabs(h->z[k] - temp_r) <= temp) { temp = std::fabs(h->z[k] - temp_r); h->o_rad = k; }
        const int r = h->adjoint ? h->o_rad : h->e_rad;
        if (r < 1 || r > nz - 2) return fail(h, SPARC_E_ARG, "source/observation row at the grid edge");
        h->delta_width = 2.0 / (h->z[r + 1] - h->z[r - 1]);
    }
    h->rhoinv.assign(n2, 0.0);
    h->c2rho0.assign(n2, 0.0);
    for (size_t q = 0; q < n2; q++) {                       // physics2d.f90:227-228
        h->rhoinv[q] = 1.0 / h->rho0[q];
        h->c2rho0[q] = h->c2[q] * h->rho0[q];
    }
    int r = host_ddz(h, h->rho0, h->gr0, nx, 1, 1, true, 1);  // :230
    if (r) return r;
    r = host_ddz(h, h->p0, h->gp0, nx, 1, 1, true, 1);        // :231
    if (r) return r;
    // x-independent profiles before the PML scales them (the column pass
    // applies the PML factor functiondecay(x, z) itself)
    h->gp0z.assign(nz, 0.0);
    h->gr0z.assign(nz, 0.0);
    for (int k = 0; k < nz; k++) {
        h->gp0z[k] = h->gp0[(size_t)k * nx];
        h->gr0z[k] = h->gr0[(size_t)k * nx];
    }
    h->fdec.assign(n2, 1.0);
    // vertical PML (physics2d.f90:288-333)
    const int nb = h->nb, nt = h->nt;
    h->azf.assign(n2, 0.0);
    h->bzf.assign(n2, 0.0);
    const double *z = h->z.data();
    const double alpha0 = kPi * (double)0.005f * diml / dimc;
    for (int k = 0; k < nz; k++) {
        for (int i = 0; i < nx; i++) {
            double alphaz = alpha0, dampz = 0.0;
            bool inpml = false;
            const size_t q = (size_t)k * nx + i;
            if (k < nb) {
                alphaz = kPi * (double)0.005f * diml / dimc * (z[k] - z[0]) / (z[nb - 1] - z[0]);
                const double rr = (z[nb - 1] - z[k]) / (z[nb - 1] - z[0]);
                dampz = 9.0 * std::sqrt(h->c2[i]) / (z[nb - 1] - z[0]) * (rr * rr);
                inpml = true;
            }
            if (k > nz - nt - 1) {
                alphaz = kPi * (double)0.005f * diml / dimc * (z[nz - 1] - z[k]) / (z[nz - 1] - z[nz - nt]);
                const double rr = (z[nz - nt] - z[k]) / (z[nz - 1] - z[nz - nt]);
                dampz = 9.0 * std::sqrt(h->c2[(size_t)(nz - 6) * nx + i]) / (z[nz - 1] - z[nz - nt]) * (rr * rr);
                inpml = true;
            }
            if (inpml) {
                const double fd = alphaz / (dampz + alphaz);
                h->fdec[q] = fd;
                h->gp0[q] *= fd;
                h->gr0[q] *= fd;
                if (i == 0) h->g[k] *= fd;
            }
            h->bzf[q] = -alphaz - dampz;
            h->azf[q] = -dampz;
        }
    }
    const int nzpml = nb + nt;
    h->az.assign((size_t)nzpml * nx, 0.0);
    h->bz.assign((size_t)nzpml * nx, 0.0);
    h->fz.assign((size_t)nzpml * nx, 1.0);
    for (int pj = 0; pj < nzpml; pj++) {
        const int k = pj < nb ? pj : nz - nt + (pj - nb);
        for (int i = 0; i < nx; i++) {
            h->az[(size_t)pj * nx + i] = h->azf[(size_t)k * nx + i];
            h->bz[(size_t)pj * nx + i] = h->bzf[(size_t)k * nx + i];
            h->fz[(size_t)pj * nx + i] = h->fdec[(size_t)k * nx + i];
        }
    }
    // horizontal PML (physics2d.f90:339-357), c of column x = 1
    const int nh = h->nh, nph = 2 * nh;
    h->axpml.assign((size_t)nph * nz, 0.0);
    h->bxpml.assign((size_t)nph * nz, 0.0);
    {
        const double *xx = h->x.data();
        const double beta0 = (double)0.005f * diml / dimc * kPi;
        for (int k = 0; k < nz; k++) {
            const double c1 = std::sqrt(h->c2[(size_t)k * nx]);
            for (int i = 0; i < nh; i++) {
                const double rr = (xx[i] - xx[nh - 1]) / (xx[nh - 1] - xx[0]);
                const double ax = -(2.25 * c1 / (xx[nh - 1] - xx[0]) * (rr * rr));
                h->axpml[(size_t)k * nph + i] = ax;
                h->bxpml[(size_t)k * nph + i] = ax - beta0 * (xx[i] - xx[0]) / (xx[nh - 1] - xx[0]);
            }
            for (int c = 0; c < nh; c++) {
                const int i = nx - nh + c;
                const double rr = (xx[nx - nh] - xx[i]) / (xx[nx - nh] - xx[nx - 1]);
                const double ax = -(2.25 * c1 / (xx[nx - 1] - xx[nx - nh]) * (rr * rr));
                h->axpml[(size_t)k * nph + nh + c] = ax;
                h->bxpml[(size_t)k * nph + nh + c] =
                    ax - beta0 * (xx[nx - 1] - xx[i]) / (xx[nx - 1] - xx[nx - nh]);
            }
        }
    }
    // flow from psi (driver.f90:94-126), reversed for the adjoint (:144-147)
    h->v0x.assign(n2, 0.0);
    h->v0z.assign(n2, 0.0);
    if (!h->flowx.empty()) {
        // compute_data: flows given directly (driver.f90:65-73)
        for (size_t q = 0; q < n2; q++) {
            h->v0x[q] = h->adjoint ? -h->flowx[q] : h->flowx[q];
            h->v0z[q] = h->adjoint ? -h->flowz[q] : h->flowz[q];
        }
    } else if (!h->psi.empty()) {
        std::vector<double> pv(n2), dz, dx;
        for (size_t q = 0; q < n2; q++)
            pv[q] = h->rho0[q] * std::sqrt(h->c2[q]) * (h->psi[q] / (diml / 1.0e8));
        r = host_ddz(h, pv, dz, nx, 1, 1, true, 1);
        if (r) return r;
        r = host_ddx(h, pv, dx, nz);
        if (r) return r;
        for (size_t q = 0; q < n2; q++) {
            h->v0x[q] = -dz[q] / h->rho0[q];
            h->v0z[q] = dx[q] / h->rho0[q];
            if (h->adjoint) { h->v0x[q] = -h->v0x[q]; h->v0z[q] = -h->v0z[q]; }
        }
    }
    // INITIALIZE_STEP (step.f90:24-41)
    {
        const double o2 = 0.5, o3 = (double)0.1665579725791151184f;
        const double o4 = (double)0.03950410250586729987f, o5 = (double)0.007810706393337838236f;
        double *b = h->betas;
        b[4] = o2;
        b[3] = o3 / b[4];
        b[2] = o4 / (b[4] * b[3]);
        b[1] = o5 / (b[4] * b[3] * b[2]);
        b[0] = 0.0;
        for (int s = 0; s < 5; s++) b[s] *= h->deltat;
    }
    // upload
    h->bg_mem.release();
    return SPARC_OK;
}

static int setup_seg(sparc_handle *h);

static int upload_background(sparc_handle *h)
{
    // grid tables were uploaded into bg_mem by setup_grid and released by
    // setup_background's caller order; re-upload everything here
    const int nz = h->nz;
    double *p = nullptr;
    if (int r = upload_x_tables(h)) return r;
    HCHECK(h, upload(h->bg_mem, &p, h->stretch, h->stream)); h->bg.stretch = p;
    HCHECK(h, upload(h->bg_mem, &p, h->unstretch, h->stream)); h->bg.unstretch = p;
    for (int ibc = 0; ibc < 4; ibc++) {
        std::vector<double> lowr, gam, rbet;
        make_tri(nz, ibc, lowr, gam, rbet);
        double *a = nullptr, *b = nullptr, *c = nullptr;
        HCHECK(h, upload(h->bg_mem, &a, lowr, h->stream));
        HCHECK(h, upload(h->bg_mem, &b, gam, h->stream));
        HCHECK(h, upload(h->bg_mem, &c, rbet, h->stream));
        h->tri[ibc] = Tri{a, b, c};
    }
    HCHECK(h, upload(h->bg_mem, &p, h->rhoinv, h->stream)); h->bg.rhoinv = p;
    {
        std::vector<double> rz(nz);
        for (int k = 0; k < nz; k++) rz[k] = h->rhoinv[(size_t)k * h->nx];
        HCHECK(h, upload(h->bg_mem, &p, rz, h->stream)); h->bg.rhoinv_z = p;
    }
    HCHECK(h, upload(h->bg_mem, &p, h->rho0, h->stream)); h->bg.rho0 = p;
    HCHECK(h, upload(h->bg_mem, &p, h->c2rho0, h->stream)); h->bg.c2rho0 = p;
    HCHECK(h, upload(h->bg_mem, &p, h->gp0, h->stream)); h->bg.gp0 = p;
    HCHECK(h, upload(h->bg_mem, &p, h->gr0, h->stream)); h->bg.gr0 = p;
    HCHECK(h, upload(h->bg_mem, &p, h->v0x, h->stream)); h->bg.v0x = p;
    HCHECK(h, upload(h->bg_mem, &p, h->v0z, h->stream)); h->bg.v0z = p;
    HCHECK(h, upload(h->bg_mem, &p, h->g, h->stream)); h->bg.g = p;
    HCHECK(h, upload(h->bg_mem, &p, h->az, h->stream)); h->bg.az = p;
    HCHECK(h, upload(h->bg_mem, &p, h->bz, h->stream)); h->bg.bz = p;
    HCHECK(h, upload(h->bg_mem, &p, h->fz, h->stream)); h->bg.fz = p;
    HCHECK(h, upload(h->bg_mem, &p, h->axpml, h->stream)); h->bg.ax = p;
    HCHECK(h, upload(h->bg_mem, &p, h->bxpml, h->stream)); h->bg.bx = p;
    const int r = h->adjoint ? h->o_rad : h->e_rad;
    std::vector<double> rs(h->rho0.begin() + (size_t)r * h->nx, h->rho0.begin() + (size_t)(r + 1) * h->nx);
    HCHECK(h, upload(h->bg_mem, &p, rs, h->stream)); h->bg.rho0_src = p;
    // kernel-mode background: dimc = 10.0**6., dimrho = 10.0**(-3.0) in float32
    // (kernel_back, initialize.f90:952-975; model_c, kernels.f90:35-42)
    {
        const size_t n2 = (size_t)h->nx * nz;
        const double kdimc = (double)1e6f, kdimrho = (double)1e-3f;
        std::vector<double> kr(n2), kc(n2);
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < h->nx; i++) {
                const size_t q = (size_t)k * h->nx + i;
                const double c = (h->c2d.empty() ? h->model[6 * (size_t)k + 1] : h->c2d[q]) / kdimc;
                kc[q] = std::pow(c, 2.0);
                kr[q] = h->model[6 * (size_t)k + 2] / kdimrho;
            }
        HCHECK(h, upload(h->bg_mem, &h->d_krho0, kr, h->stream));
        HCHECK(h, upload(h->bg_mem, &h->d_kc2, kc, h->stream));
    }
    int r2 = setup_seg(h);
    if (r2) return r2;
    HCHECK(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

// Segmented solve tables (SegTab) for the column pass: P segments of m rows
// (m from SPARC_COLSEG_SIZES, the last segment padded with identity rows and
// holding at least 3 real rows).
// Segment lengths of the column pass usable for nz rows: the segment count
// fits kSegMax and the last (padded) segment holds at least 3 real rows.
static std::vector<int> seg_lengths(int N)
{
    static const int sizes[] = {
#define X(v) v,
        SPARC_COLSEG_SIZES(X)
#undef X
    };
    std::vector<int> out;
    for (int v : sizes) {
        const int p = (N + v - 1) / v;
        if (p <= kSegMax && N - (p - 1) * v >= 3) out.push_back(v);
    }
    return out;
}

// Segment lengths the autotuner times for nz rows: the usable ones up to
// kTuneMaxMS, whose column-pass kernels hold their state in registers (longer
// segments spill to scratch), or the shortest usable one when none is that
// short.  SPARC_COLSEG_MS still forces any usable length.
constexpr int kTuneMaxMS = 12;
static std::vector<int> tune_lengths(int N)
{
    const std::vector<int> all = seg_lengths(N);
    std::vector<int> out;
    for (int v : all)
        if (v <= kTuneMaxMS) out.push_back(v);
    if (out.empty() && !all.empty()) out.push_back(all[0]);
    return out;
}

// Segmented solve tables (SegTab) of segment length m for the column pass:
// P segments of m rows, the last padded with identity rows.
static int build_seg(sparc_handle *h, int m, SegTab &sg)
{
    const int N = h->nz;
    const int P = (N + m - 1) / m;
    const int nzp = P * m;
    sg.P = P;
    sg.m = m;
    sg.nzp = nzp;
    std::vector<double> ztab((size_t)kZT * nzp + kCB, 0.0);
    {
        double c1, c2;
        dbyd2_coefs(N, c1, c2, ztab.data() + (size_t)kZT * nzp);
    }
    // z profiles; rho0, d rho0/dz, d p0/dz and 1/rho0 come from the 1D model
    // and are the same in every column
    for (int j = 0; j < N; j++) {
        const size_t q = (size_t)j * h->nx;
        ztab[(size_t)0 * nzp + j] = h->stretch[j];
        ztab[(size_t)1 * nzp + j] = h->g[j];
        ztab[(size_t)17 * nzp + j] = h->rho0[q];
        ztab[(size_t)18 * nzp + j] = h->gr0z[j];
        ztab[(size_t)19 * nzp + j] = h->gp0z[j];
        ztab[(size_t)20 * nzp + j] = h->rhoinv[q];
        for (int i = 1; i < h->nx; i++)
            if (h->rho0[q + i] != h->rho0[q] || h->rhoinv[q + i] != h->rhoinv[q] ||
                h->gr0[q + i] != h->gr0z[j] * h->fdec[q + i] || h->gp0[q + i] != h->gp0z[j] * h->fdec[q + i])
                return fail(h, SPARC_E_ARG, "background density / pressure vary along x");
    }
    SegTables t;
    if (!build_seg_tables(N, m, t)) return fail(h, SPARC_E_ARG, "singular interface system");
    std::copy(t.slots.begin(), t.slots.end(), ztab.begin() + (size_t)2 * nzp);   // rows 2-16
    sg.hw = t.hw;
    const std::vector<double> &band = t.band;
    double *p = nullptr;
    HCHECK(h, upload(h->bg_mem, &p, ztab, h->stream));
    sg.ztab = p;
    HCHECK(h, upload(h->bg_mem, &p, band, h->stream));
    sg.Minv = p;
    std::vector<double> stage(band);
    {
        // the segment tables of the first, an interior and the last segment,
        // then the profiles
        std::vector<double> kt;
        if (!seg_kind_tables(t, kt)) return fail(h, SPARC_E_ARG, "interior segment tables differ");
        static_assert(kZS == 15, "segment-table rows");
        stage.insert(stage.end(), kt.begin(), kt.end());
        for (int r = 0; r < kZF; r++)
            stage.insert(stage.end(), ztab.begin() + (size_t)stage_prow(r) * nzp,
                         ztab.begin() + (size_t)(stage_prow(r) + 1) * nzp);
        stage.insert(stage.end(), ztab.begin() + (size_t)kZT * nzp, ztab.begin() + (size_t)kZT * nzp + kCB);
    }
    HCHECK(h, upload(h->bg_mem, &p, stage, h->stream));
    sg.stage = p;
    return SPARC_OK;
}

// Default column-pass tables: the segment length SPARC_COLSEG_MS when it fits
// this nz, else the shortest usable one (the autotuner may pick another).
static int setup_seg(sparc_handle *h)
{
    HCHECK(h, colseg_init());
    const std::vector<int> ms = seg_lengths(h->nz);
    if (ms.empty()) return fail(h, SPARC_E_ARG, "nz not supported by the segmented column pass");
    int m = ms[0];
    h->seg_forced = false;
    if (const char *e = std::getenv("SPARC_COLSEG_MS")) {
        const int v = std::atoi(e);
        for (int w : ms)
            if (w == v) { m = v; h->seg_forced = true; }
    }
    h->seg_cands.clear();
    h->tuned = false;
    SegTab sg{};
    const int r = build_seg(h, m, sg);
    if (r) return r;
    h->seg = sg;
    h->seg_ms = m;
    h->seg_cands.push_back(sg);
    // untuned block width: the widest block that fits
    h->row_wide = 0;
    if (const char *e = std::getenv("SPARC_ROW_WIDE")) {
        const int v = std::atoi(e);
        h->row_wide = (v == 1 || v == 2) && h->nx == 256 ? v : 0;
    }
    h->seg_tc = 0;
    for (int tc : {1, 2, 4, 8, 16})
        if (colsplit_fits(tc, h->d, sg)) h->seg_tc = tc;
    if (!h->seg_tc) return fail(h, SPARC_E_ARG, "no column-pass block shape fits this grid");
    return SPARC_OK;
}

static int alloc_state(sparc_handle *h)
{
    const size_t f = (size_t)h->S * h->nz * h->nx;
    const size_t pz = (size_t)2 * h->S * h->d.nzpml * h->nx;
    const size_t px = (size_t)2 * h->S * h->nz * h->d.nph;
    DevArena &m = h->st_mem;
    State &s = h->st;
    for (int i = 0; i < 4; i++) {
        HCHECK(h, m.alloc(&s.a[i], f));
        HCHECK(h, m.alloc(&s.t[i], f));
    }
    HCHECK(h, m.alloc(&s.vx_unf, f));
    HCHECK(h, m.alloc(&s.pz_base, pz));
    HCHECK(h, m.alloc(&s.pz_tmp, pz));
    HCHECK(h, m.alloc(&s.px_base, px));
    HCHECK(h, m.alloc(&s.px_tmp, px));
    HCHECK(h, m.alloc(&s.dxixdx, f));
    HCHECK(h, m.alloc(&s.dxvx, f));
    HCHECK(h, m.alloc(&s.dxvz, f));
    HCHECK(h, m.alloc(&s.p, f));
    HCHECK(h, m.alloc(&s.rvx, f));
    HCHECK(h, m.alloc(&s.topbc, (size_t)h->S * h->nx));
    HCHECK(h, m.alloc(&s.w_rho, f));
    HCHECK(h, m.alloc(&s.discard, 256));
    HCHECK(h, m.alloc(&h->d_scratch, f));
    HCHECK(h, m.alloc(&h->d_scratch2, f));
    double *t = nullptr;
    HCHECK(h, m.alloc(&t, 1));
    h->tctr = (int *)t;
    return SPARC_OK;
}

static int zero_state(sparc_handle *h)
{
    const size_t f = (size_t)h->S * h->nz * h->nx;
    const size_t pz = (size_t)2 * h->S * h->d.nzpml * h->nx;
    const size_t px = (size_t)2 * h->S * h->nz * h->d.nph;
    State &s = h->st;
    for (int i = 0; i < 4; i++) {
        HCHECK(h, launch_fill(s.a[i], f, 0.0, h->stream));
        HCHECK(h, launch_fill(s.t[i], f, 0.0, h->stream));
    }
    HCHECK(h, launch_fill(s.vx_unf, f, 0.0, h->stream));
    HCHECK(h, launch_fill(s.pz_base, pz, 0.0, h->stream));
    HCHECK(h, launch_fill(s.pz_tmp, pz, 0.0, h->stream));
    HCHECK(h, launch_fill(s.px_base, px, 0.0, h->stream));
    HCHECK(h, launch_fill(s.px_tmp, px, 0.0, h->stream));
    HCHECK(h, launch_fill(s.topbc, (size_t)h->S * h->nx, 0.0, h->stream));
    h->vx_pending = false;
    return SPARC_OK;
}

// ---------------------------------------------------------------------------
// One step: R1 C1 R C2 R C3 R C4 R C5 R5 (see sparc_kernels.hip).
// Runs launch(); while profiling, brackets it with events of the given pass
// kind (0 column, 1 last column, 2 row, 3 first row, 4 last row).
template <typename F>
static int timed_launch(sparc_handle *h, int kind, F launch)
{
    if (!h->profiling) return launch();
    const size_t k = 2 * h->marks.size();
    if (k + 2 > h->evpool.size()) return fail(h, SPARC_E_STATE, "profiling event pool exhausted");
    sparc_handle::Mark m{h->evpool[k], h->evpool[k + 1], kind};
    HCHECK(h, hipEventRecord(m.a, h->stream));
    const int r = launch();
    if (r) return r;
    HCHECK(h, hipEventRecord(m.b, h->stream));
    h->marks.push_back(m);
    return SPARC_OK;
}

static int enqueue_step(sparc_handle *h, int toff, bool filt_vx, const Sink &sk)
{
    const Dims &d = h->d;
    State &s = h->st;
    Clock clk{h->tctr, toff, h->steps};
    const double *b = h->betas;   // b[1..4] = betas(2..5)
    {
        RowArgs ra{};
        ra.xi_x = s.a[0];
        ra.v_z = s.a[3];
        ra.v_x_out = nullptr;
        ra.bx1 = b[1];
        ra.last1 = 0;
        ra.filt_vx = filt_vx ? 1 : 0;
        const int r = timed_launch(h, 3, [&] {
            HCHECK(h, launch_row(0, d, h->bg, s, ra, clk, sk, h->stream, h->row_wide));
            return SPARC_OK;
        });
        if (r) return r;
    }
    for (int stg = 1; stg <= 5; stg++) {
        ColArgs ca{};
        const double *const *T = (stg == 1) ? (const double *const *)s.a : (const double *const *)s.t;
        for (int f = 0; f < 4; f++) {
            ca.T[f] = T[f];
            ca.TN[f] = s.t[f];
        }
        ca.last = stg == 5;
        ca.bnext = stg == 5 ? h->deltat : b[stg];
        int r = timed_launch(h, stg == 5 ? 1 : 0, [&] {
            HCHECK(h, launch_col(h, h->seg_tc, h->seg_ms, h->seg, ca, clk, sk));
            return SPARC_OK;
        });
        if (r) return r;
        RowArgs ra{};
        if (stg < 5) {
            ra.xi_x = s.t[0];
            ra.v_z = s.t[3];
            ra.v_x_out = s.t[2];
            ra.bx2 = b[stg];
            ra.last2 = 0;
            ra.last1 = (stg + 1 == 5);
            ra.bx1 = ra.last1 ? h->deltat : b[stg + 1];
            r = timed_launch(h, 2, [&] {
                HCHECK(h, launch_row(1, d, h->bg, s, ra, clk, sk, h->stream, h->row_wide));
                return SPARC_OK;
            });
        } else {
            ra.v_x_out = s.vx_unf;
            ra.bx2 = h->deltat;
            ra.last2 = 1;
            r = timed_launch(h, 4, [&] {
                HCHECK(h, launch_row(2, d, h->bg, s, ra, clk, sk, h->stream, h->row_wide));
                return SPARC_OK;
            });
        }
        if (r) return r;
    }
    if (sk.norm2 && !h->profiling)
        HCHECK(h, launch_vz_norm2(s.a[3], d.S, (size_t)d.nz * d.nx, clk, sk, h->stream));
    return SPARC_OK;
}

static Sink empty_sink(sparc_handle *h)
{
    Sink sk{};
    sk.orow = h->o_rad;
    sk.adjoint = h->adjoint;
    sk.nfield = h->adjoint ? 4 : 6;
    sk.norm2 = h->norm2;
    sk.nnorm = h->nnorm;
    sk.tnorm0 = h->tnorm0;
    return sk;
}

// ---------------------------------------------------------------------------
extern "C" {

const char *sparc_last_error(const sparc_handle *h)
{
    if (!h) return g_create_err.c_str();
    return h->err.c_str();
}

int sparc_create(const sparc_config *cfg, sparc_handle **out)
{
    if (!cfg || !out || !cfg->model) {
        g_create_err = "sparc_create: null argument";
        return SPARC_E_ARG;
    }
    const int nx = cfg->nx, nz = cfg->nz;
    if (nx < 16 || nx > 1024) {
        g_create_err = "sparc_create: nx must be in [16, 1024]";
        return SPARC_E_ARG;
    }
    if (nz < 16 || cfg->npmlbot < 1 || cfg->npmltop < 1 || cfg->npmlbot + cfg->npmltop >= nz ||
        cfg->npmlhor < 2 || 2 * cfg->npmlhor >= nx || cfg->nsrc < 1 || nz < 6 + 5) {
        g_create_err = "sparc_create: grid / PML sizes out of range";
        return SPARC_E_ARG;
    }
    if ((double)cfg->nsrc * nz * nx >= (double)(1 << 27)) {
        // the column pass addresses [2][S][nz][nx] arrays with 32-bit byte offsets
        g_create_err = "sparc_create: nsrc * nz * nx must stay below 2^27 points";
        return SPARC_E_ARG;
    }
    if (!(cfg->timestep > 0) || !(cfg->xlength > 0) || cfg->outputcad < cfg->timestep ||
        cfg->cadforcing < cfg->timestep) {
        g_create_err = "sparc_create: timestep / cadences out of range";
        return SPARC_E_ARG;
    }
    sparc_handle *h = new sparc_handle();
    h->nx = nx; h->nz = nz; h->S = cfg->nsrc;
    h->nb = cfg->npmlbot; h->nt = cfg->npmltop; h->nh = cfg->npmlhor;
    h->adjoint = cfg->adjoint ? 1 : 0;
    h->xlength = cfg->xlength; h->timestep = cfg->timestep; h->cadforcing = cfg->cadforcing;
    h->outputcad = cfg->outputcad; h->obsheight = cfg->obsheight; h->excitdep = cfg->excitdep;
    h->nupeak = cfg->nupeak; h->nuwidth = cfg->nuwidth;
    h->model.assign(cfg->model, cfg->model + 6 * (size_t)nz);
    if (cfg->c2d) h->c2d.assign(cfg->c2d, cfg->c2d + (size_t)nx * nz);
    if (cfg->psi) h->psi.assign(cfg->psi, cfg->psi + (size_t)nx * nz);
    h->d = Dims{nx, nz, h->S, h->nb, h->nt, h->nb + h->nt, h->nh, 2 * h->nh};
    h->steps = (int)std::floor(h->outputcad / h->timestep);      // driver.f90:214
    h->cfs = (int)std::floor(h->cadforcing / h->timestep);       // initialize.f90:991
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_err = "sparc_create: no HIP device";
        delete h;
        return SPARC_E_HIP;
    }
    (void)hipGetDevice(&h->device);
    {
        const hipError_t e = scratch_address(h->device, &h->scratch_dev);
        if (e != hipSuccess) {
            g_create_err = std::string("sparc_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            delete h;
            return SPARC_E_HIP;
        }
    }
    if (hipStreamCreate(&h->stream) != hipSuccess) {
        g_create_err = "sparc_create: hipStreamCreate failed";
        delete h;
        return SPARC_E_HIP;
    }
    h->own_stream = true;
    int r = setup_grid(h);
    if (!r) r = setup_background(h);
    if (!r) r = upload_background(h);
    if (!r) r = alloc_state(h);
    if (!r) r = zero_state(h);
    if (!r && hipStreamSynchronize(h->stream) != hipSuccess) r = fail(h, SPARC_E_HIP, "set-up sync");
    if (r) {
        g_create_err = "sparc_create: " + h->err;
        sparc_destroy(h);
        return r;
    }
    *out = h;
    return SPARC_OK;
}

int sparc_destroy(sparc_handle *h)
{
    if (!h) return SPARC_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->bg_mem.release();
    h->st_mem.release();
    h->src_mem.release();
    h->kern_mem.release();
    h->kpair_mem.release();
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SPARC_OK;
}

int sparc_set_stream(sparc_handle *h, void *stream)
{
    if (!h) return SPARC_E_ARG;
    HCHECK(h, hipStreamSynchronize(h->stream));
    if (stream) {
        if (h->own_stream) (void)hipStreamDestroy(h->stream);
        h->stream = (hipStream_t)stream;
        h->own_stream = false;
    } else if (!h->own_stream) {
        HCHECK(h, hipStreamCreate(&h->stream));
        h->own_stream = true;
    }
    return SPARC_OK;
}

int sparc_query_int(const sparc_handle *h, const char *name, int *v)
{
    if (!h || !name || !v) return SPARC_E_ARG;
    const std::string n(name);
    if (n == "e_rad") *v = h->e_rad;
    else if (n == "o_rad") *v = h->o_rad;
    else if (n == "steps") *v = h->steps;
    else if (n == "cadforcing_step") *v = h->cfs;
    else if (n == "forcing_length") *v = h->forcing_length;
    else if (n == "num_steps") *v = h->num_steps;
    else if (n == "nsrc") *v = h->S;
    else if (n == "nx") *v = h->nx;
    else if (n == "nz") *v = h->nz;
    else if (n == "nzpml") *v = h->d.nzpml;
    else if (n == "nph") *v = h->d.nph;
    else if (n == "colseg_tc") *v = h->seg_tc;
    else if (n == "row_wide") *v = h->row_wide;
    else if (n == "colseg_ms") *v = h->seg_ms;
    else if (n == "colseg_p") *v = h->seg.P;
    else if (n == "colseg_hw") *v = h->seg.hw;
    else return SPARC_E_ARG;
    return SPARC_OK;
}

int sparc_query_double(const sparc_handle *h, const char *name, double *v)
{
    if (!h || !name || !v) return SPARC_E_ARG;
    const std::string n(name);
    if (n == "deltat") *v = h->deltat;
    else if (n == "stretchx") *v = h->stretchx;
    else if (n == "delta_width") *v = h->delta_width;
    else if (n.size() == 5 && n.compare(0, 4, "beta") == 0 && n[4] >= '1' && n[4] <= '5') *v = h->betas[n[4] - '1'];
    else return SPARC_E_ARG;
    return SPARC_OK;
}

int sparc_get_background(const sparc_handle *h, const char *name, double *out, int cap, int *n)
{
    if (!h || !name) return SPARC_E_ARG;
    const std::string s(name);
    const std::vector<double> *v = nullptr;
    if (s == "z") v = &h->z;
    else if (s == "stretch") v = &h->stretch;
    else if (s == "unstretch") v = &h->unstretch;
    else if (s == "g") v = &h->g;
    else if (s == "rho0") v = &h->rho0;
    else if (s == "rhoinv") v = &h->rhoinv;
    else if (s == "c2rho0") v = &h->c2rho0;
    else if (s == "c2") v = &h->c2;
    else if (s == "gradp0_z") v = &h->gp0;
    else if (s == "gradrho0_z") v = &h->gr0;
    else if (s == "v0x") v = &h->v0x;
    else if (s == "v0z") v = &h->v0z;
    else if (s == "az") v = &h->azf;
    else if (s == "bzpml") v = &h->bzf;
    else if (s == "axpml") v = &h->axpml;
    else if (s == "bxpml") v = &h->bxpml;
    else if (s == "fwdsource") v = &h->fwdsource;
    else if (s == "eyekx") v = &h->eyekx;
    else return SPARC_E_ARG;
    if (n) *n = (int)v->size();
    if (out) {
        if ((size_t)cap < v->size()) return SPARC_E_ARG;
        std::memcpy(out, v->data(), v->size() * sizeof(double));
    }
    return SPARC_OK;
}

int sparc_set_model(sparc_handle *h, const double *c2d, const double *psi)
{
    if (!h) return SPARC_E_ARG;
    HCHECK(h, hipStreamSynchronize(h->stream));
    const size_t n2 = (size_t)h->nx * h->nz;
    if (c2d) h->c2d.assign(c2d, c2d + n2); else h->c2d.clear();
    if (psi) h->psi.assign(psi, psi + n2); else h->psi.clear();
    h->flowx.clear();
    h->flowz.clear();
    h->bg_mem.release();
    int r = setup_grid(h);
    if (!r) r = setup_background(h);
    if (!r) r = upload_background(h);
    return r;
}

int sparc_set_flow(sparc_handle *h, const double *vx, const double *vz)
{
    if (!h) return SPARC_E_ARG;
    if (!vx != !vz) return fail(h, SPARC_E_ARG, "sparc_set_flow needs both components or neither");
    HCHECK(h, hipStreamSynchronize(h->stream));
    const size_t n2 = (size_t)h->nx * h->nz;
    h->flowx.clear();
    h->flowz.clear();
    if (vx) {
        const double dimc = 1e6;
        h->flowx.resize(n2);
        h->flowz.resize(n2);
        for (size_t q = 0; q < n2; q++) {
            if (!std::isfinite(vx[q]) || !std::isfinite(vz[q]))
                return fail(h, SPARC_E_ARG, "sparc_set_flow: non-finite flow value");
            h->flowx[q] = vx[q] * 1e2 / dimc;       // m/s -> cm/s -> units of dimc
            h->flowz[q] = vz[q] * 1e2 / dimc;
        }
    }
    h->bg_mem.release();
    int r = setup_grid(h);
    if (!r) r = setup_background(h);
    if (!r) r = upload_background(h);
    return r;
}

static void set_times(sparc_handle *h, int maxtime)
{
    h->maxtime = maxtime;
    const int cfs = h->cfs > 0 ? h->cfs : 1;
    h->forcing_length = maxtime / cfs + 1;                   // initialize.f90:994
    const int ns = (int)std::floor((double)maxtime * h->timestep / h->cadforcing) + 2;
    h->num_steps = 2 * (ns / 2) + 2;                         // driver.f90:217-218
}

int sparc_set_forward_sources(sparc_handle *h, const double *xsrc, int maxtime)
{
    if (!h || !xsrc || maxtime < 0) return SPARC_E_ARG;
    if (h->adjoint) return fail(h, SPARC_E_STATE, "forward sources on an adjoint handle");
    HCHECK(h, hipStreamSynchronize(h->stream));
    set_times(h, maxtime);
    const int nx = h->nx;
    // FORWARD_SOURCE (driver.f90:1120-1145)
    const int nfwd = maxtime + 8;
    h->fwdsource.assign(nfwd, 0.0);
    for (int k = 0; k < nfwd; k++) {
        const double ttime = (double)(float)k * h->timestep;
        const double dd = ttime - (double)0.5f / h->nupeak;
        double v = std::cos(2.0 * kPi * h->nupeak * ttime) * std::exp(-(2.0 * (dd * dd) * (h->nuwidth * h->nuwidth)));
        if (std::fabs(v) < (double)0.00001f) v = 0.0;
        h->fwdsource[k] = v;
    }
    const double sc = (double)1e-5f;
    for (int k = 0; k < nfwd; k++) h->fwdsource[k] = h->fwdsource[k] * sc;
    // x profile per source (driver.f90:536-553)
    std::vector<double> vr((size_t)h->S * nx);
    const double e8 = (double)1e-8f;
    const double dx = (h->x[1] - h->x[0]) * h->xlength * e8;
    const double k43 = (double)0.43f * kPi;
    std::vector<double> xd(nx);
    for (int b = 0; b < h->S; b++) {
        int imin = 0;
        for (int i = 0; i < nx; i++) {
            xd[i] = (h->x[i] - 0.5) * h->xlength * e8 - xsrc[b];
            if (std::fabs(xd[i]) < std::fabs(xd[imin])) imin = i;
        }
        for (int i = 0; i < nx; i++) {
            const double sincx = (i == imin) ? 1.0 : std::sin(k43 * xd[i]) / (k43 * xd[i]);
            vr[(size_t)b * nx + i] = std::exp(-(xd[i] * xd[i]) / (400.0 * (dx * dx))) * sincx;
        }
    }
    h->src_mem.release();
    double *dvr = nullptr, *dfw = nullptr;
    HCHECK(h, upload(h->src_mem, &dvr, vr, h->stream));
    HCHECK(h, upload(h->src_mem, &dfw, h->fwdsource, h->stream));
    HCHECK(h, hipStreamSynchronize(h->stream));
    Src &s = h->src;
    s = Src{};
    s.mode = 1;
    s.row = h->e_rad;
    s.limit = h->forcing_length * h->cfs;
    s.delta_width = h->delta_width;
    s.vr = dvr;
    s.nvr = 1;
    s.fwd = dfw;
    s.nfwd = nfwd;
    s.timestep = h->timestep;
    s.cadforcing = h->cadforcing;
    s.num_steps = h->num_steps;
    h->have_src = true;
    return SPARC_OK;
}

int sparc_set_forward_profile(sparc_handle *h, const double *vr, const double *zprof, const double *series,
                              int nseries, int maxtime)
{
    if (!h || !vr || !zprof || !series || nseries < 1 || maxtime < 0) return SPARC_E_ARG;
    if (h->adjoint) return fail(h, SPARC_E_STATE, "forward sources on an adjoint handle");
    int zlo = -1, zhi = -1;
    for (int k = 0; k < h->nz; k++) {
        if (!std::isfinite(zprof[k])) return fail(h, SPARC_E_ARG, "non-finite source z profile");
        if (zprof[k] != 0.0) {
            if (zlo < 0) zlo = k;
            zhi = k;
        }
    }
    for (size_t q = 0; q < (size_t)h->S * h->nx; q++)
        if (!std::isfinite(vr[q])) return fail(h, SPARC_E_ARG, "non-finite source x profile");
    for (int k = 0; k < nseries; k++)
        if (!std::isfinite(series[k])) return fail(h, SPARC_E_ARG, "non-finite source time series");
    HCHECK(h, hipStreamSynchronize(h->stream));
    set_times(h, maxtime);
    h->fwdsource.assign(series, series + nseries);
    h->src_mem.release();
    double *dvr = nullptr, *dfw = nullptr, *dzp = nullptr;
    HCHECK(h, upload(h->src_mem, &dvr, std::vector<double>(vr, vr + (size_t)h->S * h->nx), h->stream));
    HCHECK(h, upload(h->src_mem, &dfw, h->fwdsource, h->stream));
    HCHECK(h, upload(h->src_mem, &dzp, std::vector<double>(zprof, zprof + h->nz), h->stream));
    HCHECK(h, hipStreamSynchronize(h->stream));
    Src &s = h->src;
    s = Src{};
    s.mode = zlo >= 0 ? 3 : 0;
    s.row = -1;
    s.limit = h->forcing_length * h->cfs;
    s.delta_width = h->delta_width;
    s.vr = dvr;
    s.nvr = 1;
    s.fwd = dfw;
    s.nfwd = nseries;
    s.timestep = h->timestep;
    s.cadforcing = h->cadforcing;
    s.num_steps = h->num_steps;
    s.zprof = dzp;
    s.zlo = zlo;
    s.zhi = zhi;
    h->have_src = true;
    return SPARC_OK;
}

int sparc_set_adjoint_sources(sparc_handle *h, const double *vr_dev, int nt, int maxtime)
{
    if (!h || !vr_dev || nt < 1 || maxtime < 0) return SPARC_E_ARG;
    if (!h->adjoint) return fail(h, SPARC_E_STATE, "adjoint sources on a forward handle");
    HCHECK(h, hipStreamSynchronize(h->stream));
    set_times(h, maxtime);
    const int nx = h->nx, fl = h->forcing_length;
    const int nvr = fl + 8;                                  // driver.f90:588-591
    h->src_mem.release();
    double *dvr = nullptr;
    HCHECK(h, h->src_mem.alloc(&dvr, (size_t)h->S * nvr * nx));
    HCHECK(h, launch_fill(dvr, (size_t)h->S * nvr * nx, 0.0, h->stream));
    const int ncopy = nt < fl ? nt : fl;
    for (int b = 0; b < h->S; b++)
        HCHECK(h, hipMemcpyAsync(dvr + (size_t)b * nvr * nx, vr_dev + (size_t)b * nt * nx,
                                 (size_t)ncopy * nx * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK(h, hipStreamSynchronize(h->stream));
    Src &s = h->src;
    s = Src{};
    s.mode = 2;
    s.row = h->o_rad;
    s.limit = h->forcing_length * h->cfs;
    s.delta_width = h->delta_width;
    s.vr = dvr;
    s.nvr = nvr;
    s.fwd = nullptr;
    s.nfwd = 0;
    s.timestep = h->timestep;
    s.cadforcing = h->cadforcing;
    s.num_steps = h->num_steps;
    h->have_src = true;
    return SPARC_OK;
}

int sparc_reset(sparc_handle *h)
{
    if (!h) return SPARC_E_ARG;
    int r = zero_state(h);
    if (r) return r;
    HCHECK(h, launch_set_int(h->tctr, 0, h->stream));
    return SPARC_OK;
}

int sparc_step(sparc_handle *h, int t)
{
    if (!h || t < 0) return SPARC_E_ARG;
    if (!h->have_src) return fail(h, SPARC_E_STATE, "sources not set");
    Sink sk = empty_sink(h);
    HCHECK(h, launch_set_int(h->tctr, t, h->stream));
    int r = enqueue_step(h, 0, false, sk);
    if (r) return r;
    HCHECK(h, launch_tick(h->tctr, 1, h->stream));
    Clock clk{h->tctr, 0, h->steps};
    HCHECK(h, launch_finalize_vx(h->d, h->st, clk, sk, h->stream));
    return SPARC_OK;
}

// Row-pass team width for nx = 256: the narrowest form whose waves still
// reach every compute unit -- 64 lanes per row (one row per wave) when the
// 32-lane form (two rows per wave) would leave compute units idle, else 32
// lanes when the 16-lane form (four rows per wave) would, else 16 lanes.
// Measured for one source: 32 lanes 10 % and 64 lanes a further 3 % faster
// gradients; for 16 sources the 16-lane form is 2 % faster than 32 lanes.
// (Timing a few steps with each width at set-up chose the 32-lane form for
// 16 sources too.)  SPARC_ROW_WIDE (0, 1 or 2) forces the choice.
static int tune_rows(sparc_handle *h)
{
    h->row_wide = 0;
    if (h->nx != 256) return SPARC_OK;
    if (const char *e = std::getenv("SPARC_ROW_WIDE")) {
        const int v = std::atoi(e);
        if (v >= 0 && v <= 2) { h->row_wide = v; return SPARC_OK; }
    }
    int dev = 0, ncu = 0;
    HCHECK(h, hipGetDevice(&dev));
    HCHECK(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const long rows = (long)h->S * h->nz;
    const long waves16 = (rows + 3) / 4, waves32 = (rows + 1) / 2;
    h->row_wide = waves32 < ncu ? 2 : (waves16 < ncu ? 1 : 0);
    return SPARC_OK;
}

// Column-pass shapes already timed in this process, by device and grid
// (nx, nz, sources, layer widths): later handles of the same grid take the
// same shape, so identical runs round identically across handles.
namespace {
struct TuneKey {
    int v[7];
    bool operator<(const TuneKey &o) const { return std::lexicographical_compare(v, v + 7, o.v, o.v + 7); }
};
struct TuneChoice { int ms, tc; };
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoice> g_tuned;
}  // namespace

// Pick the segment length and column-block width of the column pass for this
// grid and batch by timing each candidate on the handle's stream (state is
// reset afterwards), or take the shape an earlier handle of the same grid
// chose in this process.  SPARC_COLSEG_MS and SPARC_COLSEG_TC in the
// environment force a choice.
static int autotune(sparc_handle *h)
{
    if (h->tuned) return SPARC_OK;
    h->tuned = true;
    const TuneKey key{{h->device, h->nx, h->nz, h->S, h->nb, h->nt, h->nh}};
    const bool forced = h->seg_forced || std::getenv("SPARC_COLSEG_TC");
    auto fits = [&](int tc, const SegTab &sg) { return colsplit_fits(tc, h->d, sg); };
    if (!forced) {
        TuneChoice c{0, 0};
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(key);
            if (it != g_tuned.end()) c = it->second;
        }
        if (c.tc) {
            SegTab sg{};
            bool have = false;
            for (const SegTab &t : h->seg_cands)
                if (t.m == c.ms) { sg = t; have = true; }
            if (!have) {
                const int r = build_seg(h, c.ms, sg);
                if (r) return r;
                h->seg_cands.push_back(sg);
            }
            h->seg = sg;
            h->seg_ms = c.ms;
            h->seg_tc = c.tc;
            return tune_rows(h);
        }
    }
    int force_tc = 0;
    if (const char *e = std::getenv("SPARC_COLSEG_TC")) {
        const int v = std::atoi(e);
        if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) force_tc = v;
    }
    // candidate segment lengths (only the forced one under SPARC_COLSEG_MS)
    std::vector<int> ms = h->seg_forced ? std::vector<int>{h->seg.m} : tune_lengths(h->nz);
    if (force_tc && ms.size() == 1) {
        if (fits(force_tc, h->seg)) h->seg_tc = force_tc;
        return tune_rows(h);
    }
    Sink sk = empty_sink(h);
    Clock clk{h->tctr, 0, h->steps};
    State &s = h->st;
    ColArgs ca{};
    for (int f = 0; f < 4; f++) { ca.T[f] = s.t[f]; ca.TN[f] = s.t[f]; }
    ca.last = 0;
    ca.bnext = h->betas[2];
    hipEvent_t e0, e1;
    HCHECK(h, hipEventCreate(&e0));
    HCHECK(h, hipEventCreate(&e1));
    // the launchable shapes, then three timing rounds over all of them (a
    // shape's time is its fastest round, so that a shape timed while the
    // device warms up or another queue runs is not passed over)
    struct Cand { SegTab sg; int tc; float t; };
    std::vector<Cand> cands;
    for (int m : ms) {
        SegTab sg{};
        bool have = false;
        for (const SegTab &c : h->seg_cands)
            if (c.m == m) { sg = c; have = true; }
        if (!have) {
            const int r = build_seg(h, m, sg);
            if (r) return r;
            h->seg_cands.push_back(sg);
        }
        for (int tc : {1, 2, 4, 8, 16}) {
            if (force_tc && tc != force_tc) continue;
            if (!fits(tc, sg)) continue;
            if (launch_col(h, tc, m, sg, ca, clk, sk) != hipSuccess) {
                (void)hipGetLastError();     // shape not launchable here (LDS): skip it
                continue;
            }
            cands.push_back(Cand{sg, tc, 1e30f});
        }
    }
    for (int round = 0; round < 3; round++)
        for (Cand &cd : cands) {
            HCHECK(h, hipEventRecord(e0, h->stream));
            for (int r = 0; r < 10; r++) HCHECK(h, launch_col(h, cd.tc, cd.sg.m, cd.sg, ca, clk, sk));
            HCHECK(h, hipEventRecord(e1, h->stream));
            HCHECK(h, hipEventSynchronize(e1));
            float t = 0.f;
            HCHECK(h, hipEventElapsedTime(&t, e0, e1));
            cd.t = std::min(cd.t, t);
        }
    float best = 1e30f;
    SegTab best_sg = h->seg;
    int best_tc = 0;
    for (const Cand &cd : cands)
        if (cd.t < best) { best = cd.t; best_sg = cd.sg; best_tc = cd.tc; }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!best_tc) return fail(h, SPARC_E_ARG, "no column-pass block shape fits this grid");
    h->seg = best_sg;
    h->seg_ms = best_sg.m;
    h->seg_tc = best_tc;
    if (!forced) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned.emplace(key, TuneChoice{best_sg.m, best_tc});
    }
    // the timed launches wrote only stage temporaries: restore the z-PML
    // stage copy, which equals its base between steps
    const size_t pz = (size_t)2 * h->S * h->d.nzpml * h->nx;
    HCHECK(h, hipMemcpyAsync(s.pz_tmp, s.pz_base, pz * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return tune_rows(h);
}

// Steps per captured graph of the time loop.
static int graph_steps()
{
    if (const char *e = std::getenv("SPARC_GRAPH_STEPS")) return std::max(0, std::atoi(e));
    return 16;
}

}  // extern "C"

// Steps t0..t1 (inclusive) of every source from the current state, which is
// at time index t0 with the v_x filter applied; leaves it at t1 + 1, filtered.
// Step t0 is launched directly, the rest in hipGraphs of G steps each, every
// kernel reading the step counter (t = counter + offset in the graph) so one
// graph serves every block of G steps.
static int run_steps(sparc_handle *h, int t0, int t1, const Sink &sk)
{
    if (t1 < t0) return SPARC_OK;
    HCHECK(h, launch_set_int(h->tctr, t0, h->stream));
    int r = enqueue_step(h, 0, false, sk);
    if (r) return r;
    HCHECK(h, launch_tick(h->tctr, 1, h->stream));
    const int nsteps = t1 - t0 + 1;
    int done = 1;
    const int G = graph_steps();
    if (G > 1 && nsteps - done >= G) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HCHECK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        int rc = SPARC_OK;
        for (int g = 0; g < G && rc == SPARC_OK; g++) rc = enqueue_step(h, g, true, sk);
        if (rc == SPARC_OK && launch_tick(h->tctr, G, h->stream) != hipSuccess) rc = SPARC_E_HIP;
        const hipError_t ce = hipStreamEndCapture(h->stream, &graph);
        if (rc) return rc;
        HCHECK(h, ce);
        HCHECK(h, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        while (nsteps - done >= G) {
            HCHECK(h, hipGraphLaunch(exec, h->stream));
            done += G;
        }
        HCHECK(h, hipStreamSynchronize(h->stream));
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
    }
    for (; done < nsteps; done++) {
        r = enqueue_step(h, 0, true, sk);
        if (r) return r;
        HCHECK(h, launch_tick(h->tctr, 1, h->stream));
    }
    Clock clk{h->tctr, 0, h->steps};
    HCHECK(h, launch_finalize_vx(h->d, h->st, clk, sk, h->stream));
    return SPARC_OK;
}

extern "C" {

int sparc_run(sparc_handle *h, int maxtime, double *vz_out, double *snaps)
{
    if (!h || maxtime < 0) return SPARC_E_ARG;
    if (!h->have_src) return fail(h, SPARC_E_STATE, "sources not set");
    if (maxtime != h->maxtime) return fail(h, SPARC_E_STATE, "maxtime differs from the sources' maxtime");
    int r = autotune(h);
    if (r) return r;
    r = sparc_reset(h);
    if (r) return r;
    Sink sk = empty_sink(h);
    const int nout = maxtime / h->steps + 1;
    sk.vz_out = vz_out;
    sk.nout = vz_out ? nout : 0;
    sk.snap = snaps;
    sk.nsnap = snaps ? nout : 0;
    return run_steps(h, 0, maxtime, sk);
}

int sparc_run_range(sparc_handle *h, int t0, int t1, double *vz_out, double *snaps, int k0, int nslots)
{
    if (!h || t0 < 0 || t1 < t0 - 1 || k0 < 0 || nslots < 0) return SPARC_E_ARG;
    if (!h->have_src) return fail(h, SPARC_E_STATE, "sources not set");
    if (t1 > h->maxtime) return fail(h, SPARC_E_ARG, "sparc_run_range beyond the sources' maxtime");
    int r = autotune(h);
    if (r) return r;
    Sink sk = empty_sink(h);
    sk.k0 = k0;
    sk.vz_out = vz_out;
    sk.nout = vz_out ? nslots : 0;
    sk.snap = snaps;
    sk.nsnap = snaps ? nslots : 0;
    return run_steps(h, t0, t1, sk);
}

int sparc_set_norm_output(sparc_handle *h, double *norm2_dev, int t0, int n)
{
    if (!h || (norm2_dev && (n < 1 || t0 < 0))) return SPARC_E_ARG;
    h->norm2 = norm2_dev;
    h->nnorm = norm2_dev ? n : 0;
    h->tnorm0 = t0;
    return SPARC_OK;
}

int sparc_checkpoint_size(const sparc_handle *h, size_t *bytes)
{
    if (!h || !bytes) return SPARC_E_ARG;
    const size_t f = (size_t)h->S * h->nz * h->nx;
    const size_t pz = (size_t)2 * h->S * h->d.nzpml * h->nx, px = (size_t)2 * h->S * h->nz * h->d.nph;
    *bytes = (4 * f + pz + px) * sizeof(double);
    return SPARC_OK;
}

int sparc_save_checkpoint(sparc_handle *h, void *dev)
{
    if (!h || !dev) return SPARC_E_ARG;
    const size_t f = (size_t)h->S * h->nz * h->nx;
    const size_t pz = (size_t)2 * h->S * h->d.nzpml * h->nx, px = (size_t)2 * h->S * h->nz * h->d.nph;
    double *o = (double *)dev;
    for (int i = 0; i < 4; i++)
        HCHECK(h, hipMemcpyAsync(o + i * f, h->st.a[i], f * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK(h, hipMemcpyAsync(o + 4 * f, h->st.pz_base, pz * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK(h, hipMemcpyAsync(o + 4 * f + pz, h->st.px_base, px * sizeof(double), hipMemcpyDeviceToDevice,
                             h->stream));
    return SPARC_OK;
}

int sparc_load_checkpoint(sparc_handle *h, const void *dev)
{
    if (!h || !dev) return SPARC_E_ARG;
    const size_t f = (size_t)h->S * h->nz * h->nx;
    const size_t pz = (size_t)2 * h->S * h->d.nzpml * h->nx, px = (size_t)2 * h->S * h->nz * h->d.nph;
    const double *o = (const double *)dev;
    for (int i = 0; i < 4; i++)
        HCHECK(h, hipMemcpyAsync(h->st.a[i], o + i * f, f * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    // the stage copies of the PML variables equal their bases between steps
    for (double *dst : {h->st.pz_base, h->st.pz_tmp})
        HCHECK(h, hipMemcpyAsync(dst, o + 4 * f, pz * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    for (double *dst : {h->st.px_base, h->st.px_tmp})
        HCHECK(h, hipMemcpyAsync(dst, o + 4 * f + pz, px * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return SPARC_OK;
}

int sparc_get_state(sparc_handle *h, double *a_dev)
{
    if (!h || !a_dev) return SPARC_E_ARG;
    const size_t f = (size_t)h->S * h->nz * h->nx;
    for (int i = 0; i < 4; i++)
        HCHECK(h, hipMemcpyAsync(a_dev + i * f, h->st.a[i], f * sizeof(double), hipMemcpyDeviceToDevice,
                                 h->stream));
    return SPARC_OK;
}

int sparc_set_state(sparc_handle *h, const double *a_dev)
{
    if (!h || !a_dev) return SPARC_E_ARG;
    const size_t f = (size_t)h->S * h->nz * h->nx;
    for (int i = 0; i < 4; i++)
        HCHECK(h, hipMemcpyAsync(h->st.a[i], a_dev + i * f, f * sizeof(double), hipMemcpyDeviceToDevice,
                                 h->stream));
    return SPARC_OK;
}

int sparc_get_pml(sparc_handle *h, double *pv, double *px)
{
    if (!h) return SPARC_E_ARG;
    const size_t pz = (size_t)2 * h->S * h->d.nzpml * h->nx;
    const size_t pxn = (size_t)2 * h->S * h->nz * h->d.nph;
    if (pv) HCHECK(h, hipMemcpyAsync(pv, h->st.pz_base, pz * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    if (px) HCHECK(h, hipMemcpyAsync(px, h->st.px_base, pxn * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return SPARC_OK;
}

int sparc_ddx(sparc_handle *h, const double *in, double *out, int nrows)
{
    if (!h || !in || !out || nrows < 1) return SPARC_E_ARG;
    HCHECK(h, launch_ddx_rows(in, out, nrows, h->nx, h->bg.tw, h->bg.dmul, h->bg.cdx, h->stretchx, h->stream));
    return SPARC_OK;
}

static int ddz_impl(sparc_handle *h, const double *in, double *out, int nbatch, int ibc, int filt)
{
    if (!h || !in || !out || nbatch < 1 || ibc < 0 || ibc > 3) return SPARC_E_ARG;
    double *work = nullptr;
    HCHECK(h, hipMalloc((void **)&work, (size_t)nbatch * h->nz * h->nx * sizeof(double)));
    HCHECK(h, launch_ddz_cols(in, out, h->nx, h->nz, nbatch, ibc, h->bg.stretch, h->cf, h->tri[ibc], work,
                              filt, h->stream));
    HCHECK(h, hipStreamSynchronize(h->stream));
    HCHECK(h, hipFree(work));
    return SPARC_OK;
}

int sparc_ddz(sparc_handle *h, const double *in, double *out, int nbatch, int ibc)
{
    return ddz_impl(h, in, out, nbatch, ibc, 0);
}

int sparc_ddz_kern(sparc_handle *h, const double *in, double *out, int nbatch, int ibc)
{
    return ddz_impl(h, in, out, nbatch, ibc, 1);
}

int sparc_ddz_noalloc(sparc_handle *h, const double *in, double *out, int nbatch, int ibc, int filt)
{
    if (!h || !in || !out || nbatch < 1 || ibc < 0 || ibc > 3 || (filt != 0 && filt != 1))
        return SPARC_E_ARG;
    const size_t need = (size_t)nbatch * h->nz * h->nx;
    if (need > kScratchDoubles)
        return fail(h, SPARC_E_ARG,
                    "sparc_ddz_noalloc: nbatch*nz*nx = " + std::to_string(need) +
                        " doubles exceeds the library scratch of " + std::to_string(kScratchDoubles));
    if (!h->scratch_dev) return fail(h, SPARC_E_STATE, "sparc_ddz_noalloc: scratch address unresolved");
    HCHECK(h, launch_ddz_cols(in, out, h->nx, h->nz, nbatch, ibc, h->bg.stretch, h->cf, h->tri[ibc],
                              h->scratch_dev, filt, h->stream));
    return SPARC_OK;
}

int sparc_kernels_begin(sparc_handle *h)
{
    if (!h) return SPARC_E_ARG;
    const size_t n = (size_t)h->S * h->nx * h->nz;
    if (!h->kbuf) {
        // D[8] + KC KX KZ H + t1 t2 + ddz work
        HCHECK(h, h->kern_mem.alloc(&h->kbuf, 15 * n));
    }
    HCHECK(h, launch_fill(h->kbuf + 8 * n, 4 * n, 0.0, h->stream));
    h->kern_pairs = 0;
    return SPARC_OK;
}

int sparc_kernels_add(sparc_handle *h, int n_pairs, const double *fwd, const double *adj)
{
    if (!h || n_pairs < 1 || !fwd || !adj) return SPARC_E_ARG;
    if (!h->kbuf) return fail(h, SPARC_E_STATE, "sparc_kernels_add before sparc_kernels_begin");
    const int S = h->S, nt = n_pairs;
    const size_t n2 = (size_t)h->nx * h->nz, n = (size_t)S * n2;
    double *KC = h->kbuf + 8 * n, *KX = KC + n, *KZ = KC + 2 * n, *H = KC + 3 * n;
    const Tri &tr1 = h->tri[1];
    // pairs per chunk: all of them, within 2 GiB of work buffers (20 fields
    // per pair); one gather, one ddx, one ddz and one accumulation launch
    // per chunk instead of 17 launches per pair
    const size_t per_pair = 20 * n * sizeof(double);
    const int kmax = (int)std::max<size_t>(1, std::min<size_t>(((size_t)2 << 30) / per_pair,
                                                               65535 / (4 * (size_t)S)));   // ddz grid.y
    const int want = std::min(nt, kmax);
    if (h->kpair_cap < want) {
        if (h->stream) HCHECK(h, hipStreamSynchronize(h->stream));
        h->kpair_mem.release();
        h->kpair = nullptr;
        h->kpair_cap = 0;
        HCHECK(h, h->kpair_mem.alloc(&h->kpair, (size_t)20 * want * n));
        h->kpair_cap = want;
    }
    for (int k0 = 0; k0 < nt; k0 += want) {
        const int Kc = std::min(want, nt - k0);
        const size_t per = (size_t)Kc * n;
        double *G = h->kpair, *Dx = G + 8 * per, *Dz = Dx + 4 * per, *work = Dz + 4 * per;
        HCHECK(h, launch_gather_pairs(fwd, adj, S, nt, k0, Kc, n2, G, h->stream));
        HCHECK(h, launch_ddx_rows(G, Dx, 4 * Kc * S * h->nz, h->nx, h->bg.tw, h->bg.dmul, h->bg.cdx, h->stretchx,
                                  h->stream));
        HCHECK(h, launch_ddz_cols(G + 4 * per, Dz, h->nx, h->nz, 4 * Kc * S, 1, h->bg.stretch, h->cf, tr1,
                                  work, 1, h->stream));
        HCHECK(h, launch_kern_accum_pairs(fwd, adj, S, nt, k0, Kc, n2, Dx, Dz, KC, KX, KZ, H, h->stream));
    }
    h->kern_pairs += nt;
    return SPARC_OK;
}

int sparc_kernels_end(sparc_handle *h, double *kc, double *kpsi, double *kvx, double *kvz, double *hess)
{
    if (!h || !kc || !kpsi || !kvx || !kvz || !hess) return SPARC_E_ARG;
    if (!h->kbuf) return fail(h, SPARC_E_STATE, "sparc_kernels_end before sparc_kernels_begin");
    const int S = h->S;
    const size_t n2 = (size_t)h->nx * h->nz, n = (size_t)S * n2;
    double *KC = h->kbuf + 8 * n, *KX = KC + n, *KZ = KC + 2 * n, *H = KC + 3 * n;
    double *t1 = h->kbuf + 12 * n, *t2 = h->kbuf + 13 * n, *work = h->kbuf + 14 * n;
    // curl (bcz = 0, kernels.f90:832) and unit scalings (:498-579)
    HCHECK(h, launch_ddz_cols(KX, t1, h->nx, h->nz, S, 0, h->bg.stretch, h->cf, h->tri[0], work, 1,
                              h->stream));
    HCHECK(h, launch_ddx_rows(KZ, t2, S * h->nz, h->nx, h->bg.tw, h->bg.dmul, h->bg.cdx, h->stretchx, h->stream));
    const double diml = kRsun, dimc = (double)1e6f, e8 = (double)1e-8f, e5 = (double)1e-5f;
    const double dimen = std::pow(diml * e8 * h->xlength * e8 * dimc * e5, -1.0) * 1.0;
    const double dimen_c = std::pow(diml * e8 * h->xlength * e8, -1.0) * 1.0;
    const int stepskern = (int)std::floor(h->outputcad / h->timestep);
    HCHECK(h, launch_kern_final(S, n2, h->d_krho0, h->d_kc2, (double)stepskern, h->timestep, dimen,
                                dimen_c, t1, t2, KC, KX, KZ, kc, kpsi, kvx, kvz, h->stream));
    HCHECK(h, hipMemcpyAsync(hess, H, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc_kernels(sparc_handle *h, int nt, const double *fwd, const double *adj, double *kc,
                  double *kpsi, double *kvx, double *kvz, double *hess)
{
    if (!h || nt < 1 || !fwd || !adj || !kc || !kpsi || !kvx || !kvz || !hess) return SPARC_E_ARG;
    int r = sparc_kernels_begin(h);
    if (!r) r = sparc_kernels_add(h, nt, fwd, adj);
    if (!r) r = sparc_kernels_end(h, kc, kpsi, kvx, kvz, hess);
    return r;
}

// ---------------------------------------------------------------------------
// Exact discrete adjoint (sparc_discrete.hip).

// The dense z-derivative matrices of ibc 1, 2, 3 (stretch included; kept rows
// at 0), their transposes and ibc 3's response to its kept last row, built by
// the library's compact-FD solve of an identity block.
static int discrete_matrices(sparc_handle *h, DevArena &ar, DiscreteCtx &c)
{
    const int nz = h->nz;
    const size_t nn = (size_t)nz * nz;
    std::vector<double> eye(nn, 0.0), m;
    for (int i = 0; i < nz; i++) eye[(size_t)i * nz + i] = 1.0;
    for (int ibc = 1; ibc <= 3; ibc++) {
        int r = host_ddz(h, eye, m, nz, 1, ibc, true, ibc);
        if (r) return r;
        std::vector<double> mt(nn);
        for (int i = 0; i < nz; i++)
            for (int j = 0; j < nz; j++) mt[(size_t)j * nz + i] = m[(size_t)i * nz + j];
        double *p = nullptr;
        HCHECK(h, upload(ar, &p, m, h->stream));
        c.M[ibc - 1] = p;
        HCHECK(h, upload(ar, &p, mt, h->stream));
        c.MT[ibc - 1] = p;
    }
    // ibc 3 with a zero input and the kept last row 1
    double *di = nullptr, *dout = nullptr, *dw = nullptr;
    HCHECK(h, ar.alloc(&di, nz));
    HCHECK(h, ar.alloc(&dout, nz));
    HCHECK(h, ar.alloc(&dw, nz));
    HCHECK(h, launch_fill(di, nz, 0.0, h->stream));
    HCHECK(h, launch_fill(dout, nz - 1, 0.0, h->stream));
    HCHECK(h, launch_fill(dout + nz - 1, 1, 1.0, h->stream));
    HCHECK(h, launch_ddz_cols(di, dout, 1, nz, 1, 3, h->bg.stretch, h->cf, h->tri[3], dw, 0, h->stream));
    c.vN3 = dout;
    return SPARC_OK;
}

int sparc_discrete_gradient(sparc_handle *h, int maxtime, const double *gvz_dev, int nt, int interval,
                            double *gpsi_dev, double *gc_dev, double *gcoef_dev)
{
    if (!h || !gvz_dev || maxtime < 0 || interval < 0) return SPARC_E_ARG;
    if (h->adjoint) return fail(h, SPARC_E_STATE, "the discrete adjoint differentiates a forward handle");
    if (!h->have_src || h->maxtime != maxtime || (h->src.mode != 1 && h->src.mode != 3))
        return fail(h, SPARC_E_STATE, "forward sources for this maxtime not set");
    if (nt != maxtime / h->steps + 1) return fail(h, SPARC_E_ARG, "gvz must hold maxtime/steps + 1 outputs");
    if (gpsi_dev && !h->flowx.empty())
        return fail(h, SPARC_E_STATE, "d/dpsi needs the flow from psi, not an explicit flow");
    const int C = interval ? std::min(interval, maxtime + 1) : maxtime + 1;
    DevArena ar;
    DiscreteCtx c{};
    c.d = h->d;
    c.bg = h->bg;
    c.src = h->src;
    c.stream = h->stream;
    for (int s = 0; s < 5; s++) c.betas[s] = h->betas[s];
    c.deltat = h->deltat;
    c.unstretch_last = h->unstretch[h->nz - 1];
    c.steps = h->steps;
    c.orow = h->o_rad;
    auto done = [&](int r) { (void)hipStreamSynchronize(h->stream); ar.release(); return r; };
    int r = discrete_matrices(h, ar, c);
    if (r) return done(r);
    DiscreteWork w{};
    const size_t S = h->S;
    w.SN = S * h->nz * h->nx;
    w.PZ1 = S * h->d.nzpml * h->nx;
    w.PX1 = S * h->nz * h->d.nph;
    w.NST = 4 * w.SN + 2 * w.PZ1 + 2 * w.PX1;
    const int nseg = (maxtime + C) / C;
    struct B { double **p; size_t n; } bufs[] = {
        {&w.U, w.NST}, {&w.Unext, w.NST}, {&w.Y, w.NST}, {&w.R, w.NST}, {&w.acc, w.NST}, {&w.Rb, w.NST},
        {&w.Yb, w.NST}, {&w.lam, w.NST}, {&w.ck, (size_t)nseg * w.NST}, {&w.seg, (size_t)C * w.NST},
        {&w.Dst, 25 * w.SN}, {&w.Dscr, 5 * w.SN}, {&w.T0, w.SN}, {&w.T1, w.SN}, {&w.T2, w.SN}, {&w.T3, w.SN},
        {&w.GZ, w.SN}, {&w.DGZ, w.SN}, {&w.RHOB, w.SN}, {&w.XZP, w.SN}, {&w.DXB, w.SN}, {&w.DZB, w.SN},
        {&w.DZ3B, w.SN}, {&w.DDXB, w.SN}, {&w.FX, 3 * w.SN}, {&w.DFX, 3 * w.SN}, {&w.FZ, 2 * w.SN},
        {&w.DFZ, 2 * w.SN}, {&w.PB, S * h->nx}};
    for (const B &b : bufs) {
        hipError_t e = ar.alloc(b.p, b.n);
        if (e != hipSuccess) { h->err = std::string("discrete adjoint work: ") + hipGetErrorString(e); return done(SPARC_E_HIP); }
    }
    {
        // the stage quantities of one segment, kept by its recompute so that
        // the transposed steps need not run their forward step again, when
        // they take at most half of the free memory (SPARC_DISCRETE_KEEP_STAGES=0
        // recomputes them in every transposed step instead)
        const size_t need = (size_t)C * 25 * w.SN * sizeof(double);
        const char *ke = std::getenv("SPARC_DISCRETE_KEEP_STAGES");
        const bool keep = !(ke && std::atoi(ke) == 0);
        size_t fr = 0, tot = 0;
        if (keep && hipMemGetInfo(&fr, &tot) == hipSuccess && need <= fr / 2 &&
            ar.alloc(&w.Dseg, (size_t)C * 25 * w.SN) != hipSuccess) {
            (void)hipGetLastError();
            w.Dseg = nullptr;
        }
    }
    double *G = gcoef_dev;
    if (!G) {
        hipError_t e = ar.alloc(&G, 3 * w.SN);
        if (e != hipSuccess) { h->err = hipGetErrorString(e); return done(SPARC_E_HIP); }
    }
    hipError_t e = discrete_gradient(c, w, maxtime, C, gvz_dev, nt, G);
    if (e == hipSuccess && (gpsi_dev || gc_dev)) {
        const size_t n2 = (size_t)h->nx * h->nz;
        std::vector<double> psi(h->psi.empty() ? std::vector<double>() : h->psi);
        double *dc2 = nullptr, *dpsi = nullptr;
        std::vector<double> c2(h->c2.begin(), h->c2.begin() + n2);
        int rr = upload(ar, &dc2, c2, h->stream) == hipSuccess ? SPARC_OK : SPARC_E_HIP;
        if (!rr && !psi.empty()) rr = upload(ar, &dpsi, psi, h->stream) == hipSuccess ? SPARC_OK : SPARC_E_HIP;
        if (rr) return done(rr);
        const double dimc = 1e6, scale = kRsun / 1.0e8;
        e = discrete_flow_chain(c, w, G, h->bg.rho0, dc2, dpsi, scale, dimc, gpsi_dev, gc_dev);
    }
    if (e != hipSuccess) { h->err = std::string("discrete adjoint: ") + hipGetErrorString(e); return done(SPARC_E_HIP); }
    return done(SPARC_OK);
}

int sparc_ddz_T(sparc_handle *h, const double *in_dev, double *out_dev, int nbatch, int ibc, double *presetN_bar_dev)
{
    if (!h || !in_dev || !out_dev || nbatch < 1 || ibc < 1 || ibc > 3) return SPARC_E_ARG;
    DevArena ar;
    DiscreteCtx c{};
    c.d = h->d;
    c.stream = h->stream;
    int r = discrete_matrices(h, ar, c);
    hipError_t e = r ? hipSuccess : discrete_apply(c, in_dev, out_dev, nbatch, ibc + 2, presetN_bar_dev);
    (void)hipStreamSynchronize(h->stream);
    ar.release();
    if (r) return r;
    HCHECK(h, e);
    return SPARC_OK;
}

int sparc_profile_kernels(sparc_handle *h, int reps, double *ms_out)
{
    if (!h || reps < 1 || !ms_out) return SPARC_E_ARG;
    if (!h->have_src) return fail(h, SPARC_E_STATE, "sources not set");
    int r = sparc_reset(h);
    if (r) return r;
    const size_t nev = (size_t)2 * 11 * (size_t)reps;
    for (size_t i = 0; i < nev; i++) {
        hipEvent_t e;
        HCHECK(h, hipEventCreate(&e));
        h->evpool.push_back(e);
    }
    h->marks.clear();
    h->profiling = true;
    Sink sk = empty_sink(h);
    for (int t = 0; t < reps && r == SPARC_OK; t++) {
        r = launch_set_int(h->tctr, t, h->stream) == hipSuccess ? SPARC_OK : SPARC_E_HIP;
        if (!r) r = enqueue_step(h, 0, t > 0, sk);
    }
    h->profiling = false;
    if (!r) r = hipStreamSynchronize(h->stream) == hipSuccess ? SPARC_OK : SPARC_E_HIP;
    double sum[5] = {0, 0, 0, 0, 0};
    int cnt[5] = {0, 0, 0, 0, 0};
    for (const auto &m : h->marks) {
        float ms = 0.f;
        if (!r && hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
            sum[m.kind] += ms;
            cnt[m.kind]++;
        }
    }
    for (int k = 0; k < 5; k++) ms_out[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
    h->marks.clear();
    for (hipEvent_t e : h->evpool) (void)hipEventDestroy(e);
    h->evpool.clear();
    if (r) return r;
    return sparc_reset(h);
}

}  // extern "C"
