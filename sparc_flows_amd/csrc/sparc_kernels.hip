// sparc_kernels.hip -- gfx950 kernels of the 2D option-17 solver.
//
// One RK stage (step.f90:48-106 with MP_QUIET_PML_2D_DISPL,
// physics2d.f90:943-1094) runs as a column pass and a row pass:
//
//   row pass R  (one FFT team per (source, z) row): dx p of stage s, the x-PML
//               of grad_x p, RHS v_x and the v_x stage update of stage s, then
//               the x-derivatives of xi_x, v_x, v_z of stage s+1 and the x-PML
//               of d xi_x / dx.
//   column pass C (sparc_colseg.hip): the compact-FD z derivatives of xi_z,
//               p, v_x, v_z, the z-PML, divergence, pressure, RHS v_z with the
//               source, the flow terms and the stage update of xi_x, xi_z, v_z;
//               after the last stage also the z filter FILTER_VARS_2D
//               (physics2d.f90:509-556).
//
// A step is R1 C1 R(1|2) C2 R(2|3) C3 R(3|4) C4 R(4|5) C5 R5; the v_x filter of
// the previous step is applied on the fly by the next R1 (or by k_finalize_vx).
#include "sparc_device.h"
#include "sparc_launch.h"

namespace sparc {

// ---------------------------------------------------------------------------
// small helpers

__device__ __forceinline__ size_t fidx(const Dims &d, int b, int k, int x)
{
    return ((size_t)b * d.nz + k) * d.nx + x;
}

// ---------------------------------------------------------------------------
// Generic z-derivative (ddz, derivatives.f90:10-24) of ncols columns of an
// [nz][ncols] array (x fastest); one lane per column.  stretch may be null
// (plain dbyd2).  For ibc 2 / 3 the preset boundary row is read from out.
// If filt is set the result is also passed through FILTER_VAR_Z
// (kernels.f90:891-928), as ddzkern does (in place, column by column).
__global__ __launch_bounds__(64) void k_ddz_cols(const double *in, double *out,
                                                 int ncols, int nz, int ibc,
                                                 const double *__restrict__ stretch, DbyCoef cf,
                                                 Tri tr, int nbatch, double *work, int filt)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int bb = blockIdx.y;
    if (x >= ncols || bb >= nbatch) return;
    const size_t s = (size_t)ncols;
    const size_t q0 = (size_t)bb * nz * ncols + x;
    const double *B = in;
    double *W = work;
    const double preset0 = out[q0];
    const double presetN = out[q0 + (size_t)(nz - 1) * s];
    FwdSweep fw;
    fw.start(B, q0, s, ibc, preset0, cf, tr, W);
    for (int j = 2; j <= nz - 3; j++) {
        const double Bn = (j + 3 < nz) ? B[q0 + (size_t)(j + 3) * s] : 0.0;
        fw.step(j, q0 + (size_t)j * s, Bn, cf, tr, W);
    }
    fw.finish(B, q0, s, nz, ibc, presetN, cf, tr, W);
    double d = fw.ud;
    out[q0 + (size_t)(nz - 1) * s] = stretch ? d * stretch[nz - 1] : d;
    for (int j = nz - 2; j >= 0; j--) {
        d = W[q0 + (size_t)j * s] - tr.gam[j + 1] * d;
        out[q0 + (size_t)j * s] = stretch ? d * stretch[j] : d;
    }
    if (filt) {
        // FILTER_VAR_Z: rows 2..nz-1 (1-based) from the unfiltered values; W
        // holds a copy so the filter reads only unfiltered neighbours.
        for (int j = 0; j < nz; j++) W[q0 + (size_t)j * s] = out[q0 + (size_t)j * s];
        for (int j = 1; j <= nz - 2; j++) {
            const size_t q = q0 + (size_t)j * s;
            double v;
            if (j <= 4 || j >= nz - 5) v = 0.5 * (W[q + s] + W[q - s]);
            else v = filt11(W, q, s);
            out[q] = v;
        }
    }
}

// The team derivative for the factorization N1 x N2, dispatched in place
// (a wrapper function around the two forms costs the 16 x 16 row kernels
// some 30 registers).
// The enclosing kernel defines kAny, N1, N2, nxr and sdm (the circulant
// column of the any-nx form, staged in LDS).
#define SPARC_FFT_DERIV(RE, IM, T, LDS, TW, DM, SX)                                            \
    do {                                                                                       \
        if constexpr (kAny)                                                                    \
            circ_deriv_row<N2>(RE, IM, T, LDS, sdm, nxr, SX);                                  \
        else if constexpr (N1 == 32 && N2 == 8)                                                \
            fft_deriv_team_256x32(RE, IM, T, LDS, (LDS) + 8 * 33, TW, DM, SX);                 \
        else if constexpr (N1 == 64 && N2 == 4)                                                \
            fft_deriv_team_256x64(RE, IM, T, LDS, (LDS) + 264, TW, DM, SX);                    \
        else                                                                                   \
            fft_deriv_team<N1, N2>(RE, IM, T, LDS, TW, DM, SX);                                \
    } while (0)

// Generic x-derivative (ddx) of nrows real rows of length nx; two rows per
// FFT team.  NX = nx for a power of two; NX = -N2 for the any-nx form (64
// lanes, rows up to 64 N2, circ_deriv_row).
template <int NX, int N1 = (NX < 0 ? 64 : FftShape<(NX < 0 ? 16 : NX)>::N1)>
__global__ __launch_bounds__(64) void k_ddx_rows(const double *__restrict__ in,
                                                 double *__restrict__ out, int nrows,
                                                 const double2 *__restrict__ tw,
                                                 const double *__restrict__ dmul, double stretchx,
                                                 const double *__restrict__ cdx, int nx)
{
    constexpr bool kAny = NX < 0;
    constexpr int N2 = kAny ? -NX : NX / N1, TEAMS = 64 / N1;
    const int nxr = kAny ? nx : NX;
    __shared__ double2 lds[TEAMS][FftLds<N1, N2>::n];
    __shared__ double sdm[kAny ? 128 * N2 : 1];
    if constexpr (kAny) {
        // visible after circ_deriv_row's first barrier
        for (int q = threadIdx.x; q < 128 * N2; q += 64) sdm[q] = q < 2 * nxr ? cdx[q] : 0.0;
    }
    const int team = threadIdx.x / N1, t = threadIdx.x % N1;
    const int r0 = 2 * (blockIdx.x * TEAMS + team), r1 = r0 + 1;
    double re[N2], im[N2];
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
        const int x = t + N1 * j2;
        const bool in_row = !kAny || x < nxr;
        re[j2] = r0 < nrows && in_row ? in[(size_t)r0 * nxr + x] : 0.0;
        im[j2] = r1 < nrows && in_row ? in[(size_t)r1 * nxr + x] : 0.0;
    }
    SPARC_FFT_DERIV(re, im, t, lds[team], tw, dmul, stretchx);
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
        const int x = t + N1 * j2;
        if (kAny && x >= nxr) continue;
        if (r0 < nrows) out[(size_t)r0 * nxr + x] = re[j2];
        if (r1 < nrows) out[(size_t)r1 * nxr + x] = im[j2];
    }
}

// ---------------------------------------------------------------------------
// Row pass.  MODE 0: first pass of a step (x-derivatives of stage 1, and the
// z filter of the previous step's v_x when filt_vx is set).  MODE 1: stage s
// dx p / v_x update + stage s+1 x-derivatives.  MODE 2: stage 5 dx p / v_x.


// N1 lanes per row: FftShape's by default; nx = 256 also runs on 32 lanes
// (two rows per wave, three-stage transform) or 64 lanes (one row per wave,
// four radix-4 stages), which are faster when the pass is latency-bound (one
// source) and slower when it is bandwidth-bound.
//
// NX = -N2: the any-nx form, one row per block of 64 lanes, rows up to
// 64 N2 (circ_deriv_row); lane slots past nx load column nx-1 and store
// nothing.
template <int NX, int MODE, int N1 = (NX < 0 ? 64 : FftShape<(NX < 0 ? 16 : NX)>::N1)>
__global__ __launch_bounds__(64) void k_row(Dims d, Bg bg, State st, RowArgs ra, Clock clk, Sink sink)
{
    constexpr bool kAny = NX < 0;
    static_assert(!kAny || N1 == 64, "the any-nx form runs one row on 64 lanes");
    constexpr int N2 = kAny ? -NX : NX / N1, TEAMS = 64 / N1;
    const int nxr = kAny ? d.nx : NX;
    __shared__ double2 lds[TEAMS][FftLds<N1, N2>::n];
    // twiddles and bin multipliers (any nx: the circulant column) staged in
    // LDS: their reads then never wait behind this pass's global stores
    __shared__ double2 stw[kAny ? 1 : NX];
    __shared__ double sdm[kAny ? 128 * N2 : NX];
    if constexpr (kAny) {
        for (int q = threadIdx.x; q < 128 * N2; q += 64) sdm[q] = q < 2 * nxr ? bg.cdx[q] : 0.0;
    } else {
        constexpr int NQ = (NX + 63) / 64;
        double2 tv[NQ];
        double dv[NQ];
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int q = min((int)threadIdx.x + 64 * u, NX - 1);
            tv[u] = bg.tw[q];
            dv[u] = bg.dmul[q];
        }
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int q = threadIdx.x + 64 * u;
            if (NX % 64 == 0 || q < NX) {      // a guard only where it can fail, which keeps the loads above it
                stw[q] = tv[u];
                sdm[q] = dv[u];
            }
        }
    }
    const int team = threadIdx.x / N1, t = threadIdx.x % N1;
    const int r = blockIdx.x * TEAMS + team;
    const bool valid = r < d.S * d.nz;
    const int rc = valid ? r : d.S * d.nz - 1;     // clamped row: loads stay in bounds
    const int b = rc / d.nz, k = rc % d.nz;
    const int nz = d.nz, nh = d.nh, nph = d.nph;
    const int time = clock_t_(clk);
    const size_t rowq = ((size_t)b * nz + k) * nxr;
    const size_t pxq0 = ((size_t)b * nz + k) * nph;             // psivx row
    const size_t pxq1 = ((size_t)(d.S + b) * nz + k) * nph;     // psi_gradp_x row
    double *const dump = st.discard + threadIdx.x;              // target of invalid-row stores
    // column of lane slot j2 (any nx: slots past the row load column nx-1)
    // and whether the slot's stores land (a real row and, any nx, a real column)
    auto colx = [&](int j2) { const int x = t + N1 * j2; return kAny ? min(x, nxr - 1) : x; };
    auto ok = [&](int j2) { return valid && (!kAny || t + N1 * j2 < nxr); };
    double re[N2], im[N2], vx[N2];
    // x-PML auxiliary data (psi stage value and base, a_x, b_x) of the lane's
    // columns in the first (j2 = 0) and last (j2 = N2-1) column block, the
    // only ones that hold PML columns when the layer is at most N1 wide:
    // loaded here, unconditionally at clamped indices, so that their latency
    // overlaps the row loads instead of following the FFTs (set 0:
    // psi_gradp_x of the v_x update, set 1: psivx of d xi_x/dx).  Only for
    // teams of at least 32 lanes (N2 <= 8): the 16-lane form holds 16
    // columns per lane, and the extra registers cost it more than the waits.
    constexpr bool kPre = !kAny && N2 <= 8;
    auto pml_index = [&](int x) { return x < nh ? x : min(max(x - nxr + 2 * nh, 0), nph - 1); };
    double pre[2][2][4] = {};
#pragma unroll
    for (int e = 0; e < (kPre ? 2 : 0); e++) {
        const int pi = pml_index(t + (e ? N1 * (N2 - 1) : 0));
#pragma unroll
        for (int set = 0; set < 2; set++) {
            const size_t q = (set == 0 ? pxq1 : pxq0) + pi;
            pre[set][e][0] = (MODE == 0 && set == 0) || (MODE == 2 && set == 1) ? 0.0 : st.px_tmp[q];
            pre[set][e][1] = (MODE == 0 && set == 0) || (MODE == 2 && set == 1) ? 0.0 : st.px_base[q];
            pre[set][e][2] = bg.ax[(size_t)k * nph + pi];
            pre[set][e][3] = bg.bx[(size_t)k * nph + pi];
        }
    }
    // x-PML data of column block j2: prefetched for the end blocks, loaded
    // here for an inner block (a layer wider than N1)
    auto pml_data = [&](int set, int j2, int pi, double &psi, double &pb, double &axv, double &bxv) {
        if (kPre && N1 >= nh && (j2 == 0 || j2 == N2 - 1)) {
            const int e = j2 == 0 ? 0 : 1;
            psi = pre[set][e][0];
            pb = pre[set][e][1];
            axv = pre[set][e][2];
            bxv = pre[set][e][3];
        } else {
            const size_t q = (set == 0 ? pxq1 : pxq0) + pi;
            psi = st.px_tmp[q];
            pb = st.px_base[q];
            axv = bg.ax[(size_t)k * nph + pi];
            bxv = bg.bx[(size_t)k * nph + pi];
        }
    };
    const double unst = bg.unstretch[nz - 1];
    // takes the prefetched values at one point (before the first FFT, after
    // the row loads are issued), so that the compiler cannot sink their loads
    // into the PML branches after the FFTs
    auto pin_pre = [&]() {
        if (!kPre) return;
        asm volatile("" : "+v"(pre[0][0][0]), "+v"(pre[0][0][1]), "+v"(pre[0][0][2]), "+v"(pre[0][0][3]),
                     "+v"(pre[0][1][0]), "+v"(pre[0][1][1]), "+v"(pre[0][1][2]), "+v"(pre[0][1][3]),
                     "+v"(pre[1][0][0]), "+v"(pre[1][0][1]), "+v"(pre[1][0][2]), "+v"(pre[1][0][3]),
                     "+v"(pre[1][1][0]), "+v"(pre[1][1][1]), "+v"(pre[1][1][2]), "+v"(pre[1][1][3]));
    };

    if (MODE == 0) {
        // v_x of stage 1 (= a_vx after the pending filter), xi_x, v_z
        if (kPre) {
            // the choice is uniform: one loop per form, every load of a loop
            // issued before the first is used
            if (ra.filt_vx) {
#pragma unroll
                for (int j2 = 0; j2 < N2; j2++) {
                    const int x = colx(j2);
                    vx[j2] = filter_vars_2d_sel(st.vx_unf, rowq + x, k, x, nxr, nz, nh);
                }
            } else {
#pragma unroll
                for (int j2 = 0; j2 < N2; j2++) vx[j2] = st.a[2][rowq + colx(j2)];
            }
#pragma unroll
            for (int j2 = 0; j2 < N2; j2++) {
                re[j2] = ra.xi_x[rowq + colx(j2)];
                im[j2] = vx[j2];
            }
        } else {
#pragma unroll
            for (int j2 = 0; j2 < N2; j2++) {
                const int x = colx(j2);
                const double v = ra.filt_vx ? filter_vars_2d_sel(st.vx_unf, rowq + x, k, x, nxr, nz, nh)
                                            : st.a[2][rowq + x];
                vx[j2] = v;
                re[j2] = ra.xi_x[rowq + x];
                im[j2] = v;
            }
        }
        if (ra.filt_vx) {
            const int tp = time - 1;
            const int sslot = sink.snap ? out_slot(tp, clk.steps, sink.k0, sink.nsnap) : -1;
            const bool snap = sslot >= 0;
            const int fslot = sink.adjoint ? 0 : 2;
#pragma unroll
            for (int j2 = 0; j2 < N2; j2++) {
                const int x = colx(j2);
                *(ok(j2) ? &st.a[2][rowq + x] : dump) = vx[j2];
                if (snap && ok(j2))
                    sink.snap[((((size_t)b * sink.nsnap + sslot) * sink.nfield + fslot) * nz + k) * nxr + x] = vx[j2];
            }
        }
        pin_pre();
        __syncthreads();   // staged twiddles visible
        SPARC_FFT_DERIV(re, im, t, lds[team], stw, sdm, bg.stretchx);
        // re = dxi_x/dx, im = dv_x/dx
        double zr[N2], zi[N2];
#pragma unroll
        for (int j2 = 0; j2 < N2; j2++) {
            zr[j2] = ra.v_z[rowq + colx(j2)];
            zi[j2] = 0.0;
        }
        SPARC_FFT_DERIV(zr, zi, t, lds[team], stw, sdm, bg.stretchx);
#pragma unroll
        for (int j2 = 0; j2 < N2; j2++) {
            const int x = colx(j2);
            *(ok(j2) ? &st.dxvx[rowq + x] : dump) = im[j2];
            *(ok(j2) ? &st.dxvz[rowq + x] : dump) = zr[j2];
        }
    } else {
        // dx p (stage s) paired with dx xi_x (stage s+1, or nothing in MODE 2)
#pragma unroll
        for (int j2 = 0; j2 < N2; j2++) {
            const int x = colx(j2);
            re[j2] = st.p[rowq + x];
            im[j2] = MODE == 1 ? ra.xi_x[rowq + x] : 0.0;
        }
        double rvx0[N2], avx[N2], rinv[N2];
#pragma unroll
        for (int j2 = 0; j2 < N2; j2++) {
            const int x = colx(j2);
            rvx0[j2] = st.rvx[rowq + x];
            avx[j2] = st.a[2][rowq + x];
            rinv[j2] = bg.rhoinv_z[k];
        }
        pin_pre();
        __syncthreads();   // staged twiddles visible
        SPARC_FFT_DERIV(re, im, t, lds[team], stw, sdm, bg.stretchx);
        // x-PML on grad_x p (physics2d.f90:1003-1020), RHS v_x (:1039 + flows)
        const int sslot = (MODE == 2 && sink.snap) ? out_slot(time, clk.steps, sink.k0, sink.nsnap) : -1;
        const bool snap = sslot >= 0;
#pragma unroll
        for (int j2 = 0; j2 < N2; j2++) {
            const int x = colx(j2);
            double gpx = re[j2];
            if (N1 * j2 < nh || N1 * (j2 + 1) > nxr - nh) {     // team-uniform: this j2 can hold PML columns
                const bool pml = x < nh || x >= nxr - nh;
                const int pi = pml_index(x);
                double psi, pb, axv, bxv;
                pml_data(0, j2, pi, psi, pb, axv, bxv);
                const double g2 = gpx + psi;
                const double rhs = axv * g2 + bxv * psi;
                const double nb = pb + ra.bx2 * rhs;
                if (pml && ok(j2)) {
                    st.px_tmp[pxq1 + pi] = nb;
                    if (ra.last2) st.px_base[pxq1 + pi] = nb;
                }
                gpx = pml ? g2 : gpx;
            }
            double rvx = -rinv[j2] * gpx + rvx0[j2];
            rvx = (k == 0 || x == 0 || x == nxr - 1) ? 0.0 : rvx;
            const double v = avx[j2] + ra.bx2 * rvx;
            vx[j2] = v;
            *(ok(j2) ? &ra.v_x_out[rowq + x] : dump) = v;
            if (snap && ok(j2))
                sink.snap[((((size_t)b * sink.nsnap + sslot) * sink.nfield +
                            (sink.adjoint ? 2 : 4)) * nz + k) * nxr + x] = rvx;
        }
        if (MODE == 1) {
            // im holds dxi_x/dx of stage s+1; now dv_x/dx, dv_z/dx of stage s+1
            double wr[N2], wi[N2];
#pragma unroll
            for (int j2 = 0; j2 < N2; j2++) {
                wr[j2] = vx[j2];
                wi[j2] = ra.v_z[rowq + colx(j2)];
            }
            SPARC_FFT_DERIV(wr, wi, t, lds[team], stw, sdm, bg.stretchx);
#pragma unroll
            for (int j2 = 0; j2 < N2; j2++) {
                const int x = colx(j2);
                *(ok(j2) ? &st.dxvx[rowq + x] : dump) = wr[j2];
                *(ok(j2) ? &st.dxvz[rowq + x] : dump) = wi[j2];
            }
#pragma unroll
            for (int j2 = 0; j2 < N2; j2++) re[j2] = im[j2];
        }
    }
    if (MODE == 2) return;
    // re = dxi_x/dx of the next stage: top boundary value for d xi_z/dz
    // (physics2d.f90:954, before the x-PML), then the x-PML (:970-986)
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
        const int x = colx(j2);
        double dx = re[j2];
        if (k == nz - 1 && ok(j2)) st.topbc[(size_t)b * nxr + x] = -dx * unst;
        if (N1 * j2 < nh || N1 * (j2 + 1) > nxr - nh) {
            const bool pml = x < nh || x >= nxr - nh;
            const int pi = pml_index(x);
            double psi, pb, axv, bxv;
            pml_data(1, j2, pi, psi, pb, axv, bxv);
            const double d2 = dx + psi;
            const double rhs = axv * d2 + bxv * psi;
            const double nb = pb + ra.bx1 * rhs;
            if (pml && ok(j2)) {
                st.px_tmp[pxq0 + pi] = nb;
                if (ra.last1) st.px_base[pxq0 + pi] = nb;
            }
            dx = pml ? d2 : dx;
        }
        *(ok(j2) ? &st.dxixdx[rowq + x] : dump) = dx;
    }
}

// v_x filter of the last step of a run (the next row pass would do it).
__global__ __launch_bounds__(64) void k_finalize_vx(Dims d, State st, Clock clk, Sink sink)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int b = blockIdx.y;
    if (x >= d.nx) return;
    const int nx = d.nx, nz = d.nz;
    const int tp = clock_t_(clk) - 1;
    const int sslot = sink.snap ? out_slot(tp, clk.steps, sink.k0, sink.nsnap) : -1;
    const bool snap = sslot >= 0;
    for (int j = 0; j < nz; j++) {
        const size_t q = ((size_t)b * nz + j) * nx + x;
        const double v = filter_vars_2d_at(st.vx_unf, q, j, x, nx, nz, d.nh);
        st.a[2][q] = v;
        if (snap)
            sink.snap[((((size_t)b * sink.nsnap + sslot) * sink.nfield +
                        (sink.adjoint ? 0 : 2)) * nz + j) * nx + x] = v;
    }
}

__global__ void k_tick(int *tctr, int inc)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) tctr[0] += inc;
}

__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

__global__ void k_fill(double *p, size_t n, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------------------
// Kernel accumulation (PRODUCE_KERNELS, kernels.f90:200-360).

// Many snapshot pairs at once.  k_gather_pairs lays out the derivative inputs
// of pairs k0 .. k0+Kc-1 as G[8][Kc][S][n2]: the x-derivative group first
// (forward xi_x, adjoint v_x, forward v_x, forward v_z), then the
// z-derivative group (forward xi_z, adjoint v_z, forward v_x, forward v_z),
// so that each group is one ddx / ddz launch.
__global__ void k_gather_pairs(const double *__restrict__ fwd, const double *__restrict__ adj, int S, int nt,
                               int k0, int Kc, size_t n2, double *__restrict__ G)
{
    const size_t per = (size_t)Kc * S * n2, n = 8 * per;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(i / per);
        const size_t r = i - (size_t)g * per;
        const int kk = (int)(r / ((size_t)S * n2));
        const size_t r2 = r - (size_t)kk * S * n2;
        const size_t b = r2 / n2, q = r2 - b * n2;
        const int k = k0 + kk;
        // group member: source, field
        const bool isadj = (g == 1 || g == 5);
        const int field = g == 0 ? 0 : g == 4 ? 1 : (g == 1 || g == 5) ? (g == 1 ? 0 : 1)
                                                : (g == 2 || g == 6) ? 2 : 3;
        const double *src = isadj ? adj + ((b * nt + (nt - 1 - k)) * 4 + field) * n2
                                  : fwd + ((b * nt + k) * 6 + field) * n2;
        G[i] = src[q];
    }
}

// Accumulation over snapshot pairs k0 .. k0+Kc-1, in order, of the hessian
// (kernels.f90:293-294), the sound-speed kernel (:301-306) and the x / z
// components of the flow kernel (:338-355; the y terms vanish in 2D), from
// the derivatives laid out by k_gather_pairs (Dx, Dz: [4][Kc][S][n2]).
// Forward pair k meets adjoint snapshot nt-1-k.  The sums stay in registers
// across the pairs; the order of the additions is that of one pair at a time.
__global__ void k_kern_accum_pairs(const double *__restrict__ fwd, const double *__restrict__ adj, int S,
                                   int nt, int k0, int Kc, size_t n2, const double *__restrict__ Dx,
                                   const double *__restrict__ Dz, double *KC, double *KX, double *KZ,
                                   double *H)
{
    const size_t n = (size_t)S * n2, per = (size_t)Kc * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t b = i / n2, q = i % n2;
        double h = H[i], kc = KC[i], kx = KX[i], kz = KZ[i];
        for (int kk = 0; kk < Kc; kk++) {
            const int k = k0 + kk;
            const double *f = fwd + (b * nt + k) * 6 * n2;
            const double *a = adj + (b * nt + (nt - 1 - k)) * 4 * n2;
            const size_t o = (size_t)kk * n + i;
            const double faccx = f[4 * n2 + q], faccz = f[5 * n2 + q];
            h = h + faccz * faccz + faccx * faccx + 0.0 * 0.0;
            const double divf = Dx[o] + Dz[o];                    // D0 + D1
            const double diva = Dx[per + o] + Dz[per + o];        // D2 + D3
            kc = kc + divf * diva;
            const double avx = a[q], avz = a[n2 + q];
            kx = Dx[2 * per + o] * avx + 0.0 * 0.0 + Dx[3 * per + o] * avz + kx;   // D4, D5
            kz = Dz[2 * per + o] * avx + 0.0 * 0.0 + Dz[3 * per + o] * avz + kz;   // D6, D7
        }
        H[i] = h;
        KC[i] = kc;
        KX[i] = kx;
        KZ[i] = kz;
    }
}

// Unit scalings and curl (kernels.f90:498-579): t1 = ddzkern(Kx, bcz = 0),
// t2 = ddxkern(Kz).
__global__ void k_kern_final(int S, size_t n2, const double *__restrict__ rho0,
                             const double *__restrict__ c2, double steps, double dt, double dimen,
                             double dimen_c, const double *__restrict__ t1,
                             const double *__restrict__ t2, const double *__restrict__ KC,
                             const double *__restrict__ KX, const double *__restrict__ KZ,
                             double *kc, double *kpsi, double *kvx, double *kvz)
{
    const size_t n = (size_t)S * n2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t q = i % n2;
        const double r = rho0[q], cc = c2[q];
        const double kp = t1[i] - t2[i];
        kpsi[i] = 2.0 * kp * dimen * 1.0 * (r * sqrt(cc)) * steps * dt * dimen;
        kvx[i] = -2.0 * KX[i] * r * steps * dt * dimen;
        kvz[i] = -2.0 * KZ[i] * r * steps * dt * dimen;
        kc[i] = -KC[i] * r * cc * steps * dt * dimen_c * 2.0;
    }
}

// ---------------------------------------------------------------------------
// launch wrappers

#define SPARC_DISPATCH_NX(nx, MACRO) \
    switch (nx) {                    \
    case 16: MACRO(16); break;       \
    case 32: MACRO(32); break;       \
    case 64: MACRO(64); break;       \
    case 128: MACRO(128); break;     \
    case 256: MACRO(256); break;     \
    case 512: MACRO(512); break;     \
    case 1024: MACRO(1024); break;   \
    default: return hipErrorInvalidValue; \
    }

hipError_t launch_ddz_cols(const double *in, double *out, int ncols, int nz, int nbatch, int ibc,
                           const double *stretch, const DbyCoef &cf, const Tri &tr, double *work,
                           int filt, hipStream_t st)
{
    dim3 grid((ncols + 63) / 64, nbatch);
    hipLaunchKernelGGL(k_ddz_cols, grid, dim3(64), 0, st, in, out, ncols, nz, ibc, stretch, cf, tr,
                       nbatch, work, filt);
    return hipGetLastError();
}

// The sum of v_z^2 over the grid after the step of clock c (TIMESTEPPING's
// convergence statistic nor1 = sqrt(sum a(:,:,:,6)^2), driver.f90:364, which
// it writes to rms_hist, :384): one block per source, a fixed summation order.
__global__ __launch_bounds__(1024) void k_vz_norm2(const double *__restrict__ vz, size_t n2, Clock c,
                                                   double *__restrict__ out, int nnorm, int t0)
{
    __shared__ double red[16];
    const int b = blockIdx.x;
    const double *v = vz + (size_t)b * n2;
    double s = 0.0;
    for (size_t q = threadIdx.x; q < n2; q += blockDim.x) s += v[q] * v[q];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
        const int k = clock_t_(c) - t0;
        if (k >= 0 && k < nnorm) out[(size_t)b * nnorm + k] = t;
    }
}

hipError_t launch_vz_norm2(const double *vz, int S, size_t n2, const Clock &c, const Sink &sk, hipStream_t st)
{
    hipLaunchKernelGGL(k_vz_norm2, dim3(S), dim3(1024), 0, st, vz, n2, c, sk.norm2, sk.nnorm, sk.tnorm0);
    return hipGetLastError();
}

hipError_t launch_ddx_rows(const double *in, double *out, int nrows, int nx, const double2 *tw,
                           const double *dmul, const double *cdx, double stretchx, hipStream_t st)
{
    const int npairs = (nrows + 1) / 2;
    if (!nx_has_fft(nx)) {
        if (nx < 16 || nx > 1024) return hipErrorInvalidValue;
        // any nx: one row pair per 64-lane block
        dim3 grid(npairs);
        switch (any_nx_slots(nx)) {
#define A(N2V)                                                                                   \
        case N2V:                                                                                \
            hipLaunchKernelGGL((k_ddx_rows<-N2V>), grid, dim3(64), 0, st, in, out, nrows, tw, dmul, \
                               stretchx, cdx, nx);                                              \
            break;
            SPARC_ANY_NX_SLOTS(A)
#undef A
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
#define L(NXV)                                                                              \
    {                                                                                       \
        constexpr int TEAMS = 64 / FftShape<NXV>::N1;                                      \
        dim3 grid((npairs + TEAMS - 1) / TEAMS);                                            \
        hipLaunchKernelGGL(k_ddx_rows<NXV>, grid, dim3(64), 0, st, in, out, nrows, tw, dmul, \
                           stretchx, cdx, nx);                                              \
    }
    SPARC_DISPATCH_NX(nx, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_row(int mode, const Dims &d, const Bg &bg, const State &s, const RowArgs &ra,
                      const Clock &c, const Sink &sk, hipStream_t st, int wide)
{
    if (wide == 2 && d.nx == 256) {
        dim3 grid(d.S * d.nz);
        if (mode == 0) hipLaunchKernelGGL((k_row<256, 0, 64>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk);
        else if (mode == 1) hipLaunchKernelGGL((k_row<256, 1, 64>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk);
        else hipLaunchKernelGGL((k_row<256, 2, 64>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk);
        return hipGetLastError();
    }
    if (wide && d.nx == 256) {
        dim3 grid((d.S * d.nz + 1) / 2);
        if (mode == 0) hipLaunchKernelGGL((k_row<256, 0, 32>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk);
        else if (mode == 1) hipLaunchKernelGGL((k_row<256, 1, 32>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk);
        else hipLaunchKernelGGL((k_row<256, 2, 32>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk);
        return hipGetLastError();
    }
    if (!nx_has_fft(d.nx)) {
        if (d.nx < 16 || d.nx > 1024) return hipErrorInvalidValue;
        // any nx: one row per 64-lane block
        dim3 grid(d.S * d.nz);
#define A(N2V)                                                                                   \
        case N2V:                                                                                \
            if (mode == 0) hipLaunchKernelGGL((k_row<-N2V, 0, 64>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk); \
            else if (mode == 1) hipLaunchKernelGGL((k_row<-N2V, 1, 64>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk); \
            else hipLaunchKernelGGL((k_row<-N2V, 2, 64>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk); \
            break;
        switch (any_nx_slots(d.nx)) {
            SPARC_ANY_NX_SLOTS(A)
        default: return hipErrorInvalidValue;
        }
#undef A
        return hipGetLastError();
    }
#define L(NXV)                                                                                   \
    {                                                                                            \
        constexpr int TEAMS = 64 / FftShape<NXV>::N1;                                           \
        dim3 grid((d.S * d.nz + TEAMS - 1) / TEAMS);                                             \
        if (mode == 0) hipLaunchKernelGGL((k_row<NXV, 0>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk); \
        else if (mode == 1) hipLaunchKernelGGL((k_row<NXV, 1>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk); \
        else hipLaunchKernelGGL((k_row<NXV, 2>), grid, dim3(64), 0, st, d, bg, s, ra, c, sk);    \
    }
    SPARC_DISPATCH_NX(d.nx, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_finalize_vx(const Dims &d, const State &s, const Clock &c, const Sink &sk,
                              hipStream_t st)
{
    dim3 grid((d.nx + 63) / 64, d.S);
    hipLaunchKernelGGL(k_finalize_vx, grid, dim3(64), 0, st, d, s, c, sk);
    return hipGetLastError();
}

hipError_t launch_tick(int *tctr, int inc, hipStream_t st)
{
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, tctr, inc);
    return hipGetLastError();
}

hipError_t launch_set_int(int *p, int v, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, p, v);
    return hipGetLastError();
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

static unsigned grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : (b ? b : 1));
}

hipError_t launch_gather_pairs(const double *fwd, const double *adj, int S, int nt, int k0, int Kc, size_t n2,
                               double *G, hipStream_t st)
{
    hipLaunchKernelGGL(k_gather_pairs, dim3(grid_for((size_t)8 * Kc * S * n2)), dim3(256), 0, st, fwd, adj, S,
                       nt, k0, Kc, n2, G);
    return hipGetLastError();
}

hipError_t launch_kern_accum_pairs(const double *fwd, const double *adj, int S, int nt, int k0, int Kc,
                                   size_t n2, const double *Dx, const double *Dz, double *KC, double *KX,
                                   double *KZ, double *H, hipStream_t st)
{
    hipLaunchKernelGGL(k_kern_accum_pairs, dim3(grid_for((size_t)S * n2)), dim3(256), 0, st, fwd, adj, S, nt,
                       k0, Kc, n2, Dx, Dz, KC, KX, KZ, H);
    return hipGetLastError();
}

hipError_t launch_kern_final(int S, size_t n2, const double *rho0, const double *c2, double steps,
                             double dt, double dimen, double dimen_c, const double *t1, const double *t2,
                             const double *KC, const double *KX, const double *KZ, double *kc,
                             double *kpsi, double *kvx, double *kvz, hipStream_t st)
{
    hipLaunchKernelGGL(k_kern_final, dim3(grid_for((size_t)S * n2)), dim3(256), 0, st, S, n2, rho0,
                       c2, steps, dt, dimen, dimen_c, t1, t2, KC, KX, KZ, kc, kpsi, kvx, kvz);
    return hipGetLastError();
}

}  // namespace sparc
