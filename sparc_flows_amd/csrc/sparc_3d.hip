// sparc_3d.hip -- gfx950 kernels and C ABI (include/sparc_hip3d.h) of the 3D
// solver with flows, option 16 (MP_FLOWS_PML_3D_DISPL, displacement.f90:377-477).
//
// One RK stage (step.f90:48-106) runs as
//   x pass       d/dx of xi_x, v_x, v_y, v_z (FFT rows; x contiguous)
//   y pass       d/dy of xi_y, v_x, v_y, v_z (FFT lines of stride nx)
//   z pass Zp    d xi_z/dz (ibc 1) with the z-PML, div, p and rho' g / rho0
//   x, y passes  d p/dx, d p/dy
//   z pass Zv    d v_x/dz, d v_y/dz, d v_z/dz (ibc 1) and d p/dz (ibc 0, z-PML)
//                as four chains of one lane; RHS v_i = -2 v0 . grad v_i
//                - d_i p / rho0, RHS v_z - rho' g / rho0 + source, RHS xi_i = v_i
// where the z passes write the next stage's input T = a + beta * RHS (and
// its z-PML variables) instead of the RHS, and the fifth stage's the updated
// state a + deltat * RHS; every other step one pass applies
// FILTER_VARS_3D_HORIZ_PML (physics.f90:836-882).
//
// Layout: fields [nz][ny][nx], x fastest.  A z pass gives each lane one (x, y)
// column and streams it bottom to top and back (the Thomas sweeps of dbyd2 /
// mtridag), so the 64 lanes of a wave touch 64 consecutive doubles per row.
// At nx = ny = 256 a pass has 65536 columns, 1024 waves, one per SIMD: the
// sweeps issue their loads in chunks of rows (and Zv runs four independent
// chains) so that each lane keeps tens of loads in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sparc_hip3d.h"
#include "sparc_device.h"
#include "sparc_host.h"
#include "sparc_launch.h"

namespace sparc {
namespace d3 {

// d/dy (ddxyz, derivatives.f90:224-247) of nplanes [NY][nx] planes: a team of
// N1 lanes transforms two neighbouring columns (x0, x0 + 1) as one complex
// line, so each lane moves 16 bytes per row; a block covers 2 * TEAMS
// consecutive columns (128 bytes, one cache line, of each row at NY = 256).  Blocks
// are renumbered so that the eight XCDs (which take the hardware block order
// round robin) each work on contiguous stretches of columns and keep their
// partial cache lines in their own L2.
// Block width per line length: the teams' LDS stays within 64 KiB.
template <int NY> struct DdyBlock { static constexpr int n = NY <= 128 ? 256 : (NY <= 512 ? 128 : 64); };

template <int NY>
__global__ __launch_bounds__(256) void k3_ddy(const double *__restrict__ in, double *__restrict__ out, int nplanes,
                                              int nx, const double2 *__restrict__ tw,
                                              const double *__restrict__ dmul, double stretch)
{
    constexpr int N1 = FftShape<NY>::N1, N2 = FftShape<NY>::N2, TEAMS = DdyBlock<NY>::n / N1;
    __shared__ double2 lds[TEAMS][FftLds<N1, N2>::n];
    const int team = threadIdx.x / N1, t = threadIdx.x % N1;
    const int hx = nx >> 1;
    int bid = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
    const long pair = (long)bid * TEAMS + team;
    const bool ok = pair < (long)nplanes * hx;
    const long pl = ok ? pair / hx : 0;
    const int x0 = ok ? 2 * (int)(pair - pl * hx) : 0;
    const size_t base = (size_t)pl * NY * nx + x0;
    double re[N2], im[N2];
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
        const double2 v = *(const double2 *)(in + base + (size_t)(t + N1 * j2) * nx);
        re[j2] = v.x;
        im[j2] = v.y;
    }
    fft_deriv_team<N1, N2>(re, im, t, lds[team], tw, dmul, stretch);
    if (ok) {
#pragma unroll
        for (int j2 = 0; j2 < N2; j2++)
            *(double2 *)(out + base + (size_t)(t + N1 * j2) * nx) = make_double2(re[j2], im[j2]);
    }
}

// The y-derivative for any ny that is not a power of two: circ_deriv_row
// along y on one column pair (x0, x0 + 1) per 64-lane block (nx even, so
// every pair is 16-byte aligned); cdy holds the derivative's circulant
// column twice over.  Grid: exactly nplanes * nx/2 blocks.
template <int N2>
__global__ __launch_bounds__(64) void k3_ddy_any(const double *__restrict__ in, double *__restrict__ out,
                                                 int nx, int ny, const double *__restrict__ cdy, double stretch)
{
    __shared__ double2 z[64 * N2];
    __shared__ double cc[128 * N2];
    for (int q = threadIdx.x; q < 128 * N2; q += 64) cc[q] = q < 2 * ny ? cdy[q] : 0.0;
    const int t = threadIdx.x, hx = nx >> 1;
    const long pl = (long)blockIdx.x / hx;
    const int x0 = 2 * (int)((long)blockIdx.x - pl * hx);
    const size_t base = (size_t)pl * ny * nx + x0;
    double re[N2], im[N2];
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
        const int y = min(t + 64 * j2, ny - 1);
        const double2 v = *(const double2 *)(in + base + (size_t)y * nx);
        re[j2] = v.x;
        im[j2] = v.y;
    }
    circ_deriv_row<N2>(re, im, t, z, cc, ny, stretch);
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
        const int y = t + 64 * j2;
        if (y < ny) *(double2 *)(out + base + (size_t)y * nx) = make_double2(re[j2], im[j2]);
    }
}

// Arguments of the two z passes.
struct Z3 {
    const double *B;            // Zp: xi_z; Zv: v_x (v_y, v_z follow at +N, +2N)
    double *W;                  // forward-sweep values, one [nz][nxy] block per chain
    size_t N;                   // points per field
    int nxy, nz, nb, nt;
    const double *stretch;      // [nz]
    DbyCoef cf;
    Tri tr1, tr0;               // dbyd2 factors for ibc 1 and ibc 0
    const double *rhoinv;       // [nz] 1/rho0
    const double *az, *bz;      // [npml][nxy]
    const double *psi;          // z-PML variables of this stage [2][npml][nxy]: d xi_z/dz, d p/dz
    // the next stage's input (step.f90:60-62) or, after the fifth stage, the
    // updated state (:107-108): out = base + beta * RHS, for the fields and
    // the z-PML variables; the fifth stage also keeps the RHS in rhs
    const double *a, *pml;      // base: state a [6][nz][nxy], pmlvars [2][npml][nxy]
    double *tnext, *pnext;      // T of the next stage (or a), its PML variables (or pmlvars)
    double beta;
    int last;
    // Zp
    const double *dxi, *dyi;    // d xi_x/dx, d xi_y/dy
    const double *c2rho0, *gp0, *gr0;   // [nz][nxy], PML factor applied
    const double *rho0, *g;     // [nz]
    double *P, *RG;             // p; rho' g / rho0
    // Zv
    const double *dxv, *dyv;    // d v_i/dx, d v_i/dy, [3][nz][nxy]
    const double *v0x, *v0y, *v0z;
    const double *dxp, *dyp;    // d p/dx, d p/dy
    double *rhs;                // scr [6][nz][nxy]
    const double *srcdep;       // [nz]
    Src src;
    int time, e_rad;
};

__device__ __forceinline__ int pml_index(int j, int nz, int nb, int nt)
{
    return j < nb ? j : (j >= nz - nt ? nb + (j - (nz - nt)) : -1);
}

// Forward elimination (dbyd2.f + mtridag.f:96-105) of NC columns of one lane
// at once, rows 2..nz-3 in chunks of U whose loads are issued together: a
// lane then has NC * U loads in flight instead of one per dependent step.
template <int NC, int U>
__device__ __forceinline__ void sweep_fwd(FwdSweep (&fw)[NC], const double *const (&B)[NC],
                                          double *const (&W)[NC], const int (&ibc)[NC], const Tri *const (&tr)[NC],
                                          size_t c, size_t s, int nz, const DbyCoef &cf)
{
#pragma unroll
    for (int k = 0; k < NC; k++) fw[k].start(B[k], c, s, ibc[k], 0.0, cf, *tr[k], W[k]);
    for (int j0 = 2; j0 <= nz - 3; j0 += U) {
        double bn[NC][U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const size_t qn = c + (size_t)min(j0 + u + 3, nz - 1) * s;
#pragma unroll
            for (int k = 0; k < NC; k++) bn[k][u] = B[k][qn];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = j0 + u;
            if (j <= nz - 3) {
#pragma unroll
                for (int k = 0; k < NC; k++)
                    fw[k].step(j, c + (size_t)j * s, j + 3 < nz ? bn[k][u] : 0.0, cf, *tr[k], W[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NC; k++) fw[k].finish(B[k], c, s, nz, ibc[k], 0.0, cf, *tr[k], W[k]);
}

// Zp: d xi_z/dz (ibc 1) with the z-PML (displacement.f90:400-409), div, p and
// rho' (:411-415), and rho' g / rho0 for RHS v_z (:458).  One lane per column.
__global__ __launch_bounds__(256) void k3_zp(Z3 z)
{
    constexpr int UF = 16, UB = 4;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= z.nxy) return;
    const size_t s = (size_t)z.nxy;
    const int nz = z.nz;
    FwdSweep fw[1];
    const double *const Bs[1] = {z.B};
    double *const Ws[1] = {z.W};
    const int ibc[1] = {1};
    const Tri *const trs[1] = {&z.tr1};
    sweep_fwd<1, UF>(fw, Bs, Ws, ibc, trs, c, s, nz, z.cf);
    double d = fw[0].ud;
    for (int j0 = nz - 1; j0 >= 0; j0 -= UB) {
        double w[UB], dxi[UB], dyi[UB], xiz[UB], c2r[UB], gp[UB], gr[UB];
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const size_t q = c + (size_t)max(j0 - u, 0) * s;
            w[u] = z.W[q];
            dxi[u] = z.dxi[q];
            dyi[u] = z.dyi[q];
            xiz[u] = z.B[q];
            c2r[u] = z.c2rho0[q];
            gp[u] = z.gp0[q];
            gr[u] = z.gr0[q];
        }
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const int j = j0 - u;
            if (j < 0) break;
            const size_t q = c + (size_t)j * s;
            if (j < nz - 1) d = w[u] - z.tr1.gam[j + 1] * d;
            double dz = d * z.stretch[j];
            const int pj = pml_index(j, nz, z.nb, z.nt);
            if (pj >= 0) {
                const size_t qp = c + (size_t)pj * s;
                const double ps = z.psi[qp];
                const double rp = z.az[qp] * dz + z.bz[qp] * ps;
                z.pnext[qp] = z.pml[qp] + z.beta * rp;
                dz = dz + ps;
            }
            const double div = (dxi[u] + dyi[u]) + dz;
            z.P[q] = -(c2r[u] * div) - xiz[u] * gp[u];
            const double rho = -(z.rho0[j] * div) - xiz[u] * gr[u];
            z.RG[q] = z.rhoinv[j] * rho * z.g[j];
        }
    }
}

// Zv: the z-derivatives of v_x, v_y, v_z (ibc 1) and of p (ibc 0, with the
// z-PML, :422-431) as four chains of one lane, then RHS v_i = -2 v0 . grad v_i
// - d_i p / rho0 (:436-452), RHS v_z - rho' g / rho0 (:458) + forcing *
// source_dep (:460-462), RHS xi_i = v_i (:466-468), rows 1 and nz zeroed
// (:470-473).
__global__ __launch_bounds__(256) void k3_zv(Z3 z)
{
    constexpr int UF = 8, UB = 2;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= z.nxy) return;
    const size_t s = (size_t)z.nxy, N = z.N;
    const int nz = z.nz;
    FwdSweep fw[4];
    const double *const Bs[4] = {z.B, z.B + N, z.B + 2 * N, z.P};
    double *const Ws[4] = {z.W, z.W + N, z.W + 2 * N, z.W + 3 * N};
    const int ibc[4] = {1, 1, 1, 0};
    const Tri *const trs[4] = {&z.tr1, &z.tr1, &z.tr1, &z.tr0};
    sweep_fwd<4, UF>(fw, Bs, Ws, ibc, trs, c, s, nz, z.cf);
    // forcing of this column (physics.f90:678-737 for the adjoint)
    double forcing = 0.0;
    if (z.time <= z.src.limit) {
        if (z.src.mode == 1)
            forcing = (z.time >= 0 && z.time < z.src.nfwd) ? z.src.vr[c] * z.src.fwd[z.time] : 0.0;
        else if (z.src.mode == 2)
            forcing = lagrange_forcing(z.src, 0, c, z.nxy, z.time);
    }
    double d[4] = {fw[0].ud, fw[1].ud, fw[2].ud, fw[3].ud};
    double *const r = z.rhs;
    for (int j0 = nz - 1; j0 >= 0; j0 -= UB) {
        double w[4][UB], dx[3][UB], dy[3][UB], v[3][UB], u0[3][UB], dxp[UB], dyp[UB], rg[UB];
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const size_t q = c + (size_t)max(j0 - u, 0) * s;
#pragma unroll
            for (int k = 0; k < 4; k++) w[k][u] = z.W[q + k * N];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                dx[i][u] = z.dxv[q + i * N];
                dy[i][u] = z.dyv[q + i * N];
                v[i][u] = z.B[q + i * N];
            }
            u0[0][u] = z.v0x[q];
            u0[1][u] = z.v0y[q];
            u0[2][u] = z.v0z[q];
            dxp[u] = z.dxp[q];
            dyp[u] = z.dyp[q];
            rg[u] = z.RG[q];
        }
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const int j = j0 - u;
            if (j < 0) break;
            const size_t q = c + (size_t)j * s;
            if (j < nz - 1) {
#pragma unroll
                for (int k = 0; k < 3; k++) d[k] = w[k][u] - z.tr1.gam[j + 1] * d[k];
                d[3] = w[3][u] - z.tr0.gam[j + 1] * d[3];
            }
            const double st = z.stretch[j];
            double gpz = d[3] * st;
            const int pj = pml_index(j, nz, z.nb, z.nt);
            if (pj >= 0) {
                const size_t qa = c + (size_t)pj * s, qp = qa + (size_t)(z.nb + z.nt) * s;
                const double ps = z.psi[qp];
                const double rp = z.az[qa] * gpz + z.bz[qa] * ps;
                z.pnext[qp] = z.pml[qp] + z.beta * rp;
                gpz = gpz + ps;
            }
            const bool edge = j == 0 || j == nz - 1;
            const double ri = z.rhoinv[j];
            double adv[3];
#pragma unroll
            for (int i = 0; i < 3; i++)
                adv[i] = -2.0 * ((u0[0][u] * dx[i][u] + u0[1][u] * dy[i][u]) + u0[2][u] * (d[i] * st));
            double rhs[6];
            rhs[3] = edge ? 0.0 : adv[0] - ri * dxp[u];
            rhs[4] = edge ? 0.0 : adv[1] - ri * dyp[u];
            double rz = adv[2] - ri * gpz;
            rz = rz - rg[u];
            const int dk = j - z.e_rad;
            if (dk >= -4 && dk <= 4) rz = rz + forcing * z.srcdep[j];
            rhs[5] = edge ? 0.0 : rz;
#pragma unroll
            for (int i = 0; i < 3; i++) rhs[i] = edge ? 0.0 : v[i][u];
            double base[6];
#pragma unroll
            for (int f = 0; f < 6; f++) base[f] = z.a[q + f * N];
#pragma unroll
            for (int f = 0; f < 6; f++) z.tnext[q + f * N] = base[f] + z.beta * rhs[f];
            if (z.last) {
#pragma unroll
                for (int f = 0; f < 6; f++) r[q + f * N] = rhs[f];
            }
        }
    }
}

// FILTER_VARS_3D_HORIZ_PML (physics.f90:855-878) on the six fields, rows
// npmlbot+1 .. nz-npmltop, in place from a register window of the unfiltered
// column.
__global__ __launch_bounds__(256) void k3_filter(double *a, int nxy, int nz, int nb, int nt)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= 6L * nxy) return;
    const int f = (int)(id / nxy), c = (int)(id - (long)f * nxy);
    const size_t s = (size_t)nxy;
    double *A = a + (size_t)f * nz * s + c;
    double w[11];                       // w[k] = a(j - 5 + k) at filtered row j
#pragma unroll
    for (int k = 0; k < 11; k++) w[k] = A[(size_t)(nb - 5 + k) * s];
    const double c0 = 0.75390625, c1 = 0.41015625, c2 = -0.234375, c3 = 0.087890625,
                 c4 = -0.01953125, c5 = 0.001953125;
    for (int j = nb; j < nz - nt; j++) {
        double v = c0 * w[5] + 0.5 * c1 * (w[4] + w[6]);
        v = v + 0.5 * c2 * (w[3] + w[7]);
        v = v + 0.5 * c3 * (w[2] + w[8]);
        v = v + 0.5 * c4 * (w[1] + w[9]);
        v = v + 0.5 * c5 * (w[0] + w[10]);
        const int jn = j + 6;
        const double un = jn < nz ? A[(size_t)jn * s] : 0.0;
        A[(size_t)j * s] = v;
#pragma unroll
        for (int k = 0; k < 10; k++) w[k] = w[k + 1];
        w[10] = un;
    }
}

// One snapshot pair of PRODUCE_KERNELS with its y terms (kernels.f90:
// 289-355): DX, DY, DZ hold the x, y and z (ddzkern) derivatives of forward
// xi_x / xi_y / xi_z (slot 0), forward v_x, v_y, v_z (slots 1-3) and adjoint
// v_x / v_y / v_z (slot 4); the sums keep the reference's order of additions.
__global__ __launch_bounds__(256) void k3_kern_accum(const double *__restrict__ f, const double *__restrict__ a,
                                                     const double *__restrict__ DX, const double *__restrict__ DY,
                                                     const double *__restrict__ DZ, double *K, size_t N)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < N; q += stride) {
        const double fax = f[q + 6 * N], fay = f[q + 7 * N], faz = f[q + 8 * N];
        K[q + 4 * N] = K[q + 4 * N] + faz * faz + fax * fax + fay * fay;           // hessian :293-294
        const double divf = (DX[q] + DY[q]) + DZ[q];                             // compute_div :843-861
        const double diva = (DX[q + 4 * N] + DY[q + 4 * N]) + DZ[q + 4 * N];
        K[q] = K[q] + divf * diva;                                               // :303-306
        const double avx = a[q], avy = a[q + N], avz = a[q + 2 * N];
        const double *D[3] = {DX, DY, DZ};
#pragma unroll
        for (int i = 0; i < 3; i++)                                              // :338-355
            K[q + (1 + i) * N] = D[i][q + N] * avx + D[i][q + 2 * N] * avy + D[i][q + 3 * N] * avz +
                                 K[q + (1 + i) * N];
    }
}

// Units (kernels.f90:498-522, 563, 579): K = [Kc, Kvx, Kvy, Kvz, H]; T1, T2 the
// two halves of the y component of CURL_KERN (:832-834) of (Kvx, 0, Kvz):
// ddzkern(Kvx) with bcz = 0 and ddxkern(Kvz).
__global__ __launch_bounds__(256) void k3_kern_final(const double *__restrict__ K, const double *__restrict__ T1,
                                                     const double *__restrict__ T2, const double *__restrict__ rho0z,
                                                     const double *__restrict__ c2, size_t N, size_t nxy,
                                                     double steps, double dt, double dimen, double dimen_c,
                                                     double *kc, double *kpsi, double *kvx, double *kvy, double *kvz,
                                                     double *hess)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < N; q += stride) {
        const double r = rho0z[q / nxy];
        const double kp = T1[q] - T2[q];
        kpsi[q] = 2.0 * kp * dimen * 1.0 * (r * sqrt(c2[q])) * steps * dt * dimen;   // :511-513
        kvx[q] = -2.0 * K[q + N] * r * steps * dt * dimen;
        kvy[q] = -2.0 * K[q + 2 * N] * r * steps * dt * dimen;
        kvz[q] = -2.0 * K[q + 3 * N] * r * steps * dt * dimen;
        kc[q] = -K[q] * r * c2[q] * steps * dt * dimen_c * 2.0;
        hess[q] = K[q + 4 * N];
    }
}

#define SPARC3D_DISPATCH(N, MACRO) \
    switch (N) {                   \
    case 16: MACRO(16); break;     \
    case 32: MACRO(32); break;     \
    case 64: MACRO(64); break;     \
    case 128: MACRO(128); break;   \
    case 256: MACRO(256); break;   \
    case 512: MACRO(512); break;   \
    case 1024: MACRO(1024); break; \
    default: return hipErrorInvalidValue; \
    }

static hipError_t launch_ddy(const double *in, double *out, int nplanes, int nx, int ny, const double2 *tw,
                             const double *dmul, const double *cdy, double stretch, hipStream_t st)
{
    if (!nx_has_fft(ny)) {
        if (ny < 16 || ny > 1024 || (nx & 1)) return hipErrorInvalidValue;
        const dim3 grid((unsigned)((long)nplanes * (nx / 2)));
        switch (any_nx_slots(ny)) {
#define A(N2V)                                                                                           \
        case N2V: hipLaunchKernelGGL((k3_ddy_any<N2V>), grid, dim3(64), 0, st, in, out, nx, ny, cdy, stretch); break;
            SPARC_ANY_NX_SLOTS(A)
#undef A
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
#define L(NYV)                                                                                     \
    {                                                                                              \
        constexpr int BLK = DdyBlock<NYV>::n, TEAMS = BLK / FftShape<NYV>::N1;                     \
        const long pairs = (long)nplanes * (nx / 2);                                               \
        hipLaunchKernelGGL(k3_ddy<NYV>, dim3((unsigned)((pairs + TEAMS - 1) / TEAMS)), dim3(BLK), 0, \
                           st, in, out, nplanes, nx, tw, dmul, stretch);                           \
    }
    SPARC3D_DISPATCH(ny, L)
#undef L
    return hipGetLastError();
}

static hipError_t launch_zp(const Z3 &z, hipStream_t st)
{
    hipLaunchKernelGGL(k3_zp, dim3((z.nxy + 255) / 256), dim3(256), 0, st, z);
    return hipGetLastError();
}

static hipError_t launch_zv(const Z3 &z, hipStream_t st)
{
    hipLaunchKernelGGL(k3_zv, dim3((z.nxy + 255) / 256), dim3(256), 0, st, z);
    return hipGetLastError();
}

static hipError_t launch_filter(double *a, int nxy, int nz, int nb, int nt, hipStream_t st)
{
    const long n = 6L * nxy;
    hipLaunchKernelGGL(k3_filter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, nxy, nz, nb, nt);
    return hipGetLastError();
}

}  // namespace d3
}  // namespace sparc

using namespace sparc;
using namespace sparc::d3;

namespace {
constexpr double kRsun3 = 6.95989467700e10;   // initialize.f90:37
constexpr double kPi3 = 3.141592653589793;
}  // namespace

struct sparc3d_handle {
    int nx = 0, ny = 0, nz = 0, nb = 0, nt = 0, adjoint = 0;
    double xlength = 0, ylength = 0, timestep = 0, cadforcing = 0, outputcad = 0, obsheight = 0, excitdep = 0;
    double nupeak = 0, nuwidth = 0;
    std::vector<double> model, c3d;
    std::string err;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    size_t nxy = 0, n = 0, np = 0;      // columns, points per field, PML points per variable
    std::vector<double> z, stretch, g, rho0z, rhoinvz, srcdep;
    int e_rad = 0, o_rad = 0, steps = 1, cfs = 1, maxtime = -1;
    double deltat = 0, stretchx = 0, stretchy = 0, betas[5] = {0, 0, 0, 0, 0};
    DbyCoef cf{};
    Tri tri[4]{};
    DevArena bg_mem, st_mem, src_mem;
    const double2 *twx = nullptr, *twy = nullptr;
    const double *cdx = nullptr, *cdy = nullptr;      // circulant columns [2n] (an n that is not a power of two)
    const double *dmx = nullptr, *dmy = nullptr, *d_stretch = nullptr, *d_rhoinv = nullptr, *d_rho0 = nullptr,
                 *d_g = nullptr, *d_srcdep = nullptr;
    double *c2rho0 = nullptr, *gp0 = nullptr, *gr0 = nullptr, *az = nullptr, *bz = nullptr;
    double *v0x = nullptr, *v0y = nullptr, *v0z = nullptr;
    double *a = nullptr, *scr = nullptr, *tmp = nullptr, *pml = nullptr, *psit = nullptr;
    double *DXI = nullptr, *DYI = nullptr, *DX3 = nullptr, *DY3 = nullptr, *P = nullptr, *RG = nullptr,
           *DXP = nullptr, *DYP = nullptr, *W = nullptr;
    Src src{};
    bool have_src = false;
    // kernel accumulation: K [5][N] (Kc, Kvx, Kvy, Kvz, H), derivatives D [3][5][N],
    // kernel_back's rho0 [nz] and c2 [N] (initialize.f90:937-979)
    DevArena kern_mem;
    double *K = nullptr, *KD = nullptr, *krho0 = nullptr, *kc2 = nullptr;
    long kern_pairs = 0;
    // instrumentation
    struct Mark { hipEvent_t a, b; int kind; };
    std::vector<hipEvent_t> evpool;
    std::vector<Mark> marks;
    bool profiling = false;
};

#define HCHECK3(h, expr)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
            return SPARC_E_HIP;                                                            \
        }                                                                                  \
    } while (0)

static thread_local std::string g_create_err3;

static int fail3(sparc3d_handle *h, int code, const std::string &msg)
{
    h->err = msg;
    return code;
}


// Launch with HIP events around it when profiling (kinds as in sparc3d_profile).
template <class F>
static int timed3(sparc3d_handle *h, int kind, F launch)
{
    if (!h->profiling) {
        HCHECK3(h, launch());
        return SPARC_OK;
    }
    hipEvent_t e0, e1;
    HCHECK3(h, hipEventCreate(&e0));
    HCHECK3(h, hipEventCreate(&e1));
    HCHECK3(h, hipEventRecord(e0, h->stream));
    HCHECK3(h, launch());
    HCHECK3(h, hipEventRecord(e1, h->stream));
    h->marks.push_back({e0, e1, kind});
    return SPARC_OK;
}

#define RUN3(kind, expr)                                       \
    do {                                                       \
        int r_ = timed3(h, kind, [&]() { return (expr); });    \
        if (r_) return r_;                                     \
    } while (0)

// ddx of nrows rows, ddy of nplanes planes (device).
static int ddx3(sparc3d_handle *h, const double *in, double *out, int nrows)
{
    RUN3(1, launch_ddx_rows(in, out, nrows, h->nx, h->twx, h->dmx, h->cdx, h->stretchx, h->stream));
    return SPARC_OK;
}

static int ddy3(sparc3d_handle *h, const double *in, double *out, int nplanes)
{
    RUN3(2, launch_ddy(in, out, nplanes, h->nx, h->ny, h->twy, h->dmy, h->cdy, h->stretchy, h->stream));
    return SPARC_OK;
}

static Z3 zargs(sparc3d_handle *h, const double *B)
{
    Z3 z{};
    z.B = B;
    z.W = h->W;
    z.N = h->n;
    z.nxy = (int)h->nxy;
    z.nz = h->nz;
    z.nb = h->nb;
    z.nt = h->nt;
    z.stretch = h->d_stretch;
    z.cf = h->cf;
    z.tr1 = h->tri[1];
    z.tr0 = h->tri[0];
    z.rhoinv = h->d_rhoinv;
    z.az = h->az;
    z.bz = h->bz;
    return z;
}

// One RK stage s (0-based) of MP_FLOWS_PML_3D_DISPL at time index t.
static int enqueue_stage(sparc3d_handle *h, int s, int t)
{
    const size_t N = h->n;
    const int nyz = h->ny * h->nz, nz = h->nz;
    // stage 1 reads a and pmlvars (step.f90:50-52); the z passes of stage s
    // write stage s+1's input T = a + beta * RHS (:60-62) into tmp / psit, and
    // those of stage 5 the update a += deltat * RHS (:107-108) into a / pmlvars
    const double *T = s == 0 ? h->a : h->tmp, *PS = s == 0 ? h->pml : h->psit;
    const bool last = s == 4;
    double *tnext = last ? h->a : h->tmp, *pnext = last ? h->pml : h->psit;
    const double beta = last ? h->deltat : h->betas[s + 1];
    int r;
    // ddxyz(xi_x, xi_y, xi_z) and the x, y derivatives of the velocity
    if ((r = ddx3(h, T, h->DXI, nyz))) return r;
    if ((r = ddx3(h, T + 3 * N, h->DX3, 3 * nyz))) return r;
    if ((r = ddy3(h, T + N, h->DYI, nz))) return r;
    if ((r = ddy3(h, T + 3 * N, h->DY3, 3 * nz))) return r;
    {
        Z3 z = zargs(h, T + 2 * N);
        z.psi = PS;
        z.pml = h->pml;
        z.pnext = pnext;
        z.beta = beta;
        z.dxi = h->DXI;
        z.dyi = h->DYI;
        z.c2rho0 = h->c2rho0;
        z.gp0 = h->gp0;
        z.gr0 = h->gr0;
        z.rho0 = h->d_rho0;
        z.g = h->d_g;
        z.P = h->P;
        z.RG = h->RG;
        RUN3(3, launch_zp(z, h->stream));
    }
    if ((r = ddx3(h, h->P, h->DXP, nyz))) return r;
    if ((r = ddy3(h, h->P, h->DYP, nz))) return r;
    {
        Z3 z = zargs(h, T + 3 * N);
        z.P = h->P;
        z.RG = h->RG;
        z.psi = PS;
        z.pml = h->pml;
        z.pnext = pnext;
        z.a = h->a;
        z.tnext = tnext;
        z.beta = beta;
        z.last = last;
        z.dxv = h->DX3;
        z.dyv = h->DY3;
        z.v0x = h->v0x;
        z.v0y = h->v0y;
        z.v0z = h->v0z;
        z.dxp = h->DXP;
        z.dyp = h->DYP;
        z.rhs = h->scr;
        z.srcdep = h->d_srcdep;
        z.src = h->src;
        z.time = t;
        z.e_rad = h->e_rad;
        RUN3(4, launch_zv(z, h->stream));
    }
    return SPARC_OK;
}

static int enqueue_step3(sparc3d_handle *h, int t)
{
    for (int s = 0; s < 5; s++) {
        int r = enqueue_stage(h, s, t);
        if (r) return r;
    }
    const int filt = (t + 1) % 2 == 0;                           // step.f90:114
    if (filt) RUN3(5, launch_filter(h->a, (int)h->nxy, h->nz, h->nb, h->nt, h->stream));
    return SPARC_OK;
}

// d/dz of a host column (set-up): dbyd2 with ibc 1, times stretch when given.
static int host_ddz3(sparc3d_handle *h, const std::vector<double> &in, std::vector<double> &out,
                     const double *stretch_dev)
{
    const int nz = h->nz;
    double *di = nullptr, *dout = nullptr, *w = nullptr;
    DevArena tmp;
    HCHECK3(h, upload(tmp, &di, in, h->stream));
    HCHECK3(h, tmp.alloc(&dout, nz));
    HCHECK3(h, tmp.alloc(&w, nz));
    HCHECK3(h, launch_ddz_cols(di, dout, 1, nz, 1, 1, stretch_dev, h->cf, h->tri[1], w, 0, h->stream));
    out.assign(nz, 0.0);
    HCHECK3(h, hipMemcpyAsync(out.data(), dout, nz * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    tmp.release();
    return SPARC_OK;
}

static int setup3(sparc3d_handle *h)
{
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const size_t nxy = h->nxy, N = h->n;
    const double dimc = 1e6, dimrho = 1e-3, diml = kRsun3, dimc2 = std::pow(dimc, 2.0);
    // x, y spectral tables (initialize.f90:194-225): exp(-2 pi i k/n), i k
    // (zero at the Nyquist bin of an even n), and for an n that is not a
    // power of two the derivative's circulant column c[m] = -sum_k dm[k]
    // sin(2 pi (k m mod n)/n), twice over (circ_deriv_row)
    auto tables = [&](int n, const double2 **tw, const double **dm, const double **cd) -> int {
        std::vector<double> t(2 * (size_t)n), m(n);
        for (int k = 0; k < n; k++) {
            const double ang = -2.0 * kPi3 * (double)k / (double)n;
            double c = std::cos(ang), s = std::sin(ang);
            if (4 * k == n || 4 * k == 3 * n) c = 0.0;
            if (2 * k == n) s = 0.0;
            t[2 * k] = c;
            t[2 * k + 1] = s;
        }
        const double norm = 1.0 / (double)n;
        for (int k = 0; k < n; k++) {
            const int kk = 2 * k < n ? k : (2 * k == n ? -1 : n - k);
            const double e = kk < 0 ? 0.0 : (double)kk * norm * 2.0 * kPi3;
            m[k] = 2 * k <= n ? e : -e;
        }
        std::vector<double> c(2 * (size_t)n, 0.0);
        if (!nx_has_fft(n)) {
            std::vector<long double> sn(n);
            for (int q = 0; q < n; q++) sn[q] = std::sin(2.0L * (long double)kPi3 * q / n);
            for (int j = 0; j < n; j++) {
                long double acc = 0.0L;
                for (int k = 0; k < n; k++) acc -= (long double)m[k] * sn[(int)(((long)k * j) % n)];
                c[j] = c[j + n] = (double)acc;
            }
        }
        double *p = nullptr;
        HCHECK3(h, upload(h->bg_mem, &p, t, h->stream));
        *tw = (const double2 *)p;
        HCHECK3(h, upload(h->bg_mem, &p, m, h->stream));
        *dm = p;
        HCHECK3(h, upload(h->bg_mem, &p, c, h->stream));
        *cd = p;
        return SPARC_OK;
    };
    int r;
    if ((r = tables(nx, &h->twx, &h->dmx, &h->cdx))) return r;
    if ((r = tables(ny, &h->twy, &h->dmy, &h->cdy))) return r;
    h->stretchx = kRsun3 / h->xlength;
    h->stretchy = kRsun3 / h->ylength;
    h->cf = make_coef(nz);
    for (int ibc = 0; ibc < 4; ibc++) {
        std::vector<double> lowr, gam, rbet;
        make_tri(nz, ibc, lowr, gam, rbet);
        double *pa = nullptr, *pb = nullptr, *pc = nullptr;
        HCHECK3(h, upload(h->bg_mem, &pa, lowr, h->stream));
        HCHECK3(h, upload(h->bg_mem, &pb, gam, h->stream));
        HCHECK3(h, upload(h->bg_mem, &pc, rbet, h->stream));
        h->tri[ibc] = Tri{pa, pb, pc};
    }
    // z and stretch = 1/dbyd2(z) (initialize.f90:466-470); deltat (:474)
    h->z.resize(nz);
    for (int k = 0; k < nz; k++) h->z[k] = h->model[6 * (size_t)k];
    std::vector<double> us;
    if ((r = host_ddz3(h, h->z, us, nullptr))) return r;
    h->stretch.resize(nz);
    for (int k = 0; k < nz; k++) h->stretch[k] = 1.0 / us[k];
    double *p = nullptr;
    HCHECK3(h, upload(h->bg_mem, &p, h->stretch, h->stream));
    h->d_stretch = p;
    h->deltat = h->timestep * dimc / kRsun3;
    // solar_data (initialize.f90:731-794)
    std::vector<double> c2(N), p0z(nz);
    h->rho0z.resize(nz);
    h->g.resize(nz);
    for (int k = 0; k < nz; k++) {
        const double *row = &h->model[6 * (size_t)k];
        h->rho0z[k] = row[2] / dimrho;
        p0z[k] = row[3] / (dimrho * dimc2);
        h->g[k] = row[4] * diml / dimc2;
        for (size_t q = 0; q < nxy; q++) {
            const double c = (h->c3d.empty() ? row[1] : h->c3d[(size_t)k * nxy + q]) / dimc;
            c2[(size_t)k * nxy + q] = c * c;
        }
    }
    // e_rad, o_rad, source_dep (physics.f90:44-60, 105)
    {
        const double t2 = 1.0 + h->excitdep / kRsun3;
        double t = 1.0;
        for (int k = 0; k < nz; k++)
            if (std::fabs(h->z[k] - t2) <= t) { t = std::fabs(h->z[k] - t2); h->e_rad = k; }
        const double tr = 1.0 + h->obsheight / kRsun3;
        t = 1.0;
        for (int k = 0; k < nz; k++)
            if (std::fabs(h->z[k] - tr) <= t) { t = std::fabs(h->z[k] - tr); h->o_rad = k; }
        h->srcdep.resize(nz);
        const double k4 = (double)4.0e7f;
        for (int k = 0; k < nz; k++) {
            const double dz = h->z[k] - h->z[h->e_rad];
            h->srcdep[k] = std::exp(-k4 * (dz * dz));
        }
    }
    // rhoinv, c2rho0, d rho0/dz, d p0/dz (physics.f90:203-208)
    h->rhoinvz.resize(nz);
    for (int k = 0; k < nz; k++) h->rhoinvz[k] = 1.0 / h->rho0z[k];
    std::vector<double> gr0z, gp0z;
    if ((r = host_ddz3(h, h->rho0z, gr0z, h->d_stretch))) return r;
    if ((r = host_ddz3(h, p0z, gp0z, h->d_stretch))) return r;
    std::vector<double> c2rho0(N), gp0(N), gr0(N);
    for (int k = 0; k < nz; k++)
        for (size_t q = 0; q < nxy; q++) {
            const size_t i = (size_t)k * nxy + q;
            c2rho0[i] = c2[i] * h->rho0z[k];
            gp0[i] = gp0z[k];
            gr0[i] = gr0z[k];
        }
    // z-PML (physics.f90:250-296)
    const int nb = h->nb, nt = h->nt, npml = nb + nt;
    std::vector<double> az((size_t)npml * nxy, 0.0), bz((size_t)npml * nxy, 0.0);
    const double *zz = h->z.data();
    const double a0 = kPi3 * (double)0.005f * diml / dimc;
    for (int k = 0; k < nz; k++) {
        const bool bot = k < nb, top = k > nz - nt - 1;
        if (!bot && !top) continue;
        const int pj = bot ? k : nb + (k - (nz - nt));
        double alphaz = a0;
        if (bot) alphaz = a0 * (zz[k] - zz[0]) / (zz[nb - 1] - zz[0]);
        if (top) alphaz = a0 * (zz[nz - 1] - zz[k]) / (zz[nz - 1] - zz[nz - nt]);
        for (size_t q = 0; q < nxy; q++) {
            double dampz = 0.0;
            if (bot) {
                const double rr = (zz[nb - 1] - zz[k]) / (zz[nb - 1] - zz[0]);
                dampz = 18.0 * std::sqrt(c2[q]) / (zz[nb - 1] - zz[0]) * (rr * rr);
            }
            if (top) {
                const double rr = (zz[nz - nt] - zz[k]) / (zz[nz - 1] - zz[nz - nt]);
                dampz = 9.0 * std::sqrt(c2[(size_t)(nz - 1) * nxy + q]) / (zz[nz - 1] - zz[nz - nt]) * (rr * rr);
            }
            const double fd = alphaz / (dampz + alphaz);
            const size_t i = (size_t)k * nxy + q;
            gp0[i] *= fd;
            gr0[i] *= fd;
            if (q == 0) h->g[k] *= fd;
            bz[(size_t)pj * nxy + q] = -alphaz - dampz;
            az[(size_t)pj * nxy + q] = -dampz;
        }
    }
    HCHECK3(h, upload(h->bg_mem, &h->c2rho0, c2rho0, h->stream));
    HCHECK3(h, upload(h->bg_mem, &h->gp0, gp0, h->stream));
    HCHECK3(h, upload(h->bg_mem, &h->gr0, gr0, h->stream));
    HCHECK3(h, upload(h->bg_mem, &h->az, az, h->stream));
    HCHECK3(h, upload(h->bg_mem, &h->bz, bz, h->stream));
    HCHECK3(h, upload(h->bg_mem, &p, h->rhoinvz, h->stream));
    h->d_rhoinv = p;
    HCHECK3(h, upload(h->bg_mem, &p, h->rho0z, h->stream));
    h->d_rho0 = p;
    HCHECK3(h, upload(h->bg_mem, &p, h->g, h->stream));
    h->d_g = p;
    HCHECK3(h, upload(h->bg_mem, &p, h->srcdep, h->stream));
    h->d_srcdep = p;
    HCHECK3(h, h->bg_mem.alloc(&h->v0x, N));
    HCHECK3(h, h->bg_mem.alloc(&h->v0y, N));
    HCHECK3(h, h->bg_mem.alloc(&h->v0z, N));
    HCHECK3(h, hipMemsetAsync(h->v0x, 0, N * sizeof(double), h->stream));
    HCHECK3(h, hipMemsetAsync(h->v0y, 0, N * sizeof(double), h->stream));
    HCHECK3(h, hipMemsetAsync(h->v0z, 0, N * sizeof(double), h->stream));
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
    // state and work buffers
    DevArena &m = h->st_mem;
    HCHECK3(h, m.alloc(&h->a, 6 * N));
    HCHECK3(h, m.alloc(&h->scr, 6 * N));
    HCHECK3(h, m.alloc(&h->tmp, 6 * N));
    HCHECK3(h, m.alloc(&h->pml, 2 * h->np));
    HCHECK3(h, m.alloc(&h->psit, 2 * h->np));
    HCHECK3(h, m.alloc(&h->DXI, N));
    HCHECK3(h, m.alloc(&h->DYI, N));
    HCHECK3(h, m.alloc(&h->DX3, 3 * N));
    HCHECK3(h, m.alloc(&h->DY3, 3 * N));
    HCHECK3(h, m.alloc(&h->P, N));
    HCHECK3(h, m.alloc(&h->RG, N));
    HCHECK3(h, m.alloc(&h->DXP, N));
    HCHECK3(h, m.alloc(&h->DYP, N));
    HCHECK3(h, m.alloc(&h->W, 4 * N));   // one block per chain of the Zv pass
    HCHECK3(h, hipMemsetAsync(h->a, 0, 6 * N * sizeof(double), h->stream));
    HCHECK3(h, hipMemsetAsync(h->scr, 0, 6 * N * sizeof(double), h->stream));
    HCHECK3(h, hipMemsetAsync(h->pml, 0, 2 * h->np * sizeof(double), h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

static void set_times3(sparc3d_handle *h, int maxtime)
{
    h->maxtime = maxtime;
    const int cfs = h->cfs > 0 ? h->cfs : 1;
    const int fl = maxtime / cfs + 1;                           // initialize.f90:994
    const int ns = (int)std::floor((double)maxtime * h->timestep / h->cadforcing) + 2;
    h->src.limit = fl * cfs;
    h->src.num_steps = 2 * (ns / 2) + 2;                        // driver.f90:217-218
    h->src.timestep = h->timestep;
    h->src.cadforcing = h->cadforcing;
}

extern "C" {

const char *sparc3d_last_error(const sparc3d_handle *h)
{
    if (!h) return g_create_err3.c_str();
    return h->err.c_str();
}

int sparc3d_create(const sparc3d_config *cfg, sparc3d_handle **out)
{
    if (!cfg || !out || !cfg->model) {
        g_create_err3 = "null argument";
        return SPARC_E_ARG;
    }
    *out = nullptr;
    if (cfg->nx < 16 || cfg->nx > 1024 || (cfg->nx & 1) || cfg->ny < 16 || cfg->ny > 1024) {
        g_create_err3 = "nx must be even and in [16, 1024], ny in [16, 1024]";
        return SPARC_E_ARG;
    }
    if (cfg->nz < 16 || cfg->npmlbot < 5 || cfg->npmltop < 5 || cfg->npmlbot + cfg->npmltop + 1 > cfg->nz) {
        g_create_err3 = "nz >= 16 and z-PML rows of at least 5 each (the z filter reaches 5 rows) are needed";
        return SPARC_E_ARG;
    }
    if (cfg->timestep <= 0 || cfg->xlength <= 0 || cfg->ylength <= 0 || cfg->outputcad < cfg->timestep ||
        cfg->cadforcing < cfg->timestep) {
        g_create_err3 = "timestep, lengths or cadences out of range";
        return SPARC_E_ARG;
    }
    sparc3d_handle *h = new sparc3d_handle();
    h->nx = cfg->nx;
    h->ny = cfg->ny;
    h->nz = cfg->nz;
    h->nb = cfg->npmlbot;
    h->nt = cfg->npmltop;
    h->adjoint = cfg->adjoint ? 1 : 0;
    h->xlength = cfg->xlength;
    h->ylength = cfg->ylength;
    h->timestep = cfg->timestep;
    h->cadforcing = cfg->cadforcing;
    h->outputcad = cfg->outputcad;
    h->obsheight = cfg->obsheight;
    h->excitdep = cfg->excitdep;
    h->nupeak = cfg->nupeak;
    h->nuwidth = cfg->nuwidth;
    h->nxy = (size_t)h->nx * h->ny;
    h->n = h->nxy * h->nz;
    h->np = (size_t)(h->nb + h->nt) * h->nxy;
    h->steps = (int)std::floor(h->outputcad / h->timestep);     // driver.f90:214
    h->cfs = (int)std::floor(h->cadforcing / h->timestep);
    h->model.assign(cfg->model, cfg->model + 6 * (size_t)h->nz);
    if (cfg->c3d) h->c3d.assign(cfg->c3d, cfg->c3d + h->n);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_err3 = "hipStreamCreate failed";
        delete h;
        return SPARC_E_HIP;
    }
    h->own_stream = true;
    const int r = setup3(h);
    if (r) {
        g_create_err3 = h->err;
        sparc3d_destroy(h);
        return r;
    }
    *out = h;
    return SPARC_OK;
}

int sparc3d_destroy(sparc3d_handle *h)
{
    if (!h) return SPARC_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto &mk : h->marks) {
        (void)hipEventDestroy(mk.a);
        (void)hipEventDestroy(mk.b);
    }
    h->bg_mem.release();
    h->st_mem.release();
    h->src_mem.release();
    h->kern_mem.release();
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SPARC_OK;
}

int sparc3d_set_stream(sparc3d_handle *h, void *stream)
{
    if (!h) return SPARC_E_ARG;
    HCHECK3(h, hipStreamSynchronize(h->stream));
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
    return SPARC_OK;
}

int sparc3d_set_flow(sparc3d_handle *h, const double *vx, const double *vy, const double *vz)
{
    if (!h) return SPARC_E_ARG;
    if ((vx == nullptr) != (vy == nullptr) || (vx == nullptr) != (vz == nullptr))
        return fail3(h, SPARC_E_ARG, "give all three flow components or none");
    const size_t N = h->n;
    HCHECK3(h, hipStreamSynchronize(h->stream));
    if (!vx) {
        HCHECK3(h, hipMemsetAsync(h->v0x, 0, N * sizeof(double), h->stream));
        HCHECK3(h, hipMemsetAsync(h->v0y, 0, N * sizeof(double), h->stream));
        HCHECK3(h, hipMemsetAsync(h->v0z, 0, N * sizeof(double), h->stream));
        HCHECK3(h, hipStreamSynchronize(h->stream));
        return SPARC_OK;
    }
    const double sgn = h->adjoint ? -1.0 : 1.0;                 // driver.f90:144-147
    std::vector<double> v(N);
    const double *src[3] = {vx, vy, vz};
    double *dst[3] = {h->v0x, h->v0y, h->v0z};
    for (int c = 0; c < 3; c++) {
        for (size_t q = 0; q < N; q++) v[q] = sgn * (src[c][q] / 1e6);
        HCHECK3(h, hipMemcpyAsync(dst[c], v.data(), N * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HCHECK3(h, hipStreamSynchronize(h->stream));
    }
    return SPARC_OK;
}

int sparc3d_set_forward_source(sparc3d_handle *h, double xsrc, double ysrc, int maxtime)
{
    if (!h || maxtime < 0) return SPARC_E_ARG;
    if (h->adjoint) return fail3(h, SPARC_E_STATE, "forward source on an adjoint handle");
    HCHECK3(h, hipStreamSynchronize(h->stream));
    set_times3(h, maxtime);
    // FORWARD_SOURCE (driver.f90:1120-1145)
    const int nfwd = maxtime + 8;
    std::vector<double> fw(nfwd);
    for (int k = 0; k < nfwd; k++) {
        const double ttime = (double)(float)k * h->timestep;
        const double dd = ttime - (double)0.5f / h->nupeak;
        double v = std::cos(2.0 * kPi3 * h->nupeak * ttime) * std::exp(-(2.0 * (dd * dd) * (h->nuwidth * h->nuwidth)));
        if (std::fabs(v) < (double)0.00001f) v = 0.0;
        fw[k] = v * (double)1e-5f;
    }
    // radial form of the x profile (driver.f90:536-553)
    const int nx = h->nx, ny = h->ny;
    const double e8 = (double)1e-8f, k43 = (double)0.43f * kPi3;
    const double dx = (1.0 / (double)nx) * h->xlength * e8;
    // NaN ysrc: the reference's vr, the x profile in every y row (a line
    // source); otherwise the profile in the horizontal distance (a point source)
    const bool line = std::isnan(ysrc);
    std::vector<double> rd(h->nxy), vr(h->nxy);
    size_t imin = 0;
    int ixmin = 0;
    for (int j = 0; j < ny; j++) {
        const double yd = line ? 0.0 : ((double)j / (double)ny - 0.5) * h->ylength * e8 - ysrc;
        for (int i = 0; i < nx; i++) {
            const double xd = ((double)i / (double)nx - 0.5) * h->xlength * e8 - xsrc;
            const size_t q = (size_t)j * nx + i;
            rd[q] = line ? xd : std::sqrt(yd * yd + xd * xd);
            if (std::fabs(rd[q]) < std::fabs(rd[imin])) imin = q;
            if (j == 0 && std::fabs(xd) < std::fabs(rd[ixmin])) ixmin = i;
        }
    }
    for (size_t q = 0; q < h->nxy; q++) {
        const bool peak = line ? (int)(q % nx) == ixmin : q == imin;
        const double sinc = peak ? 1.0 : std::sin(k43 * rd[q]) / (k43 * rd[q]);
        vr[q] = std::exp(-(rd[q] * rd[q]) / (400.0 * (dx * dx))) * sinc;
    }
    h->src_mem.release();
    double *dvr = nullptr, *dfw = nullptr;
    HCHECK3(h, upload(h->src_mem, &dvr, vr, h->stream));
    HCHECK3(h, upload(h->src_mem, &dfw, fw, h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    h->src.mode = 1;
    h->src.row = h->e_rad;
    h->src.vr = dvr;
    h->src.nvr = 1;
    h->src.fwd = dfw;
    h->src.nfwd = nfwd;
    h->have_src = true;
    return SPARC_OK;
}

int sparc3d_set_adjoint_source(sparc3d_handle *h, const double *vr_dev, int nt, int maxtime)
{
    if (!h || !vr_dev || nt < 1 || maxtime < 0) return SPARC_E_ARG;
    if (!h->adjoint) return fail3(h, SPARC_E_STATE, "adjoint source on a forward handle");
    HCHECK3(h, hipStreamSynchronize(h->stream));
    set_times3(h, maxtime);
    const int fl = maxtime / (h->cfs > 0 ? h->cfs : 1) + 1;
    const int nvr = fl + 8;                                     // driver.f90:588-591
    h->src_mem.release();
    double *dvr = nullptr;
    HCHECK3(h, h->src_mem.alloc(&dvr, (size_t)nvr * h->nxy));
    HCHECK3(h, hipMemsetAsync(dvr, 0, (size_t)nvr * h->nxy * sizeof(double), h->stream));
    const int ncopy = nt < fl ? nt : fl;
    HCHECK3(h, hipMemcpyAsync(dvr, vr_dev, (size_t)ncopy * h->nxy * sizeof(double), hipMemcpyDeviceToDevice,
                              h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    h->src.mode = 2;
    h->src.row = h->o_rad;
    h->src.vr = dvr;
    h->src.nvr = nvr;
    h->src.fwd = nullptr;
    h->src.nfwd = 0;
    h->have_src = true;
    return SPARC_OK;
}

int sparc3d_reset(sparc3d_handle *h)
{
    if (!h) return SPARC_E_ARG;
    HCHECK3(h, hipMemsetAsync(h->a, 0, 6 * h->n * sizeof(double), h->stream));
    HCHECK3(h, hipMemsetAsync(h->scr, 0, 6 * h->n * sizeof(double), h->stream));
    HCHECK3(h, hipMemsetAsync(h->pml, 0, 2 * h->np * sizeof(double), h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_step(sparc3d_handle *h, int t)
{
    if (!h || t < 0) return SPARC_E_ARG;
    if (!h->have_src) return fail3(h, SPARC_E_STATE, "no source set");
    int r = enqueue_step3(h, t);
    if (r) return r;
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_run_range(sparc3d_handle *h, int t0, int t1, double *vz_out, int k0, int nslots)
{
    if (!h || t0 < 0 || t1 < t0 || (vz_out && nslots < 1)) return SPARC_E_ARG;
    if (!h->have_src) return fail3(h, SPARC_E_STATE, "no source set");
    const size_t plane = h->nxy;
    for (int t = t0; t <= t1; t++) {
        int r = enqueue_step3(h, t);
        if (r) return r;
        if (vz_out && t % h->steps == 0) {                      // driver.f90:388-419
            const int k = t / h->steps - k0;
            if (k >= 0 && k < nslots)
                HCHECK3(h, hipMemcpyAsync(vz_out + (size_t)k * plane, h->a + 5 * h->n + (size_t)h->o_rad * plane,
                                          plane * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        }
    }
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_run(sparc3d_handle *h, int maxtime, double *vz_out)
{
    if (!h || maxtime < 0) return SPARC_E_ARG;
    return sparc3d_run_range(h, 0, maxtime, vz_out, 0, maxtime / h->steps + 1);
}

int sparc3d_checkpoint_size(const sparc3d_handle *h, size_t *bytes)
{
    if (!h || !bytes) return SPARC_E_ARG;
    *bytes = (6 * h->n + 2 * h->np) * sizeof(double);
    return SPARC_OK;
}

int sparc3d_save_checkpoint(sparc3d_handle *h, void *dev)
{
    if (!h || !dev) return SPARC_E_ARG;
    double *d = (double *)dev;
    HCHECK3(h, hipMemcpyAsync(d, h->a, 6 * h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK3(h, hipMemcpyAsync(d + 6 * h->n, h->pml, 2 * h->np * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return SPARC_OK;
}

int sparc3d_load_checkpoint(sparc3d_handle *h, const void *dev)
{
    if (!h || !dev) return SPARC_E_ARG;
    const double *d = (const double *)dev;
    HCHECK3(h, hipMemcpyAsync(h->a, d, 6 * h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK3(h, hipMemcpyAsync(h->pml, d + 6 * h->n, 2 * h->np * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return SPARC_OK;
}

int sparc3d_get_snapshot(sparc3d_handle *h, double *dev)
{
    if (!h || !dev) return SPARC_E_ARG;
    const size_t N = h->n;
    if (!h->adjoint) {
        HCHECK3(h, hipMemcpyAsync(dev, h->a, 6 * N * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        HCHECK3(h, hipMemcpyAsync(dev + 6 * N, h->scr + 3 * N, 3 * N * sizeof(double), hipMemcpyDeviceToDevice,
                                  h->stream));
    } else {
        HCHECK3(h, hipMemcpyAsync(dev, h->a + 3 * N, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    }
    return SPARC_OK;
}

int sparc3d_kernels_begin(sparc3d_handle *h)
{
    if (!h) return SPARC_E_ARG;
    const size_t N = h->n, nxy = h->nxy;
    HCHECK3(h, hipStreamSynchronize(h->stream));
    if (!h->K) {
        HCHECK3(h, h->kern_mem.alloc(&h->K, 5 * N));
        HCHECK3(h, h->kern_mem.alloc(&h->KD, 15 * N));
        // kernel_back (initialize.f90:937-979): 10.0**6., 10.0**(-3.0) as REAL
        const double dimc = (double)1e6f, dimrho = (double)1e-3f;
        std::vector<double> r(h->nz), c2(N);
        for (int k = 0; k < h->nz; k++) {
            const double *row = &h->model[6 * (size_t)k];
            r[k] = row[2] / dimrho;
            for (size_t q = 0; q < nxy; q++) {
                const double c = (h->c3d.empty() ? row[1] : h->c3d[(size_t)k * nxy + q]) / dimc;
                c2[(size_t)k * nxy + q] = std::pow(c, 2.0);
            }
        }
        HCHECK3(h, upload(h->kern_mem, &h->krho0, r, h->stream));
        HCHECK3(h, upload(h->kern_mem, &h->kc2, c2, h->stream));
    }
    HCHECK3(h, hipMemsetAsync(h->K, 0, 5 * N * sizeof(double), h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    h->kern_pairs = 0;
    return SPARC_OK;
}

int sparc3d_kernels_add(sparc3d_handle *h, const double *fwd, const double *adj)
{
    if (!h || !fwd || !adj) return SPARC_E_ARG;
    if (!h->K) return fail3(h, SPARC_E_STATE, "sparc3d_kernels_add before sparc3d_kernels_begin");
    const size_t N = h->n;
    const int nyz = h->ny * h->nz, nz = h->nz;
    double *DX = h->KD, *DY = DX + 5 * N, *DZ = DY + 5 * N;
    int r;
    // ddxkern, ddykern (:665-691 and the y analogue) of forward xi_x / xi_y,
    // forward v and adjoint v_x / v_y
    if ((r = ddx3(h, fwd, DX, nyz))) return r;
    if ((r = ddx3(h, fwd + 3 * N, DX + N, 3 * nyz))) return r;
    if ((r = ddx3(h, adj, DX + 4 * N, nyz))) return r;
    if ((r = ddy3(h, fwd + N, DY, nz))) return r;
    if ((r = ddy3(h, fwd + 3 * N, DY + N, 3 * nz))) return r;
    if ((r = ddy3(h, adj + N, DY + 4 * N, nz))) return r;
    // ddzkern (:648-661): dbyd2 ibc 1, stretch, FILTER_VAR_Z
    const double *zin[5] = {fwd + 2 * N, fwd + 3 * N, fwd + 4 * N, fwd + 5 * N, adj + 2 * N};
    for (int i = 0; i < 5; i++)
        HCHECK3(h, launch_ddz_cols(zin[i], DZ + i * N, (int)h->nxy, nz, 1, 1, h->d_stretch, h->cf, h->tri[1], h->W,
                                   1, h->stream));
    const unsigned blocks = (unsigned)std::min<size_t>((N + 255) / 256, 16384);
    hipLaunchKernelGGL(k3_kern_accum, dim3(blocks), dim3(256), 0, h->stream, fwd, adj, DX, DY, DZ, h->K, N);
    HCHECK3(h, hipGetLastError());
    h->kern_pairs++;
    return SPARC_OK;
}

int sparc3d_kernels_end(sparc3d_handle *h, double *kc, double *kpsi, double *kvx, double *kvy, double *kvz,
                        double *hess)
{
    if (!h || !kc || !kpsi || !kvx || !kvy || !kvz || !hess) return SPARC_E_ARG;
    if (!h->K) return fail3(h, SPARC_E_STATE, "sparc3d_kernels_end before sparc3d_kernels_begin");
    const double diml = kRsun3, dimc = (double)1e6f, e8 = (double)1e-8f, e5 = (double)1e-5f;
    const double dimen = std::pow(diml * e8, -3.0) * std::pow(dimc * e5, -1.0);   // kernels.f90:503
    const double dimen_c = std::pow(diml * e8, -3.0);                             // :563
    const int stepskern = (int)std::floor(h->outputcad / h->timestep);
    const size_t N = h->n;
    // the psi kernel's curl (CURL_KERN :832-834, bcz = 0) into the derivative scratch
    double *T1 = h->KD, *T2 = h->KD + N;
    HCHECK3(h, launch_ddz_cols(h->K + N, T1, (int)h->nxy, h->nz, 1, 0, h->d_stretch, h->cf, h->tri[0], h->W, 1,
                               h->stream));
    int r;
    if ((r = ddx3(h, h->K + 3 * N, T2, h->ny * h->nz))) return r;
    const unsigned blocks = (unsigned)std::min<size_t>((N + 255) / 256, 16384);
    hipLaunchKernelGGL(k3_kern_final, dim3(blocks), dim3(256), 0, h->stream, h->K, T1, T2, h->krho0, h->kc2, N,
                       h->nxy, (double)stepskern, h->timestep, dimen, dimen_c, kc, kpsi, kvx, kvy, kvz, hess);
    HCHECK3(h, hipGetLastError());
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_get_state(sparc3d_handle *h, double *a_dev)
{
    if (!h || !a_dev) return SPARC_E_ARG;
    HCHECK3(h, hipMemcpyAsync(a_dev, h->a, 6 * h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_set_state(sparc3d_handle *h, const double *a_dev)
{
    if (!h || !a_dev) return SPARC_E_ARG;
    HCHECK3(h, hipMemcpyAsync(h->a, a_dev, 6 * h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_get_pml(sparc3d_handle *h, double *pml_dev)
{
    if (!h || !pml_dev) return SPARC_E_ARG;
    HCHECK3(h, hipMemcpyAsync(pml_dev, h->pml, 2 * h->np * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_get_scr(sparc3d_handle *h, double *scr_dev)
{
    if (!h || !scr_dev) return SPARC_E_ARG;
    HCHECK3(h, hipMemcpyAsync(scr_dev, h->scr, 6 * h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_ddx(sparc3d_handle *h, const double *in, double *out, int nplanes)
{
    if (!h || !in || !out || nplanes < 1) return SPARC_E_ARG;
    int r = ddx3(h, in, out, nplanes * h->ny);
    if (r) return r;
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_ddy(sparc3d_handle *h, const double *in, double *out, int nplanes)
{
    if (!h || !in || !out || nplanes < 1) return SPARC_E_ARG;
    int r = ddy3(h, in, out, nplanes);
    if (r) return r;
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_ddz(sparc3d_handle *h, const double *in, double *out, int nfields, int ibc)
{
    if (!h || !in || !out || nfields < 1 || ibc < 0 || ibc > 3) return SPARC_E_ARG;
    for (int f = 0; f < nfields; f++)   // the work buffer holds one field
        HCHECK3(h, launch_ddz_cols(in + f * h->n, out + f * h->n, (int)h->nxy, h->nz, 1, ibc, h->d_stretch, h->cf,
                                   h->tri[ibc], h->W, 0, h->stream));
    HCHECK3(h, hipStreamSynchronize(h->stream));
    return SPARC_OK;
}

int sparc3d_query_int(const sparc3d_handle *h, const char *name, int *v)
{
    if (!h || !name || !v) return SPARC_E_ARG;
    const std::string n = name;
    if (n == "nx") *v = h->nx;
    else if (n == "ny") *v = h->ny;
    else if (n == "nz") *v = h->nz;
    else if (n == "e_rad") *v = h->e_rad;
    else if (n == "o_rad") *v = h->o_rad;
    else if (n == "steps") *v = h->steps;
    else if (n == "npml") *v = h->nb + h->nt;
    else if (n == "maxtime") *v = h->maxtime;
    else return SPARC_E_ARG;
    return SPARC_OK;
}

int sparc3d_query_double(const sparc3d_handle *h, const char *name, double *v)
{
    if (!h || !name || !v) return SPARC_E_ARG;
    const std::string n = name;
    if (n == "deltat") *v = h->deltat;
    else if (n == "stretchx") *v = h->stretchx;
    else if (n == "stretchy") *v = h->stretchy;
    else return SPARC_E_ARG;
    return SPARC_OK;
}

int sparc3d_profile(sparc3d_handle *h, int reps, double *ms_out)
{
    if (!h || reps < 1 || !ms_out) return SPARC_E_ARG;
    if (!h->have_src) return fail3(h, SPARC_E_STATE, "no source set");
    int r0 = sparc3d_reset(h);
    if (r0) return r0;
    h->profiling = true;
    int r = SPARC_OK;
    for (int t = 0; t < reps && !r; t++) r = enqueue_step3(h, t);
    h->profiling = false;
    if (r) return r;
    HCHECK3(h, hipStreamSynchronize(h->stream));
    double sum[8] = {0}, cnt[8] = {0};
    for (auto &mk : h->marks) {
        float ms = 0.f;
        HCHECK3(h, hipEventElapsedTime(&ms, mk.a, mk.b));
        sum[mk.kind] += ms;
        cnt[mk.kind] += 1;
        (void)hipEventDestroy(mk.a);
        (void)hipEventDestroy(mk.b);
    }
    h->marks.clear();
    for (int k = 0; k < 8; k++) ms_out[k] = cnt[k] > 0 ? sum[k] / cnt[k] : 0.0;
    return SPARC_OK;
}

}  // extern "C"
