// sparc_discrete.hip -- exact discrete adjoint of the 2D option-17 step.
//
// The reference's adjoint (driver.f90:144-147, kernels.f90:286-355) is the
// continuous recipe: a time-reversed forward run with -v0 whose snapshots are
// correlated with the forward ones.  This file is its exact counterpart for
// gradients: the transpose of the discrete step of step.f90:43-119 (five
// low-storage RK stages of MP_QUIET_PML_2D_DISPL, physics2d.f90:943-1094, the
// z filter FILTER_VARS_2D, physics2d.f90:509-556, and the PML auxiliaries),
// run backwards over a recomputed forward trajectory, with the gradient of a
// functional of the vz traces with respect to c^2 rho0, v0x and v0z
// accumulated stage by stage, then mapped to the stream function and the
// sound speed.
//
// Building blocks (all batched over the S sources, fields [S][nz][nx]):
//   * x-derivatives: the library's FFT row kernel; its transpose is -Dx (the
//     multiplier i k is odd and the Nyquist bin is dropped);
//   * z-derivatives and their transposes: the dbyd2 map of each ibc as a
//     dense nz x nz matrix (built by the library's own compact-FD solve of an
//     identity block), applied by an LDS-tiled kernel (k_dz_dense);
//   * pointwise kernels for the PML, pressure, density and flow terms of the
//     forward stage (k_fwd_b, k_fwd_c) and of its transpose (k_adj_a,
//     k_adj_b, k_adj_c), the z filter and its transpose.
// The forward stages here are unfused and recomputed from the trajectory; they
// agree with the fused solver to rounding.
#include "sparc_device.h"
#include "sparc_launch.h"

namespace sparc {

namespace {

constexpr int kTX = 64, kTJ = 32, kTI = 64;

// out[b][j][x] = sum_i M[j][i] in[b][i][x] + vN[j] * preScale * pre[b * preStride + x]
// (the last term only when vN is not null).  Block: 64 columns x 32 output
// rows, the input tile and the matrix tile staged in LDS.
__global__ __launch_bounds__(256) void k_dz_dense(const double *__restrict__ in, double *__restrict__ out, int nz,
                                                  int nx, const double *__restrict__ M,
                                                  const double *__restrict__ vN, const double *__restrict__ pre,
                                                  size_t preStride, double preScale)
{
    __shared__ double Ms[kTJ][kTI + 1];
    __shared__ double Is[kTI][kTX];
    const int tx = threadIdx.x & (kTX - 1), rg = threadIdx.x >> 6;
    const int x = blockIdx.x * kTX + tx, j0 = blockIdx.y * kTJ, b = blockIdx.z;
    const double *inb = in + (size_t)b * nz * nx;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i0 = 0; i0 < nz; i0 += kTI) {
        for (int r = rg; r < kTI; r += 4) {
            const int i = i0 + r;
            Is[r][tx] = (i < nz && x < nx) ? inb[(size_t)i * nx + x] : 0.0;
        }
        for (int u = 0; u < kTJ * kTI / 256; u++) {
            const int e = threadIdx.x + 256 * u, jj = e / kTI, ii = e % kTI;
            const int j = j0 + jj, i = i0 + ii;
            Ms[jj][ii] = (j < nz && i < nz) ? M[(size_t)j * nz + i] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int ii = 0; ii < kTI; ii++) {
            const double v = Is[ii][tx];
#pragma unroll
            for (int r = 0; r < 8; r++) acc[r] += Ms[rg * 8 + r][ii] * v;
        }
        __syncthreads();
    }
    if (x >= nx) return;
    const double pv = vN ? preScale * pre[(size_t)b * preStride + x] : 0.0;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int j = j0 + rg * 8 + r;
        if (j < nz) out[((size_t)b * nz + j) * nx + x] = acc[r] + (vN ? vN[j] * pv : 0.0);
    }
}

// out[b][x] = sum_k w[k] in[b][k][x]
__global__ void k_colsum_w(const double *__restrict__ in, const double *__restrict__ w, double *__restrict__ out,
                           int nz, int nx, int nb)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb * nx) return;
    const int b = q / nx, x = q - b * nx;
    double s = 0.0;
    for (int k = 0; k < nz; k++) s += w[k] * in[((size_t)b * nz + k) * nx + x];
    out[q] = s;
}

// Geometry of a point q of an [S][nz][nx] field: source, row, column, its
// z-PML row (-1 outside the layers) and x-PML column (-1 outside).
struct Pt {
    int b, k, i, pj, ci;
};
__device__ __forceinline__ Pt point(size_t q, const Dims &d)
{
    Pt p;
    const size_t n2 = (size_t)d.nz * d.nx;
    p.b = (int)(q / n2);
    const int r = (int)(q - (size_t)p.b * n2);
    p.k = r / d.nx;
    p.i = r - p.k * d.nx;
    p.pj = p.k < d.npmlbot ? p.k : (p.k >= d.nz - d.npmltop ? d.npmlbot + (p.k - (d.nz - d.npmltop)) : -1);
    p.ci = p.i < d.nh ? p.i : (p.i >= d.nx - d.nh ? d.nh + (p.i - (d.nx - d.nh)) : -1);
    return p;
}

// Full-state views: a [4][S][nz][nx], pz [2][S][nzpml][nx], px [2][S][nz][nph].
struct Views {
    size_t SN, PZ1, PX1;   // one field of a, pz, px
};
__device__ __forceinline__ size_t pzi(const Dims &d, const Pt &p) { return ((size_t)p.b * d.nzpml + p.pj) * d.nx + p.i; }
__device__ __forceinline__ size_t pxi(const Dims &d, const Pt &p) { return ((size_t)p.b * d.nz + p.k) * d.nph + p.ci; }

// Forward stage, first pointwise part (physics2d.f90:957-998, :1043): z-PML of
// dxi_z/dz, x-PML of dxi_x/dx, div, p, rho'.  dzxi includes the kept last
// row (:954).  Writes DIV, P, RHO and the RHS of psivz, psivx.
__global__ void k_fwd_b(Dims d, Bg bg, Views v, const double *__restrict__ Y, const double *__restrict__ dxxi,
                        const double *__restrict__ dzxi, double *__restrict__ DIV, double *__restrict__ P,
                        double *__restrict__ RHO, double *__restrict__ R)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= v.SN) return;
    const Pt p = point(q, d);
    const size_t r2 = (size_t)p.k * d.nx + p.i;
    const double *ypz = Y + 4 * v.SN, *ypx = ypz + 2 * v.PZ1;
    double *rpz = R + 4 * v.SN, *rpx = rpz + 2 * v.PZ1;
    double dz = dzxi[q], dx = dxxi[q];
    if (p.pj >= 0) {
        const size_t e = pzi(d, p), c = (size_t)p.pj * d.nx + p.i;
        const double psi = ypz[e];
        rpz[e] = bg.az[c] * dz + bg.bz[c] * psi;
        dz = dz + psi;
    }
    if (p.ci >= 0) {
        const size_t e = pxi(d, p), c = (size_t)p.k * d.nph + p.ci;
        const double psi = ypx[e];
        dx = dx + psi;
        rpx[e] = bg.ax[c] * dx + bg.bx[c] * psi;
    }
    const double div = dx + dz, xiz = Y[v.SN + q];
    DIV[q] = div;
    P[q] = -(bg.c2rho0[r2] * div) - bg.gp0[r2] * xiz;
    RHO[q] = -(bg.rho0[r2] * div) - bg.gr0[r2] * xiz;
}

// Forward stage, second pointwise part (physics2d.f90:1003-1092): PMLs of the
// pressure gradient, RHS v_x, v_z, xi, the source of time t, the flow terms,
// the zeroed rows and columns.  DV = [dvx/dx, dvz/dx, dvx/dz, dvz/dz].
__global__ void k_fwd_c(Dims d, Bg bg, Views v, Src src, int t, const double *__restrict__ Y,
                        const double *__restrict__ gpx_in, const double *__restrict__ gpz_in,
                        const double *__restrict__ RHO, const double *__restrict__ DV, double *__restrict__ R)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= v.SN) return;
    const Pt p = point(q, d);
    const size_t r2 = (size_t)p.k * d.nx + p.i;
    const double *ypz = Y + 4 * v.SN, *ypx = ypz + 2 * v.PZ1;
    double *rpz = R + 4 * v.SN, *rpx = rpz + 2 * v.PZ1;
    double gpx = gpx_in[q], gpz = gpz_in[q];
    if (p.ci >= 0) {
        const size_t e = v.PX1 + pxi(d, p), c = (size_t)p.k * d.nph + p.ci;
        const double psi = ypx[e];
        gpx = gpx + psi;
        rpx[e] = bg.ax[c] * gpx + bg.bx[c] * psi;
    }
    if (p.pj >= 0) {
        const size_t e = v.PZ1 + pzi(d, p), c = (size_t)p.pj * d.nx + p.i;
        const double psi = ypz[e];
        rpz[e] = bg.az[c] * gpz + bg.bz[c] * psi;
        gpz = gpz + psi;
    }
    double rvx = -bg.rhoinv[r2] * gpx;
    double rvz = -bg.rhoinv[r2] * (gpz + RHO[q] * bg.g[p.k]);
    double rxx = Y[2 * v.SN + q], rxz = Y[3 * v.SN + q];
    if (t <= src.limit) {                                   // :1051-1061
        const double fs = (t >= 0 && t < src.nfwd) ? src.fwd[t] : 0.0;
        if (src.mode == 1 && p.k == src.row)
            rvz += src.vr[(size_t)p.b * d.nx + p.i] * fs * src.delta_width / bg.rho0_src[p.i];
        else if (src.mode == 3)
            rvz += src.vr[(size_t)p.b * d.nx + p.i] * fs * src.zprof[p.k];
    }
    rvx = rvx - 2.0 * bg.v0x[r2] * DV[q];                   // :1065-1079
    rvx = rvx - 2.0 * bg.v0z[r2] * DV[2 * v.SN + q];
    rvz = rvz - 2.0 * bg.v0x[r2] * DV[v.SN + q];
    rvz = rvz - 2.0 * bg.v0z[r2] * DV[3 * v.SN + q];
    const bool zero = p.k == 0 || p.i == 0 || p.i == d.nx - 1;  // :1083-1092
    R[q] = zero ? 0.0 : rxx;
    R[v.SN + q] = zero ? 0.0 : rxz;
    R[2 * v.SN + q] = zero ? 0.0 : rvx;
    R[3 * v.SN + q] = zero ? 0.0 : rvz;
}

// out = U + beta R over the whole state (n doubles); beta 0 copies.
__global__ void k_axpy(const double *__restrict__ U, const double *__restrict__ R, double beta,
                       double *__restrict__ out, size_t n)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q < n) out[q] = U[q] + beta * R[q];
}

// acc += Yb, Rb = beta Yb
__global__ void k_accum(double *__restrict__ acc, double *__restrict__ Rb, const double *__restrict__ Yb,
                        double beta, size_t n)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double y = Yb[q];
    acc[q] += y;
    Rb[q] = beta * y;
}

// One of the filtered rows of FILTER_VARS_2D (physics2d.f90:509-556):
// rows 5..nz-6 (0-based) take the 11-point filter, rows nz-5..nz-2 the
// two-point mean, for columns nh-1 .. nx-nh.
__device__ __constant__ double kFC[6] = {0.75390625, 0.41015625, -0.234375, 0.087890625, -0.01953125,
                                         0.001953125};

// a_new = filter(U + dt R) on the 4 fields, pz/px += dt R (the last stage's
// update, step.f90:107-112); A = out state.
__global__ void k_update_filter(Dims d, Views v, const double *__restrict__ U, const double *__restrict__ R,
                                double dt, double *__restrict__ A)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t n4 = 4 * v.SN, nall = n4 + 2 * v.PZ1 + 2 * v.PX1;
    if (q >= nall) return;
    if (q >= n4) { A[q] = U[q] + dt * R[q]; return; }
    const size_t f = q / v.SN, qq = q - f * v.SN;
    const Pt p = point(qq, d);
    const size_t base = f * v.SN + ((size_t)p.b * d.nz) * d.nx + p.i;
    auto u = [&](int k) { const size_t e = base + (size_t)k * d.nx; return U[e] + dt * R[e]; };
    double val = u(p.k);
    if (p.i >= d.nh - 1 && p.i <= d.nx - d.nh && p.k >= 5 && p.k <= d.nz - 2) {
        if (p.k <= d.nz - 6) {
            val = kFC[0] * u(p.k);
            for (int m = 1; m <= 5; m++) val += 0.5 * kFC[m] * (u(p.k - m) + u(p.k + m));
        } else {
            val = 0.5 * (u(p.k + 1) + u(p.k - 1));
        }
    }
    A[q] = val;
}

// Transpose of the filter on the 4 fields (pz, px copied): out = F^T lam.
__global__ void k_filter_T(Dims d, Views v, const double *__restrict__ lam, double *__restrict__ out)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t n4 = 4 * v.SN, nall = n4 + 2 * v.PZ1 + 2 * v.PX1;
    if (q >= nall) return;
    if (q >= n4) { out[q] = lam[q]; return; }
    const size_t f = q / v.SN, qq = q - f * v.SN;
    const Pt p = point(qq, d);
    const size_t base = f * v.SN + ((size_t)p.b * d.nz) * d.nx + p.i;
    const int j = p.k, nz = d.nz;
    const bool xin = p.i >= d.nh - 1 && p.i <= d.nx - d.nh;
    if (!xin) { out[q] = lam[q]; return; }
    double val = (j >= 5 && j <= nz - 2) ? 0.0 : lam[q];    // rows the filter replaces
    for (int k = max(5, j - 5); k <= min(nz - 6, j + 5); k++) {
        const int m = abs(k - j);
        val += (m == 0 ? kFC[0] : 0.5 * kFC[m]) * lam[base + (size_t)k * d.nx];
    }
    for (int k = max(nz - 5, j - 1); k <= min(nz - 2, j + 1); k++)
        if (abs(k - j) == 1) val += 0.5 * lam[base + (size_t)k * d.nx];
    out[q] = val;
}

// Transpose of k_fwd_c's RHS (and of the flow terms' inputs):
// FX = [-2 v0x Rvx, -2 v0x Rvz, adjoint of dp/dx], FZ = [-2 v0z Rvx,
// -2 v0z Rvz], GZ = adjoint of the raw dp/dz, RHOB = adjoint of rho';
// writes the adjoints of psi_gradp_x and psip into Yb.
__global__ void k_adj_a(Dims d, Bg bg, Views v, const double *__restrict__ Rb, double *__restrict__ FX,
                        double *__restrict__ FZ, double *__restrict__ GZ, double *__restrict__ RHOB,
                        double *__restrict__ Yb)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= v.SN) return;
    const Pt p = point(q, d);
    const size_t r2 = (size_t)p.k * d.nx + p.i;
    const bool zero = p.k == 0 || p.i == 0 || p.i == d.nx - 1;
    const double rvx = zero ? 0.0 : Rb[2 * v.SN + q], rvz = zero ? 0.0 : Rb[3 * v.SN + q];
    const double *rpz = Rb + 4 * v.SN, *rpx = rpz + 2 * v.PZ1;
    double *ypz = Yb + 4 * v.SN, *ypx = ypz + 2 * v.PZ1;
    FX[q] = -2.0 * bg.v0x[r2] * rvx;
    FX[v.SN + q] = -2.0 * bg.v0x[r2] * rvz;
    FZ[q] = -2.0 * bg.v0z[r2] * rvx;
    FZ[v.SN + q] = -2.0 * bg.v0z[r2] * rvz;
    double gpx = -bg.rhoinv[r2] * rvx;
    if (p.ci >= 0) {
        const size_t e = v.PX1 + pxi(d, p), c = (size_t)p.k * d.nph + p.ci;
        const double tot = gpx + bg.ax[c] * rpx[e];
        ypx[e] = tot + bg.bx[c] * rpx[e];
        gpx = tot;
    }
    FX[2 * v.SN + q] = gpx;
    double gpz = -bg.rhoinv[r2] * rvz;
    RHOB[q] = -bg.rhoinv[r2] * bg.g[p.k] * rvz;
    if (p.pj >= 0) {
        const size_t e = v.PZ1 + pzi(d, p), c = (size_t)p.pj * d.nx + p.i;
        ypz[e] = gpz + bg.bz[c] * rpz[e];
        gpz = gpz + bg.az[c] * rpz[e];
    }
    GZ[q] = gpz;
}

// Second transposed part: v adjoints, p adjoint, div adjoint through both
// PMLs of the displacement derivatives, the gradient contributions of this
// stage: G = [d/d(c^2 rho0), d/dv0x, d/dv0z] += pairs with the stage's
// forward quantities D = [div, dvx/dx, dvz/dx, dvx/dz, dvz/dz].
__global__ void k_adj_b(Dims d, Bg bg, Views v, const double *__restrict__ Rb, const double *__restrict__ DFX,
                        const double *__restrict__ DFZ, const double *__restrict__ DGZ,
                        const double *__restrict__ RHOB, const double *__restrict__ D, double *__restrict__ Yb,
                        double *__restrict__ XZP, double *__restrict__ DXB, double *__restrict__ DZB,
                        double *__restrict__ G)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= v.SN) return;
    const Pt p = point(q, d);
    const size_t r2 = (size_t)p.k * d.nx + p.i;
    const bool zero = p.k == 0 || p.i == 0 || p.i == d.nx - 1;
    const double rxx = zero ? 0.0 : Rb[q], rxz = zero ? 0.0 : Rb[v.SN + q];
    const double rvx = zero ? 0.0 : Rb[2 * v.SN + q], rvz = zero ? 0.0 : Rb[3 * v.SN + q];
    const double *rpz = Rb + 4 * v.SN, *rpx = rpz + 2 * v.PZ1;
    double *ypz = Yb + 4 * v.SN, *ypx = ypz + 2 * v.PZ1;
    Yb[2 * v.SN + q] = -DFX[q] + DFZ[q] + rxx;                 // dx^T = -dx
    Yb[3 * v.SN + q] = -DFX[v.SN + q] + DFZ[v.SN + q] + rxz;
    const double pb = -DFX[2 * v.SN + q] + DGZ[q];
    const double rb = RHOB[q];
    const double divb = -bg.rho0[r2] * rb - bg.c2rho0[r2] * pb;
    XZP[q] = -bg.gr0[r2] * rb - bg.gp0[r2] * pb;
    double dxb = divb, dzb = divb;
    if (p.ci >= 0) {
        const size_t e = pxi(d, p), c = (size_t)p.k * d.nph + p.ci;
        const double tot = divb + bg.ax[c] * rpx[e];
        ypx[e] = tot + bg.bx[c] * rpx[e];
        dxb = tot;
    }
    if (p.pj >= 0) {
        const size_t e = pzi(d, p), c = (size_t)p.pj * d.nx + p.i;
        ypz[e] = divb + bg.bz[c] * rpz[e];
        dzb = divb + bg.az[c] * rpz[e];
    }
    DXB[q] = dxb;
    DZB[q] = dzb;
    G[q] += -pb * D[q];
    G[v.SN + q] += -2.0 * (rvx * D[v.SN + q] + rvz * D[2 * v.SN + q]);
    G[2 * v.SN + q] += -2.0 * (rvx * D[3 * v.SN + q] + rvz * D[4 * v.SN + q]);
}

// The kept last row of dxi_z/dz is -dxi_x/dx unstretch (:954): its adjoint
// goes back into the last row of the dxi_x/dx adjoint.
__global__ void k_adj_preset(const double *__restrict__ pb, double *__restrict__ DXB, int nz, int nx, int nb,
                             double unstretch_last)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb * nx) return;
    const int b = q / nx, x = q - b * nx;
    DXB[((size_t)b * nz + nz - 1) * nx + x] += -unstretch_last * pb[q];
}

// Yb xi_x = -dx(DXB), Yb xi_z = XZP + Dz3^T DZB
__global__ void k_adj_c(Views v, const double *__restrict__ DDXB, const double *__restrict__ XZP,
                        const double *__restrict__ DZ3B, double *__restrict__ Yb)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= v.SN) return;
    Yb[q] = -DDXB[q];
    Yb[v.SN + q] = XZP[q] + DZ3B[q];
}

// lam[vz][b][o_rad][x] += gvz[b][k][x]
__global__ void k_inject(double *__restrict__ lam, const double *__restrict__ gvz, int k, int nt, int orow,
                         int nz, int nx, int nb, size_t SN)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb * nx) return;
    const int b = q / nx, x = q - b * nx;
    lam[3 * SN + ((size_t)b * nz + orow) * nx + x] += gvz[((size_t)b * nt + k) * nx + x];
}

// gradient of the flow map (driver.f90:94-126): inputs Gx, Gz (d/dv0x,
// d/dv0z) -> WX = Gx / rho0, WZ = Gz / rho0 (for Dz1^T and dx^T)
__global__ void k_flow_in(Views v, int n2, const double *__restrict__ rho0, const double *__restrict__ G,
                          double *__restrict__ WX, double *__restrict__ WZ)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= v.SN) return;
    const double r = rho0[q % n2];
    WX[q] = G[v.SN + q] / r;
    WZ[q] = G[2 * v.SN + q] / r;
}

// d pv = -Dz1^T WX - dx WZ; gpsi = d pv rho0 c / (Rsun/1e8); gc = 2 c rho0 G2
// / dimc + d pv rho0 psi / (Rsun/1e8) / dimc, with c = sqrt(c2) in code units
__global__ void k_flow_out(Views v, int n2, const double *__restrict__ rho0, const double *__restrict__ c2,
                           const double *__restrict__ psi, double scale, double dimc,
                           const double *__restrict__ G, const double *__restrict__ DZWX,
                           const double *__restrict__ DXWZ, double *__restrict__ gpsi, double *__restrict__ gc)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= v.SN) return;
    const size_t r2 = q % n2;
    const double gpv = -DZWX[q] - DXWZ[q];
    const double c = sqrt(c2[r2]);
    if (gpsi) gpsi[q] = gpv * rho0[r2] * c / scale;
    if (gc) gc[q] = 2.0 * c * rho0[r2] * G[q] / dimc + (psi ? gpv * rho0[r2] * psi[r2] / scale / dimc : 0.0);
}

inline dim3 grid1(size_t n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

}  // namespace

#define DCHK(expr)                                   \
    do {                                             \
        hipError_t e_ = (expr);                      \
        if (e_ != hipSuccess) return e_;             \
    } while (0)

static hipError_t dz_dense(const DiscreteCtx &c, const double *in, double *out, int nbatch, const double *M,
                           const double *vN, const double *pre, size_t preStride, double preScale)
{
    dim3 grid((c.d.nx + kTX - 1) / kTX, (c.d.nz + kTJ - 1) / kTJ, nbatch);
    hipLaunchKernelGGL(k_dz_dense, grid, dim3(256), 0, c.stream, in, out, c.d.nz, c.d.nx, M, vN, pre, preStride,
                       preScale);
    return hipGetLastError();
}

static hipError_t ddx(const DiscreteCtx &c, const double *in, double *out, int nrows)
{
    return launch_ddx_rows(in, out, nrows, c.d.nx, c.bg.tw, c.bg.dmul, c.bg.cdx, c.bg.stretchx, c.stream);
}

// One forward RHS evaluation of the stage state Y at time t: R (full state)
// and the stage quantities D = [div, dvx/dx, dvz/dx, dvx/dz, dvz/dz].
static hipError_t fwd_rhs(const DiscreteCtx &c, const DiscreteWork &w, const double *Y, int t, double *R,
                          double *D)
{
    const Dims &d = c.d;
    const Views v{w.SN, w.PZ1, w.PX1};
    const int S = d.S;
    // dxi_x/dx, dv_x/dx, dv_z/dx
    DCHK(ddx(c, Y, w.T0, S * d.nz));
    DCHK(ddx(c, Y + 2 * w.SN, D + w.SN, 2 * S * d.nz));
    // dxi_z/dz (ibc 3, kept last row -dxi_x/dx unstretch), dv_x/dz, dv_z/dz (ibc 1)
    DCHK(dz_dense(c, Y + w.SN, w.T1, S, c.M[2], c.vN3, w.T0 + (size_t)(d.nz - 1) * d.nx, (size_t)d.nz * d.nx,
                  -c.unstretch_last));
    DCHK(dz_dense(c, Y + 2 * w.SN, D + 3 * w.SN, 2 * S, c.M[0], nullptr, nullptr, 0, 0.0));
    hipLaunchKernelGGL(k_fwd_b, grid1(w.SN), dim3(256), 0, c.stream, d, c.bg, v, Y, w.T0, w.T1, D, w.T2, w.T3, R);
    DCHK(hipGetLastError());
    // dp/dx, dp/dz (ibc 2, its first row 0)
    DCHK(ddx(c, w.T2, w.T0, S * d.nz));
    DCHK(dz_dense(c, w.T2, w.T1, S, c.M[1], nullptr, nullptr, 0, 0.0));
    hipLaunchKernelGGL(k_fwd_c, grid1(w.SN), dim3(256), 0, c.stream, d, c.bg, v, c.src, t, Y, w.T0, w.T1, w.T3,
                       D + w.SN, R);
    return hipGetLastError();
}

// One forward step U -> Unew (Unew != U), keeping the stage quantities in
// Dst[5] when not null.
static hipError_t fwd_step(const DiscreteCtx &c, const DiscreteWork &w, const double *U, int t, double *Unew,
                           double *Dst)
{
    for (int s = 0; s < 5; s++) {
        hipLaunchKernelGGL(k_axpy, grid1(w.NST), dim3(256), 0, c.stream, U, w.R, s == 0 ? 0.0 : c.betas[s], w.Y,
                           w.NST);
        DCHK(hipGetLastError());
        DCHK(fwd_rhs(c, w, w.Y, t, w.R, Dst ? Dst + (size_t)s * 5 * w.SN : w.Dscr));
    }
    const Views v{w.SN, w.PZ1, w.PX1};
    hipLaunchKernelGGL(k_update_filter, grid1(w.NST), dim3(256), 0, c.stream, c.d, v, U, w.R, c.deltat, Unew);
    return hipGetLastError();
}

// Transpose of one step at time t: lam (adjoint of the state after the step)
// -> adjoint of U_t, accumulating the gradient into G.  Dkept: the step's
// stage quantities [5][5][SN] kept by the segment recompute, or null to
// recompute them from U here.
static hipError_t bwd_step(const DiscreteCtx &c, const DiscreteWork &w, const double *U, int t, double *lam,
                           double *G, const double *Dkept)
{
    const Dims &d = c.d;
    const Views v{w.SN, w.PZ1, w.PX1};
    const int S = d.S;
    const double *Dst = Dkept;
    if (!Dst) {
        DCHK(fwd_step(c, w, U, t, w.Unext, w.Dst));          // the stage quantities of this step
        Dst = w.Dst;
    }
    hipLaunchKernelGGL(k_filter_T, grid1(w.NST), dim3(256), 0, c.stream, d, v, lam, w.acc);
    DCHK(hipGetLastError());
    hipLaunchKernelGGL(k_axpy, grid1(w.NST), dim3(256), 0, c.stream, w.acc, w.acc, c.deltat - 1.0, w.Rb, w.NST);
    DCHK(hipGetLastError());
    for (int s = 4; s >= 0; s--) {
        hipLaunchKernelGGL(k_adj_a, grid1(w.SN), dim3(256), 0, c.stream, d, c.bg, v, w.Rb, w.FX, w.FZ, w.GZ, w.RHOB,
                           w.Yb);
        DCHK(hipGetLastError());
        DCHK(ddx(c, w.FX, w.DFX, 3 * S * d.nz));
        DCHK(dz_dense(c, w.FZ, w.DFZ, 2 * S, c.MT[0], nullptr, nullptr, 0, 0.0));
        DCHK(dz_dense(c, w.GZ, w.DGZ, S, c.MT[1], nullptr, nullptr, 0, 0.0));
        hipLaunchKernelGGL(k_adj_b, grid1(w.SN), dim3(256), 0, c.stream, d, c.bg, v, w.Rb, w.DFX, w.DFZ, w.DGZ,
                           w.RHOB, Dst + (size_t)s * 5 * w.SN, w.Yb, w.XZP, w.DXB, w.DZB, G);
        DCHK(hipGetLastError());
        hipLaunchKernelGGL(k_colsum_w, grid1((size_t)S * d.nx), dim3(256), 0, c.stream, w.DZB, c.vN3, w.PB, d.nz,
                           d.nx, S);
        DCHK(hipGetLastError());
        DCHK(dz_dense(c, w.DZB, w.DZ3B, S, c.MT[2], nullptr, nullptr, 0, 0.0));
        hipLaunchKernelGGL(k_adj_preset, grid1((size_t)S * d.nx), dim3(256), 0, c.stream, w.PB, w.DXB, d.nz, d.nx,
                           S, c.unstretch_last);
        DCHK(hipGetLastError());
        DCHK(ddx(c, w.DXB, w.DDXB, S * d.nz));
        hipLaunchKernelGGL(k_adj_c, grid1(w.SN), dim3(256), 0, c.stream, v, w.DDXB, w.XZP, w.DZ3B, w.Yb);
        DCHK(hipGetLastError());
        hipLaunchKernelGGL(k_accum, grid1(w.NST), dim3(256), 0, c.stream, w.acc, w.Rb, w.Yb,
                           s > 0 ? c.betas[s] : 0.0, w.NST);
        DCHK(hipGetLastError());
    }
    return hipMemcpyAsync(lam, w.acc, w.NST * sizeof(double), hipMemcpyDeviceToDevice, c.stream);
}

hipError_t discrete_gradient(const DiscreteCtx &c, const DiscreteWork &w, int maxtime, int interval,
                             const double *gvz, int nt, double *G)
{
    const Dims &d = c.d;
    const int S = d.S;
    const int C = interval;
    const int nseg = (maxtime + C) / C;                       // segments start at 0, C, 2C, ...
    // forward pass from the zero state, keeping the state at every segment start
    DCHK(launch_fill(w.U, w.NST, 0.0, c.stream));
    for (int t = 0; t <= maxtime; t++) {
        if (t % C == 0)
            DCHK(hipMemcpyAsync(w.ck + (size_t)(t / C) * w.NST, w.U, w.NST * sizeof(double),
                                hipMemcpyDeviceToDevice, c.stream));
        DCHK(fwd_step(c, w, w.U, t, w.Unext, nullptr));
        DCHK(hipMemcpyAsync(w.U, w.Unext, w.NST * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    }
    DCHK(launch_fill(w.lam, w.NST, 0.0, c.stream));
    DCHK(launch_fill(G, 3 * w.SN, 0.0, c.stream));
    for (int j = nseg - 1; j >= 0; j--) {
        const int t0 = j * C, t1 = std::min(maxtime, t0 + C - 1);
        // states U_t of the segment
        DCHK(hipMemcpyAsync(w.seg, w.ck + (size_t)j * w.NST, w.NST * sizeof(double), hipMemcpyDeviceToDevice,
                            c.stream));
        // states U_t0..U_t1 and, when w.Dseg holds a segment, the stage
        // quantities of steps t0..t1 (the state after step t1 is not needed)
        const size_t D1 = (size_t)25 * w.SN;
        for (int t = t0; t < t1 + (w.Dseg ? 1 : 0); t++)
            DCHK(fwd_step(c, w, w.seg + (size_t)(t - t0) * w.NST, t,
                          t < t1 ? w.seg + (size_t)(t - t0 + 1) * w.NST : w.Unext,
                          w.Dseg ? w.Dseg + (size_t)(t - t0) * D1 : nullptr));
        for (int t = t1; t >= t0; t--) {
            if (t % c.steps == 0) {                            // output k after step t
                hipLaunchKernelGGL(k_inject, grid1((size_t)S * d.nx), dim3(256), 0, c.stream, w.lam, gvz,
                                   t / c.steps, nt, c.orow, d.nz, d.nx, S, w.SN);
                DCHK(hipGetLastError());
            }
            DCHK(bwd_step(c, w, w.seg + (size_t)(t - t0) * w.NST, t, w.lam, G,
                          w.Dseg ? w.Dseg + (size_t)(t - t0) * D1 : nullptr));
        }
    }
    return hipSuccess;
}

hipError_t discrete_flow_chain(const DiscreteCtx &c, const DiscreteWork &w, const double *G, const double *rho0,
                               const double *c2, const double *psi, double scale, double dimc, double *gpsi,
                               double *gc)
{
    const Dims &d = c.d;
    const Views v{w.SN, w.PZ1, w.PX1};
    const int n2 = d.nz * d.nx;
    hipLaunchKernelGGL(k_flow_in, grid1(w.SN), dim3(256), 0, c.stream, v, n2, rho0, G, w.FX, w.FZ);
    DCHK(hipGetLastError());
    DCHK(dz_dense(c, w.FX, w.DFX, d.S, c.MT[0], nullptr, nullptr, 0, 0.0));   // Dz1^T (Gx / rho0)
    DCHK(ddx(c, w.FZ, w.DFZ, d.S * d.nz));                                   // dx (Gz / rho0)
    hipLaunchKernelGGL(k_flow_out, grid1(w.SN), dim3(256), 0, c.stream, v, n2, rho0, c2, psi, scale, dimc, G,
                       w.DFX, w.DFZ, gpsi, gc);
    return hipGetLastError();
}

hipError_t discrete_apply(const DiscreteCtx &c, const double *in, double *out, int nbatch, int which,
                          double *presetN_bar)
{
    // which: 0..2 Dz of ibc 1..3, 3..5 their transposes
    if (which < 0 || which > 5) return hipErrorInvalidValue;
    const double *M = which < 3 ? c.M[which] : c.MT[which - 3];
    DCHK(dz_dense(c, in, out, nbatch, M, nullptr, nullptr, 0, 0.0));
    if (which == 5 && presetN_bar) {
        hipLaunchKernelGGL(k_colsum_w, grid1((size_t)nbatch * c.d.nx), dim3(256), 0, c.stream, in, c.vN3,
                           presetN_bar, c.d.nz, c.d.nx, nbatch);
        DCHK(hipGetLastError());
    }
    return hipSuccess;
}

}  // namespace sparc
