// sparc_device.h -- device-side data structures and building blocks shared by
// the solver kernels: the register/LDS FFT used by the x-derivative and the
// compact-FD (dbyd2) coefficient set used by the z-derivative.
//
// Layout in HBM: every field of every source is an [S][nz][nx] array of
// doubles, x fastest (the reference's column-major a(nx,1,nz) per source).
#pragma once
#include <hip/hip_runtime.h>

namespace sparc {

// ---------------------------------------------------------------------------
// Problem geometry.
struct Dims {
    int nx, nz, S;            // grid and number of sources in the batch
    int npmlbot, npmltop, nzpml;
    int nh, nph;              // horizontal PML width and 2*nh
};

// Background arrays shared by every source of a handle (device pointers).
struct Bg {
    const double *rhoinv, *rho0, *c2rho0, *gp0, *gr0, *v0x, *v0z;  // [nz][nx]
    const double *rhoinv_z;                                        // [nz]: 1/rho0 is x-independent
    const double *g, *stretch, *unstretch;                 // [nz]
    const double *az, *bz;                                 // [nzpml][nx]
    const double *fz;                                      // [nzpml][nx] PML factor of d p0/dz, d rho0/dz
    const double *ax, *bx;                                 // [nz][nph]
    const double *rho0_src;                                // [nx] rho0 at the injection row
    const double *eyekx;                                   // [nx/2+1]
    const double *dmul;                                    // [nx] bin multiplier (see fft_deriv_team)
    const double2 *tw;                                     // [nx] exp(-2 pi i k/nx)
    const double *cdx;                                     // [2 nx] circulant derivative column, twice (circ_deriv_row)
    double stretchx;
};

// The dbyd2 right-hand-side coefficients: dbyd2.f keeps them as default
// REAL (float32); they are stored here already widened to double.
struct DbyCoef {
    double c1, c2;                 // interior 6th order (dbyd2.f:85-86)
    double b1, b2, b3, b4, b5;     // rows 2 and N2-1, 5th order (dbyd2.f:103-107)
    double f1, f2, f3, f4;         // rows 1 and N2 (dbyd2.f:121-124)
    double alpha2, beta1, cnst;    // preconditioning (dbyd2.f:156-162)
};

// Thomas factorization of the constant dbyd2 matrix, one set per ibc.
// lowr[j]: sub-diagonal, gam[j] = uppr[j-1]/betad[j-1], rbet[j] = 1/betad[j].
struct Tri {
    const double *lowr, *gam, *rbet;
};

// Segmented (SPIKE) form of the same matrices for the column pass: the column
// is cut into P <= kSegMax segments of exactly m rows (rows >= nz of the last
// segment are identity padding).  Each segment is solved locally (rbl, gl:
// Thomas factors restarted at the segment start) and corrected with the spikes
// V (response to the row above the segment) and W (to the row below), weighted
// by the interface values from the exact inverse Minv of the 2P x 2P
// interface system.  ztab rows (each nzp = P*m long): 0 stretch, 1 g, then
// for slot 0 (ibc 1), 1 (ibc 2), 2 (ibc 3): lowr, rbl, gl, V, W (rows 2-16),
// then the x-independent profiles rho0, d rho0/dz, d p0/dz, 1/rho0 (17-20).
constexpr int kSegMax = 64;
constexpr int kZT = 21;     // then the kCB boundary-row coefficients (sparc_seg.h)
// The column pass keeps in LDS the kZS segment-table rows of the three dbyd2
// sets (ztab rows 2-16) for three segments only, [kZS][3][m]: the first, an
// interior one and the last.  Every interior segment's tables are the same
// (the coefficients are constant away from the boundary rows and each
// segment's factorization restarts at its start; the host checks it).  Then
// the kZF epilogue profiles stretch, g, rho0, d rho0/dz, d p0/dz, 1/rho0,
// [kZF][nzp] (ztab rows stage_prow(r)).
constexpr int kZS = 15, kZF = 6;
__host__ __device__ inline int stage_prow(int r) { return r == 0 ? 0 : (r == 1 ? 1 : 15 + r); }
struct SegTab {
    int P, m, hw, nzp;              // segments, rows per segment, Minv window, P*m
    const double *ztab;             // [kZT][nzp], then [kCB]
    const double *Minv;             // [3][2P][4hw+4]: row k of each inverse, entries 2(k/2-hw)+w
    const double *stage;            // the column pass's LDS image: Minv, [kZS][3][m] segment tables, [kZF][nzp] profiles, [kCB]
};

// Prognostic state and work arrays of a batch (device pointers).
struct State {
    double *a[4];           // xi_x, xi_z, v_x, v_z   [S][nz][nx]
    double *t[4];           // RK stage state
    double *vx_unf;         // v_x after the final update, before the z filter
    double *pz_base, *pz_tmp;  // [2][S][nzpml][nx]  (psivz, psip)
    double *px_base, *px_tmp;  // [2][S][nz][nph]    (psivx, psi_gradp_x)
    double *dxixdx, *dxvx, *dxvz, *p, *rvx;  // [S][nz][nx]
    double *topbc;          // [S][nx]
    double *w_rho;          // partial RHS v_z between column-pass phases [S][nz][nx]
    double *discard;        // >= 256 doubles: target of stores from padding rows
};

// Source term of the RHS (physics2d.f90:1051-1061).
struct Src {
    int mode;               // 0 none, 1 forward, 2 adjoint, 3 forward with a z profile
    int row;                // e_rad (forward) or o_rad (adjoint), 0-based; -1 in mode 3
    int limit;              // forcing_length * cadforcing_step
    double delta_width;
    const double *vr;       // forward: [S][nx]; adjoint: [S][nvr][nx]
    int nvr;
    const double *fwd;      // forward time series [nfwd]
    int nfwd;
    double timestep, cadforcing;
    int num_steps;
    // mode 3: RHS v_z(x, z) += vr[s][x] * fwd[t] * zprof[z] on rows zlo..zhi
    const double *zprof;    // [nz], zero outside zlo..zhi
    int zlo, zhi;
};

// Mode-3 source amplitude vr * fwd(t) of one column, for a segment of rows
// a .. a+ms-1; 0 when the segment misses the profile's rows or the source
// has ended.
__device__ __forceinline__ double profile_amp(const Src &src, int b, int x, int nx, int a, int ms,
                                              int time)
{
    if (src.mode != 3 || a > src.zhi || a + ms - 1 < src.zlo || time > src.limit) return 0.0;
    const double fs = (time >= 0 && time < src.nfwd) ? src.fwd[time] : 0.0;
    return src.vr[(size_t)b * nx + x] * fs;
}

// Per-launch time information. t = *tctr + toff.
struct Clock {
    const int *tctr;
    int toff;
    int steps;              // output cadence in steps
};

// Snapshot / output sink of a run (device pointers, may be null).
struct Sink {
    double *vz_out;         // [S][nout][nx]
    int nout;
    double *snap;           // [S][nsnap][nfield][nz][nx]
    int nsnap, nfield;      // forward: 6 (xix, xiz, vx, vz, accx, accz); adjoint: 4 (vx, vz, accx, accz)
    int orow;               // o_rad
    int adjoint;
    int k0;                 // output index held by slot 0 of vz_out / snap
    double *norm2;          // [S][nnorm] (may be null): sum of v_z^2 over the grid after step t,
    int nnorm, tnorm0;      //   in slot t - tnorm0 (TIMESTEPPING's nor1 squared, driver.f90:364)
};

// Slot of time index t in an output buffer of n slots starting at output
// index k0 (outputs are the multiples of steps), or -1.
__device__ __forceinline__ int out_slot(int t, int steps, int k0, int n)
{
    if (t < 0 || t % steps) return -1;
    const int k = t / steps - k0;
    return (k >= 0 && k < n) ? k : -1;
}

// Arguments of one row pass (k_row).
struct RowArgs {
    const double *xi_x;     // stage s+1 xi_x (MODE 0: a; MODE 1: temp)
    const double *v_z;      // stage s+1 v_z
    double *v_x_out;        // MODE 1: temp v_x of stage s+1; MODE 2: vx_unf
    double bx2;             // beta_{s+1} (or deltat when last2) for the dx p part
    int last2;
    double bx1;             // beta_{s+2} (or deltat when last1) for the x-derivative part
    int last1;
    int filt_vx;            // MODE 0: apply the pending v_x filter
};

// Arguments of one column pass (k_col).
struct ColArgs {
    const double *T[4];     // stage state (xi_x unused)
    double *TN[4];          // next stage state (last stage: a_new, unfiltered)
    double bnext;           // beta_{s+1}, or deltat at the last stage
    int last;
};

// ---------------------------------------------------------------------------
// Twiddles of the 32-point circle, exp(2 pi i k/32) = kC32[k] + i kS32[k].
__device__ constexpr double kC32[32] = {
    1.0, 0.98078528040323043, 0.92387953251128674, 0.83146961230254524,
    0.70710678118654757, 0.55557023301960229, 0.38268343236508984, 0.19509032201612833,
    0.0, -0.19509032201612833, -0.38268343236508984, -0.55557023301960229,
    -0.70710678118654757, -0.83146961230254524, -0.92387953251128674, -0.98078528040323043,
    -1.0, -0.98078528040323043, -0.92387953251128674, -0.83146961230254524,
    -0.70710678118654757, -0.55557023301960229, -0.38268343236508984, -0.19509032201612833,
    0.0, 0.19509032201612833, 0.38268343236508984, 0.55557023301960229,
    0.70710678118654757, 0.83146961230254524, 0.92387953251128674, 0.98078528040323043};
__device__ constexpr double kS32[32] = {
    0.0, 0.19509032201612833, 0.38268343236508984, 0.55557023301960229,
    0.70710678118654757, 0.83146961230254524, 0.92387953251128674, 0.98078528040323043,
    1.0, 0.98078528040323043, 0.92387953251128674, 0.83146961230254524,
    0.70710678118654757, 0.55557023301960229, 0.38268343236508984, 0.19509032201612833,
    0.0, -0.19509032201612833, -0.38268343236508984, -0.55557023301960229,
    -0.70710678118654757, -0.83146961230254524, -0.92387953251128674, -0.98078528040323043,
    -1.0, -0.98078528040323043, -0.92387953251128674, -0.83146961230254524,
    -0.70710678118654757, -0.55557023301960229, -0.38268343236508984, -0.19509032201612833};

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }
__host__ __device__ constexpr int bitrev(int i, int bits)
{
    return bits == 0 ? 0 : ((i & 1) << (bits - 1)) | bitrev(i >> 1, bits - 1);
}

// In-register N-point DFT (N in {2,4,8,16,32}); sign -1 forward, +1 inverse,
// unnormalized.  Radix-2 decimation in time; all indices are compile-time.
template <int N, int SIGN>
__device__ __forceinline__ void dft_reg(double (&re)[N], double (&im)[N])
{
    constexpr int B = ilog2c(N);
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int j = bitrev(i, B);
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
#pragma unroll
    for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
        for (int k = 0; k < len / 2; k++) {
            const int q = k * (32 / len);
            const double wr = kC32[q];
            const double wi = SIGN * kS32[q];
#pragma unroll
            for (int s = 0; s < N; s += len) {
                const int a = s + k, b = s + k + len / 2;
                double tr, ti;
                if (k == 0) { tr = re[b]; ti = im[b]; }
                else { tr = re[b] * wr - im[b] * wi; ti = re[b] * wi + im[b] * wr; }
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] = re[a] + tr; im[a] = im[a] + ti;
            }
        }
    }
}

// Spectral x-derivative of two real rows at once (z = f + i g), following
// ddx (derivatives.f90:28-58): forward DFT, multiply bin k by i*eyekx(k)
// (negative frequencies by the conjugate, the Nyquist bin by zero: FFTW's
// c2r drops its imaginary part), inverse DFT, times stretchx.
//
// nx = N1*N2.  A team of N1 lanes; lane t holds z[t + N1*j2], j2 < N2, in
// (re, im) and gets D(z) back at the same positions.  lds points at the
// team's N2*(N1+1) double2 slots.  Every lane of the block must call this
// (it contains block barriers).
template <int N1, int N2>
__device__ __forceinline__ void fft_deriv_team(double (&re)[N2], double (&im)[N2], int t,
                                               double2 *lds, const double2 *__restrict__ tw,
                                               const double *__restrict__ dmul, double stretchx)
{
    constexpr int LD = N1 + 1;
    // step 1: DFT over j2, twiddle W_n^(t*k2), transpose through LDS
    dft_reg<N2, -1>(re, im);
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
        const double2 w = tw[t * k2];
        const double r = re[k2] * w.x - im[k2] * w.y;
        const double i = re[k2] * w.y + im[k2] * w.x;
        lds[k2 * LD + t] = make_double2(r, i);
    }
    __syncthreads();
    // step 2 (lanes t < N2): DFT over j1 -> X[t + N2*k1]; times i*dmul[k]
    // (dmul[k] = eyekx(k) below the Nyquist bin, 0 at it, -eyekx(nx-k) above);
    // inverse DFT over k1; conjugate twiddle; back to LDS
    if (N1 == N2 || t < N2) {
        double ar[N1], ai[N1];
#pragma unroll
        for (int j1 = 0; j1 < N1; j1++) {
            const double2 v = lds[t * LD + j1];
            ar[j1] = v.x; ai[j1] = v.y;
        }
        dft_reg<N1, -1>(ar, ai);
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) {
            const double w = dmul[t + N2 * k1];
            const double r = ar[k1];
            ar[k1] = -(ai[k1] * w);
            ai[k1] = r * w;
        }
        dft_reg<N1, +1>(ar, ai);
#pragma unroll
        for (int j1 = 0; j1 < N1; j1++) {
            const double2 w = tw[j1 * t];   // conj: exp(+2 pi i j1 t/n)
            const double r = ar[j1] * w.x + ai[j1] * w.y;
            const double i = ai[j1] * w.x - ar[j1] * w.y;
            lds[t * LD + j1] = make_double2(r, i);
        }
    }
    __syncthreads();
    // step 3: inverse DFT over k2 -> x[t + N1*j2]
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
        const double2 v = lds[k2 * LD + t];
        re[k2] = v.x; im[k2] = v.y;
    }
    __syncthreads();
    dft_reg<N2, +1>(re, im);
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
        re[j2] = re[j2] * stretchx;
        im[j2] = im[j2] * stretchx;
    }
}

// The same derivative for nx = 256 on a team of 32 lanes (two rows per
// wave), lane t holding z[t + 32 m], m < 8.  Three-stage Cooley-Tukey,
// 256 = 8 x (4 x 8) with n = t + 32 m, t = a + 8 b, k = k2 + 8 (c + 4 d):
// an 8-point DFT over m in each lane; two 4-point DFTs over b per lane; an
// 8-point DFT over a, the bin multiplier, and the inverse 8-point DFT over
// d in each lane; then the inverse stages in reverse order.  Every lane
// does the same work in every stage (the 16 x 16 form leaves no lane idle
// either, but runs one row per 16 lanes).  A and B are the team's two LDS
// buffers of 8 x 33 slots, used alternately so that each exchange needs
// one barrier.
__device__ __forceinline__ void fft_deriv_team_256x32(double (&re)[8], double (&im)[8], int t, double2 *A,
                                                      double2 *B, const double2 *__restrict__ tw,
                                                      const double *__restrict__ dmul, double stretchx)
{
    constexpr int LD = 33;
    // stage 1: DFT over m, twiddle W_256^(t k2) -> A[k2][t]
    dft_reg<8, -1>(re, im);
#pragma unroll
    for (int k2 = 0; k2 < 8; k2++) {
        const double2 w = tw[t * k2];
        A[k2 * LD + t] = make_double2(re[k2] * w.x - im[k2] * w.y, re[k2] * w.y + im[k2] * w.x);
    }
    __syncthreads();
    // stage 2: for (k2, a) = (p / 8, p % 8), p = 2t, 2t+1: DFT over b of
    // A[k2][a + 8b], twiddle W_32^(a c) -> B[k2][8c + a]
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int p = 2 * t + r, k2 = p >> 3, a = p & 7;
        double yr[4], yi[4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const double2 v = A[k2 * LD + a + 8 * b];
            yr[b] = v.x; yi[b] = v.y;
        }
        dft_reg<4, -1>(yr, yi);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const double2 w = tw[8 * a * c];
            B[k2 * LD + 8 * c + a] = make_double2(yr[c] * w.x - yi[c] * w.y, yr[c] * w.y + yi[c] * w.x);
        }
    }
    __syncthreads();
    // stage 3 at (k2, c) = (t / 4, t % 4): DFT over a -> X[k2 + 8c + 32d];
    // times i*dmul[k]; inverse DFT over d -> a; twiddle W_32^(-a c) -> A[k2][8c + a]
    {
        const int k2 = t >> 2, c = t & 3;
        double zr[8], zi[8];
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const double2 v = B[k2 * LD + 8 * c + a];
            zr[a] = v.x; zi[a] = v.y;
        }
        dft_reg<8, -1>(zr, zi);
#pragma unroll
        for (int d = 0; d < 8; d++) {
            const double w = dmul[k2 + 8 * c + 32 * d];
            const double r = zr[d];
            zr[d] = -(zi[d] * w);
            zi[d] = r * w;
        }
        dft_reg<8, +1>(zr, zi);
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const double2 w = tw[8 * a * c];   // conj: exp(+2 pi i a c / 32)
            A[k2 * LD + 8 * c + a] = make_double2(zr[a] * w.x + zi[a] * w.y, zi[a] * w.x - zr[a] * w.y);
        }
    }
    __syncthreads();
    // stage 2 inverse: for (k2, a): inverse DFT over c of A[k2][8c + a] -> B[k2][a + 8b]
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int p = 2 * t + r, k2 = p >> 3, a = p & 7;
        double yr[4], yi[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const double2 v = A[k2 * LD + 8 * c + a];
            yr[c] = v.x; yi[c] = v.y;
        }
        dft_reg<4, +1>(yr, yi);
#pragma unroll
        for (int b = 0; b < 4; b++) B[k2 * LD + a + 8 * b] = make_double2(yr[b], yi[b]);
    }
    __syncthreads();
    // stage 1 inverse: twiddle W_256^(-t k2), inverse DFT over k2 -> m
#pragma unroll
    for (int k2 = 0; k2 < 8; k2++) {
        const double2 v = B[k2 * LD + t];
        const double2 w = tw[t * k2];
        re[k2] = v.x * w.x + v.y * w.y;
        im[k2] = v.y * w.x - v.x * w.y;
    }
    dft_reg<8, +1>(re, im);
#pragma unroll
    for (int m = 0; m < 8; m++) {
        re[m] = re[m] * stretchx;
        im[m] = im[m] * stretchx;
    }
}

// LDS slots (double2) a team needs for the derivative of an nx-point row.
// The same derivative for nx = 256 on 64 lanes (one row per wave), in four
// radix-4 stages: with x = t + 64 m and k = k2 + 4p + 16r + 64g,
//   S1 (lane t):            DFT over m, twiddle W_256^(t k2)      -> A[k2][t]
//   S2 (lane (k2, t'), t = t' + 16 v): DFT over v, W_64^(t' p)     -> B[k2][t'][p]
//   S3 (lane (k2, p, s), t' = s + 4 w): DFT over w, W_16^(s r)    -> A[k2][p][s][r]
//   S4 (lane (k2, p, r)):   DFT over s -> X[k]; times i*dmul[k]; inverse over g
// and the inverse stages back to lane t holding x = t + 64 m.  A and B are the
// wave's two LDS buffers of 256 slots (one pad slot per 32), used alternately
// so that each exchange needs one barrier.
__device__ __forceinline__ int fft64_ix(int e) { return e + (e >> 5); }

__device__ __forceinline__ void fft_deriv_team_256x64(double (&re)[4], double (&im)[4], int t, double2 *A,
                                                      double2 *B, const double2 *__restrict__ tw,
                                                      const double *__restrict__ dmul, double stretchx)
{
    // S1
    dft_reg<4, -1>(re, im);
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) {
        const double2 w = tw[(t * k2) & 255];
        A[fft64_ix(k2 * 64 + t)] = make_double2(re[k2] * w.x - im[k2] * w.y, re[k2] * w.y + im[k2] * w.x);
    }
    __syncthreads();
    const int k2 = t >> 4;
    {   // S2
        const int tp = t & 15;
        double yr[4], yi[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double2 a = A[fft64_ix(k2 * 64 + tp + 16 * v)];
            yr[v] = a.x; yi[v] = a.y;
        }
        dft_reg<4, -1>(yr, yi);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const double2 w = tw[(4 * tp * p) & 255];
            B[fft64_ix(k2 * 64 + tp * 4 + p)] = make_double2(yr[p] * w.x - yi[p] * w.y, yr[p] * w.y + yi[p] * w.x);
        }
    }
    __syncthreads();
    const int p = (t >> 2) & 3, s = t & 3;     // S3 and IS3: (k2, p, s); S4: (k2, p, r = s)
    {   // S3
        double yr[4], yi[4];
#pragma unroll
        for (int w4 = 0; w4 < 4; w4++) {
            const double2 a = B[fft64_ix(k2 * 64 + (s + 4 * w4) * 4 + p)];
            yr[w4] = a.x; yi[w4] = a.y;
        }
        dft_reg<4, -1>(yr, yi);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const double2 w = tw[(16 * s * r) & 255];
            A[fft64_ix(k2 * 64 + p * 16 + s * 4 + r)] =
                make_double2(yr[r] * w.x - yi[r] * w.y, yr[r] * w.y + yi[r] * w.x);
        }
    }
    __syncthreads();
    {   // S4, the derivative, inverse over g
        const int r = s;
        double xr[4], xi[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double2 a = A[fft64_ix(k2 * 64 + p * 16 + q * 4 + r)];
            xr[q] = a.x; xi[q] = a.y;
        }
        dft_reg<4, -1>(xr, xi);
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const double w = dmul[k2 + 4 * p + 16 * r + 64 * g];
            const double x0 = xr[g];
            xr[g] = -(xi[g] * w);
            xi[g] = x0 * w;
        }
        dft_reg<4, +1>(xr, xi);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double2 w = tw[(16 * q * r) & 255];   // conj: exp(+2 pi i q r / 16)
            B[fft64_ix(k2 * 64 + p * 16 + q * 4 + r)] =
                make_double2(xr[q] * w.x + xi[q] * w.y, xi[q] * w.x - xr[q] * w.y);
        }
    }
    __syncthreads();
    {   // IS3: inverse over r at (k2, p, s) -> t' = s + 4 w
        double yr[4], yi[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const double2 a = B[fft64_ix(k2 * 64 + p * 16 + s * 4 + r)];
            yr[r] = a.x; yi[r] = a.y;
        }
        dft_reg<4, +1>(yr, yi);
#pragma unroll
        for (int w4 = 0; w4 < 4; w4++) {
            const int tp = s + 4 * w4;
            const double2 w = tw[(4 * tp * p) & 255];
            A[fft64_ix(k2 * 64 + tp * 4 + p)] = make_double2(yr[w4] * w.x + yi[w4] * w.y, yi[w4] * w.x - yr[w4] * w.y);
        }
    }
    __syncthreads();
    {   // IS2: inverse over p at (k2, t') -> t = t' + 16 v
        const int tp = t & 15;
        double yr[4], yi[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double2 a = A[fft64_ix(k2 * 64 + tp * 4 + q)];
            yr[q] = a.x; yi[q] = a.y;
        }
        dft_reg<4, +1>(yr, yi);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int tt = tp + 16 * v;
            const double2 w = tw[(tt * k2) & 255];
            B[fft64_ix(k2 * 64 + tt)] = make_double2(yr[v] * w.x + yi[v] * w.y, yi[v] * w.x - yr[v] * w.y);
        }
    }
    __syncthreads();
    // IS1: inverse over k2 at lane t -> x = t + 64 m
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double2 a = B[fft64_ix(q * 64 + t)];
        re[q] = a.x; im[q] = a.y;
    }
    dft_reg<4, +1>(re, im);
#pragma unroll
    for (int m = 0; m < 4; m++) {
        re[m] = re[m] * stretchx;
        im[m] = im[m] * stretchx;
    }
}

// The same derivative for any nx (the form for an nx that is not a power of
// two): the linear map that fft_deriv_team applies, multiplying bin k by
// i*dmul[k] (for odd nx there is no Nyquist bin to drop), written as the
// circulant product (D z)[x] = sum_y c[(x - y) mod nx] z[y] with c the map's
// first column, then times stretchx.  O(nx^2) per row pair against the
// FFT's O(nx log nx): a general-grid fallback, not a tuned path.
//
// A team of 64 lanes (one row pair z = f + i g per block); lane t holds
// z[t + 64 j2], j2 < N2, 64 N2 >= nx; slots with t + 64 j2 >= nx carry
// nothing and get garbage.  cc: c repeated, cc[i] = c[i mod nx] for
// i < 2 nx (padded to 128 N2 entries), so that x - y + nx, in [1, 2 nx) for
// a real slot, indexes it without a wrap.  z: at least 64 N2 double2.
// Every lane of the block must call this (it contains block barriers).
template <int N2>
__device__ __forceinline__ void circ_deriv_row(double (&re)[N2], double (&im)[N2], int t, double2 *z,
                                               const double *cc, int nx, double stretchx)
{
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
        const int x = t + 64 * j2;
        if (x < nx) z[x] = make_double2(re[j2], im[j2]);
    }
    __syncthreads();
    double ar[N2], ai[N2];
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) ar[j2] = ai[j2] = 0.0;
#pragma unroll 2
    for (int y = 0; y < nx; y++) {
        const double2 v = z[y];            // the same address on every lane: a broadcast
#pragma unroll
        for (int j2 = 0; j2 < N2; j2++) {
            const double c = cc[t + 64 * j2 - y + nx];   // < 64 N2 + nx <= 128 N2
            ar[j2] += c * v.x;
            ai[j2] += c * v.y;
        }
    }
    __syncthreads();   // z free for the next call
#pragma unroll
    for (int j2 = 0; j2 < N2; j2++) {
        re[j2] = ar[j2] * stretchx;
        im[j2] = ai[j2] * stretchx;
    }
}

template <int N1, int N2> struct FftLds { static constexpr int n = N2 * (N1 + 1); };
template <> struct FftLds<32, 8> { static constexpr int n = 2 * 8 * 33; };
template <> struct FftLds<64, 4> { static constexpr int n = 2 * 264; };

// Factorization N1 x N2 of the power-of-two row lengths (N1 >= N2).  Any
// other nx runs the circulant form, a kernel template argument NX = -N2
// (rows of at most 64 N2 on 64 lanes; any_nx_slots picks N2).
template <int NX> struct FftShape;
template <> struct FftShape<16> { static constexpr int N1 = 4, N2 = 4; };
template <> struct FftShape<32> { static constexpr int N1 = 8, N2 = 4; };
template <> struct FftShape<64> { static constexpr int N1 = 8, N2 = 8; };
template <> struct FftShape<128> { static constexpr int N1 = 16, N2 = 8; };
template <> struct FftShape<256> { static constexpr int N1 = 16, N2 = 16; };
template <> struct FftShape<512> { static constexpr int N1 = 32, N2 = 16; };
template <> struct FftShape<1024> { static constexpr int N1 = 32, N2 = 32; };
__host__ __device__ inline bool nx_has_fft(int nx) { return nx >= 16 && nx <= 1024 && (nx & (nx - 1)) == 0; }
// Slots per lane of the circulant form: the smallest instantiated N2 with
// 64 N2 >= nx (SPARC_ANY_NX_SLOTS lists them), so that few slots are idle.
#define SPARC_ANY_NX_SLOTS(X) X(2) X(3) X(4) X(5) X(6) X(8) X(10) X(12) X(16)
__host__ __device__ inline int any_nx_slots(int nx)
{
    const int need = (nx + 63) / 64;
    return need <= 2 ? 2 : (need <= 6 ? need : (need <= 8 ? 8 : (need <= 10 ? 10 : (need <= 12 ? 12 : 16))));
}

// ---------------------------------------------------------------------------
// Compact-FD forward elimination (dbyd2.f + mtridag.f:96-105) of one column,
// streaming the right-hand side; the window holds B(j-2..j+2).
struct FwdSweep {
    double m2, m1, c0, p1, p2, ud;

    __device__ __forceinline__ void start(const double *__restrict__ B, size_t q0, size_t s, int ibc,
                                          double preset0, const DbyCoef &cf, const Tri &tr,
                                          double *__restrict__ W)
    {
        const double B0 = B[q0], B1 = B[q0 + s], B2 = B[q0 + 2 * s], B3 = B[q0 + 3 * s],
                     B4 = B[q0 + 4 * s];
        const double R1 = cf.b1 * B0 + cf.b2 * B1 + cf.b3 * B2 + cf.b4 * B3 + cf.b5 * B4;
        double R0;
        if (ibc == 1 || ibc == 3) R0 = cf.f1 * B0 + cf.f2 * B1 + cf.f3 * B2 + cf.f4 * B3;
        else if (ibc == 0) R0 = 0.0;
        else R0 = preset0;
        if (ibc & 1) R0 = (cf.alpha2 * R0 - cf.beta1 * R1) * cf.cnst;
        const double u0 = R0 * tr.rbet[0];
        W[q0] = u0;
        ud = (R1 - tr.lowr[1] * u0) * tr.rbet[1];
        W[q0 + s] = ud;
        m2 = B0; m1 = B1; c0 = B2; p1 = B3; p2 = B4;
    }
    // interior row j (2 <= j <= N-3); Bn = B(j+3) (unused past the end)
    __device__ __forceinline__ void step(int j, size_t qj, double Bn, const DbyCoef &cf,
                                         const Tri &tr, double *__restrict__ W)
    {
        const double R = cf.c1 * (p1 - m1) + cf.c2 * (p2 - m2);
        ud = (R - tr.lowr[j] * ud) * tr.rbet[j];
        W[qj] = ud;
        m2 = m1; m1 = c0; c0 = p1; p1 = p2; p2 = Bn;
    }
    // rows N-2 and N-1
    __device__ __forceinline__ void finish(const double *__restrict__ B, size_t q0, size_t s, int N,
                                           int ibc, double presetN, const DbyCoef &cf,
                                           const Tri &tr, double *__restrict__ W)
    {
        const double Bl = B[q0 + (size_t)(N - 1) * s], Bl1 = B[q0 + (size_t)(N - 2) * s],
                     Bl2 = B[q0 + (size_t)(N - 3) * s], Bl3 = B[q0 + (size_t)(N - 4) * s],
                     Bl4 = B[q0 + (size_t)(N - 5) * s];
        const double Rm = -(cf.b1 * Bl + cf.b2 * Bl1 + cf.b3 * Bl2 + cf.b4 * Bl3 + cf.b5 * Bl4);
        double RN;
        if (ibc == 1 || ibc == 2) {
            RN = -(cf.f1 * Bl + cf.f2 * Bl1 + cf.f3 * Bl2 + cf.f4 * Bl3);
            RN = (cf.alpha2 * RN - cf.beta1 * Rm) * cf.cnst;
        } else if (ibc == 0) {
            RN = 0.0;
        } else {
            RN = presetN;
        }
        ud = (Rm - tr.lowr[N - 2] * ud) * tr.rbet[N - 2];
        W[q0 + (size_t)(N - 2) * s] = ud;
        ud = (RN - tr.lowr[N - 1] * ud) * tr.rbet[N - 1];
        W[q0 + (size_t)(N - 1) * s] = ud;
    }
};

// ---------------------------------------------------------------------------
// Helpers shared by the pass kernels.

// 11-point z filter (physics2d.f90:517-533, kernels.f90:911-918) at row k of
// a column whose row stride is s.
__device__ __forceinline__ double filt11(const double *__restrict__ a, size_t q, size_t s)
{
    return 0.75390625 * a[q] + 0.5 * 0.41015625 * (a[q - s] + a[q + s]) +
           0.5 * -0.234375 * (a[q - 2 * s] + a[q + 2 * s]) +
           0.5 * 0.087890625 * (a[q - 3 * s] + a[q + 3 * s]) +
           0.5 * -0.01953125 * (a[q - 4 * s] + a[q + 4 * s]) +
           0.5 * 0.001953125 * (a[q - 5 * s] + a[q + 5 * s]);
}

// FILTER_VARS_2D value of row k, column x (HORIZONTAL_PMLS: columns
// npmlhor..nx-npmlhor+1 and rows 6..nz-1, 1-based, are rewritten).
__device__ __forceinline__ double filter_vars_2d_at(const double *__restrict__ a, size_t q, int k,
                                                    int x, int nx, int nz, int nh)
{
    if (k < 5 || k > nz - 2 || x < nh - 1 || x > nx - nh) return a[q];
    if (k <= nz - 6) return filt11(a, q, (size_t)nx);
    return 0.5 * (a[q + nx] + a[q - nx]);
}

// The same value with all eleven rows loaded (row indices clamped to the
// column) and the form selected afterwards, so that a wave whose lanes take
// different forms issues one set of loads instead of one per form.
__device__ __forceinline__ double filter_vars_2d_sel(const double *__restrict__ a, size_t q, int k,
                                                     int x, int nx, int nz, int nh)
{
    const size_t q0 = q - (size_t)k * nx;       // row 0 of this column
    double w[11];
#pragma unroll
    for (int u = 0; u < 11; u++) w[u] = a[q0 + (size_t)min(max(k - 5 + u, 0), nz - 1) * nx];
    const double f11 = 0.75390625 * w[5] + 0.5 * 0.41015625 * (w[4] + w[6]) +
                       0.5 * -0.234375 * (w[3] + w[7]) +
                       0.5 * 0.087890625 * (w[2] + w[8]) +
                       0.5 * -0.01953125 * (w[1] + w[9]) +
                       0.5 * 0.001953125 * (w[0] + w[10]);
    const double f2 = 0.5 * (w[6] + w[4]);
    const bool keep = k < 5 || k > nz - 2 || x < nh - 1 || x > nx - nh;
    return keep ? w[5] : (k <= nz - 6 ? f11 : f2);
}

__device__ __forceinline__ int clock_t_(const Clock &c) { return c.tctr[0] + c.toff; }

// LAGRANGE_INTERP (physics.f90:678-737) for column x of source b.
__device__ __forceinline__ double lagrange_forcing(const Src &src, int b, int x, int nx, int time)
{
    const double cad = src.cadforcing;
    double xt = (double)time * src.timestep / cad;
    const int time1 = (int)floor(xt) + 1;
    xt = time * src.timestep;
    double xs[7];
#pragma unroll
    for (int q = 0; q < 7; q++) xs[q] = (double)(time1 - 4 + q) * cad;
    const double *vr = src.vr + (size_t)b * src.nvr * nx + x;
    // the seven slices' values (clamped slice indices), all loaded before the
    // first is used so that they cost one memory round trip: the empty asm
    // takes all seven at one point, which keeps the compiler from turning the
    // zeroing of slices outside 1..nvr below into branches that wait for each
    // load in turn
    double vs[7];
#pragma unroll
    for (int q = 0; q < 7; q++) {
        const int slice = time1 - 3 + q;  // 1-based vr slice
        const int sl = min(max(slice, 1), max(src.nvr, 1));
        vs[q] = vr[(size_t)(sl - 1) * nx];
    }
    asm volatile("" : "+v"(vs[0]), "+v"(vs[1]), "+v"(vs[2]), "+v"(vs[3]), "+v"(vs[4]), "+v"(vs[5]), "+v"(vs[6]));
#pragma unroll
    for (int q = 0; q < 7; q++) {
        const int slice = time1 - 3 + q;
        vs[q] = (slice >= 1 && slice <= src.nvr) ? vs[q] : 0.0;
    }
    double f = 0.0;
#pragma unroll
    for (int q = 0; q < 7; q++) {
        bool ok = true;
        if (q <= 3) ok = xs[q] >= 0.0;
        if (q == 5) ok = time1 < src.num_steps;
        if (q == 6) ok = time1 < src.num_steps - 1;
        double LC = 0.0;
        if (ok) {
            double den = 1.0;
            bool first = true;
#pragma unroll
            for (int r = 0; r < 7; r++) {
                if (r == q) continue;
                if (first) { den = xs[q] - xs[r]; first = false; }
                else den = den * (xs[q] - xs[r]);
            }
            LC = vs[q] / den;
        }
        double tq = LC;
#pragma unroll
        for (int r = 0; r < 7; r++)
            if (r != q) tq = tq * (xt - xs[r]);
        f = (q == 0) ? tq : f + tq;
    }
    return f;
}

}  // namespace sparc
