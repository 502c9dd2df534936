// sparc_launch.h -- host-side launchers of the solver kernels.
#pragma once
#include "sparc_device.h"

namespace sparc {

hipError_t launch_ddz_cols(const double *in, double *out, int ncols, int nz, int nbatch, int ibc,
                           const double *stretch, const DbyCoef &cf, const Tri &tr, double *work,
                           int filt, hipStream_t st);
// x-derivative of nrows rows of length nx: the FFT form for a power of two,
// the circulant product (cdx, Bg::cdx's layout) for any other nx in [16, 1024].
hipError_t launch_ddx_rows(const double *in, double *out, int nrows, int nx, const double2 *tw,
                           const double *dmul, const double *cdx, double stretchx, hipStream_t st);
// Row pass; at nx = 256 wide = 1 runs 32 lanes per row (two rows per wave),
// wide = 2 runs 64 lanes per row (one row per wave).
hipError_t launch_row(int mode, const Dims &d, const Bg &bg, const State &s, const RowArgs &ra,
                      const Clock &c, const Sink &sk, hipStream_t st, int wide);
// Segmented column pass, one lane per (column, segment, field); tc in
// {1, 2, 4, 8, 16} columns per block, ms = rows per segment, one of
// SPARC_COLSEG_SIZES.
#define SPARC_COLSEG_SIZES(X) X(4) X(6) X(8) X(10) X(12) X(16) X(19) X(20) X(24)
hipError_t colseg_init();
// Whether a column-pass block of tc columns fits: its threads (3 lane groups
// of at most 128) and its dynamic LDS (after colseg_init, the device maximum).
bool colsplit_fits(int tc, const Dims &d, const SegTab &sg);
hipError_t launch_colsplit(int tc, int ms, const Dims &d, const Bg &bg, const DbyCoef &cf, const SegTab &sg,
                           const State &s, const ColArgs &ca, const Src &src, const Clock &c, const Sink &sk,
                           hipStream_t st);
hipError_t launch_finalize_vx(const Dims &d, const State &s, const Clock &c, const Sink &sk,
                              hipStream_t st);
hipError_t launch_tick(int *tctr, int inc, hipStream_t st);
// sk.norm2 slot of the step at clock c: the sum of v_z^2 of each source.
hipError_t launch_vz_norm2(const double *vz, int S, size_t n2, const Clock &c, const Sink &sk, hipStream_t st);
hipError_t launch_set_int(int *p, int v, hipStream_t st);
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t st);

hipError_t launch_gather_pairs(const double *fwd, const double *adj, int S, int nt, int k0, int Kc, size_t n2,
                               double *G, hipStream_t st);
hipError_t launch_kern_accum_pairs(const double *fwd, const double *adj, int S, int nt, int k0, int Kc,
                                   size_t n2, const double *Dx, const double *Dz, double *KC, double *KX,
                                   double *KZ, double *H, hipStream_t st);
hipError_t launch_kern_final(int S, size_t n2, const double *rho0, const double *c2, double steps,
                             double dt, double dimen, double dimen_c, const double *t1, const double *t2,
                             const double *KC, const double *KX, const double *KZ, double *kc,
                             double *kpsi, double *kvx, double *kvz, hipStream_t st);

// Exact discrete adjoint (sparc_discrete.hip).  Full states are
// [4][S][nz][nx] fields, then [2][S][nzpml][nx] (psivz, psip), then
// [2][S][nz][nph] (psivx, psi_gradp_x).
struct DiscreteCtx {
    Dims d;
    Bg bg;
    Src src;
    hipStream_t stream;
    const double *M[3], *MT[3];    // the z-derivative of ibc 1, 2, 3 as nz x nz matrices, and their transposes
    const double *vN3;             // ibc 3's response to its kept last row [nz]
    double betas[5], deltat, unstretch_last;
    int steps, orow;
};
struct DiscreteWork {
    size_t SN, PZ1, PX1, NST;      // one [S][nz][nx] field, one pz field, one px field, a full state
    double *U, *Unext, *Y, *R, *acc, *Rb, *Yb, *lam;          // NST each
    double *ck, *seg;                                         // segment starts, one segment of states
    double *Dst, *Dscr;                                       // [5][5][SN], [5][SN]
    double *Dseg;                                             // [interval][5][5][SN] or null
    double *T0, *T1, *T2, *T3, *GZ, *DGZ, *RHOB, *XZP, *DXB, *DZB, *DZ3B, *DDXB;   // SN each
    double *FX, *DFX, *FZ, *DFZ;                              // 3, 3, 2, 2 SN
    double *PB;                                               // [S][nx]
};
// dJ/d(c^2 rho0), dJ/dv0x, dJ/dv0z into G [3][S][nz][nx] for J = sum <gvz, vz>
// (gvz [S][nt][nx]) over a forward run of maxtime+1 steps from the zero
// state, recomputing segments of `interval` steps from checkpoints.
hipError_t discrete_gradient(const DiscreteCtx &c, const DiscreteWork &w, int maxtime, int interval,
                             const double *gvz, int nt, double *G);
// dJ/dpsi, dJ/dc [S][nz][nx] from G through c^2 rho0 and the flow map.
hipError_t discrete_flow_chain(const DiscreteCtx &c, const DiscreteWork &w, const double *G, const double *rho0,
                               const double *c2, const double *psi, double scale, double dimc, double *gpsi,
                               double *gc);
// out = Dz_ibc in (which 0..2) or Dz_ibc^T in (which 3..5) for nbatch fields;
// for which 5 also the adjoint of the kept last row presetN_bar [nbatch][nx].
hipError_t discrete_apply(const DiscreteCtx &c, const double *in, double *out, int nbatch, int which,
                          double *presetN_bar);

}  // namespace sparc
