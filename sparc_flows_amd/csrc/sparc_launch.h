// sparc_launch.h -- host-side launchers of the solver kernels.
#pragma once
#include "sparc_device.h"

namespace sparc {

hipError_t launch_ddz_cols(const double *in, double *out, int ncols, int nz, int nbatch, int ibc,
                           const double *stretch, const DbyCoef &cf, const Tri &tr, double *work,
                           int filt, hipStream_t st);
hipError_t launch_ddx_rows(const double *in, double *out, int nrows, int nx, const double2 *tw,
                           const double *eyekx, double stretchx, hipStream_t st);
// Row pass; at nx = 256 wide = 1 runs 32 lanes per row (two rows per wave),
// wide = 2 runs 64 lanes per row (one row per wave).
hipError_t launch_row(int mode, const Dims &d, const Bg &bg, const State &s, const RowArgs &ra,
                      const Clock &c, const Sink &sk, hipStream_t st, int wide);
// Segmented column pass, one lane per (column, segment, field); tc in
// {1, 2, 4, 8, 16} columns per block, ms = rows per segment, one of
// SPARC_COLSEG_SIZES.
#define SPARC_COLSEG_SIZES(X) X(4) X(6) X(8) X(10) X(12) X(16) X(19) X(24) X(32)
hipError_t colseg_init();
hipError_t launch_colsplit(int tc, int ms, const Dims &d, const Bg &bg, const DbyCoef &cf, const SegTab &sg,
                           const State &s, const ColArgs &ca, const Src &src, const Clock &c, const Sink &sk,
                           hipStream_t st);
hipError_t launch_finalize_vx(const Dims &d, const State &s, const Clock &c, const Sink &sk,
                              hipStream_t st);
hipError_t launch_tick(int *tctr, int inc, hipStream_t st);
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

}  // namespace sparc
