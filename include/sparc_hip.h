/*
 * sparc_hip.h -- C ABI of the MI355X (gfx950) implementation of SPARC's 2D
 * time-domain solver for the linearized Euler equations with background
 * flows (option 17: MP_QUIET_PML_2D_DISPL).
 *
 * The reference has no library interface: its callers (data_forward.py:14-18,
 * full.py:17-29,44-59, linesearch.py:15-19) start one `sparc SS LL` process
 * per source and phase and exchange data through files.  Each entry point
 * below replaces one piece of that program; the comment names the reference
 * code it stands for.
 *
 * Memory: arrays named "host" are read on the host during the call; arrays
 * named "device" are HIP device pointers (HBM), e.g. from torch tensors.
 * Layouts are C row-major: [nz][nx] fields are the reference's column-major
 * (nx,1,nz) arrays.  All floating point is double.
 *
 * Errors: every function returns 0 on success or a negative SPARC_E_* code;
 * sparc_last_error() gives the message.  Calls on one handle are not thread
 * safe; work is queued on the handle's stream (sparc_set_stream).
 */
#ifndef SPARC_HIP_H
#define SPARC_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPARC_OK 0
#define SPARC_E_ARG (-1)      /* invalid argument / unsupported configuration */
#define SPARC_E_HIP (-2)      /* HIP runtime error */
#define SPARC_E_STATE (-3)    /* call out of order (e.g. run before sources) */
#define SPARC_E_NONFINITE (-4) /* non-finite values detected in the state */

/* Runtime values of the params.i constants read by the 2D path
 * (params.i:25-28 grid/timestep, :37 cadforcing, :66 nupeak/nuwidth,
 * :89 npmlhor, :122 npmlbot/npmltop, :139 obsheight, :143 outputcad,
 * :152 excitdep), the background model file rows (solar_data,
 * initialize.f90:743-747) and the optional 2D sound speed and stream function
 * (model_c_lsLL.fits / model_psi_lsLL.fits, driver.f90:94-126). */
typedef struct sparc_config {
    int nx, nz;               /* nx in [16, 1024]: FFT rows for a power of two, the circulant form otherwise */
    int npmlbot, npmltop, npmlhor;
    int nsrc;                 /* sources advanced together in one batch */
    int adjoint;              /* 0: forward ('spectral'), 1: adjoint ('adjoint') */
    double xlength;           /* cm */
    double timestep;          /* s */
    double cadforcing;        /* s */
    double outputcad;         /* s */
    double obsheight;         /* cm */
    double excitdep;          /* cm */
    double nupeak, nuwidth;   /* values as the reference stores them */
    const double *model;      /* host [nz][6]: z/Rsun, c, rho, p, g, Gamma (cgs) */
    const double *c2d;        /* host [nz][nx] sound speed (cm/s) or NULL */
    const double *psi;        /* host [nz][nx] stream function (Mm) or NULL */
} sparc_config;

typedef struct sparc_handle sparc_handle;

/* Initialize_all + solar_data + mp_Initialize_RHS_2d + INITIALIZE_STEP and
 * the flow from psi (initialize.f90:178-474, 731-794; physics2d.f90:23-405;
 * step.f90:24-41; driver.f90:94-147).  Allocates the handle's HBM state. */
int sparc_create(const sparc_config *cfg, sparc_handle **out);
/* Frees the handle (end of the `sparc` process, driver.f90:318-334). */
int sparc_destroy(sparc_handle *h);
/* Message of the last error on h (h may be NULL for sparc_create errors). */
const char *sparc_last_error(const sparc_handle *h);
/* Queue work on an existing hipStream_t (NULL: the handle's own stream). */
int sparc_set_stream(sparc_handle *h, void *hip_stream);
/* Integer / double facts of the handle: "e_rad", "o_rad", "steps",
 * "cadforcing_step", "forcing_length", "nsrc", "nx", "nz", "nzpml", "nph",
 * "colseg_tc", "colseg_ms", "colseg_p", "colseg_hw" (column-pass launch shape) / "deltat",
 * "stretchx", "delta_width", "beta2".."beta5". */
int sparc_query_int(const sparc_handle *h, const char *name, int *value);
int sparc_query_double(const sparc_handle *h, const char *name, double *value);
/* Copies a named background array to host memory (z, stretch, unstretch, g,
 * rho0, rhoinv, c2rho0, gradp0_z, gradrho0_z, v0x, v0z, az, bzpml, axpml,
 * bxpml, fwdsource); *n gets the element count. */
int sparc_get_background(const sparc_handle *h, const char *name, double *host_out, int cap,
                         int *n);
/* Replace the 2D sound speed and stream function and redo the background
 * set-up (model_c / model_psi of a new iterate, driver.f90:94-126). */
int sparc_set_model(sparc_handle *h, const double *c2d_host, const double *psi_host);
/* Set the background flow directly instead of from psi: host [nz][nx]
 * v0_x, v0_z in m/s, as the compute_data branch reads true_vx.fits and
 * true_vz.fits and scales them by 1e2/dimc (driver.f90:65-73).  An adjoint
 * handle reverses the flow (driver.f90:144-147).  Both NULL restores the
 * psi flow of the last sparc_set_model/sparc_create. */
int sparc_set_flow(sparc_handle *h, const double *vx_host, const double *vz_host);

/* Forward sources: DETERMINE_STATUS forward branch (driver.f90:521-571) and
 * FORWARD_SOURCE (driver.f90:1120-1145) for x positions xsrc (Mm, one per
 * source as in master.pixels), for time indices 0..maxtime. */
int sparc_set_forward_sources(sparc_handle *h, const double *xsrc_host, int maxtime);
/* A separable forward source in place of FORWARD_SOURCE's one-row form
 * (physics2d.f90:1056-1059 injects vr * fwdsource(time+1) * delta_width /
 * rho0 at row e_rad only): RHS v_z(x, z) += vr[s][x] * series[time] *
 * zprof[z] for time <= forcing_length * cadforcing_step, with zprof already
 * holding the 1/rho0 and any z normalisation.  All arrays are host:
 * vr [nsrc][nx], zprof [nz], series [nseries] (0 beyond its end).  Used for
 * a force distributed in z, such as a Gaussian. */
int sparc_set_forward_profile(sparc_handle *h, const double *vr_host, const double *zprof_host,
                              const double *series_host, int nseries, int maxtime);
/* Adjoint sources: DETERMINE_STATUS adjoint branch (driver.f90:584-594).
 * vr: device [nsrc][nt][nx] (source.fits of each source), nt slices at
 * cadence cadforcing, interpolated by LAGRANGE_INTERP (physics.f90:678-737). */
int sparc_set_adjoint_sources(sparc_handle *h, const double *vr_dev, int nt, int maxtime);

/* TIMESTEPPING (driver.f90:337-419) for time indices 0..maxtime of every
 * source: resets the state, then steps.  Outputs (device, may be NULL):
 *   vz_out [nsrc][nout][nx], nout = maxtime/steps + 1: a(:,o_rad,v_z) after
 *          each step whose index is a multiple of steps (vz_cc.bin records);
 *   snaps  [nsrc][nout][nf][nz][nx]: the partial-state snapshots
 *          (WRITE_OUT_PARTIAL_STATE, all_modules.f90:750-784): forward
 *          nf = 6 {xi_x, xi_z, v_x, v_z, acc_x, acc_z}; adjoint nf = 4
 *          {v_x, v_z, acc_x, acc_z}. */
int sparc_run(sparc_handle *h, int maxtime, double *vz_out_dev, double *snaps_dev);
/* Restart the state at zero and set the step counter (time0 = 0 start). */
int sparc_reset(sparc_handle *h);
/* Part of TIMESTEPPING (driver.f90:337-419): steps t0..t1 (inclusive) from
 * the current state, which must be the state at time index t0 (after
 * sparc_reset for t0 = 0, sparc_load_checkpoint, or a previous call that
 * ended at t0 - 1).  Outputs as sparc_run, but slot j of vz_out_dev / snaps_dev
 * (nslots slots each, may be NULL) holds output index k0 + j, i.e. time index
 * (k0 + j) * steps; other outputs are not written.  The reference restarts a
 * run from its dumped state the same way (init = time0 + 1, driver.f90:236-237,
 * 350). */
int sparc_run_range(sparc_handle *h, int t0, int t1, double *vz_out_dev, double *snaps_dev, int k0,
                    int nslots);
/* Checkpoints of the full solver state (fields, z- and x-PML variables) for
 * checkpoint-and-recompute; they replace the reference's full-state dumps
 * (WRITE_OUT_FULL_STATE, all_modules.f90:623-694).  A checkpoint taken after
 * sparc_run / sparc_run_range ended at time index t - 1 restarts at t. */
/* TIMESTEPPING's convergence statistic (driver.f90:364, written to rms_hist
 * every step at :384): for the runs that follow, norm2_dev (device
 * [nsrc][n]) gets sum(v_z^2) over the grid after each step t in t0 .. t0+n-1
 * (slot t - t0; nor1 is its square root).  NULL switches it off.  The array
 * must outlive those runs. */
int sparc_set_norm_output(sparc_handle *h, double *norm2_dev, int t0, int n);
int sparc_checkpoint_size(const sparc_handle *h, size_t *bytes);
int sparc_save_checkpoint(sparc_handle *h, void *dev);
int sparc_load_checkpoint(sparc_handle *h, const void *dev);
/* One STEP (step.f90:43-119) at time index t of every source. */
int sparc_step(sparc_handle *h, int t);
/* State of every source: device [4][nsrc][nz][nx], fields xi_x, xi_z, v_x, v_z
 * (a(:,:,:,1/3/4/6)). get finalizes the pending z filter of v_x first. */
int sparc_get_state(sparc_handle *h, double *a_dev);
int sparc_set_state(sparc_handle *h, const double *a_dev);
/* Scratch of the last RHS (scr(:,:,:,1/3/4/6)) is not kept; the PML
 * auxiliaries are: pmlvars device [2][nsrc][nzpml][nx] (psivz, psip),
 * pmlvarsx device [2][nsrc][nz][2*npmlhor] (psivx, psi_gradp_x). */
int sparc_get_pml(sparc_handle *h, double *pmlvars_dev, double *pmlvarsx_dev);

/* ddx (derivatives.f90:28-58) of nrows rows of length nx: device in/out. */
int sparc_ddx(sparc_handle *h, const double *in_dev, double *out_dev, int nrows);
/* ddz (derivatives.f90:10-24) of nbatch [nz][nx] fields with boundary mode
 * ibc (dbyd2.f:17-23, 0..3).  For ibc 2 / 3 the first / last row of out is
 * the caller-set boundary value, as in dbyd2. */
int sparc_ddz(sparc_handle *h, const double *in_dev, double *out_dev, int nbatch, int ibc);
/* ddzkern (kernels.f90:648-661): ddz followed by FILTER_VAR_Z. */
int sparc_ddz_kern(sparc_handle *h, const double *in_dev, double *out_dev, int nbatch, int ibc);
/* No counterpart in the reference.  sparc_ddz (filt 0) or sparc_ddz_kern
 * (filt 1) without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the Thomas work
 * buffer is a library-owned device array of 1048576 doubles (8 MiB) per GPU,
 * kept until the process ends.  nbatch*nz*nx above that returns SPARC_E_ARG.
 * The work is queued on the handle's stream and not waited for.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (other streams,
 * other handles, other threads) race. */
int sparc_ddz_noalloc(sparc_handle *h, const double *in_dev, double *out_dev, int nbatch, int ibc,
                      int filt);

/* PRODUCE_KERNELS (kernels.f90:13-612) for every source of a pair of runs:
 * fwd_snaps device [nsrc][nt][6][nz][nx] and adj_snaps [nsrc][nt][4][nz][nx]
 * as written by sparc_run (label order 0, steps, ...); forward label
 * (k-1)*steps is paired with adjoint label adj_end-(k-1)*steps.  Uses the
 * kernel-mode background (kernel_back, initialize.f90:937-979).  Outputs
 * device [nsrc][nz][nx]: kc (kernel_c), kpsi (kernel_psi), kvx, kvz, hess. */
int sparc_kernels(sparc_handle *h, int nt, const double *fwd_snaps_dev,
                  const double *adj_snaps_dev, double *kc_dev, double *kpsi_dev, double *kvx_dev,
                  double *kvz_dev, double *hess_dev);
/* The same accumulation in pieces (PRODUCE_KERNELS' loop over snapshot
 * pairs, kernels.f90:200-360, then its scalings, :495-580): begin zeroes the
 * accumulators; add accumulates n pairs from fwd [nsrc][n][6][nz][nx] and
 * adj [nsrc][n][4][nz][nx], pairing fwd slot i with adj slot n-1-i; end
 * writes the scaled kernels as sparc_kernels does. */
int sparc_kernels_begin(sparc_handle *h);
int sparc_kernels_add(sparc_handle *h, int n, const double *fwd_snaps_dev, const double *adj_snaps_dev);
int sparc_kernels_end(sparc_handle *h, double *kc_dev, double *kpsi_dev, double *kvx_dev,
                      double *kvz_dev, double *hess_dev);

/* Exact discrete adjoint.  No counterpart in the reference: its adjoint
 * (driver.f90:144-147, kernels.f90:286-355) is the continuous recipe, a
 * time-reversed run with -v0 correlated with the forward snapshots; this is
 * the transpose of the discrete step itself (step.f90:43-119 with
 * MP_QUIET_PML_2D_DISPL, physics2d.f90:943-1094, and FILTER_VARS_2D,
 * :509-556), for gradients that finite differences of the solver reproduce.
 * For a forward handle whose sources are set for maxtime: the gradient of
 * J = sum_{s,k,x} gvz[s][k][x] vz[s][k][x] (vz the traces sparc_run writes,
 * nt = maxtime/steps + 1 outputs; gvz device [S][nt][nx], e.g. dL/dvz of any
 * misfit L) over the run from the zero state.  The forward trajectory is
 * recomputed from a checkpoint every `interval` steps (0: one segment, every
 * state of the run kept).  Outputs (device [S][nz][nx] per source, each may
 * be NULL): gpsi = dJ/dpsi (model_psi units; needs the flow from psi),
 * gc = dJ/dc (cm/s); gcoef [3][S][nz][nx] = dJ/d(c^2 rho0), dJ/dv0x,
 * dJ/dv0z in code units.  The c-dependence of the PML coefficients (rows 1
 * and nz-5, column 1) is not part of gc. */
int sparc_discrete_gradient(sparc_handle *h, int maxtime, const double *gvz_dev, int nt, int interval,
                            double *gpsi_dev, double *gc_dev, double *gcoef_dev);
/* The transpose of sparc_ddz with ibc 1..3 (kept rows at 0), for nbatch
 * [nz][nx] fields; for ibc 3, presetN_bar (device [nbatch][nx], may be NULL)
 * gets the adjoint of the kept last row. */
int sparc_ddz_T(sparc_handle *h, const double *in_dev, double *out_dev, int nbatch, int ibc,
                double *presetN_bar_dev);

/* Instrumentation (no reference counterpart; the reference prints per-step
 * MPI_WTIME, driver.f90:355-381): average device time in ms of one launch of
 * each pass, timed with HIP events recorded around every launch of reps real
 * time steps (t = 0..reps-1 from a zero state, sources as set) on the
 * handle's stream: ms_out[0] column pass, [1] last-stage column pass (with
 * the z filter), [2] row pass, [3] first row pass (with the pending v_x
 * filter), [4] last-stage row pass.  ms_out holds 5 values.  Resets the state
 * before and after. */
int sparc_profile_kernels(sparc_handle *h, int reps, double *ms_out);

#ifdef __cplusplus
}
#endif
#endif
