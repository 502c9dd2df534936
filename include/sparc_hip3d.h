/*
 * sparc_hip3d.h -- C ABI of the MI355X (gfx950) implementation of SPARC's 3D
 * time-domain solver with flows (option 16: MP_FLOWS_PML_3D_DISPL,
 * displacement.f90:377-477), the 3D counterpart of sparc_hip.h.
 *
 * The reference runs this path as the same `sparc SS LL` program with 3D
 * params.i values (step.f90:43-119 dispatches option 16).  Each entry point
 * below replaces one piece of it; the comment names the reference code.  Where
 * the reference cannot run the path as written (a forward source that kernel
 * mode never sets, driver.f90:358-360; a flow derived from a 2D psi only,
 * driver.f90:94-126) the entry point says what it does instead.
 *
 * Memory and errors as in sparc_hip.h: "host" arrays are read during the
 * call, "device" arrays are HIP device pointers; layouts are C row-major,
 * fields [nz][ny][nx] (the reference's column-major (nx,ny,nz)); functions
 * return SPARC_OK or a negative SPARC_E_* code, and sparc3d_last_error()
 * gives the message.
 */
#ifndef SPARC_HIP3D_H
#define SPARC_HIP3D_H

#include "sparc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Runtime values of the params.i constants of a 3D build (nx, ny, nz,
 * xlength, ylength, timestep, cadences, PML rows, heights, source
 * frequencies), the background model file rows (solar_data,
 * initialize.f90:743-747) and an optional 3D sound speed. */
typedef struct sparc3d_config {
    int nx, ny, nz;           /* nx even, ny any, both in [16, 1024] (FFT for powers of two, circulant form otherwise); nz >= 16 */
    int npmlbot, npmltop;     /* z-PML rows, each >= 5 (the z filter reaches 5 rows) */
    int adjoint;              /* 0: forward, 1: adjoint */
    double xlength, ylength;  /* cm */
    double timestep;          /* s */
    double cadforcing;        /* s */
    double outputcad;         /* s */
    double obsheight;         /* cm */
    double excitdep;          /* cm */
    double nupeak, nuwidth;
    const double *model;      /* host [nz][6]: z/Rsun, c, rho, p, g, Gamma (cgs) */
    const double *c3d;        /* host [nz][ny][nx] sound speed (cm/s) or NULL */
} sparc3d_config;

typedef struct sparc3d_handle sparc3d_handle;

/* Initialize_all + solar_data + mp_Initialize_RHS (physics.f90:34-404: e_rad,
 * o_rad, source_dep, d rho0/dz, d p0/dz, the z-PML profiles) + INITIALIZE_STEP
 * (step.f90:24-41). */
int sparc3d_create(const sparc3d_config *cfg, sparc3d_handle **out);
int sparc3d_destroy(sparc3d_handle *h);
const char *sparc3d_last_error(const sparc3d_handle *h);
/* Queue the handle's work on this HIP stream (default stream priority only). */
int sparc3d_set_stream(sparc3d_handle *h, void *stream);

/* Background flow (v0_x, v0_y, v0_z), host [nz][ny][nx] in cm/s, as the
 * compute_data branch reads true_vx/true_vz in m/s (driver.f90:65-73); an
 * adjoint handle reverses it (driver.f90:144-147).  NULL, NULL, NULL: no flow. */
int sparc3d_set_flow(sparc3d_handle *h, const double *vx, const double *vy, const double *vz);

/* Forward source: forcing(x, y) = vr(x, y) * fwdsource(t+1), injected as
 * forcing * source_dep(k) on rows |k - e_rad| <= 4 (displacement.f90:459-461),
 * with fwdsource from FORWARD_SOURCE (driver.f90:1120-1145).  ysrc_mm NaN: vr
 * as DETERMINE_STATUS builds it, the x profile at xsrc repeated in every y row
 * (driver.f90:536-553, a line source); otherwise the same profile in the
 * horizontal distance to (xsrc, ysrc) (a point source).  Sets maxtime. */
int sparc3d_set_forward_source(sparc3d_handle *h, double xsrc_mm, double ysrc_mm, int maxtime);
/* Adjoint source (DETERMINE_STATUS adjoint branch + LAGRANGE_INTERP,
 * driver.f90:584-627, physics.f90:678-737): device [nt][ny][nx] slices at the
 * forcing cadence.  Sets maxtime. */
int sparc3d_set_adjoint_source(sparc3d_handle *h, const double *vr_dev, int nt, int maxtime);

/* Zero the state and PML variables. */
int sparc3d_reset(sparc3d_handle *h);
/* One STEP (step.f90:43-119) at time index t: five RK stages of
 * MP_FLOWS_PML_3D_DISPL, the update, and FILTER_VARS_3D_HORIZ_PML
 * (physics.f90:836-882) when mod(t+1, 2) == 0. */
int sparc3d_step(sparc3d_handle *h, int t);
/* TIMESTEPPING (driver.f90:337-419) for t = 0..maxtime from the current
 * state: vz at o_rad after every step with t % steps == 0 goes to device
 * vz_out [nout][ny][nx] (nout = maxtime/steps + 1; may be NULL). */
int sparc3d_run(sparc3d_handle *h, int maxtime, double *vz_out);

/* Time indices t0..t1 from the current state (a restart from a checkpoint,
 * driver.f90:236-237); the vz output of time t (t % steps == 0) goes to slot
 * t/steps - k0 of device vz_out [nslots][ny][nx] when it exists. */
int sparc3d_run_range(sparc3d_handle *h, int t0, int t1, double *vz_out, int k0, int nslots);
/* Full solver state (a and the z-PML variables) to / from a device buffer of
 * sparc3d_checkpoint_size bytes: in-HBM checkpoints instead of the
 * reference's dumped full state. */
int sparc3d_checkpoint_size(const sparc3d_handle *h, size_t *bytes);
int sparc3d_save_checkpoint(sparc3d_handle *h, void *dev);
int sparc3d_load_checkpoint(sparc3d_handle *h, const void *dev);
/* The partial snapshot of the current time (WRITE_OUT_PARTIAL_STATE,
 * all_modules.f90:750-784) to device memory: forward [9][nz][ny][nx] xi_x,
 * xi_y, xi_z, v_x, v_y, v_z and the last stage's RHS of v_x, v_y, v_z
 * (acc); adjoint [3][nz][ny][nx] v_x, v_y, v_z. */
int sparc3d_get_snapshot(sparc3d_handle *h, double *dev);

/* PRODUCE_KERNELS with its y terms (kernels.f90:13-612 with test_in_2d
 * false): begin zeroes the sums; add accumulates one forward snapshot (9
 * fields, device) with the adjoint snapshot (3 fields, device) it pairs with
 * (the caller pairs forward output k with adjoint output nt-1-k, :215, :256);
 * end writes Kc, K_psi, Kvx, Kvy, Kvz and the hessian, [nz][ny][nx] device,
 * in the reference's units (:498-522, 563, 579) on kernel_back's background
 * (initialize.f90:937-979); K_psi is the y component of CURL_KERN
 * (:815-836, bcz = 0) of (Kvx, 0, Kvz) scaled as :511-513.  Uses the
 * handle's 3D sound speed for c2. */
int sparc3d_kernels_begin(sparc3d_handle *h);
int sparc3d_kernels_add(sparc3d_handle *h, const double *fwd, const double *adj);
int sparc3d_kernels_end(sparc3d_handle *h, double *kc, double *kpsi, double *kvx, double *kvy, double *kvz,
                        double *hess);

/* State a [6][nz][ny][nx] (xi_x, xi_y, xi_z, v_x, v_y, v_z), device. */
int sparc3d_get_state(sparc3d_handle *h, double *a_dev);
int sparc3d_set_state(sparc3d_handle *h, const double *a_dev);
/* z-PML variables [2][npmlbot+npmltop][ny][nx] (d xi_z/dz, d p/dz), device. */
int sparc3d_get_pml(sparc3d_handle *h, double *pml_dev);
/* The RHS (scr) of the last stage [6][nz][ny][nx], device. */
int sparc3d_get_scr(sparc3d_handle *h, double *scr_dev);

/* Single operators of ddxyz (derivatives.f90:175-305) on device [n][ny][nx]
 * blocks: d/dx and d/dy by FFT (Nyquist dropped), d/dz by dbyd2 (ibc 0..3)
 * times stretch, on n consecutive [nz][ny][nx] fields for ddz. */
int sparc3d_ddx(sparc3d_handle *h, const double *in, double *out, int nplanes);
int sparc3d_ddy(sparc3d_handle *h, const double *in, double *out, int nplanes);
int sparc3d_ddz(sparc3d_handle *h, const double *in, double *out, int nfields, int ibc);

/* Integer / double facts of the set-up: "nx", "ny", "nz", "e_rad", "o_rad",
 * "steps", "npml", "maxtime"; "deltat", "stretchx", "stretchy". */
int sparc3d_query_int(const sparc3d_handle *h, const char *name, int *v);
int sparc3d_query_double(const sparc3d_handle *h, const char *name, double *v);

/* Instrumentation: reset the state, run `reps` steps from t = 0 with HIP
 * events around every kernel (the state is left after those steps) and
 * return the mean milliseconds per launch of each pass kind in ms_out[8]:
 * (unused), x-derivative, y-derivative, z pass of xi_z (div, p, rho'), z
 * pass of v_x, v_y, v_z and p (the RHS and the next stage's input), z filter
 * (every other step), and two unused slots. */
int sparc3d_profile(sparc3d_handle *h, int reps, double *ms_out);

#ifdef __cplusplus
}
#endif

#endif /* SPARC_HIP3D_H */
