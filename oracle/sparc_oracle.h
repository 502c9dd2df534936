/*
 * sparc_oracle.h -- CPU restatement of SPARC's 2D option-17 solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in sparc_flows_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * It restates, in plain single-threaded C, the reference Fortran files
 *   step.f90, physics2d.f90 (init, FILTER_VARS_2D, MP_QUIET_PML_2D_DISPL),
 *   derivatives.f90 (ddx, ddz), dbyd2.f, mtridag.f, initialize.f90
 *   (grid, solar_data, stretch, kernel_back), driver.f90 (flow from psi,
 *   FORWARD_SOURCE, DETERMINE_STATUS source profile, TIMESTEPPING),
 *   physics.f90 (LAGRANGE_INTERP) and kernels.f90 (PRODUCE_KERNELS),
 * keeping the reference's float32-widened literals.  Each function in
 * sparc_oracle.c cites the reference lines it follows.
 *
 * Layout: every 2D field is stored [nz][nx] (x fastest), the same memory
 * order as the reference's column-major a(nx,1,nz).  Index k (0-based) here
 * is Fortran index k+1.
 */
#ifndef SPARC_ORACLE_H
#define SPARC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Runtime copy of the params.i constants that the option-17 path reads. */
typedef struct orc_params {
    int nx, nz;            /* grid (ny = 1) */
    int npmlbot, npmltop;  /* vertical PML widths */
    int npmlhor;           /* horizontal PML width */
    int adjoint;           /* 0 = forward ('spectral'), 1 = adjoint */
    double xlength;        /* cm */
    double timestep;       /* s */
    double cadforcing;     /* s */
    double outputcad;      /* s */
    double obsheight;      /* cm */
    double excitdep;       /* cm */
    double nupeak;         /* Hz, as stored by the reference (float32-widened) */
    double nuwidth;        /* as stored by the reference (float32-widened) */
} orc_params;

typedef struct orc_model orc_model;

/* Background model rows: bg[k*6 + {0..5}] = z/Rsun, c, rho, p, g, Gamma (cgs).
 * c2d: optional [nz][nx] sound speed in cm/s (NULL = use the 1D column).
 * psi: optional [nz][nx] stream function in Mm (model_psi file; NULL = no flow). */
orc_model *orc_create(const orc_params *p, const double *bg, const double *c2d,
                      const double *psi);
void orc_destroy(orc_model *m);

/* Flow in m/s [nz][nx] as the compute_data branch reads it (true_vx/true_vz,
 * driver.f90:65-73), replacing the flow from psi; reversed for an adjoint. */
int orc_set_flow(orc_model *m, const double *vx_ms, const double *vz_ms);

/* Forward source at x_src (Mm), fwdsource table of length maxtime+8. */
int orc_set_forward_source(orc_model *m, double xloc, int maxtime);
/* Separable forward source: RHS v_z += vr[x] * series[t] * zprof[z]. */
int orc_set_forward_profile(orc_model *m, const double *vr, const double *zprof,
                            const double *series, int nseries, int maxtime);
/* Adjoint source vr[nt][nx], nt = forcing_length slices. */
int orc_set_adjoint_source(orc_model *m, const double *vr, int nt, int maxtime);

/* One call of STEP at m->time = t. */
void orc_step(orc_model *m, int t);
/* TIMESTEPPING: i = 0..maxtime.  vz_out[nout][nx] (nout = maxtime/steps+1)
 * gets a(:,o_rad,vz) when mod(i,steps)==0.  snaps (may be NULL) gets per
 * output: forward {xix, xiz, vx, vz, acc_x, acc_z}, adjoint {vx, vz, acc_x,
 * acc_z}, each [nz][nx]. Returns the number of outputs. */
int orc_run(orc_model *m, int maxtime, double *vz_out, double *snaps);

/* State access: fields {xi_x, xi_z, v_x, v_z} [4][nz][nx]. */
void orc_get_state(const orc_model *m, double *a4);
void orc_set_state(orc_model *m, const double *a4);
/* PML auxiliaries: pmlvars [2][nzpml][nx], pmlvarsx [2][nz][2*npmlhor]. */
void orc_get_pml(const orc_model *m, double *pmlvars, double *pmlvarsx);
void orc_get_scr(const orc_model *m, double *scr4);

/* Derived quantities, for tests. */
int orc_info_int(const orc_model *m, const char *name);
double orc_info_double(const orc_model *m, const char *name);
/* Copies a named 1D/2D array (z, stretch, unstretch, g, rho0, rhoinv,
 * c2rho0, gradp0_z, gradrho0_z, v0x, v0z, az, bzpml, axpml, bxpml,
 * fwdsource, vr, c2). Returns element count, or -1. */
int orc_get_array(const orc_model *m, const char *name, double *out, int cap);

/* Stand-alone operators. */
void orc_dbyd2(double *A, const double *B, int N1, int N2, int ibc);
void orc_ddx(const orc_model *m, const double *var, double *dvar, int nrows);
void orc_ddz(const orc_model *m, const double *var, double *dvar, int ibc);
/* One RHS evaluation (MP_QUIET_PML_2D_DISPL) on an explicit stage state.
 * temp4 [4][nz][nx], psipml [2][nzpml][nx], psipmlx [2][nz][2*npmlhor];
 * outputs scr4, scrpml, scrpmlx. */
void orc_rhs(orc_model *m, int t, const double *temp4, const double *psipml,
             const double *psipmlx, double *scr4, double *scrpml, double *scrpmlx);
void orc_filter_vars_2d(const orc_model *m, double *a4);
void orc_filter_var_z(int nx, int nz, double *var);
/* LAGRANGE_INTERP forcing at time index t into forcing[nx]. */
void orc_lagrange(orc_model *m, int t, double *forcing);
/* FORWARD_SOURCE: fills fwd[nt]. */
void orc_forward_source_series(double timestep, double nupeak, double nuwidth,
                               int nt, double *fwd);
/* Complex FFT used by ddx (sign -1 forward, +1 inverse, unnormalized). */
void orc_fft(double *re, double *im, int n, int sign);

/* Exact discrete adjoint (transpose of orc_step; not a reference routine).
 * orc_ddz_T: out = Dz_ibc^T in (stretch included); for ibc 3, presetN_bar[nx]
 * (may be NULL) gets the adjoint of the kept last row.  orc_ddx_T = -orc_ddx.
 * orc_rhs_T: transpose of orc_rhs without its source; pbar (may be NULL)
 * gets the adjoint of p.  orc_filter_vars_2d_T: transpose of the z filter.
 * orc_discrete_gradient: dJ/d(c^2 rho0), dJ/dv0x, dJ/dv0z [nz][nx] of
 * J = sum_k <gvz[k], vz_out[k]> over a forward run from the zero state;
 * lam0 (may be NULL) gets the adjoint of the initial state [4][nz][nx]. */
void orc_ddz_T(orc_model *m, const double *in, double *out, int ibc, double *presetN_bar);
void orc_ddx_T(const orc_model *m, const double *in, double *out, int nrows);
void orc_rhs_T(orc_model *m, const double *Rb4, const double *Rbpz, const double *Rbpx,
               double *Yb4, double *Ybpz, double *Ybpx, double *pbar);
void orc_filter_vars_2d_T(const orc_model *m, double *a4);
int orc_discrete_gradient(orc_model *m, int maxtime, const double *gvz, double *g_c2rho0,
                          double *g_v0x, double *g_v0z, double *lam0);

/* PRODUCE_KERNELS (kernels.f90:13-612) for one source.
 * fwd: [nt_kern][6][nz][nx] forward snapshots ordered by label 0,steps,...
 * adj: [nt_kern][4][nz][nx] adjoint snapshots ordered by label 0,steps,...
 * (the pairing adj_end - (k-1)*steps is done inside).
 * Outputs [nz][nx]: kc, kpsi, kvx, kvz, hess. */
int orc_kernels(const orc_params *p, const double *bg, const double *c2d,
                int nt_kern, const double *fwd, const double *adj,
                double *kc, double *kpsi, double *kvx, double *kvz, double *hess);

#ifdef __cplusplus
}
#endif
#endif
