/*
 * ps3d_oracle.h -- CPU restatement of the ps3D model time step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only to check or time the reference algorithm.  The product
 * library (ps3d_amd/libps3d.so) never links or calls it.
 *
 * Field layout follows for_src/main_module.f90:118-131 on one rank:
 * a(-1:nx+2, -1:ny+2, -1:nz+2), i fastest (Fortran column-major), halo onx=2.
 * Tendency arrays du..db hold 3 such slots back to back (main_module.f90:124-127).
 * FS is complex (nx/2+1, ny, nz), stored as interleaved (re, im) doubles,
 * kx fastest (main_module.f90:141 with fstart=(1,1,1), fend=(nx/2+1,ny,nz)).
 */
#ifndef PS3D_ORACLE_H
#define PS3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int32_t nx, ny, nz;
    int32_t enable_nonlinear;
    int32_t enable_vertical_boundaries;
    int32_t enable_AB_3_order;
    double Lx, Ly, Lz, dx, dy, dz, dt;
    double f0, N0, Ro, dsqr, eps_ab;
    double Ah, Av, Kh, Kv, Ahbi, Avbi, Khbi, Kvbi;
} orc_params;

typedef struct orc_model orc_model;

/* field ids for orc_ptr */
enum {
    ORC_U = 0, ORC_V, ORC_W, ORC_B, ORC_P,
    ORC_DU, ORC_DV, ORC_DW, ORC_DB,          /* 3 slots each */
    ORC_FLUX_EAST, ORC_FLUX_NORTH, ORC_FLUX_TOP,
    ORC_FS,                                  /* complex, interleaved */
    ORC_KSQR,                                /* (nx/2+1, ny, nz) or (nx/2+1, ny, 1) */
    ORC_KX, ORC_KY, ORC_KZ,
    ORC_NFIELDS
};

orc_model *orc_create(const orc_params *prm);
void orc_destroy(orc_model *m);
double *orc_ptr(orc_model *m, int field);
int64_t orc_field_len(const orc_model *m, int field);   /* in doubles */

/* advection.f90:2-48 / :53-184 (tendency written into the given slot arrays) */
void orc_buoyancy_advection_2nd(orc_model *m, double *db_loc);
void orc_momentum_advection_2nd(orc_model *m, double *du_loc, double *dv_loc, double *dw_loc);
/* mixing.f90:225-274 / :96-158 / :164-217, in place on p1 */
void orc_harmonic(orc_model *m, double *p1, double Hmix, double Vmix);
void orc_biharmonic_h(orc_model *m, double *p1, double Hmix);
void orc_biharmonic_v(orc_model *m, double *p1, double Vmix);
/* parallel_mpi.f90:70-81 single-rank periodic branch (z, then y, then x) */
void orc_border_exchg_3d(orc_model *m, double *a);
/* integrate.f90:12-111: everything before the call to pressure_solver */
void orc_integrate_pre(orc_model *m, int itt, int tau, int taum1, int taum2);
/* dt, Ro and the biharmonic coefficients for the next steps */
void orc_set_dynamic(orc_model *m, double dt, double Ro, double Ahbi, double Avbi, double Khbi, double Kvbi);
/* set_forcing (integrate.f90:82) hook of orc_integrate_pre; NULL: empty */
typedef void (*orc_forcing_fn)(orc_model *m, void *user, int itt, int tau);
void orc_set_forcing(orc_model *m, orc_forcing_fn fn, void *user);
/* for_src/config.f90:224-237 wavemaker, whole arrays; re/im: 4 entries */
void orc_wavemaker(orc_model *m, int itt, double omega, const double *const *re, const double *const *im);
/* lee_wave1.f90:239-314 background-flow advection into slot tau (b_back NULL: 0) */
void orc_background_flow(orc_model *m, int tau, const double *u_back, const double *b_back);
/* steady body force f[0..3] added to du..db at slot tau (NULL entries: none) */
void orc_body_force(orc_model *m, int tau, const double *const *f);
/* both as the model's hook (arrays stay the caller's) */
void orc_set_background(orc_model *m, const double *u_back, const double *b_back, const double *const *f);
/* orc_wavemaker as the model's hook (amplitude arrays stay the caller's) */
void orc_set_wavemaker(orc_model *m, double omega, const double *const *re, const double *const *im);
/* pressure_solver.f90:2-120 */
void orc_pressure_solver(orc_model *m);
/* integrate.f90:2-115 */
void orc_integrate(orc_model *m, int itt, int tau, int taum1, int taum2);
/* main.f90:139-161 loop body without diagnostics; state = {itt, tau, taum1, taum2} */
void orc_run(orc_model *m, int nsteps, int32_t *state);

/* my_fft.f90 transforms, including their scalings.  qp: real (nx,ny,nz)
   interior array; qs: complex (nx/2+1,ny,nz) interleaved. */
void orc_my_fft_2d(orc_model *m, const double *qp, double *qs);
void orc_my_ifft_2d(orc_model *m, const double *qs, double *qp);
void orc_my_fft_3d(orc_model *m, const double *qp, double *qs);
void orc_my_ifft_3d(orc_model *m, const double *qs, double *qp);
/* the spectral step of pressure_solver.f90:30-71 applied to FS in place */
void orc_spectral_solve(orc_model *m, double *fs);

/* pressure_solver.f90:126-162; d and x complex interleaved, length n */
void orc_solve_tridiag(const double *a, const double *b, const double *c,
                       const double *d, double *x, int n);

/* plain complex DFT helper (any n) used by the transforms, exposed for tests:
   sign = -1 forward (e^{-i...}), +1 backward; unnormalised. */
void orc_cfft(int n, double *data, int sign);

/* wavenumber_initialisation (my_fft.f90:87-156, ksqr part) with hat_ksqr
   (eigenvectors.f90:224-229) */
void orc_wavenumbers(orc_model *m);
double orc_hat_ksqr(double k, double dx);

#ifdef __cplusplus
}
#endif
#endif
