/*
 * ps3d.h -- C-ABI of the MI355X (gfx950) implementation of the ps3D model
 * time step.  Plain C types only; every function returns 0 on success or a
 * PS3D_ERR_* code, with a message available from ps3d_last_error().  No C++
 * exception crosses this boundary.
 *
 * Field layout (the contract with the Fortran driver's main_module,
 * for_src/main_module.f90:118-131 on one rank): every 3-D field is the full
 * array a(-1:nx+2, -1:ny+2, -1:nz+2) including the onx=2 halo, Fortran
 * column-major, i fastest.  Tendencies du..db are (.., 1:3) arrays of three
 * such slots.  FS is complex(8) (nx/2+1, ny, nz), kx fastest.  ksqr is
 * (nx/2+1, ny, 1) with a rigid lid, else (nx/2+1, ny, nz).
 *
 * Ownership: the caller owns host arrays; the library owns device buffers
 * and never frees caller memory.  One model = one HIP stream; calls on one
 * model must not overlap from several host threads.
 *
 * Reference routines are cited as file:line of the ps3D sources (for_src/).
 */
#ifndef PS3D_H
#define PS3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PS3D_OK            0
#define PS3D_ERR_ARG       1   /* invalid argument / unsupported size */
#define PS3D_ERR_HIP       2   /* HIP runtime error */
#define PS3D_ERR_STATE     3   /* call not valid in the model's state */
#define PS3D_ERR_INTERNAL  4

/* precision of a model */
#define PS3D_FP64   0   /* plain FP64 path: the drop-in */
#define PS3D_MIXED  1   /* FP32 stencils/FFTs, FP64 solve/accumulation */

/* field ids for ps3d_set_field / ps3d_get_field */
#define PS3D_U           0
#define PS3D_V           1
#define PS3D_W           2
#define PS3D_B           3
#define PS3D_P           4
#define PS3D_DU          5   /* (.., 3): all three tendency slots */
#define PS3D_DV          6
#define PS3D_DW          7
#define PS3D_DB          8
#define PS3D_FS          9   /* complex (nx/2+1, ny, nz) */
#define PS3D_KSQR       10

/* friction operators, mixing.f90 */
#define PS3D_HARMONIC      0   /* harmonic(p1,Hmix,Vmix)   mixing.f90:225-274 */
#define PS3D_BIHARMONIC_H  1   /* biharmonic_h(p1,Hmix)    mixing.f90:96-158  */
#define PS3D_BIHARMONIC_V  2   /* biharmonic_v(p1,Vmix)    mixing.f90:164-217 */

/* Parameters of main_module (main_module.f90:16-59) that the step reads.
   Logicals are 0/1.  Every value is taken as given: the library derives
   nothing the driver has already computed (so single-precision literals of
   a config survive exactly). */
typedef struct ps3d_params {
    int32_t nx, ny, nz;
    int32_t enable_nonlinear;
    int32_t enable_vertical_boundaries;
    int32_t enable_AB_3_order;
    double Lx, Ly, Lz, dx, dy, dz, dt;
    double f0, N0, Ro, dsqr, eps_ab;
    double Ah, Av, Kh, Kv, Ahbi, Avbi, Khbi, Kvbi;
    int32_t precision;            /* PS3D_FP64 or PS3D_MIXED */
    int32_t reserved[7];          /* zero */
} ps3d_params;

typedef struct ps3d_model ps3d_model;

/* Last error message of the calling thread ("" if none); copies at most
   len-1 bytes plus a terminating NUL into buf; returns the message length. */
int ps3d_last_error(char *buf, int len);

/* Library version string, e.g. "ps3d-amd 0.1 gfx950". */
const char *ps3d_version(void);

/* Allocate a model and its device mirror.  Replaces allocate_main_module
   (main_module.f90:112-144) and wavenumber_initialisation
   (my_fft.f90:87-156) for the step.  ksqr: the driver's ksqr array
   (main_module.f90:100) or NULL, in which case the library computes it with
   hat_ksqr (eigenvectors.f90:224-229) exactly as my_fft.f90:130-155 does.
   stream: a hipStream_t to run on, or NULL for a library-owned stream.
   Sizes: nx, ny even and >= 4 (nz >= 2 with a rigid lid, even and >= 4
   otherwise), every size a product of 2, 3 and 5 (the mixed-radix FFT),
   main.f90:45-55 being the reference's own size check. */
int ps3d_create(const ps3d_params *prm, const double *ksqr, void *stream,
                ps3d_model **out);
int ps3d_destroy(ps3d_model *m);

/* Change the parameters that callers of integrate change between steps:
   main_module's dt, Ro and friction coefficients as diag_opt_balance.f90
   sets them around its backward and forward integrations (:190-205, :259-
   292: dt = -|dt|, Ro ramped, Ahbi..Kvbi zeroed), and eps_ab, f0, N0,
   enable_nonlinear, enable_AB_3_order.  The grid, domain, spacing, dsqr,
   boundary form and precision must equal the model's (PS3D_ERR_ARG
   otherwise).  Takes effect at the next step. */
int ps3d_set_params(ps3d_model *m, const ps3d_params *prm);

/* Host <-> device copies of whole main_module arrays (halo included).
   Synchronous with respect to the host.  ps3d_set_field of PS3D_KSQR
   replaces the wavenumber table (and the solver factors derived from it). */
int ps3d_set_field(ps3d_model *m, int field, const double *host);
int ps3d_get_field(ps3d_model *m, int field, double *host);
/* number of doubles of a field's host array (complex counts as 2) */
int ps3d_field_size(const ps3d_model *m, int field, int64_t *n);

/* ---- the step ------------------------------------------------------- */

/* integrate.f90:2-115 for a config whose set_forcing is empty: tendencies,
   friction, Adams-Bashforth, halo exchange, pressure_solver.  itt and the
   tau slots (1..3) are the driver's (main.f90:112-113, 156-160). */
int ps3d_integrate(ps3d_model *m, int itt, int tau, int taum1, int taum2);

/* main.f90:139-161 without diagnostics: nsteps calls of integrate with the
   tau rotation.  state = {itt, tau, taum1, taum2}, updated on return.
   The steps are queued back to back on the model's stream and the call
   returns when they are queued.  use_graph = 1: past the first two steps
   (and the mixed form's cold start and correction steps) the steps run as
   replays of a hipGraph of six steps (the period of the tau rotation, the
   buffer parity and the mixed form's increment ring), captured from the
   model's state on first use and again whenever that state's signature
   (buffers, constants, launch shapes, options) changes; the same bits as
   queued steps.  Steps with a forcing hook or the wavemaker, and steps
   whose kernels are being timed, are always queued directly. */
int ps3d_run(ps3d_model *m, int nsteps, int32_t *state, int use_graph);

/* wait for the model's stream */
int ps3d_synchronize(ps3d_model *m);

/* Whether an unforced FP64 step forms the tendencies (integrate.f90:12-54)
   inside the fused friction + Adams-Bashforth kernel (on = 1) or in a
   kernel of their own that writes slot tau for the friction kernel to read
   (on = 0, the default).  Both give the same bits; the fused form moves a
   third fewer bytes (it reads the state once and never reads slot tau back)
   but is limited by its instruction issue and runs slower on MI355X.  No
   counterpart in the reference, whose loops are fixed: a switch for tests
   and measurements. */
int ps3d_set_fused_tendency(ps3d_model *m, int on);

/* Radix plans of a mixed-form model's FP32 transforms along x, y and z
   (plans[0..2]): 0 = radix-4 passes first, then 2, 3, 5 (the plan of every
   FP64 transform, which the oracle follows bit for bit); 1 = radix-8 passes
   first (fewer passes through LDS; lengths without a factor 8 run plan 0
   either way).  ps3d_tune picks them per transform length by timing.  A
   FP64 model accepts only 0.  The reference leaves the factorisation to
   FFTW's run-time planner behind P3DFFT (my_fft.f90:5-82, p3dfft_setup);
   this is its counterpart for the device transforms. */
int ps3d_set_fft_plan(ps3d_model *m, const int32_t plans[3]);
int ps3d_get_fft_plan(ps3d_model *m, int32_t plans[3]);

/* ---- forcing: the set_forcing hook of integrate.f90:82 ----------------
   A forcing hook runs where the reference calls set_forcing: after the
   friction block (:59-78) and before the Adams-Bashforth update (:87-102),
   inside ps3d_integrate and ps3d_run.  fn is called on the host with the
   model's stream, the driver's itt and tau.  It may launch its own kernels
   on that stream on the device arrays (ps3d_device_ptr): the device forcing
   path; or synchronise the stream and go through host copies
   (ps3d_get_field / ps3d_set_field: u, v, w, b and du..db at tau hold the
   post-friction state there): the host round trip.  A nonzero return stops
   the step with PS3D_ERR_ARG.  While a hook or the wavemaker is set the
   step takes the unfused friction kernels.  FP64 form only; fn = NULL
   removes the hook.  Replaces the config's set_forcing
   (the for_config files, e.g. lee_wave1.f90:232-331) for GPU runs. */
typedef int (*ps3d_forcing_fn)(ps3d_model *m, void *user, int itt, int tau, void *stream);
int ps3d_set_forcing(ps3d_model *m, ps3d_forcing_fn fn, void *user);

/* Device address of a field array (Fortran layout with halo, as the host
   array): u, v, w, b, p (slot ignored) or du..db at slot 1..3.  u..b move
   between two buffers from step to step: query inside the hook. */
int ps3d_device_ptr(ps3d_model *m, int field, int slot, void **ptr);

/* The wavemaker forcing of for_src/config.f90:224-237 (lee_wave1.f90:320-323
   with omega = -omega_e) on the device, before any hook:
     q = q + dt*(cos(omega*t)*q_r + sin(omega*t)*q_i),  t = dt*itt,
   for q = u, v, w, b.  re[f], im[f] (f = 0..3: u, v, w, b) are host arrays
   of a field's layout, copied to the device; a NULL pair leaves q unforced.
   cos and sin are evaluated on the host, as the Fortran driver does. */
int ps3d_set_wavemaker(ps3d_model *m, double omega, const double *const re[4], const double *const im[4]);
int ps3d_clear_wavemaker(ps3d_model *m);

/* Advection by a background flow (u_back, 0, 0) and of a background
   buoyancy b_back, lee_wave1.f90:239-314, on the device: each step adds
   -Ro * (flux divergences) to du..db at the tau slot, before any other
   forcing.  u_back, b_back: host arrays of a field's layout (with halo),
   copied; b_back NULL: zero; u_back NULL: off.  Both forms: the terms are
   formed from the post-friction state, as the reference's set_forcing sees
   it (the mixed form then runs friction, forcing and Adams-Bashforth as
   three passes instead of its fused kernel). */
int ps3d_set_background_flow(ps3d_model *m, const double *u_back, const double *b_back);

/* A steady body force: f[q] (q = 0..3: du, dv, dw, db; host arrays of a
   field's layout, copied) added to the tendency slot tau each step, after
   the background advection.  NULL entries (or f = NULL): none. */
int ps3d_set_body_force(ps3d_model *m, const double *const f[4]);

/* ---- phases, for drivers whose set_forcing does work ------------------
   integrate.f90 in pieces: tendency (:12-54), friction (:59-78), then the
   driver's set_forcing on host mirrors, then ab_update (:87-102) with the
   halo exchange (:106-111), then pressure_solver (:113). */
int ps3d_tendency(ps3d_model *m, int tau);
/* advection.f90:2-48 into db(:,:,:,tau) interior only */
int ps3d_buoyancy_advection(ps3d_model *m, int tau);
/* advection.f90:53-184 into du,dv,dw(:,:,:,tau) interior only */
int ps3d_momentum_advection(ps3d_model *m, int tau);
/* the friction block integrate.f90:59-78 */
int ps3d_friction(ps3d_model *m);
/* one friction call on field (PS3D_U..PS3D_B), op PS3D_HARMONIC.. */
int ps3d_mixing(ps3d_model *m, int op, int field, double hmix, double vmix);
/* integrate.f90:87-111 */
int ps3d_ab_update(ps3d_model *m, int itt, int tau, int taum1, int taum2);
/* border_exchg_3D (parallel_mpi.f90:70-81), one rank, on a device field */
int ps3d_border_exchg(ps3d_model *m, int field);
/* pressure_solver.f90:2-120 */
int ps3d_pressure_solver(ps3d_model *m);

/* ---- building blocks, host buffers in and out (for tests and drivers) */

/* my_fft_2D / my_fft_3D (my_fft.f90:170-178, 197-205): qp real
   (nx,ny,nz) -> qs complex (nx/2+1,ny,nz) as interleaved doubles;
   three_d = 0 transforms x,y ('ffn'), 1 transforms x,y,z ('fft'). */
int ps3d_my_fft(ps3d_model *m, int three_d, const double *qp, double *qs);
/* my_ifft_2D / my_ifft_3D (my_fft.f90:159-167, 184-192) */
int ps3d_my_ifft(ps3d_model *m, int three_d, const double *qs, double *qp);
/* the spectral step of pressure_solver.f90:30-71 on complex qs in place:
   per-column solve_tridiag (pressure_solver.f90:126-162) and the zero-mode
   recurrence with a rigid lid, else the division by -ksqr */
int ps3d_spectral_solve(ps3d_model *m, double *qs);

/* ---- the FFT library of my_fft.f90, on one GPU -------------------------
   my_fft.f90 (:5-82, :159-252) reaches its transforms through P3DFFT's
   p3dfft_setup / p3dfft_get_dims / p3dfft_ftran_r2c / p3dfft_btran_c2r.
   One GPU holds the whole domain, so these are plain on-device transforms
   (the pencil transposes become LDS transposes inside the FFT kernels),
   with P3DFFT's semantics: unnormalised, e^{-ikx} forward; op is a string
   of three characters, 'f' = transform, 'n' = none, per dimension in
   execution order: forward "fft" (x, y, z) or "ffn" (x, y), backward "fft"
   or "nff" (y, x).  qp: real (nx, ny, nz); qs: complex (nx/2+1, ny, nz) as
   interleaved doubles, x fastest (P3DFFT's spectral layout on one rank).
   Host buffers; synchronous.  nx/2 and ny (and nz for a z transform) must
   be products of 2, 3 and 5.  ps3d_amd/fortran/p3dfft_gpu.f90 is the
   Fortran module p3dfft over these entry points. */
typedef struct ps3d_fft ps3d_fft;
int ps3d_fft_create(int nx, int ny, int nz, void *stream, ps3d_fft **out);
int ps3d_fft_destroy(ps3d_fft *f);
int ps3d_fft_forward(ps3d_fft *f, const char *op, const double *qp, double *qs);
int ps3d_fft_backward(ps3d_fft *f, const char *op, const double *qs, double *qp);

/* ---- launch-shape autotuner ------------------------------------------ */

/* Times the candidate launch shapes of this model's step on the live model
   (levels per work-group of the tendency and friction + AB kernels; the
   form of the Thomas solve: whole columns in registers, the real and
   imaginary parts on two lanes, columns per wavefront on the fly or with
   factor tables; in the mixed form also the radix plan of each FP32
   transform axis) and keeps the fastest; the choice is cached per (stage,
   grid, boundary, precision) in cache_path (NULL or "": none).
   next_tau is the slot the next step writes (it is used as scratch).  The
   model's state is unchanged and results do not depend on the shapes.  No
   counterpart in the reference (its loop shapes are fixed by the compiler). */
int ps3d_tune(ps3d_model *m, int next_tau, const char *cache_path, int reps);
/* the shapes in use, "stage shape value" lines; returns the text length */
int ps3d_tuning(ps3d_model *m, char *buf, int len);

/* ---- diagnostics ----------------------------------------------------- */

/* Normal-mode balancing of the model's current u, v, w, b after Warn et al.
   to third order: diag_balance (diag_balance.f90:89-362) with the discrete
   normal modes of eigenvectors.f90:6-218 and, under a rigid lid, the
   vertical-mode transforms of my_fft.f90:112-127, :211-252; FP64 on the
   device.  out[16] receives u0, v0, w0, b0, u1 .. b1, u2 .. b2, u3 .. b3
   (full arrays with halo, as module_diag_balance; NULL entries skipped).
   kx, ky, kz are main_module's wavenumbers (NULL: computed as
   wavenumber_initialisation does, my_fft.f90:87-110).  filter_width < 0:
   enable_diag_balance_filter off, else diag_balance_filter_width.  The
   model's fields, tendency slots and p are left as they were.  FP64 models. */
int ps3d_diag_balance(ps3d_model *m, int filter_width, const double *kx, const double *ky, const double *kz,
                      double *const *out);

/* Optimal balance, diag_opt_balance.f90:55-179 (FP64): from the model's
   current state, iterate {integrate backward n_end = int(period/dt)+1 Euler
   steps with Ro ramped to 0 and no biharmonic friction; project on the
   geostrophic mode (or average the linear model: time_average); integrate
   forward with Ro ramped back; restore the geostrophic base point (or the
   time-average boundary condition)} until the norm of the change
   (opt_balance_norm, :338-363) is below tol after at least two iterations.
   out[4] (nullable entries): host arrays u_bal, v_bal, w_bal, b_bal with
   halo; norms (nullable): max_iterations entries, norm_diff per iteration;
   *iterations: iterations run.  The model's state (fields, tendency slots,
   p), dt, Ro and friction are as before on return.  kx, ky, kz as for
   ps3d_diag_balance (NULL: my_fft.f90:87-110). */
typedef struct ps3d_opt_balance_params {
    double period;           /* opt_balance_period (s) */
    double average;          /* opt_balance_average (s) */
    double tol;              /* opt_balance_tol */
    int32_t max_iterations;  /* opt_balance_max_Itts */
    int32_t average_times;   /* opt_balance_average_times */
    int32_t time_average;    /* enable_diag_opt_time_ave */
    int32_t reserved;        /* 0 */
} ps3d_opt_balance_params;
int ps3d_opt_balance(ps3d_model *m, const ps3d_opt_balance_params *ob, const double *kx, const double *ky,
                     const double *kz, double *const out[4], double *norms, int32_t *iterations);

/* diagnose.F90:11-53 on the device: out[8] = KE, PE (per grid point), max
   |u_h|, max |w|, horizontal and vertical CFL, horizontal and vertical
   Peclet numbers; *nan_seen = 1 if |u_h| was NaN (diagnose.F90:42-45, where
   the reference aborts).  Fields stay on the device. */
int ps3d_diagnose(ps3d_model *m, double *out, int32_t *nan_seen);

/* The same monitors without any allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc), which ps3d_diagnose's first call on a model does.  u, v, w, b
   are caller-owned FP64 device arrays in the field layout above (halo
   included, nx x ny x nz from prm; prm's dt, dx, dz, dsqr, N0 and friction
   coefficients enter as in ps3d_diagnose); stream: a hipStream_t or NULL.
   The partial sums go to a fixed device scratch of 2049 records per GPU (one
   per work-group of at most 2048, which stride over any grid, plus the
   total), held until the process ends; grids with an extent above 2^20 are
   refused.  Synchronous: returns after the 8 values are on the host.  Calls
   that overlap on one GPU (from several threads or streams) race on the
   scratch.  Same bits as ps3d_diagnose on the same fields. */
int ps3d_diagnose_noalloc(const ps3d_params *prm, const double *u, const double *v, const double *w,
                          const double *b, void *stream, double *out, int32_t *nan_seen);

/* ---- restart records (restart.f90:2-82, one rank) ---------------------- */

/* write_restart / read_restart of the model's whole state (u, v, w, b, p and
   the three tendency slots of du..db, halos included) in the reference's
   unformatted sequential record layout: (nx, ny, nz, itt, time, taum2, taum1,
   tau), (is_pe .. ke_pe), (u, v, w, b, p, du, dv, dw, db).  read checks the
   grid and that the file comes from one rank.  FP64 models. */
int ps3d_write_restart(ps3d_model *m, const char *path, int itt, double time, int tau, int taum1, int taum2);
int ps3d_read_restart(ps3d_model *m, const char *path, int32_t *itt, double *time, int32_t *tau, int32_t *taum1,
                      int32_t *taum2);
/* the same records from / to host arrays: state = u, v, w, b, p (H doubles
   each) then du, dv, dw, db (3 H each), H = (nx+4)(ny+4)(nz+4) */
int ps3d_restart_write_host(const char *path, int nx, int ny, int nz, int itt, double time, int tau, int taum1,
                            int taum2, const double *state);
int ps3d_restart_read_host(const char *path, int nx, int ny, int nz, int32_t *itt, double *time, int32_t *tau,
                           int32_t *taum1, int32_t *taum2, double *state);

/* ---- measurement ----------------------------------------------------- */

/* Time every launch of the named kernel with HIP events on the model's
   stream (NULL turns it off).  Names are those of ps3d_kernel_names(). */
int ps3d_profile_kernel(ps3d_model *m, const char *name);
/* launches and total milliseconds recorded since the last call */
int ps3d_profile_read(ps3d_model *m, int64_t *launches, double *total_ms);
/* comma-separated names of the model's kernels */
const char *ps3d_kernel_names(void);

/* HBM rates of hand-written streaming kernels on the current GPU, GB/s:
   out[0] copy (read + write bytes counted), out[1] copy with nontemporal
   loads and stores, out[2] read only, out[3] nontemporal read, out[4] write
   only, out[5] nontemporal write; out[6] the work-group count used.  bytes:
   size of each of the two buffers allocated (a multiple of 16, >= 1 MiB);
   the best of reps launches after one warm-up; grid: work-groups (0: eight
   per CU, grid-strided).  No counterpart in the reference: it measures the
   achievable bandwidth that the step's kernels are compared with. */
int ps3d_hbm_probe(int64_t bytes, int reps, int grid, double *out);

/* Vector FMA issue rate on the current GPU, TFLOP/s: out[0] FP32, out[1]
   FP64 (eight independent FMA chains per lane, the best of reps launches
   after one warm-up).  FP32 here is compiled, like every kernel of the
   library, without the packed FP32 instructions, so it shows the rate the
   mixed form's FP32 stencils issue at.  No counterpart in the reference. */
int ps3d_fma_probe(int reps, double *out);

/* ---- batched ensemble ------------------------------------------------
   No counterpart in the reference, which advances one model per process
   (main.f90:138-162); an ensemble there is a set of separate runs.  Here
   members of different sizes and parameters advance together: each stage of
   the step (tendency, friction + AB, halo, the transforms, Thomas or the
   periodic z solve, projection) is one launch over all members, and a step
   can be replayed as one captured hipGraph.  Every member computes exactly
   what a single model with the same parameters computes (the fused-friction
   configurations: Avbi = Kvbi = 0; FFT lengths among the compiled-in ones,
   column lengths ny (and nz without a lid) <= 512).  The members share one
   precision form: PS3D_FP64, or PS3D_MIXED (rigid-lid members of at most 213
   levels), whose first run starts every member's pressure history as a
   single mixed model's first step does; that step and the next one (which
   add a correction solve) are queued directly, later ones replay the graph. */
typedef struct ps3d_ensemble ps3d_ensemble;

int ps3d_ensemble_create(int n, const ps3d_params *prms, void *stream, ps3d_ensemble **out);
int ps3d_ensemble_destroy(ps3d_ensemble *e);
/* field ids and layouts as for ps3d_set_field / ps3d_get_field */
int ps3d_ensemble_set_field(ps3d_ensemble *e, int member, int field, const double *host);
int ps3d_ensemble_get_field(ps3d_ensemble *e, int member, int field, double *host);
/* Forcing of one member, as ps3d_set_background_flow / ps3d_set_body_force
   set it on a single model (lee_wave1.f90:239-314, a steady body force):
   background-flow advection and body force formed on the device from the
   post-friction state (integrate.f90:82).  Forced members run the friction
   chain into a third set of buffers, then a halo copy and one forcing +
   Adams-Bashforth stage (each one launch over the forced members), and stay
   bit-identical to a single model of their form with the same forcing (in
   the mixed form the third buffers hold (hi, lo) pairs, and the cold start
   forms the forcing on the pre-friction state, as a single mixed model
   does). */
int ps3d_ensemble_set_background_flow(ps3d_ensemble *e, int member, const double *u_back, const double *b_back);
int ps3d_ensemble_set_body_force(ps3d_ensemble *e, int member, const double *const f[4]);
/* Radix plans of a member's FP32 transforms (as ps3d_set_fft_plan; PS3D_MIXED
   ensembles); ps3d_ensemble_tune picks them per transform length.  A member
   with the same plans as a single mixed model gives that model's bits. */
int ps3d_ensemble_set_fft_plan(ps3d_ensemble *e, int member, const int32_t plans[3]);
int ps3d_ensemble_get_fft_plan(ps3d_ensemble *e, int member, int32_t plans[3]);
/* nsteps steps of every member; state = {itt, tau, taum1, taum2}, shared by
   the members and advanced as by ps3d_run; use_graph replays a captured step */
int ps3d_ensemble_run(ps3d_ensemble *e, int nsteps, int32_t *state, int use_graph);
int ps3d_ensemble_synchronize(ps3d_ensemble *e);
/* Autotuner: times the candidate tile shapes of the k-marching stages
   (levels per work-group of tendency and friction + AB) per member shape on
   the live ensemble, keeps the fastest, and caches the choice per
   (stage, nx x ny x nz, boundary, precision) in cache_path (NULL: no file).
   Results do not depend on the tile shapes. */
int ps3d_ensemble_tune(ps3d_ensemble *e, const char *cache_path, int reps);
/* the tile shapes in use, one "stage shape levels" line per member shape,
   copied as by ps3d_last_error; returns the text's length, or minus an
   error code */
int ps3d_ensemble_describe(ps3d_ensemble *e, char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif /* PS3D_H */
