// ps3d_internal.h -- device-side data structures of the ps3D step on gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ps3d.h"

namespace ps3d {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define PS3D_HIP_CHECK(x)                                                              \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            throw ::ps3d::Error(PS3D_ERR_HIP, std::string(#x) + " failed: " +          \
                                                  hipGetErrorString(e_));              \
    } while (0)

#define PS3D_REQUIRE(cond, msg)                                                        \
    do {                                                                               \
        if (!(cond)) throw ::ps3d::Error(PS3D_ERR_ARG, (msg));                         \
    } while (0)

// message of the calling thread's last failed C-ABI call
std::string &error_slot();

// runs f, turning exceptions into PS3D_ERR_* codes and a message: no C++
// exception crosses the C-ABI
template <class F>
int guard(F &&f)
{
    try {
        f();
        error_slot().clear();
        return PS3D_OK;
    } catch (const Error &e) {
        error_slot() = e.what();
        return e.code;
    } catch (const std::exception &e) {
        error_slot() = e.what();
        return PS3D_ERR_INTERNAL;
    } catch (...) {
        error_slot() = "unknown error";
        return PS3D_ERR_INTERNAL;
    }
}

constexpr int ONX = 2;   // halo width, main_module.f90:93

// column tiles of the k-marching kernels (stencil.hip; the ensemble's block
// counts use the same): tendency TEND_TX x TEND_TY, friction + AB FRIC_TX x FRIC_TY
#ifndef PS3D_FRIC_TY
#define PS3D_FRIC_TY 8
#endif
#ifndef PS3D_TEND_TX
#define PS3D_TEND_TX 64
#endif
#ifndef PS3D_TEND_TY
#define PS3D_TEND_TY 8
#endif
constexpr int TEND_TX = PS3D_TEND_TX, TEND_TY = PS3D_TEND_TY;
constexpr int FRIC_TX = 64, FRIC_TY = PS3D_FRIC_TY;

// Extents of one model.  Physical fields are (X, Y, Z) = (nx+4, ny+4, nz+4)
// with Fortran index i in -1..nx+2 at offset i+1 (i fastest).
struct Grid {
    int nx, ny, nz;
    int64_t X, Y, Z, XY, H;   // padded extents, plane and field sizes
    int64_t N;                // nx*ny*nz
    int nkx;                  // nx/2+1
    int64_t S;                // nkx*ny*nz complex spectral points
};

// Scalars the kernels read, in the form they use them.  Every value comes
// from the driver's main_module unchanged (no re-derivation).
struct StepConsts {
    int nx, ny, nz;
    int64_t X, XY;
    int lid;           // enable_vertical_boundaries
    int nonlinear;     // enable_nonlinear
    double dx, dy, dz, dt, Ro, dsqr, f0, N0;
    double rdx, rdy, rdz, rdt;   // 1/dx ... rounded to nearest, for div_rn
};

// exp(-2 pi i q/n) for the FFT tables, computed on the host.  sincos() is
// called explicitly: a compiler may or may not fuse separate cos() and sin()
// calls into it, and the two paths can differ in the last bit.
inline double2 host_twiddle(int q, int n)
{
    const double ang = 2.0 * M_PI * (double)q / (double)n;
    double s, c;
    ::sincos(ang, &s, &c);
    return make_double2(c, -s);
}

// a / b, correctly rounded, from rb = RN(1/b): q = RN(a*rb) is within one ulp
// of a/b, the residual a - q*b is exact in one FMA, and one correction step
// q + residual*rb rounds to RN(a/b) (Markstein's theorem).  Three
// instructions instead of the ~10 of a general IEEE division, same bits as
// a / b for all finite normal results (a zero quotient comes out as +0).
__device__ __forceinline__ double div_rn(double a, double b, double rb)
{
    const double q = a * rb;
    const double e = __builtin_fma(-q, b, a);
    return __builtin_fma(e, rb, q);
}

// Division by a grid constant in a stencil of element type T: the FP64 path
// keeps the reference's correctly rounded quotient, the FP32 path of the
// mixed form multiplies by the reciprocal.
template <class T>
__device__ __forceinline__ T sdiv(T a, double b, double rb);
template <>
__device__ __forceinline__ double sdiv<double>(double a, double b, double rb) { return div_rn(a, b, rb); }
template <>
__device__ __forceinline__ float sdiv<float>(float a, double, double rb) { return a * (float)rb; }

// Writes val, the new value at interior point (i, j, k), to every halo point
// whose fully wrapped index is (i, j, k): what border_exchg_3D
// (parallel_mpi.f90:70-211, one rank) copies there.  x and y are periodic;
// z is periodic, or with neumann_z the halo levels take copies of levels 1
// and nz (pressure_solver.f90:82-93).  Producers call it for the points they
// write, so no separate exchange pass is needed.  Requires nx, ny, nz >= 4
// (halo_images_ok): a point then has at most one image per axis (two in z
// under Neumann), and only lanes next to a face take the stores.
// Blocks are dealt round-robin to the 8 XCDs; renumbering them so that each
// XCD walks a contiguous run of work items keeps neighbouring tiles, which
// share apron rows, in one L2.
__device__ __forceinline__ int xcd_remap(int bid, int n)
{
    const int per = n / 8;
    return bid < per * 8 ? (bid % 8) * per + bid / 8 : bid;
}

inline bool halo_images_ok(int nx, int ny, int nz) { return nx >= 4 && ny >= 4 && nz >= 4; }

template <class T>
__device__ __forceinline__ void write_images(const StepConsts &c, T *a, int i, int j, int k, T val, int neumann_z)
{
    const int ox = i <= 2 ? c.nx : (i >= c.nx - 1 ? -c.nx : 0);
    const int oy = j <= 2 ? c.ny : (j >= c.ny - 1 ? -c.ny : 0);
    int oz;
    if (neumann_z) oz = k == 1 ? -1 : (k == c.nz ? 1 : 0);
    else oz = k <= 2 ? c.nz : (k >= c.nz - 1 ? -c.nz : 0);
    if ((ox | oy | oz) == 0) return;
    T *p = a + (int64_t)(i + 1) + c.X * (int64_t)(j + 1) + c.XY * (int64_t)(k + 1);
    const int64_t sx = ox, sy = (int64_t)oy * c.X;
    if (ox) p[sx] = val;
    if (oy) {
        p[sy] = val;
        if (ox) p[sx + sy] = val;
    }
    if (oz) {
        const int nzi = neumann_z ? 2 : 1;   // Neumann: two copies, at distance 1 and 2
        for (int r = 1; r <= nzi; ++r) {
            T *q = p + (int64_t)(oz * r) * c.XY;
            q[0] = val;
            if (ox) q[sx] = val;
            if (oy) {
                q[sy] = val;
                if (ox) q[sx + sy] = val;
            }
        }
    }
}

// complex type of a real type
template <class T> struct cplx_of;
template <> struct cplx_of<double> { using type = double2; };
template <> struct cplx_of<float> { using type = float2; };

// One mixed-radix FFT length: factors (radix 4, 2, 3, 5) and its table
// tw[q] = exp(-2 pi i q / n), q = 0..n-1, in device memory.
constexpr int MAX_FAC = 16;
struct FftPlan {
    int n;
    int nf;
    int fac[MAX_FAC];
    const double2 *tw;
    const float2 *twf;   // the same table rounded to FP32 (mixed form)
};

struct Model {
    ps3d_params prm{};
    Grid g{};
    StepConsts c{};
    hipStream_t stream = nullptr;
    bool own_stream = false;

    // prognostic fields (canonical buffers), halo included
    double *u = nullptr, *v = nullptr, *w = nullptr, *b = nullptr, *p = nullptr;
    // second buffers for the out-of-place friction passes
    double *alt[4] = {nullptr, nullptr, nullptr, nullptr};
    // tendencies, 3 slots of H each (main_module.f90:124-127)
    double *du = nullptr, *dv = nullptr, *dw = nullptr, *db = nullptr;
    double2 *FS = nullptr;       // spectral array (nkx, ny, nz)
    double *ksqr = nullptr;      // (nkx, ny) rigid lid, (nkx, ny, nz) periodic
    double *tri_cp = nullptr;    // Thomas factors c'(k) per column (rigid lid)
    double *tri_rm = nullptr;    // 1/m(k) per column (1/b(1) at k=1)
    double *aloc = nullptr;      // real (nx, ny, nz) scratch for the building-block API

    FftPlan px{}, py{}, pz{};    // lengths nx/2, ny, nz
    double2 *tw_r2c = nullptr;   // exp(-2 pi i k / nx), k = 0..nx/2
    std::vector<void *> allocations;

    // per-field "current buffer" during the friction block (0 = canonical)
    int cur[4] = {0, 0, 0, 0};
    // use the fused friction + Adams-Bashforth kernel where it applies
    bool fuse_friction = true;
    // form the tendency inside it (unforced FP64 steps): reads the state
    // once and never reads slot tau back, but measured slower than the
    // separate tendency kernel (ps3d_set_fused_tendency), so off by default
    bool fuse_tendency = false;
    // producers write the halo images of what they update (halo_images_ok)
    bool fuse_halo = false;
    // launch shapes (results do not depend on them; set by ps3d_tune):
    // levels per work-group of the k-marching kernels (0: heuristic) and
    // the form of the Thomas solve (launch_thomas, launch_thomas_f: 1 whole
    // columns in registers; 2 the real and imaginary parts of a column on two
    // lanes; 16, 32, 64 columns per wave on the fly; -16, -32, -64 the factor
    // tables; 0: the heuristic)
    int tend_kc = 0, fric_kc = 0, thomas_cpw = 0;
    int tfric_kc = 0;   // the same for the kernel with the tendency fused in
    // the heuristic's first choice: whole columns in registers where nz is
    // one of the compiled-in column lengths (k_thomas_reg)
    bool thomas_reg = true;
    // ps3d_run's captured steps: a hipGraph of six steps (the period of the
    // tau rotation, the buffer parity and the pressure-increment ring) and
    // the signature of the state it was captured from
    hipGraph_t step_graph = nullptr;
    hipGraphExec_t step_exec = nullptr;
    std::vector<uint64_t> graph_key;
    // radix plans of the x, y and z transforms (0: radix 4 first, the FP64
    // plan; 1: radix 8 first); FP64 transforms always use plan 0, FP32 ones
    // (mixed form) the plan set here (ps3d_set_fft_plan, ps3d_tune)
    int fft_plan[3] = {0, 0, 0};

    // mixed form (PS3D_MIXED): prognostic fields as FP32 (hi, lo) pairs whose
    // sum carries the FP64-accumulated state.  The pressure is carried in
    // FP64 (pint) and each step solves, in FP32, only for its increment dp;
    // the FP32 tendency slots hold the tendencies minus the gradient of the
    // pressure they were computed with, so every FP32 quantity is small and
    // the hydrostatic / geostrophic cancellations happen in FP64.
    bool mixed = false;
    float *fh[4] = {}, *fl[4] = {};   // canonical (hi, lo) of u, v, w, b
    float *ah[4] = {}, *al[4] = {};   // second buffers
    float *tf[4] = {};                // pressure-referenced tendencies, 3 slots each
    double *pint = nullptr;           // pressure p_n (FP64), halo included
    float *dpr[3] = {};               // ring of pressure increments dp_n, dp_n-1, dp_n-2
    int dp_cur = 0;                   // dpr index of the latest increment
    bool cold = true;                 // no pressure history yet: the next step starts one
    int refine = 0;                   // steps left that add a second (correction) solve
    float2 *FSf = nullptr;            // spectral array (of the increment)
    double2 *dp64 = nullptr;          // forward-sweep values of the Thomas solve
    double *stage = nullptr;          // 3H doubles for host conversions
    float2 *twf_r2c = nullptr;
    bool lo_stale[4] = {};            // a step since the joined field's halo was last rebuilt
    std::vector<double> p_host;       // p as set by the host, returned until the next step

    // profiling of one kernel by name
    std::string prof_name;
    std::vector<hipEvent_t> prof_ev;   // pairs (start, stop)
    int64_t prof_launches = 0;
    double prof_ms = 0.0;

    // forcing (integrate.f90:82): the caller's hook and the built-in wavemaker
    ps3d_forcing_fn forcing = nullptr;
    void *forcing_user = nullptr;
    double *wm_re[4] = {}, *wm_im[4] = {};   // amplitudes, allocated on first use
    bool wm_field[4] = {};                    // fields the wavemaker forces
    double wm_omega = 0.0;
    bool wm_on = false;
    // background flow (u_back, b_back) and body force (per field)
    double *bg_u = nullptr, *bg_b = nullptr;
    bool bg_on = false, bg_b_on = false;
    double *bf[4] = {};
    bool bf_on[4] = {};

    // work arrays of the balancing diagnostic (balance.hip), made on first use
    struct BalanceWork *bal = nullptr;
    void *mon = nullptr;              // partial sums of the monitors (monitor.hip)

    double *field_buf(int id) const;
    double *cur_buf(int f) const { return cur[f] ? alt[f] : field_buf(f); }
};

// ---- launchers (stencil.hip) -------------------------------------------
void launch_tendency(Model &m, int tau);
void launch_buoyancy_advection(Model &m, int tau);
void launch_momentum_advection(Model &m, int tau);
void launch_friction_block(Model &m);
void launch_mixing(Model &m, int op, int field, double hmix, double vmix, int zero_top);
void launch_ab_exchange(Model &m, int itt, int tau, int taum1, int taum2);
// the wavemaker on the current buffers: q += dt*(cs*re + sn*im), interior
void launch_wavemaker(Model &m, double cs, double sn);
// background-flow advection and body force into the tendency slot tau (both
// forms; mixed_fields: the FP32 hi fields to advect, default the canonical ones)
void launch_background(Model &m, int tau, float *const *mixed_fields = nullptr);
// mixed step with background forcing: friction, forcing on the post-friction
// state, then Adams-Bashforth (the fused kernel's terms in three passes)
void launch_friction_forcing_ab_mixed(Model &m, int itt, int tau, int taum1, int taum2, double qc1, const float *d1,
                                      double qc2, const float *d2);
bool fused_friction_ok(const Model &m);
// with_tendency: the tendency of integrate.f90:12-54 formed in the same
// kernel (k_tend_fric_ab) instead of read from slot tau, which it writes
void launch_friction_ab(Model &m, int itt, int tau, int taum1, int taum2, bool with_tendency = false);
void launch_border_exchg(Model &m, double *a, int neumann_z);
void launch_project_exchange(Model &m);

// ---- launchers (fft.hip) ------------------------------------------------
void fft_plan_init(Model &m, FftPlan &pl, int n);
void fft_r2c_x(Model &m, const double *src, int src_is_field, int fuse_div, double2 *dst);
void fft_c2c_y(Model &m, double2 *data, int sign, double scale1, double scale2);
void fft_c2r_x(Model &m, const double2 *src, double *dst, int dst_is_field, double inv_scale);
void fft_z_periodic_solve(Model &m, double2 *data);
void fft_c2c_z(Model &m, double2 *data, int sign, double s1, double s2, double s3);

void fft_r2c_x_f(Model &m, const float *u, const float *v, const float *w, float2 *dst);
void fft_c2r_x_f(Model &m, const float2 *src, float *dst, double inv_scale);
void fft_c2c_y_f(Model &m, float2 *data, int sign, double scale1, double scale2);
void fft_z_periodic_solve_f(Model &m, float2 *data);
bool fft_sizes_compiled(int n);
std::string fft_sizes_text();

// ---- launchers (tridiag.hip) --------------------------------------------
void tridiag_factor(Model &m);
void launch_thomas(Model &m, double2 *data);
void launch_thomas_f(Model &m, float2 *data);

void launch_friction_ab_raw(Model &m, int itt, int tau, int taum1, int taum2, bool with_tendency = false);
// the step takes k_tend_fric_ab (tendency, friction and AB in one kernel)
bool fused_tendency_active(const Model &m);
// forced FP64 step: the fused friction chain alone into the second buffers
// (halos pre-friction), marked current for the forcing and launch_ab_exchange
void launch_friction_fused_only(Model &m);
// then the background flow / body force and the AB update in one pass
void launch_background_ab(Model &m, int itt, int tau, int taum1, int taum2);
void launch_advection_fields(Model &m, const double *u, const double *v, const double *w, const double *b,
                             double *du, double *dv, double *dw, double *db);
// one whole step, integrate.f90:2-115 (api.hip)
void model_step(Model &m, int itt, int tau, int taum1, int taum2);
void launch_divide_ksqr(Model &m, double2 *data);
// frees the host side of the balancing work (device arrays are in allocations)
void balance_release(Model &m);
// the single-model autotuner (tune.hip)
void tune_model(Model &m, int next_tau, const char *cache_path, int reps);
// tuning cache files of "stage shape value" lines: read (missing file: empty;
// malformed line: error) and written atomically (temporary file, rename)
std::map<std::string, int> tune_cache_read(const char *path);
void tune_cache_write(const char *path, const std::map<std::string, int> &cache);
std::string describe_tuning(const Model &m);

// monitors and restart records (monitor.hip)
bool monitor(Model &m, double *out8);
// the same on caller-owned FP64 device fields, partial records in a static
// device scratch (no allocation); stream may be null (the null stream)
bool monitor_noalloc(const ps3d_params &p, const double *u, const double *v, const double *w, const double *b,
                     hipStream_t stream, double *out8);
void restart_write_host(const char *path, int nx, int ny, int nz, int32_t itt, double time, int32_t tau,
                        int32_t taum1, int32_t taum2, const double *state);
void restart_read_host(const char *path, int nx, int ny, int nz, int32_t *itt, double *time, int32_t *tau,
                       int32_t *taum1, int32_t *taum2, double *state);
void restart_io(Model &m, const char *path, bool write, int32_t *itt, double *time, int32_t *tau, int32_t *taum1,
                int32_t *taum2);
// the balancing diagnostic (balance.hip); out[16] host arrays, nullable
void diag_balance(Model &m, int filter_width, double *const *out, const double *kx, const double *ky,
                  const double *kz);
void opt_balance(Model &m, const ps3d_opt_balance_params &ob, double *const *out, double *norms, int *iterations,
                 const double *kx, const double *ky, const double *kz);

// ---- mixed form (stencil.hip) -------------------------------------------
// tendency with the reference pressure pint (null: zero)
void launch_tendency_mixed(Model &m, int tau, const double *pint);
// friction + AB into the second buffers; qc1, qc2 weight the last two
// pressure increments (see the step in api.hip); swap: make them canonical
void launch_friction_ab_mixed(Model &m, int itt, int tau, int taum1, int taum2, double qc1, const float *d1,
                              double qc2, const float *d2, bool swap);
void launch_project_exchange_f(Model &m, float *dp);
void launch_halo_f(Model &m, float *const *arrays, const int *neumann, int n);
void launch_convert(Model &m, int dir, int64_t n, const double *src_d, double *dst_d, const float *src_f,
                    float *dst_f, const float *src_lo, float *dst_lo);

// ---- profiling helpers (api.hip) ----------------------------------------
struct ProfScope {
    Model &m;
    bool on;
    hipEvent_t e1{}, e2{};
    ProfScope(Model &mm, const char *name);
    ~ProfScope();
};

// ---- batched ensemble (ensemble.hip) --------------------------------------
// Members of different sizes and parameters advance together: every stage of
// the step is ONE launch over all members, whose blocks are dealt to the
// members by a prefix table (EnsMap) and which read the member's constants,
// buffers and tile shapes from a device-side table (EnsMember).  The step
// counters live in device memory (EnsState) and are advanced by the last
// launch of the step, so one captured hipGraph replays every step.
constexpr int ENS_MAX = 256;

struct EnsMap {
    int n;                       // members
    int start[ENS_MAX + 1];      // first block of each member; start[n] = total
};

struct EnsMember {
    StepConsts c;
    int64_t H;
    double *f[4][2];             // u, v, w, b: the two buffers of each field
    double *tend[4];             // du..db, 3 slots of H each
    double *p;
    double2 *FS;
    const double *ksqr, *cp, *rm;
    const double2 *twx, *twy, *twz, *wk;   // FFT tables: nx/2, ny, nz, r2c
    double Hm[2], Vm[2], Hbi[2];            // friction of (u, v, w) and of b
    int harm[2], bih[2];
    double eps_ab, fxa, lxy, lxyz, rlxy, rlxyz;   // rl* = 1 / l*, rounded on the host
    int ab3, n2, nkx;
    int tend_kc, tend_nchunk, tend_tx, tend_ty;   // tile shapes (tunable)
    int fric_kc, fric_nchunk, fric_tx, fric_ty;
    // mixed form (PS3D_MIXED members): (hi, lo) of u, v, w, b in two
    // buffers, the FP32 tendency slots, the FP64 pressure, the ring of
    // pressure increments, the FP32 spectral array and FFT tables
    float *fh[4][2], *fl[4][2];
    float *tf[4];
    double *pint;
    float *dpr[3];
    float2 *FSf;
    const float2 *twxf, *twyf, *twzf, *wkf;
    int fft_plan[3];   // radix plans of the FP32 x, y, z transforms (0, 1)
    // forcing: background flow u_back, b_back and the body force (null:
    // none); forced members run friction into the third buffers (FP64: tmp;
    // mixed: the (hi, lo) pair tmph, tmpl), then the halo copy and the
    // forcing + AB stages
    int forced;
    double *tmp[4];
    float *tmph[4], *tmpl[4];
    const double *ub, *bb, *bf[4];
};

// member of block g: the last member whose first block is <= g (members
// without blocks in a stage have empty ranges and are skipped)
__device__ __forceinline__ int ens_find(const EnsMap &m, int g)
{
    int lo = 0, hi = m.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (m.start[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct EnsState {
    int itt, tau, taum1, taum2;
    int par;                     // which buffer of f[][] holds the state
    int dp_cur;                  // mixed form: dpr index of the latest increment
};

// stage launchers (one launch each over all members)
void ens_tendency(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st);
void ens_friction_ab(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st);
void ens_project(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st);
// forced FP64 members: pre-friction halos into the third buffers (4 fields x
// halo points per member), then forcing + AB (interior points per member)
void ens_halo_copy(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, bool mixed);
void ens_background_ab(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st);
// mixed form: the forcing on the pre-friction state into slot tau (the cold
// start's provisional step), and forcing + Adams-Bashforth from the third
// (hi, lo) buffers into the state's next buffers (the steady step)
void ens_background_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st);
void ens_background_ab_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st);
void ens_advance(hipStream_t s, EnsState *st);
void ens_r2c(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, size_t lds);
// nmax: the longest column among the members (selects the compiled variant)
void ens_cols_y(hipStream_t s, const EnsMember *mem, const EnsMap &map, int sign, int nmax, size_t lds);
void ens_zsolve(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nmax, size_t lds);
void ens_c2r(hipStream_t s, const EnsMember *mem, const EnsMap &map, size_t lds);
// Thomas stage: ens_thomas_cols() columns per work-group; nz_max: most levels
// of a rigid-lid member (sizes the LDS of the on-the-fly form)
// reg: the register form (k_ens_thomas_reg, 64 columns per work-group), for
// ensembles whose rigid-lid members all have thomas_reg_ok levels; split
// (when not reg): the real and imaginary parts on two lanes
// (k_ens_thomas_split, 32 columns per work-group), where they all have
// thomas_split_ok levels
int ens_thomas_cols(bool reg, bool split);
// nz levels have a compiled-in register Thomas solve, and the spectral array
// (fs_bytes) is addressed with 32-bit offsets
bool thomas_reg_ok(int nz, uint64_t fs_bytes);
// the same for the form with the real and imaginary parts on two lanes
// (k_thomas_split), which has taller compiled-in lengths
bool thomas_split_ok(int nz, uint64_t fs_bytes);
// a compiled-in FP32 transform length whose plan 1 differs from plan 0
bool fft_plan_choice(int n);
void ens_thomas(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nz_max, bool reg, bool split);
// the same stages for members of the mixed form.  cold: the provisional
// Euler step of the cold start (tendency without the reference pressure,
// friction + Euler); which: the increment a solve writes and a projection
// applies (0: dpr[0], the cold start's estimate; 1: the new increment
// dpr[dp_cur + 1]; 2: the correction dpr[dp_cur + 2])
void ens_tendency_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int cold);
void ens_friction_ab_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int cold);
void ens_project_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int which);
// pint = the estimate dpr[0] (images included); then the three increments zeroed
void ens_cold_pint(hipStream_t s, const EnsMember *mem, const EnsMap &map);
// per-point maps over the H points of each member: zero the ring (what = 0)
// or add the correction to the new increment (what = 1)
void ens_points_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int what);
void ens_r2c_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, size_t lds);
void ens_cols_y_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, int sign, int nmax, size_t lds);
void ens_zsolve_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nmax, size_t lds);
void ens_c2r_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int which, size_t lds);
void ens_thomas_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nz_max, bool reg, bool split);
// block counts and LDS bytes of the batched FFT stages for one member
int ens_r2c_blocks(const EnsMember &e);
int ens_cols_blocks(const EnsMember &e, int n_len, int lines);
size_t ens_x_lds(int n2);
size_t ens_cols_lds(int n);
bool ens_fft_size_ok(int n, bool column_pass);

}  // namespace ps3d

// the opaque handle of the C-ABI
struct ps3d_model : public ps3d::Model {};
