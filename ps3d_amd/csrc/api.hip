// api.hip -- the C-ABI of include/ps3d.h: model set-up, host <-> device
// copies, the step and its phases, building blocks and measurement.
#include "ps3d_internal.h"

#include <cmath>
#include <cstring>
#include <string>

using namespace ps3d;

namespace {

bool is_235(int n)
{
    if (n < 1) return false;
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

template <class T>
T *dalloc(Model &m, int64_t n)
{
    void *p = nullptr;
    PS3D_HIP_CHECK(hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    PS3D_HIP_CHECK(hipMemsetAsync(p, 0, sizeof(T) * (size_t)(n > 0 ? n : 1), m.stream));
    m.allocations.push_back(p);
    return static_cast<T *>(p);
}

// eigenvectors.f90:224-229
double hat_ksqr(double k, double d) { return 2.0 * (1.0 - std::cos(k * d)) / (d * d); }

// my_fft.f90:87-156, the ksqr part
std::vector<double> make_ksqr(const ps3d_params &p)
{
    const double pi = std::acos(0.0) * 2.0;   // main_module.f90:116
    const int nkx = p.nx / 2 + 1;
    std::vector<double> kx(nkx), ky(p.ny), kz(p.nz);
    for (int i = 1; i <= nkx; ++i) kx[i - 1] = (double)(2 * (i - 1)) * pi / p.Lx;
    for (int j = 1; j <= (p.ny + 1) / 2; ++j) ky[j - 1] = (double)(2 * (j - 1)) * pi / p.Ly;
    for (int j = (p.ny + 1) / 2 + 1; j <= p.ny; ++j) ky[j - 1] = -((double)(2 * (p.ny - j + 1)) * pi / p.Ly);
    for (int k = 1; k <= (p.nz + 1) / 2; ++k) kz[k - 1] = (double)(2 * (k - 1)) * pi / p.Lz;
    for (int k = (p.nz + 1) / 2 + 1; k <= p.nz; ++k) kz[k - 1] = -((double)(2 * (p.nz - k + 1)) * pi / p.Lz);
    std::vector<double> ks;
    if (p.enable_vertical_boundaries) {
        ks.resize((size_t)nkx * p.ny);
        for (int j = 0; j < p.ny; ++j)
            for (int i = 0; i < nkx; ++i) ks[i + (size_t)nkx * j] = hat_ksqr(kx[i], p.dx) + hat_ksqr(ky[j], p.dy);
    } else {
        ks.resize((size_t)nkx * p.ny * p.nz);
        for (int k = 0; k < p.nz; ++k)
            for (int j = 0; j < p.ny; ++j)
                for (int i = 0; i < nkx; ++i)
                    ks[i + (size_t)nkx * (j + (size_t)p.ny * k)] =
                        hat_ksqr(kx[i], p.dx) + hat_ksqr(ky[j], p.dy) + hat_ksqr(kz[k], p.dz) / p.dsqr;
    }
    return ks;
}

int64_t ksqr_len(const Model &m)
{
    return m.prm.enable_vertical_boundaries ? (int64_t)m.g.nkx * m.g.ny : m.g.S;
}

int64_t field_len(const Model &m, int field)
{
    switch (field) {
    case PS3D_U: case PS3D_V: case PS3D_W: case PS3D_B: case PS3D_P: return m.g.H;
    case PS3D_DU: case PS3D_DV: case PS3D_DW: case PS3D_DB: return 3 * m.g.H;
    case PS3D_FS: return 2 * m.g.S;
    case PS3D_KSQR: return ksqr_len(m);
    default: return -1;
    }
}

void *field_dev(Model &m, int field)
{
    switch (field) {
    case PS3D_U: case PS3D_V: case PS3D_W: case PS3D_B: return m.cur_buf(field);
    case PS3D_P: return m.p;
    case PS3D_DU: return m.du;
    case PS3D_DV: return m.dv;
    case PS3D_DW: return m.dw;
    case PS3D_DB: return m.db;
    case PS3D_FS: return m.FS;
    case PS3D_KSQR: return m.ksqr;
    default: return nullptr;
    }
}

void check_model(const ps3d_model *h)
{
    if (!h) throw Error(PS3D_ERR_ARG, "null model");
}

// the building-block entry points work on the FP64 buffers of the drop-in
void require_fp64(const Model &m, const char *what)
{
    if (m.mixed)
        throw Error(PS3D_ERR_STATE, std::string(what) + ": not available on a PS3D_MIXED model (its step runs "
                                                       "only as a whole: ps3d_integrate / ps3d_run)");
}

// set / get of a mixed model.  u, v, w, b go through the FP64 staging
// buffer and the (hi, lo) split; p is the FP64 pressure (a value the host
// sets is returned until the next step and, as in the reference, not read by
// the step).  The tendency slots and FS hold this form's own quantities
// (pressure-referenced tendencies, the spectrum of the pressure increment)
// and are not exchanged with the host.
const char *kMixedFields = "PS3D_MIXED: only u, v, w, b, p and ksqr are exchanged with the host";

void set_field_mixed(Model &m, int field, const double *host, int64_t l)
{
    if (field == PS3D_P) {
        m.p_host.assign(host, host + l);
        return;
    }
    if (field > PS3D_B) throw Error(PS3D_ERR_STATE, kMixedFields);
    PS3D_HIP_CHECK(hipMemcpyAsync(m.stage, host, sizeof(double) * l, hipMemcpyHostToDevice, m.stream));
    launch_convert(m, 0, l, m.stage, nullptr, nullptr, m.fh[field], nullptr, m.fl[field]);
}

void get_field_mixed(Model &m, int field, double *host, int64_t l)
{
    if (field == PS3D_P) {
        if (!m.p_host.empty()) std::memcpy(host, m.p_host.data(), sizeof(double) * l);
        else PS3D_HIP_CHECK(hipMemcpyAsync(host, m.pint, sizeof(double) * l, hipMemcpyDeviceToHost, m.stream));
        return;
    }
    if (field > PS3D_B) throw Error(PS3D_ERR_STATE, kMixedFields);
    launch_convert(m, 1, l, nullptr, m.stage, m.fh[field], nullptr, m.fl[field], nullptr);
    // after a step, the joined field's halo is rebuilt from its interior as
    // border_exchg_3D would (the same values as the halo images of the hi and
    // lo parts that the step's kernels write, summed)
    if (m.lo_stale[field]) launch_border_exchg(m, m.stage, 0);
    PS3D_HIP_CHECK(hipMemcpyAsync(host, m.stage, sizeof(double) * l, hipMemcpyDeviceToHost, m.stream));
}

void check_tau(int tau, int taum1, int taum2)
{
    auto ok = [](int t) { return t >= 1 && t <= 3; };
    PS3D_REQUIRE(ok(tau) && ok(taum1) && ok(taum2), "tau slots must be 1..3");
    PS3D_REQUIRE(tau != taum1 && tau != taum2 && taum1 != taum2, "tau slots must differ");
}

// FP32 Poisson solve of pressure_solver.f90:16-73 on the hi parts of u, v, w
// (their divergence / dt), result into dst (interior; halo left to the caller)
void solve_mixed(Model &m, const float *u, const float *v, const float *w, float *dst)
{
    const ps3d_params &p = m.prm;
    fft_r2c_x_f(m, u, v, w, m.FSf);
    if (p.enable_vertical_boundaries) {
        fft_c2c_y_f(m, m.FSf, -1, p.dx, p.dy);
        launch_thomas_f(m, m.FSf);
        fft_c2c_y_f(m, m.FSf, +1, 0.0, 0.0);
        fft_c2r_x_f(m, m.FSf, dst, p.Lx * p.Ly);
    } else {
        fft_c2c_y_f(m, m.FSf, -1, 0.0, 0.0);
        fft_z_periodic_solve_f(m, m.FSf);
        fft_c2c_y_f(m, m.FSf, +1, 0.0, 0.0);
        fft_c2r_x_f(m, m.FSf, dst, p.Lx * p.Ly * p.Lz);
    }
}

// Starts the pressure history of a mixed model: an estimate p_est of this
// step's pressure from a provisional Euler step (tendency without a reference
// pressure, friction + AB into the second buffers, FP32 solve), held in
// pint with zero increments, so that the step that follows is referenced to
// p_est and its FP32 quantities are small from the first step on.
bool background_on(const Model &m)
{
    return m.bg_on || m.bf_on[0] || m.bf_on[1] || m.bf_on[2] || m.bf_on[3];
}

void cold_start_mixed(Model &m, int itt, int tau, int taum1, int taum2)
{
    launch_tendency_mixed(m, tau, nullptr);
    if (background_on(m)) launch_background(m, tau);   // provisional step: pre-friction state is close enough
    launch_friction_ab_mixed(m, 0, tau, taum1, taum2, 0.0, nullptr, 0.0, nullptr, false);
    (void)itt;
    float *est = m.dpr[0];
    {   // the divergence reads the provisional fields' halos
        float *hal[3] = {m.ah[0], m.ah[1], m.ah[2]};
        launch_halo_f(m, hal, nullptr, 3);
    }
    solve_mixed(m, m.ah[0], m.ah[1], m.ah[2], est);
    launch_halo_f(m, &est, &m.prm.enable_vertical_boundaries, 1);
    launch_convert(m, 3, m.g.H, nullptr, m.pint, est, nullptr, nullptr, nullptr);
    for (int q = 0; q < 3; ++q) launch_convert(m, 4, m.g.H, nullptr, nullptr, nullptr, m.dpr[q], nullptr, nullptr);
    m.dp_cur = 0;
    m.cold = false;
    m.refine = 2;
}

// One step of the mixed form.  With p_{n-1} the FP64 pressure of the last
// step, the tendency slot of step q holds s_q = d_q - G p_{q-1} (G the
// gradient of the projection, w scaled by 1/dsqr through dt/dsqr), and the
// AB sum sum_q c_q d_q with sum_q c_q = 1 gives
//   u* - dt G p_{n-1} = u + friction + dt sum_q c_q s_q - dt G Q,
//   Q = (c1 + c2) dp_{n-1} + c2 dp_{n-2},  dp_m = p_m - p_{m-1},
// so that L dp_n = div(u* - dt G p_{n-1}) / dt, u_n = u' - dt G dp_n and
// p_n = p_{n-1} + dp_n are the reference's pressure_solver.f90:16-110
// rearranged exactly.  The balanced parts (Coriolis, buoyancy, G p) are
// formed in FP64; what FP32 stores and transforms is small.
void integrate_mixed(Model &m, int itt, int tau, int taum1, int taum2)
{
    const ps3d_params &p = m.prm;
    if (m.cold) cold_start_mixed(m, itt, tau, taum1, taum2);
    launch_tendency_mixed(m, tau, m.pint);
    double qc1 = 0.0, qc2 = 0.0;
    if (itt == 0) {
    } else if (itt == 1 || !p.enable_AB_3_order) {
        qc1 = -(0.5 + p.eps_ab);                  // c1; c2 = 0
    } else {
        qc1 = -16.0 / 12.0 + 5.0 / 12.0;          // c1 + c2
        qc2 = 5.0 / 12.0;                         // c2
    }
    const float *d1 = m.dpr[m.dp_cur];
    const float *d2 = m.dpr[(m.dp_cur + 2) % 3];
    if (background_on(m))   // set_forcing after the friction block, as integrate.f90:82
        launch_friction_forcing_ab_mixed(m, itt, tau, taum1, taum2, qc1, qc1 != 0.0 ? d1 : nullptr, qc2,
                                         qc2 != 0.0 ? d2 : nullptr);
    else
        launch_friction_ab_mixed(m, itt, tau, taum1, taum2, qc1, qc1 != 0.0 ? d1 : nullptr, qc2,
                                 qc2 != 0.0 ? d2 : nullptr, true);
    const int nxt = (m.dp_cur + 1) % 3;
    solve_mixed(m, m.fh[0], m.fh[1], m.fh[2], m.dpr[nxt]);
    launch_project_exchange_f(m, m.dpr[nxt]);
    if (m.refine > 0) {
        // The first steps' increments are large (the first projection
        // removes the initial state's divergence and the next one takes that
        // pressure back), and the FP32 solve's error is relative to the
        // increment: one correction solve on the projected fields, its
        // result added to the projection, to pint and to the increment.
        float *corr = m.dpr[(nxt + 1) % 3];   // dp_{n-2}: not needed any more
        solve_mixed(m, m.fh[0], m.fh[1], m.fh[2], corr);
        launch_project_exchange_f(m, corr);
        launch_convert(m, 5, m.g.H, nullptr, nullptr, corr, m.dpr[nxt], nullptr, nullptr);
        m.refine -= 1;
    }
    m.dp_cur = nxt;
    for (int f = 0; f < 4; ++f) m.lo_stale[f] = true;
    m.p_host.clear();
}

void pressure(Model &m)
{
    const ps3d_params &p = m.prm;
    fft_r2c_x(m, nullptr, 0, 1, m.FS);
    if (p.enable_vertical_boundaries) {
        fft_c2c_y(m, m.FS, -1, p.dx, p.dy);
        launch_thomas(m, m.FS);
        fft_c2c_y(m, m.FS, +1, 0.0, 0.0);
        fft_c2r_x(m, m.FS, m.p, 1, p.Lx * p.Ly);
    } else {
        fft_c2c_y(m, m.FS, -1, 0.0, 0.0);
        fft_z_periodic_solve(m, m.FS);
        fft_c2c_y(m, m.FS, +1, 0.0, 0.0);
        fft_c2r_x(m, m.FS, m.p, 1, p.Lx * p.Ly * p.Lz);
    }
    launch_project_exchange(m);
}

// set_forcing (integrate.f90:82): background flow and body force, the
// wavemaker, then the caller's hook
void apply_forcing(Model &m, int itt, int tau)
{
    if (background_on(m)) launch_background(m, tau);
    if (m.wm_on) {
        const double t = m.prm.dt * itt;   // config.f90:232
        launch_wavemaker(m, std::cos(m.wm_omega * t), std::sin(m.wm_omega * t));
    }
    if (m.forcing) {
        const int rc = m.forcing(static_cast<ps3d_model *>(&m), m.forcing_user, itt, tau, m.stream);
        PS3D_REQUIRE(rc == 0, "forcing hook returned " + std::to_string(rc));
    }
}

void integrate(Model &m, int itt, int tau, int taum1, int taum2)
{
    if (m.mixed) {
        integrate_mixed(m, itt, tau, taum1, taum2);
        return;
    }
    const bool forced = m.forcing || m.wm_on || background_on(m);
    if (fused_tendency_active(m)) {
        launch_friction_ab(m, itt, tau, taum1, taum2, true);   // tendency formed inside
        pressure(m);
        return;
    }
    launch_tendency(m, tau);
    if (fused_friction_ok(m) && m.fuse_friction && !forced) {
        launch_friction_ab(m, itt, tau, taum1, taum2);
    } else if (fused_friction_ok(m) && m.fuse_friction) {
        // set_forcing sees the post-friction state (integrate.f90:82)
        launch_friction_fused_only(m);
        if (!m.wm_on && !m.forcing) {
            launch_background_ab(m, itt, tau, taum1, taum2);   // background flow / body force only
        } else {
            apply_forcing(m, itt, tau);
            launch_ab_exchange(m, itt, tau, taum1, taum2);
        }
    } else {
        launch_friction_block(m);
        if (forced) apply_forcing(m, itt, tau);
        launch_ab_exchange(m, itt, tau, taum1, taum2);
    }
    pressure(m);
}

// ---- graph replay of ps3d_run ---------------------------------------------
// Six steps repeat the whole host-side state of a step: the tau rotation
// (period 3), the buffer parity of the friction pass (2) and the mixed
// form's ring of pressure increments (3).  Captured once from a state, the
// six steps replay from every later state with the same signature.
constexpr int kGraphSteps = 6;

void rotate(int32_t *st)
{
    st[1] = st[1] % 3 + 1;
    st[2] = st[2] % 3 + 1;
    st[3] = st[3] % 3 + 1;
    st[0] += 1;
}

// host-side work per step (a forcing hook, the time-dependent wavemaker) or
// per-kernel timing events cannot be replayed
bool graph_capable(const Model &m)
{
    return !m.forcing && !m.wm_on && m.prof_name.empty() && m.stream != nullptr;
}

// the AB order no longer changes; the mixed form's history is started
bool graph_steady(const Model &m, const int32_t *st)
{
    return st[0] >= 2 && (!m.mixed || (!m.cold && m.refine == 0));
}

uint64_t fnv(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t q = 0; q < n; ++q) h = (h ^ b[q]) * 1099511628211ull;
    return h;
}

// everything a captured step depends on: the tau slots, the buffers the
// fields are in, the ring index, the constants, launch shapes and options
std::vector<uint64_t> graph_signature(const Model &m, const int32_t *st)
{
    std::vector<uint64_t> k = {(uint64_t)st[1], (uint64_t)st[2], (uint64_t)st[3], (uint64_t)m.dp_cur};
    for (int f = 0; f < 4; ++f) k.push_back((uint64_t)m.cur[f]);
    for (const void *p : {(const void *)m.u, (const void *)m.v, (const void *)m.w, (const void *)m.b,
                          (const void *)m.alt[0], (const void *)m.alt[1], (const void *)m.alt[2],
                          (const void *)m.alt[3]})
        k.push_back((uint64_t)(uintptr_t)p);
    for (int f = 0; f < 4; ++f) {
        k.push_back((uint64_t)(uintptr_t)m.fh[f]);
        k.push_back((uint64_t)(uintptr_t)m.ah[f]);
        k.push_back((uint64_t)(uintptr_t)m.fl[f]);
        k.push_back((uint64_t)(uintptr_t)m.al[f]);
    }
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, &m.c, sizeof(m.c));
    h = fnv(h, &m.prm, sizeof(m.prm));
    const int opts[] = {m.tend_kc, m.fric_kc, m.tfric_kc, m.thomas_cpw, (int)m.thomas_reg, (int)m.fuse_tendency,
                        (int)m.fuse_friction, (int)m.fuse_halo, m.fft_plan[0], m.fft_plan[1], m.fft_plan[2],
                        (int)m.bg_on, (int)m.bg_b_on, (int)m.bf_on[0], (int)m.bf_on[1], (int)m.bf_on[2],
                        (int)m.bf_on[3]};
    h = fnv(h, opts, sizeof(opts));
    k.push_back(h);
    return k;
}

void drop_graph(Model &m)
{
    if (m.step_exec) (void)hipGraphExecDestroy(m.step_exec);
    if (m.step_graph) (void)hipGraphDestroy(m.step_graph);
    m.step_exec = nullptr;
    m.step_graph = nullptr;
    m.graph_key.clear();
}

// capture six steps from state st (which is left as it is: the captured
// steps have not run); the host-side state goes round its period and comes
// back to where it was, which the signature checks
void capture_steps(Model &m, const int32_t *st)
{
    drop_graph(m);
    const std::vector<uint64_t> key = graph_signature(m, st);
    int32_t cs[4] = {st[0], st[1], st[2], st[3]};
    PS3D_HIP_CHECK(hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
        for (int q = 0; q < kGraphSteps; ++q) {
            integrate(m, cs[0], cs[1], cs[2], cs[3]);
            rotate(cs);
        }
    } catch (...) {
        (void)hipStreamEndCapture(m.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    PS3D_HIP_CHECK(hipStreamEndCapture(m.stream, &g));
    if (graph_signature(m, st) != key) {
        (void)hipGraphDestroy(g);
        throw Error(PS3D_ERR_STATE, "graph capture: the step's state has no period of six steps");
    }
    m.step_graph = g;
    PS3D_HIP_CHECK(hipGraphInstantiate(&m.step_exec, g, nullptr, nullptr, 0));
    m.graph_key = key;
}

const char *kKernelNames =
    "tendency,friction_ab,tend_friction_ab,harmonic,biharmonic_h,biharmonic_v,ab_update,halo,fft_x_r2c,fft_y_fwd,thomas,"
    "fft_y_inv,fft_x_c2r,z_solve,fft_z_fwd,fft_z_inv,project";

}  // namespace

namespace ps3d {

void model_step(Model &m, int itt, int tau, int taum1, int taum2)
{
    integrate(m, itt, tau, taum1, taum2);
}

std::string &error_slot()
{
    thread_local std::string msg;
    return msg;
}

ProfScope::ProfScope(Model &mm, const char *name) : m(mm), on(false)
{
    if (!m.prof_name.empty() && m.prof_name == name) {
        on = true;
        PS3D_HIP_CHECK(hipEventCreate(&e1));
        PS3D_HIP_CHECK(hipEventCreate(&e2));
        PS3D_HIP_CHECK(hipEventRecord(e1, m.stream));
    }
}

ProfScope::~ProfScope()
{
    if (on) {
        (void)hipEventRecord(e2, m.stream);
        m.prof_ev.push_back(e1);
        m.prof_ev.push_back(e2);
    }
}

}  // namespace ps3d


extern "C" {

int ps3d_last_error(char *buf, int len)
{
    const std::string &g_err = error_slot();
    const int n = (int)g_err.size();
    if (buf && len > 0) {
        const int c = n < len - 1 ? n : len - 1;
        std::memcpy(buf, g_err.data(), (size_t)c);
        buf[c] = '\0';
    }
    return n;
}

const char *ps3d_version(void) { return "ps3d-amd 0.1 gfx950"; }

const char *ps3d_kernel_names(void) { return kKernelNames; }

int ps3d_create(const ps3d_params *prm, const double *ksqr, void *stream, ps3d_model **out)
{
    return guard([&] {
        PS3D_REQUIRE(prm && out, "null argument");
        *out = nullptr;
        const ps3d_params &p = *prm;
        PS3D_REQUIRE(p.precision == PS3D_FP64 || p.precision == PS3D_MIXED,
                     "precision must be PS3D_FP64 or PS3D_MIXED");
        const bool mixed = p.precision == PS3D_MIXED;
        PS3D_REQUIRE(p.nx >= 2 && p.nx % 2 == 0, "nx must be even and >= 2 (main.f90:45-55)");
        PS3D_REQUIRE(p.ny >= 2 && p.ny % 2 == 0, "ny must be even and >= 2 (main.f90:45-55)");
        PS3D_REQUIRE(p.nz >= 2, "nz must be >= 2");
        PS3D_REQUIRE(p.enable_vertical_boundaries || p.nz % 2 == 0, "nz must be even without a rigid lid");
        PS3D_REQUIRE(is_235(p.nx / 2) && is_235(p.ny) && (p.enable_vertical_boundaries || is_235(p.nz)),
                     "FFT sizes (nx/2, ny, nz) must be products of 2, 3 and 5");
        PS3D_REQUIRE(p.ny <= 1024 && p.nz <= 1024 && p.nx <= 8192, "grid too large for the FFT kernels");
        PS3D_REQUIRE(p.dx > 0 && p.dy > 0 && p.dz > 0 && p.dt != 0 && p.dsqr > 0, "bad grid spacing or dt");
        if (mixed) {
            PS3D_REQUIRE(!(p.Avbi > 0.) && !(p.Kvbi > 0.),
                         "PS3D_MIXED: vertical biharmonic friction (Avbi, Kvbi > 0) is only in the FP64 form");
            PS3D_REQUIRE(fft_sizes_compiled(p.nx / 2) && fft_sizes_compiled(p.ny) &&
                             (p.enable_vertical_boundaries || fft_sizes_compiled(p.nz)),
                         "PS3D_MIXED: nx/2, ny (and nz without a rigid lid) must be compiled-in FFT lengths (" +
                             fft_sizes_text() + ")");
        }
        auto *m = new ps3d_model();
        try {
            m->prm = p;
            Grid &g = m->g;
            g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
            g.X = p.nx + 2 * ONX; g.Y = p.ny + 2 * ONX; g.Z = p.nz + 2 * ONX;
            g.XY = g.X * g.Y;
            g.H = g.XY * g.Z;
            g.N = (int64_t)p.nx * p.ny * p.nz;
            g.nkx = p.nx / 2 + 1;
            g.S = (int64_t)g.nkx * p.ny * p.nz;
            StepConsts &c = m->c;
            c.nx = p.nx; c.ny = p.ny; c.nz = p.nz;
            c.X = g.X; c.XY = g.XY;
            c.lid = p.enable_vertical_boundaries;
            c.nonlinear = p.enable_nonlinear;
            c.dx = p.dx; c.dy = p.dy; c.dz = p.dz; c.dt = p.dt;
            c.Ro = p.Ro; c.dsqr = p.dsqr; c.f0 = p.f0; c.N0 = p.N0;
            c.rdx = 1.0 / p.dx; c.rdy = 1.0 / p.dy; c.rdz = 1.0 / p.dz; c.rdt = 1.0 / p.dt;
            m->fuse_halo = halo_images_ok(p.nx, p.ny, p.nz);
            if (stream) {
                m->stream = static_cast<hipStream_t>(stream);
            } else {
                PS3D_HIP_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
                m->own_stream = true;
            }
            if (mixed) {
                m->mixed = true;
                for (int f = 0; f < 4; ++f) {
                    m->fh[f] = dalloc<float>(*m, g.H);
                    m->fl[f] = dalloc<float>(*m, g.H);
                    m->ah[f] = dalloc<float>(*m, g.H);
                    m->al[f] = dalloc<float>(*m, g.H);
                    m->tf[f] = dalloc<float>(*m, 3 * g.H);
                }
                m->pint = dalloc<double>(*m, g.H);
                for (int q = 0; q < 3; ++q) m->dpr[q] = dalloc<float>(*m, g.H);
                m->FSf = dalloc<float2>(*m, g.S);
                m->stage = dalloc<double>(*m, 3 * g.H);
                // Thomas scratch for columns whose sweep does not fit the LDS
                // of the on-the-fly form (32 columns x nz x 24 B > 64 KiB)
                if (p.enable_vertical_boundaries && 32 * 24 * (int64_t)p.nz > 64 * 1024)
                    m->dp64 = dalloc<double2>(*m, g.S);
            } else {
                m->u = dalloc<double>(*m, g.H);
                m->v = dalloc<double>(*m, g.H);
                m->w = dalloc<double>(*m, g.H);
                m->b = dalloc<double>(*m, g.H);
                m->p = dalloc<double>(*m, g.H);
                for (int f = 0; f < 4; ++f) m->alt[f] = dalloc<double>(*m, g.H);
                m->du = dalloc<double>(*m, 3 * g.H);
                m->dv = dalloc<double>(*m, 3 * g.H);
                m->dw = dalloc<double>(*m, 3 * g.H);
                m->db = dalloc<double>(*m, 3 * g.H);
                m->FS = dalloc<double2>(*m, g.S);
                m->aloc = dalloc<double>(*m, g.N);
            }
            m->ksqr = dalloc<double>(*m, ksqr_len(*m));
            if (p.enable_vertical_boundaries) {
                m->tri_cp = dalloc<double>(*m, g.S);
                m->tri_rm = dalloc<double>(*m, g.S);
            }
            fft_plan_init(*m, m->px, p.nx / 2);
            fft_plan_init(*m, m->py, p.ny);
            if (!p.enable_vertical_boundaries) fft_plan_init(*m, m->pz, p.nz);
            {
                std::vector<double2> wk(g.nkx);
                for (int k = 0; k < g.nkx; ++k) wk[k] = host_twiddle(k, p.nx);
                m->tw_r2c = dalloc<double2>(*m, g.nkx);
                PS3D_HIP_CHECK(hipMemcpyAsync(m->tw_r2c, wk.data(), sizeof(double2) * g.nkx,
                                              hipMemcpyHostToDevice, m->stream));
                std::vector<float2> wkf(g.nkx);
                for (int k = 0; k < g.nkx; ++k) wkf[k] = make_float2((float)wk[k].x, (float)wk[k].y);
                m->twf_r2c = dalloc<float2>(*m, g.nkx);
                PS3D_HIP_CHECK(hipMemcpyAsync(m->twf_r2c, wkf.data(), sizeof(float2) * g.nkx,
                                              hipMemcpyHostToDevice, m->stream));
                PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
            }
            std::vector<double> ks;
            const double *ksrc = ksqr;
            if (!ksrc) {
                ks = make_ksqr(p);
                ksrc = ks.data();
            }
            PS3D_HIP_CHECK(hipMemcpyAsync(m->ksqr, ksrc, sizeof(double) * ksqr_len(*m), hipMemcpyHostToDevice,
                                          m->stream));
            tridiag_factor(*m);
            PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
        } catch (...) {
            ps3d_destroy(m);
            throw;
        }
        *out = m;
    });
}

int ps3d_destroy(ps3d_model *m)
{
    return guard([&] {
        if (!m) return;
        if (m->stream) (void)hipStreamSynchronize(m->stream);
        drop_graph(*m);
        balance_release(*m);
        for (auto ev : m->prof_ev) (void)hipEventDestroy(ev);
        for (void *p : m->allocations) (void)hipFree(p);
        if (m->own_stream && m->stream) (void)hipStreamDestroy(m->stream);
        delete m;
    });
}

int ps3d_set_params(ps3d_model *m, const ps3d_params *prm)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(prm, "null argument");
        const ps3d_params &o = m->prm, &p = *prm;
        // what the grid, the wavenumber tables and the solver factors were
        // built from stays fixed for the model's lifetime
        PS3D_REQUIRE(p.nx == o.nx && p.ny == o.ny && p.nz == o.nz && p.Lx == o.Lx && p.Ly == o.Ly &&
                         p.Lz == o.Lz && p.dx == o.dx && p.dy == o.dy && p.dz == o.dz && p.dsqr == o.dsqr &&
                         p.enable_vertical_boundaries == o.enable_vertical_boundaries &&
                         p.precision == o.precision,
                     "ps3d_set_params: only dt, Ro, eps_ab, f0, N0, the friction coefficients, "
                     "enable_nonlinear and enable_AB_3_order can change after ps3d_create");
        PS3D_REQUIRE(p.dt != 0.0, "dt must be nonzero");
        PS3D_REQUIRE(!m->mixed || (!(p.Avbi > 0.) && !(p.Kvbi > 0.)),
                     "PS3D_MIXED: vertical biharmonic friction (Avbi, Kvbi > 0) is only in the FP64 form");
        for (int f = 0; f < 4; ++f)
            PS3D_REQUIRE(m->cur[f] == 0, "ps3d_set_params: friction results not yet folded in by ab_update");
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
        ps3d_params &q = m->prm;
        q.dt = p.dt; q.Ro = p.Ro; q.eps_ab = p.eps_ab; q.f0 = p.f0; q.N0 = p.N0;
        q.Ah = p.Ah; q.Av = p.Av; q.Kh = p.Kh; q.Kv = p.Kv;
        q.Ahbi = p.Ahbi; q.Avbi = p.Avbi; q.Khbi = p.Khbi; q.Kvbi = p.Kvbi;
        q.enable_nonlinear = p.enable_nonlinear;
        q.enable_AB_3_order = p.enable_AB_3_order;
        StepConsts &c = m->c;
        c.dt = p.dt; c.rdt = 1.0 / p.dt;
        c.Ro = p.Ro; c.f0 = p.f0; c.N0 = p.N0;
        c.nonlinear = p.enable_nonlinear;
    });
}

int ps3d_field_size(const ps3d_model *m, int field, int64_t *n)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(n, "null argument");
        const int64_t l = field_len(*m, field);
        PS3D_REQUIRE(l >= 0, "unknown field id");
        *n = l;
    });
}

int ps3d_set_field(ps3d_model *m, int field, const double *host)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(host, "null host pointer");
        const int64_t l = field_len(*m, field);
        PS3D_REQUIRE(l >= 0, "unknown field id");
        if (m->mixed && field != PS3D_KSQR) {
            set_field_mixed(*m, field, host, l);
            if (field <= PS3D_B) m->lo_stale[field] = false;
        } else {
            PS3D_HIP_CHECK(hipMemcpyAsync(field_dev(*m, field), host, sizeof(double) * l, hipMemcpyHostToDevice,
                                          m->stream));
        }
        if (field == PS3D_KSQR) tridiag_factor(*m);
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

int ps3d_get_field(ps3d_model *m, int field, double *host)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(host, "null host pointer");
        const int64_t l = field_len(*m, field);
        PS3D_REQUIRE(l >= 0, "unknown field id");
        if (m->mixed && field != PS3D_KSQR)
            get_field_mixed(*m, field, host, l);
        else
            PS3D_HIP_CHECK(hipMemcpyAsync(host, field_dev(*m, field), sizeof(double) * l, hipMemcpyDeviceToHost,
                                          m->stream));
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

int ps3d_integrate(ps3d_model *m, int itt, int tau, int taum1, int taum2)
{
    return guard([&] {
        check_model(m);
        check_tau(tau, taum1, taum2);
        PS3D_REQUIRE(itt >= 0, "itt must be >= 0");
        integrate(*m, itt, tau, taum1, taum2);
    });
}

int ps3d_run(ps3d_model *m, int nsteps, int32_t *st, int use_graph)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(st && nsteps >= 0, "bad arguments");
        int s = 0;
        auto eager = [&] {
            check_tau(st[1], st[2], st[3]);
            integrate(*m, st[0], st[1], st[2], st[3]);
            rotate(st);
            ++s;
        };
        if (use_graph) {
            // steps that cannot be replayed run first: the first two (Euler,
            // AB2) and the mixed form's cold start and correction steps
            while (s < nsteps && graph_capable(*m) && !graph_steady(*m, st)) eager();
            if (graph_capable(*m) && nsteps - s >= kGraphSteps) {
                if (!m->step_exec || m->graph_key != graph_signature(*m, st)) {
                    // one eager step first, so that whatever the step sets up
                    // on first use is in place before the capture
                    eager();
                    if (nsteps - s >= kGraphSteps) capture_steps(*m, st);
                }
                while (m->step_exec && nsteps - s >= kGraphSteps) {
                    PS3D_HIP_CHECK(hipGraphLaunch(m->step_exec, m->stream));
                    for (int q = 0; q < kGraphSteps; ++q) rotate(st);
                    s += kGraphSteps;
                }
            }
        }
        while (s < nsteps) eager();
    });
}

int ps3d_synchronize(ps3d_model *m)
{
    return guard([&] {
        check_model(m);
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

int ps3d_tendency(ps3d_model *m, int tau)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_tendency");
        PS3D_REQUIRE(tau >= 1 && tau <= 3, "tau must be 1..3");
        launch_tendency(*m, tau);
    });
}

int ps3d_buoyancy_advection(ps3d_model *m, int tau)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_buoyancy_advection");
        PS3D_REQUIRE(tau >= 1 && tau <= 3, "tau must be 1..3");
        launch_buoyancy_advection(*m, tau);
    });
}

int ps3d_momentum_advection(ps3d_model *m, int tau)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_momentum_advection");
        PS3D_REQUIRE(tau >= 1 && tau <= 3, "tau must be 1..3");
        launch_momentum_advection(*m, tau);
    });
}

int ps3d_friction(ps3d_model *m)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_friction");
        launch_friction_block(*m);
    });
}

int ps3d_mixing(ps3d_model *m, int op, int field, double hmix, double vmix)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_mixing");
        PS3D_REQUIRE(field >= PS3D_U && field <= PS3D_B, "mixing field must be u, v, w or b");
        PS3D_REQUIRE(op >= PS3D_HARMONIC && op <= PS3D_BIHARMONIC_V, "unknown mixing operator");
        launch_mixing(*m, op, field, hmix, vmix, 0);
    });
}

int ps3d_ab_update(ps3d_model *m, int itt, int tau, int taum1, int taum2)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_ab_update");
        check_tau(tau, taum1, taum2);
        launch_ab_exchange(*m, itt, tau, taum1, taum2);
    });
}

int ps3d_border_exchg(ps3d_model *m, int field)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_border_exchg");
        PS3D_REQUIRE(field >= PS3D_U && field <= PS3D_P, "border_exchg field must be u, v, w, b or p");
        launch_border_exchg(*m, field == PS3D_P ? m->p : m->cur_buf(field), 0);
    });
}

int ps3d_pressure_solver(ps3d_model *m)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_pressure_solver");
        for (int f = 0; f < 4; ++f)
            PS3D_REQUIRE(m->cur[f] == 0, "pressure_solver: friction results not yet folded in by ab_update");
        pressure(*m);
    });
}

int ps3d_my_fft(ps3d_model *m, int three_d, const double *qp, double *qs)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_my_fft");
        PS3D_REQUIRE(qp && qs, "null buffer");
        const ps3d_params &p = m->prm;
        PS3D_REQUIRE(!three_d || !p.enable_vertical_boundaries,
                     "3-D transform needs a periodic model (no z plan with a rigid lid)");
        PS3D_HIP_CHECK(hipMemcpyAsync(m->aloc, qp, sizeof(double) * m->g.N, hipMemcpyHostToDevice, m->stream));
        fft_r2c_x(*m, m->aloc, 0, 0, m->FS);
        if (three_d) {
            fft_c2c_y(*m, m->FS, -1, 0.0, 0.0);
            fft_c2c_z(*m, m->FS, -1, p.dx, p.dy, p.dz);
        } else {
            fft_c2c_y(*m, m->FS, -1, p.dx, p.dy);
        }
        PS3D_HIP_CHECK(hipMemcpyAsync(qs, m->FS, sizeof(double2) * m->g.S, hipMemcpyDeviceToHost, m->stream));
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

int ps3d_my_ifft(ps3d_model *m, int three_d, const double *qs, double *qp)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_my_ifft");
        PS3D_REQUIRE(qp && qs, "null buffer");
        const ps3d_params &p = m->prm;
        PS3D_REQUIRE(!three_d || !p.enable_vertical_boundaries,
                     "3-D transform needs a periodic model (no z plan with a rigid lid)");
        PS3D_HIP_CHECK(hipMemcpyAsync(m->FS, qs, sizeof(double2) * m->g.S, hipMemcpyHostToDevice, m->stream));
        if (three_d) fft_c2c_z(*m, m->FS, +1, 0.0, 0.0, 0.0);
        fft_c2c_y(*m, m->FS, +1, 0.0, 0.0);
        fft_c2r_x(*m, m->FS, m->aloc, 0, three_d ? p.Lx * p.Ly * p.Lz : p.Lx * p.Ly);
        PS3D_HIP_CHECK(hipMemcpyAsync(qp, m->aloc, sizeof(double) * m->g.N, hipMemcpyDeviceToHost, m->stream));
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

int ps3d_spectral_solve(ps3d_model *m, double *qs)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_spectral_solve");
        PS3D_REQUIRE(qs, "null buffer");
        PS3D_HIP_CHECK(hipMemcpyAsync(m->FS, qs, sizeof(double2) * m->g.S, hipMemcpyHostToDevice, m->stream));
        if (m->prm.enable_vertical_boundaries) launch_thomas(*m, m->FS);
        else launch_divide_ksqr(*m, m->FS);
        PS3D_HIP_CHECK(hipMemcpyAsync(qs, m->FS, sizeof(double2) * m->g.S, hipMemcpyDeviceToHost, m->stream));
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

// my_fft.f90:87-110 for the wavenumber tables the caller does not pass
static void default_wavenumbers(const Model &m, const double *&kx, const double *&ky, const double *&kz,
                                std::vector<double> &wx, std::vector<double> &wy, std::vector<double> &wz)
{
    const ps3d_params &p = m.prm;
    const double pi = std::acos(0.0) * 2.0;
    if (!kx) {
        wx.resize(m.g.nkx);
        for (int i = 1; i <= m.g.nkx; ++i) wx[i - 1] = (double)(2 * (i - 1)) * pi / p.Lx;
        kx = wx.data();
    }
    if (!ky) {
        wy.resize(p.ny);
        for (int j = 1; j <= (p.ny + 1) / 2; ++j) wy[j - 1] = (double)(2 * (j - 1)) * pi / p.Ly;
        for (int j = (p.ny + 1) / 2 + 1; j <= p.ny; ++j) wy[j - 1] = -((double)(2 * (p.ny - j + 1)) * pi / p.Ly);
        ky = wy.data();
    }
    if (!kz) {
        wz.resize(p.nz);
        if (p.enable_vertical_boundaries) {
            for (int k = 1; k <= p.nz; ++k) wz[k - 1] = (double)(float)(k - 1) * pi / p.Lz;
        } else {
            for (int k = 1; k <= (p.nz + 1) / 2; ++k) wz[k - 1] = (double)(2 * (k - 1)) * pi / p.Lz;
            for (int k = (p.nz + 1) / 2 + 1; k <= p.nz; ++k) wz[k - 1] = -((double)(2 * (p.nz - k + 1)) * pi / p.Lz);
        }
        kz = wz.data();
    }
}

int ps3d_diag_balance(ps3d_model *m, int filter_width, const double *kx, const double *ky, const double *kz,
                      double *const *out)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_diag_balance");
        PS3D_REQUIRE(out, "null output table");
        std::vector<double> wx, wy, wz;
        default_wavenumbers(*m, kx, ky, kz, wx, wy, wz);
        diag_balance(*m, filter_width, out, kx, ky, kz);
    });
}

int ps3d_opt_balance(ps3d_model *m, const ps3d_opt_balance_params *ob, const double *kx, const double *ky,
                     const double *kz, double *const out[4], double *norms, int32_t *iterations)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_opt_balance");
        PS3D_REQUIRE(ob, "null parameters");
        std::vector<double> wx, wy, wz;
        default_wavenumbers(*m, kx, ky, kz, wx, wy, wz);
        int its = 0;
        opt_balance(*m, *ob, out, norms, &its, kx, ky, kz);
        if (iterations) *iterations = its;
    });
}

int ps3d_tune(ps3d_model *m, int next_tau, const char *cache_path, int reps)
{
    return guard([&] {
        check_model(m);
        tune_model(*m, next_tau, (cache_path && *cache_path) ? cache_path : nullptr, reps);
    });
}

int ps3d_tuning(ps3d_model *m, char *buf, int len)
{
    int n = 0;
    const int rc = guard([&] {
        check_model(m);
        const std::string s = describe_tuning(*m);
        n = (int)s.size();
        if (buf && len > 0) {
            const int c = n < len - 1 ? n : len - 1;
            std::memcpy(buf, s.data(), (size_t)c);
            buf[c] = '\0';
        }
    });
    return rc == PS3D_OK ? n : -rc;
}

int ps3d_set_forcing(ps3d_model *m, ps3d_forcing_fn fn, void *user)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_set_forcing");
        m->forcing = fn;
        m->forcing_user = fn ? user : nullptr;
    });
}

int ps3d_device_ptr(ps3d_model *m, int field, int slot, void **ptr)
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_device_ptr");
        PS3D_REQUIRE(ptr, "null output");
        if (field >= PS3D_U && field <= PS3D_B) {
            *ptr = m->cur_buf(field);
        } else if (field == PS3D_P) {
            *ptr = m->p;
        } else if (field >= PS3D_DU && field <= PS3D_DB) {
            PS3D_REQUIRE(slot >= 1 && slot <= 3, "tendency slot must be 1..3");
            double *t[4] = {m->du, m->dv, m->dw, m->db};
            *ptr = t[field - PS3D_DU] + (int64_t)(slot - 1) * m->g.H;
        } else {
            throw Error(PS3D_ERR_ARG, "device_ptr: field must be u, v, w, b, p or du..db");
        }
    });
}

int ps3d_set_wavemaker(ps3d_model *m, double omega, const double *const re[4], const double *const im[4])
{
    return guard([&] {
        check_model(m);
        require_fp64(*m, "ps3d_set_wavemaker");
        PS3D_REQUIRE(re && im, "null amplitude table");
        PS3D_REQUIRE(std::isfinite(omega), "omega must be finite");
        bool any = false;
        for (int f = 0; f < 4; ++f) {
            PS3D_REQUIRE((re[f] == nullptr) == (im[f] == nullptr),
                         "wavemaker: give both or neither of re[f] and im[f]");
            m->wm_field[f] = re[f] != nullptr;
            if (!re[f]) continue;
            any = true;
            if (!m->wm_re[f]) {
                m->wm_re[f] = dalloc<double>(*m, m->g.H);
                m->wm_im[f] = dalloc<double>(*m, m->g.H);
            }
            PS3D_HIP_CHECK(hipMemcpyAsync(m->wm_re[f], re[f], sizeof(double) * (size_t)m->g.H,
                                          hipMemcpyHostToDevice, m->stream));
            PS3D_HIP_CHECK(hipMemcpyAsync(m->wm_im[f], im[f], sizeof(double) * (size_t)m->g.H,
                                          hipMemcpyHostToDevice, m->stream));
        }
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
        m->wm_omega = omega;
        m->wm_on = any;
    });
}

int ps3d_set_background_flow(ps3d_model *m, const double *u_back, const double *b_back)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(u_back || !b_back, "background flow: b_back needs u_back");
        const size_t bytes = sizeof(double) * (size_t)m->g.H;
        if (u_back) {
            if (!m->bg_u) m->bg_u = dalloc<double>(*m, m->g.H);
            PS3D_HIP_CHECK(hipMemcpyAsync(m->bg_u, u_back, bytes, hipMemcpyHostToDevice, m->stream));
        }
        if (b_back) {
            if (!m->bg_b) m->bg_b = dalloc<double>(*m, m->g.H);
            PS3D_HIP_CHECK(hipMemcpyAsync(m->bg_b, b_back, bytes, hipMemcpyHostToDevice, m->stream));
        }
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
        m->bg_on = u_back != nullptr;
        m->bg_b_on = b_back != nullptr;
    });
}

int ps3d_set_body_force(ps3d_model *m, const double *const f[4])
{
    return guard([&] {
        check_model(m);
        const size_t bytes = sizeof(double) * (size_t)m->g.H;
        for (int q = 0; q < 4; ++q) {
            const double *src = f ? f[q] : nullptr;
            if (src) {
                if (!m->bf[q]) m->bf[q] = dalloc<double>(*m, m->g.H);
                PS3D_HIP_CHECK(hipMemcpyAsync(m->bf[q], src, bytes, hipMemcpyHostToDevice, m->stream));
            }
            m->bf_on[q] = src != nullptr;
        }
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

int ps3d_clear_wavemaker(ps3d_model *m)
{
    return guard([&] {
        check_model(m);
        m->wm_on = false;
        for (int f = 0; f < 4; ++f) m->wm_field[f] = false;
    });
}

int ps3d_set_fused_tendency(ps3d_model *m, int on)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(on == 0 || on == 1, "on must be 0 or 1");
        m->fuse_tendency = on != 0;
    });
}

int ps3d_set_fft_plan(ps3d_model *m, const int32_t plans[3])
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(plans, "null plans");
        for (int q = 0; q < 3; ++q) {
            PS3D_REQUIRE(plans[q] == 0 || plans[q] == 1, "radix plans are 0 or 1");
            PS3D_REQUIRE(m->mixed || plans[q] == 0,
                         "FP64 transforms keep plan 0 (the factorisation the oracle reproduces bit for bit)");
        }
        for (int q = 0; q < 3; ++q) m->fft_plan[q] = plans[q];
    });
}

int ps3d_get_fft_plan(ps3d_model *m, int32_t plans[3])
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(plans, "null plans");
        for (int q = 0; q < 3; ++q) plans[q] = m->fft_plan[q];
    });
}

int ps3d_diagnose(ps3d_model *m, double *out, int32_t *nan_seen)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(out, "null output");
        const bool nan = monitor(*m, out);
        if (nan_seen) *nan_seen = nan ? 1 : 0;
    });
}

int ps3d_diagnose_noalloc(const ps3d_params *prm, const double *u, const double *v, const double *w, const double *b,
                          void *stream, double *out, int32_t *nan_seen)
{
    return guard([&] {
        PS3D_REQUIRE(prm && out, "null argument");
        PS3D_REQUIRE(u && v && w && b, "null field");
        const bool nan = monitor_noalloc(*prm, u, v, w, b, static_cast<hipStream_t>(stream), out);
        if (nan_seen) *nan_seen = nan ? 1 : 0;
    });
}

int ps3d_write_restart(ps3d_model *m, const char *path, int itt, double time, int tau, int taum1, int taum2)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(path, "null path");
        check_tau(tau, taum1, taum2);
        int32_t it = itt, t0 = tau, t1 = taum1, t2 = taum2;
        double tm = time;
        restart_io(*m, path, true, &it, &tm, &t0, &t1, &t2);
    });
}

int ps3d_read_restart(ps3d_model *m, const char *path, int32_t *itt, double *time, int32_t *tau, int32_t *taum1,
                      int32_t *taum2)
{
    return guard([&] {
        check_model(m);
        PS3D_REQUIRE(path && itt && time && tau && taum1 && taum2, "null argument");
        restart_io(*m, path, false, itt, time, tau, taum1, taum2);
    });
}

int ps3d_restart_write_host(const char *path, int nx, int ny, int nz, int itt, double time, int tau, int taum1,
                            int taum2, const double *state)
{
    return guard([&] {
        PS3D_REQUIRE(path && state, "null argument");
        PS3D_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad grid size");
        restart_write_host(path, nx, ny, nz, itt, time, tau, taum1, taum2, state);
    });
}

int ps3d_restart_read_host(const char *path, int nx, int ny, int nz, int32_t *itt, double *time, int32_t *tau,
                           int32_t *taum1, int32_t *taum2, double *state)
{
    return guard([&] {
        PS3D_REQUIRE(path && state && itt && time && tau && taum1 && taum2, "null argument");
        PS3D_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad grid size");
        restart_read_host(path, nx, ny, nz, itt, time, tau, taum1, taum2, state);
    });
}

int ps3d_profile_kernel(ps3d_model *m, const char *name)
{
    return guard([&] {
        check_model(m);
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
        for (auto ev : m->prof_ev) (void)hipEventDestroy(ev);
        m->prof_ev.clear();
        m->prof_launches = 0;
        m->prof_ms = 0.0;
        m->prof_name = name ? name : "";
    });
}

int ps3d_profile_read(ps3d_model *m, int64_t *launches, double *total_ms)
{
    return guard([&] {
        check_model(m);
        PS3D_HIP_CHECK(hipStreamSynchronize(m->stream));
        for (size_t i = 0; i + 1 < m->prof_ev.size(); i += 2) {
            float ms = 0.f;
            PS3D_HIP_CHECK(hipEventElapsedTime(&ms, m->prof_ev[i], m->prof_ev[i + 1]));
            m->prof_ms += ms;
            m->prof_launches += 1;
        }
        for (auto ev : m->prof_ev) (void)hipEventDestroy(ev);
        m->prof_ev.clear();
        if (launches) *launches = m->prof_launches;
        if (total_ms) *total_ms = m->prof_ms;
        m->prof_launches = 0;
        m->prof_ms = 0.0;
    });
}

}  // extern "C"
