// ensemble.hip -- batched ensemble of ps3D models: the runtime around the
// batched stage kernels (member table, block maps, step state, hipGraph
// replay) and the autotuner of the k-marching tile shapes.
//
// Each member is an ordinary model (buffers, FFT tables, ksqr, Thomas
// factors), all of one form, FP64 or mixed; the ensemble owns a device table
// that points into them and runs one launch per stage over all of them.  The
// bodies of the batched kernels are those of the single-model kernels, so
// every member's result is the same, bit for bit, as a single model's.
#include "ps3d_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

using namespace ps3d;

namespace {

// S_PTS: 256 points per work-group over the H points of a member
enum Stage { S_TEND, S_FRIC, S_R2C, S_COLS, S_THOMAS, S_ZSOLVE, S_PROJ, S_PTS, S_HCOPY, S_BGAB, S_NSTAGE };

// tile shapes of the k-marching kernels (stencil.hip)
constexpr int kTTX = TEND_TX, kTTY = TEND_TY, kFTX = FRIC_TX, kFTY = FRIC_TY;

int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace

struct ps3d_ensemble {
    std::vector<ps3d_model *> mem;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<EnsMember> tab;
    EnsMember *dev_tab = nullptr;
    EnsState *dev_state = nullptr;
    EnsState state{0, 3, 2, 1, 0, 0};   // the device state once the queued steps have run
    EnsMap maps[S_NSTAGE]{};
    size_t lds_x = 0, lds_y = 0, lds_z = 0;
    int ny_max = 0, nz_max = 0;         // longest y and (periodic) z columns
    int nzl_max = 0;                    // most levels of a rigid-lid member (Thomas LDS)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::map<std::string, int> tuned;   // "stage shape" -> levels per work-group
    // mixed form: the first step starts the pressure history (cold start),
    // and it and the next one add a correction solve (as a single model)
    bool mixed = false, cold = true;
    int refine = 0;
    // every rigid-lid member's Thomas solve has the register form
    bool thomas_reg = false;
    bool thomas_split = false;   // (when not thomas_reg) k_ens_thomas_split
    std::vector<void *> tmp_alloc;      // third buffers of forced members
};

namespace {

std::string shape_key(const ps3d_params &p)
{
    std::ostringstream o;
    o << p.nx << "x" << p.ny << "x" << p.nz << (p.enable_vertical_boundaries ? ":lid" : ":periodic")
      << (p.precision == PS3D_MIXED ? ":mixed" : ":fp64");
    return o.str();
}

// levels per work-group -> (kc, nchunk)
void set_chunks(int nz, int kc, int &okc, int &onch)
{
    kc = std::max(1, std::min(kc, nz));
    onch = cdiv(nz, kc);
    okc = cdiv(nz, onch);
}

// the single-model launchers' choices (stencil.hip), the untuned default
int default_tend_kc(const Grid &g)
{
    const int tiles = cdiv(g.nx, kTTX) * cdiv(g.ny, kTTY);
    int nchunk = cdiv(512, tiles);
    const int maxch = g.nz / 8 > 1 ? g.nz / 8 : 1;
    nchunk = std::min(nchunk, maxch);
    return cdiv(g.nz, nchunk);
}

int default_fric_kc(const Grid &g)
{
    const int tiles = cdiv(g.nx, kFTX) * cdiv(g.ny, kFTY);
    int nchunk = cdiv(2048 * 512 / (kFTX * kFTY), tiles * 4);
    nchunk = std::max(1, std::min(nchunk, g.nz));
    return cdiv(g.nz, nchunk);
}

void fill_member(EnsMember &e, ps3d_model &m)
{
    const ps3d_params &p = m.prm;
    const Grid &g = m.g;
    e = EnsMember{};
    e.c = m.c;
    e.H = g.H;
    double *canon[4] = {m.u, m.v, m.w, m.b};
    double *tend[4] = {m.du, m.dv, m.dw, m.db};
    for (int q = 0; q < 4; ++q) {
        e.f[q][0] = canon[q];
        e.f[q][1] = m.alt[q];
        e.tend[q] = tend[q];
    }
    e.p = m.p;
    e.FS = m.FS;
    e.ksqr = m.ksqr;
    e.cp = m.tri_cp;
    e.rm = m.tri_rm;
    e.twx = m.px.tw;
    e.twy = m.py.tw;
    e.twz = m.pz.tw;
    e.wk = m.tw_r2c;
    e.Hm[0] = p.Ah; e.Vm[0] = p.Av; e.Hbi[0] = p.Ahbi;
    e.harm[0] = (p.Ah > 0. || p.Av > 0.); e.bih[0] = p.Ahbi > 0.;
    e.Hm[1] = p.Kh; e.Vm[1] = p.Kv; e.Hbi[1] = p.Khbi;
    e.harm[1] = (p.Kh > 0. || p.Kv > 0.); e.bih[1] = p.Khbi > 0.;
    e.eps_ab = p.eps_ab;
    e.ab3 = p.enable_AB_3_order;
    e.fxa = 1.0 / (p.dz * p.dz) / p.dsqr;   // tridiag.hip: the factors' fxa
    e.lxy = p.Lx * p.Ly;
    e.lxyz = p.Lx * p.Ly * p.Lz;
    e.rlxy = 1.0 / e.lxy;
    e.rlxyz = 1.0 / e.lxyz;
    e.n2 = p.nx / 2;
    e.nkx = g.nkx;
    e.tend_tx = cdiv(g.nx, kTTX);
    e.tend_ty = cdiv(g.ny, kTTY);
    e.fric_tx = cdiv(g.nx, kFTX);
    e.fric_ty = cdiv(g.ny, kFTY);
    if (m.mixed) {
        for (int q = 0; q < 4; ++q) {
            e.fh[q][0] = m.fh[q];
            e.fh[q][1] = m.ah[q];
            e.fl[q][0] = m.fl[q];
            e.fl[q][1] = m.al[q];
            e.tf[q] = m.tf[q];
        }
        e.pint = m.pint;
        for (int q = 0; q < 3; ++q) e.dpr[q] = m.dpr[q];
        e.FSf = m.FSf;
        e.twxf = m.px.twf;
        e.twyf = m.py.twf;
        e.twzf = m.pz.twf;
        e.wkf = m.twf_r2c;
    }
    for (int q = 0; q < 3; ++q) e.fft_plan[q] = m.mixed ? m.fft_plan[q] : 0;
}

void build_maps(ps3d_ensemble &E)
{
    const int n = (int)E.mem.size();
    for (int s = 0; s < S_NSTAGE; ++s) {
        E.maps[s].n = n;
        int64_t acc = 0;
        for (int i = 0; i < n; ++i) {
            const EnsMember &e = E.tab[i];
            const Grid &g = E.mem[i]->g;
            const bool lid = e.c.lid != 0;
            int64_t nb = 0;
            switch (s) {
            case S_TEND: nb = (int64_t)e.tend_tx * e.tend_ty * e.tend_nchunk; break;
            case S_FRIC: nb = (int64_t)e.fric_tx * e.fric_ty * e.fric_nchunk * 4; break;
            case S_R2C: nb = ens_r2c_blocks(e); break;
            case S_COLS: nb = ens_cols_blocks(e, g.ny, g.nz); break;
            case S_THOMAS: nb = lid ? cdiv((int64_t)g.nkx * g.ny, ens_thomas_cols(E.thomas_reg, E.thomas_split)) : 0; break;
            case S_ZSOLVE: nb = lid ? 0 : ens_cols_blocks(e, g.nz, g.ny); break;
            case S_PROJ: nb = (int64_t)cdiv(g.nx, 64) * cdiv(g.ny, 4) * g.nz; break;
            case S_PTS: nb = E.mixed ? cdiv(g.H, 256) : 0; break;
            case S_HCOPY: nb = e.forced ? 4 * cdiv(g.H - g.N, 256) : 0; break;
            case S_BGAB: nb = e.forced ? (int64_t)cdiv(g.nx, 64) * cdiv(g.ny, 4) * g.nz : 0; break;
            }
            E.maps[s].start[i] = (int)acc;
            acc += nb;
        }
        PS3D_REQUIRE(acc < (int64_t)1 << 31, "ensemble too large for one launch per stage");
        E.maps[s].start[n] = (int)acc;
    }
    E.lds_x = E.lds_y = E.lds_z = 0;
    E.ny_max = E.nz_max = E.nzl_max = 0;
    for (int i = 0; i < n; ++i) {
        const Grid &g = E.mem[i]->g;
        if (E.tab[i].c.lid) E.nzl_max = std::max(E.nzl_max, g.nz);
        E.lds_x = std::max(E.lds_x, ens_x_lds(g.nx / 2));
        E.lds_y = std::max(E.lds_y, ens_cols_lds(g.ny));
        E.ny_max = std::max(E.ny_max, g.ny);
        if (!E.tab[i].c.lid) {
            E.lds_z = std::max(E.lds_z, ens_cols_lds(g.nz));
            E.nz_max = std::max(E.nz_max, g.nz);
        }
    }
}

void upload_table(ps3d_ensemble &E)
{
    PS3D_HIP_CHECK(hipMemcpyAsync(E.dev_tab, E.tab.data(), sizeof(EnsMember) * E.tab.size(), hipMemcpyHostToDevice,
                                  E.stream));
    PS3D_HIP_CHECK(hipStreamSynchronize(E.stream));
    if (E.exec) {
        (void)hipGraphExecDestroy(E.exec);
        E.exec = nullptr;
    }
    if (E.graph) {
        (void)hipGraphDestroy(E.graph);
        E.graph = nullptr;
    }
}

void apply_tiles(ps3d_ensemble &E, int stage, const std::string &key, int kc)
{
    for (size_t i = 0; i < E.mem.size(); ++i) {
        if (shape_key(E.mem[i]->prm) != key) continue;
        EnsMember &e = E.tab[i];
        if (stage == S_TEND) set_chunks(e.c.nz, kc, e.tend_kc, e.tend_nchunk);
        else set_chunks(e.c.nz, kc, e.fric_kc, e.fric_nchunk);
    }
}

void launch_stage(ps3d_ensemble &E, int s)
{
    switch (s) {
    case S_TEND:
        if (E.mixed) ens_tendency_f(E.stream, E.dev_tab, E.maps[S_TEND], E.dev_state, 0);
        else ens_tendency(E.stream, E.dev_tab, E.maps[S_TEND], E.dev_state);
        break;
    case S_FRIC:
        if (E.mixed) ens_friction_ab_f(E.stream, E.dev_tab, E.maps[S_FRIC], E.dev_state, 0);
        else ens_friction_ab(E.stream, E.dev_tab, E.maps[S_FRIC], E.dev_state);
        break;
    default: break;
    }
}

// mixed form: the FP32 Poisson solve of the new state's divergence into the
// increment selected by which (see ens_c2r_f)
void enqueue_solve_f(ps3d_ensemble &E, int which)
{
    hipStream_t s = E.stream;
    ens_r2c_f(s, E.dev_tab, E.maps[S_R2C], E.dev_state, E.lds_x);
    ens_cols_y_f(s, E.dev_tab, E.maps[S_COLS], -1, E.ny_max, E.lds_y);
    ens_thomas_f(s, E.dev_tab, E.maps[S_THOMAS], E.nzl_max, E.thomas_reg, E.thomas_split);
    ens_zsolve_f(s, E.dev_tab, E.maps[S_ZSOLVE], E.nz_max, E.lds_z);
    ens_cols_y_f(s, E.dev_tab, E.maps[S_COLS], +1, E.ny_max, E.lds_y);
    ens_c2r_f(s, E.dev_tab, E.maps[S_R2C], E.dev_state, which, E.lds_x);
}

// mixed form, start of the pressure history of every member (the single
// model's cold start): a provisional Euler step into the second buffers, its
// FP32 pressure as the reference pressure pint, zero increments
void enqueue_cold_mixed(ps3d_ensemble &E)
{
    hipStream_t s = E.stream;
    ens_tendency_f(s, E.dev_tab, E.maps[S_TEND], E.dev_state, 1);
    // forced members: the forcing on the pre-friction state (a single
    // model's provisional step)
    ens_background_f(s, E.dev_tab, E.maps[S_BGAB], E.dev_state);
    ens_friction_ab_f(s, E.dev_tab, E.maps[S_FRIC], E.dev_state, 1);
    enqueue_solve_f(E, 0);
    ens_cold_pint(s, E.dev_tab, E.maps[S_PROJ]);
    ens_points_f(s, E.dev_tab, E.maps[S_PTS], E.dev_state, 0);
}

// one step of the mixed form for every member (integrate_mixed of a single
// model); refine: the correction solve of the first two steps
void enqueue_step_mixed(ps3d_ensemble &E, bool refine)
{
    hipStream_t s = E.stream;
    ens_tendency_f(s, E.dev_tab, E.maps[S_TEND], E.dev_state, 0);
    ens_friction_ab_f(s, E.dev_tab, E.maps[S_FRIC], E.dev_state, 0);   // writes the halo images too
    // forced members (their friction went to the third (hi, lo) buffers):
    // the pre-friction hi halos, then forcing + Adams-Bashforth
    ens_halo_copy(s, E.dev_tab, E.maps[S_HCOPY], E.dev_state, true);
    ens_background_ab_f(s, E.dev_tab, E.maps[S_BGAB], E.dev_state);
    enqueue_solve_f(E, 1);
    ens_project_f(s, E.dev_tab, E.maps[S_PROJ], E.dev_state, 1);
    if (refine) {
        enqueue_solve_f(E, 2);
        ens_project_f(s, E.dev_tab, E.maps[S_PROJ], E.dev_state, 2);
        ens_points_f(s, E.dev_tab, E.maps[S_PTS], E.dev_state, 1);
    }
    ens_advance(s, E.dev_state);
}

// integrate.f90:2-115 for every member: the single-model step, stage by stage
void enqueue_step(ps3d_ensemble &E)
{
    if (E.mixed) {
        enqueue_step_mixed(E, false);
        return;
    }
    hipStream_t s = E.stream;
    ens_tendency(s, E.dev_tab, E.maps[S_TEND], E.dev_state);
    ens_friction_ab(s, E.dev_tab, E.maps[S_FRIC], E.dev_state);   // writes the halo images too
    // forced members (their friction went to the third buffers): the
    // pre-friction halos, then set_forcing (integrate.f90:82) and AB
    ens_halo_copy(s, E.dev_tab, E.maps[S_HCOPY], E.dev_state, false);
    ens_background_ab(s, E.dev_tab, E.maps[S_BGAB], E.dev_state);
    ens_r2c(s, E.dev_tab, E.maps[S_R2C], E.dev_state, E.lds_x);
    ens_cols_y(s, E.dev_tab, E.maps[S_COLS], -1, E.ny_max, E.lds_y);
    ens_thomas(s, E.dev_tab, E.maps[S_THOMAS], E.nzl_max, E.thomas_reg, E.thomas_split);
    ens_zsolve(s, E.dev_tab, E.maps[S_ZSOLVE], E.nz_max, E.lds_z);
    ens_cols_y(s, E.dev_tab, E.maps[S_COLS], +1, E.ny_max, E.lds_y);
    ens_c2r(s, E.dev_tab, E.maps[S_R2C], E.lds_x);
    ens_project(s, E.dev_tab, E.maps[S_PROJ], E.dev_state);       // and of u, v, w, p
    ens_advance(s, E.dev_state);
}

void check_member(const ps3d_ensemble *e, int member)
{
    if (!e) throw Error(PS3D_ERR_ARG, "null ensemble");
    PS3D_REQUIRE(member >= 0 && member < (int)e->mem.size(), "member index out of range");
}

// a mixed member's state is in its second buffers when par is 1: exchange
// the buffer roles around a single-model call that reads or writes it
template <class F>
void with_parity(ps3d_model *m, int par, F &&fn)
{
    auto flip = [&] {
        for (int q = 0; q < 4; ++q) {
            std::swap(m->fh[q], m->ah[q]);
            std::swap(m->fl[q], m->al[q]);
        }
    };
    if (par) flip();
    try {
        fn();
    } catch (...) {
        if (par) flip();
        throw;
    }
    if (par) flip();
}

void destroy(ps3d_ensemble *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    for (ps3d_model *m : e->mem) ps3d_destroy(m);
    for (void *p : e->tmp_alloc) (void)hipFree(p);
    if (e->dev_tab) (void)hipFree(e->dev_tab);
    if (e->dev_state) (void)hipFree(e->dev_state);
    if (e->own_stream && e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

// a member's forcing as its model holds it (set by the single-model calls)
void refresh_forcing(ps3d_ensemble &E, int i)
{
    ps3d_model &m = *E.mem[i];
    EnsMember &e = E.tab[i];
    e.ub = m.bg_on ? m.bg_u : nullptr;
    e.bb = (m.bg_on && m.bg_b_on) ? m.bg_b : nullptr;
    bool any = e.ub != nullptr;
    for (int q = 0; q < 4; ++q) {
        e.bf[q] = m.bf_on[q] ? m.bf[q] : nullptr;
        any = any || e.bf[q];
    }
    if (any && !E.mixed && !e.tmp[0]) {
        for (int q = 0; q < 4; ++q) {
            void *p = nullptr;
            PS3D_HIP_CHECK(hipMalloc(&p, sizeof(double) * (size_t)m.g.H));
            E.tmp_alloc.push_back(p);
            e.tmp[q] = static_cast<double *>(p);
        }
    }
    if (any && E.mixed && !e.tmph[0]) {   // the (hi, lo) pair of the mixed form
        for (int q = 0; q < 8; ++q) {
            void *p = nullptr;
            PS3D_HIP_CHECK(hipMalloc(&p, sizeof(float) * (size_t)m.g.H));
            E.tmp_alloc.push_back(p);
            (q < 4 ? e.tmph[q] : e.tmpl[q - 4]) = static_cast<float *>(p);
        }
    }
    e.forced = any ? 1 : 0;
    build_maps(E);
    upload_table(E);
}

// Radix plans of a mixed ensemble's FP32 transforms, per axis and transform
// length: each candidate is set on every member with that length and the
// stages that run it are timed over the whole ensemble (x: the forward x
// pass; y: the forward x pass and both y passes; z: those and the periodic
// z solve).  They write only the spectral scratch FSf, which every solve
// starts by overwriting.
void tune_fft_plans(ps3d_ensemble &E, std::map<std::string, int> &cache, int reps, hipEvent_t t0, hipEvent_t t1)
{
    const char *axis_name[3] = {"ens_fft_plan_x", "ens_fft_plan_y", "ens_fft_plan_z"};
    auto len_of = [&](size_t i, int ax) {
        const Grid &g = E.mem[i]->g;
        return ax == 0 ? g.nx / 2 : (ax == 1 ? g.ny : g.nz);
    };
    auto uses_axis = [&](size_t i, int ax) { return ax < 2 || !E.tab[i].c.lid; };
    auto set_plan = [&](int ax, int len, int plan) {
        for (size_t i = 0; i < E.mem.size(); ++i)
            if (uses_axis(i, ax) && len_of(i, ax) == len) {
                E.mem[i]->fft_plan[ax] = plan;
                E.tab[i].fft_plan[ax] = plan;
            }
    };
    auto run = [&](int ax) {
        hipStream_t s = E.stream;
        ens_r2c_f(s, E.dev_tab, E.maps[S_R2C], E.dev_state, E.lds_x);
        if (ax >= 1) ens_cols_y_f(s, E.dev_tab, E.maps[S_COLS], -1, E.ny_max, E.lds_y);
        if (ax == 2) ens_zsolve_f(s, E.dev_tab, E.maps[S_ZSOLVE], E.nz_max, E.lds_z);
        if (ax >= 1) ens_cols_y_f(s, E.dev_tab, E.maps[S_COLS], +1, E.ny_max, E.lds_y);
    };
    for (int ax = 0; ax < 3; ++ax) {
        std::vector<int> lens;
        for (size_t i = 0; i < E.mem.size(); ++i) {
            const int L = len_of(i, ax);
            if (uses_axis(i, ax) && fft_plan_choice(L) && std::find(lens.begin(), lens.end(), L) == lens.end())
                lens.push_back(L);
        }
        for (int L : lens) {
            const std::string key = std::string(axis_name[ax]) + " " + std::to_string(L) + ":mixed";
            auto hit = cache.find(key);
            if (hit != cache.end()) {
                set_plan(ax, L, hit->second);
                E.tuned[key] = hit->second;
                continue;
            }
            constexpr int kRounds = 3;   // interleaved rounds, the median
            std::vector<float> tm[2];
            for (int round = 0; round < kRounds; ++round)
                for (int plan : {0, 1}) {
                    set_plan(ax, L, plan);
                    upload_table(E);
                    run(ax);   // warm-up
                    PS3D_HIP_CHECK(hipEventRecord(t0, E.stream));
                    for (int r = 0; r < reps; ++r) run(ax);
                    PS3D_HIP_CHECK(hipEventRecord(t1, E.stream));
                    PS3D_HIP_CHECK(hipEventSynchronize(t1));
                    float ms = 0.f;
                    PS3D_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
                    tm[plan].push_back(ms);
                }
            for (auto &v : tm) std::sort(v.begin(), v.end());
            const int best = tm[1][kRounds / 2] < tm[0][kRounds / 2] ? 1 : 0;
            set_plan(ax, L, best);
            upload_table(E);
            E.tuned[key] = best;
            cache[key] = best;
        }
    }
}


}  // namespace

extern "C" {

int ps3d_ensemble_set_background_flow(ps3d_ensemble *e, int member, const double *u_back, const double *b_back)
{
    return guard([&] {
        check_member(e, member);
        PS3D_HIP_CHECK(hipStreamSynchronize(e->stream));
        const int rc = ps3d_set_background_flow(e->mem[member], u_back, b_back);
        if (rc != PS3D_OK) throw Error(rc, error_slot());
        refresh_forcing(*e, member);
    });
}

int ps3d_ensemble_set_body_force(ps3d_ensemble *e, int member, const double *const f[4])
{
    return guard([&] {
        check_member(e, member);
        PS3D_HIP_CHECK(hipStreamSynchronize(e->stream));
        const int rc = ps3d_set_body_force(e->mem[member], f);
        if (rc != PS3D_OK) throw Error(rc, error_slot());
        refresh_forcing(*e, member);
    });
}

int ps3d_ensemble_set_fft_plan(ps3d_ensemble *e, int member, const int32_t plans[3])
{
    return guard([&] {
        check_member(e, member);
        PS3D_HIP_CHECK(hipStreamSynchronize(e->stream));
        const int rc = ps3d_set_fft_plan(e->mem[member], plans);
        if (rc != PS3D_OK) throw Error(rc, error_slot());
        for (int q = 0; q < 3; ++q) e->tab[member].fft_plan[q] = e->mem[member]->fft_plan[q];
        upload_table(*e);
    });
}

int ps3d_ensemble_get_fft_plan(ps3d_ensemble *e, int member, int32_t plans[3])
{
    return guard([&] {
        check_member(e, member);
        PS3D_REQUIRE(plans, "null plans");
        for (int q = 0; q < 3; ++q) plans[q] = e->tab[member].fft_plan[q];
    });
}

int ps3d_ensemble_create(int n, const ps3d_params *prms, void *stream, ps3d_ensemble **out)
{
    return guard([&] {
        PS3D_REQUIRE(prms && out, "null argument");
        *out = nullptr;
        PS3D_REQUIRE(n >= 1 && n <= ENS_MAX, "ensemble size must be 1.." + std::to_string(ENS_MAX));
        for (int i = 0; i < n; ++i) {
            const ps3d_params &p = prms[i];
            const std::string who = "member " + std::to_string(i) + ": ";
            PS3D_REQUIRE(p.precision == PS3D_FP64 || p.precision == PS3D_MIXED, who + "unknown precision form");
            PS3D_REQUIRE(p.precision == prms[0].precision, who + "the members of an ensemble share one precision form");
            PS3D_REQUIRE(p.precision != PS3D_MIXED || !p.enable_vertical_boundaries ||
                             (sizeof(double2) + sizeof(double)) * (size_t)ens_thomas_cols(false, false) * (size_t)p.nz <=
                                 160 * 1024,
                         who + "mixed ensembles: rigid-lid members of at most 213 levels");
            PS3D_REQUIRE(!(p.Avbi > 0.) && !(p.Kvbi > 0.),
                         who + "vertical biharmonic friction (Avbi, Kvbi > 0) is not in the batched step");
            PS3D_REQUIRE(p.nx >= 2 && p.ny >= 2 && ens_fft_size_ok(p.nx / 2, false) && ens_fft_size_ok(p.ny, true) &&
                             (p.enable_vertical_boundaries || ens_fft_size_ok(p.nz, true)),
                         who + "nx/2, ny (and nz without a lid) must be compiled-in FFT lengths (" + fft_sizes_text() +
                             "), ny and nz at most 512");
            // the batched friction and projection write the halo images themselves
            PS3D_REQUIRE(halo_images_ok(p.nx, p.ny, p.nz), who + "nx, ny and nz must be at least 4");
        }
        auto *E = new ps3d_ensemble();
        try {
            if (stream) {
                E->stream = static_cast<hipStream_t>(stream);
            } else {
                PS3D_HIP_CHECK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
                E->own_stream = true;
            }
            E->mixed = prms[0].precision == PS3D_MIXED;
            E->mem.assign(n, nullptr);
            E->tab.resize(n);
            for (int i = 0; i < n; ++i) {
                ps3d_model *m = nullptr;
                const int rc = ps3d_create(&prms[i], nullptr, E->stream, &m);
                if (rc != PS3D_OK) throw Error(rc, "member " + std::to_string(i) + ": " + error_slot());
                E->mem[i] = m;
                fill_member(E->tab[i], *m);
                set_chunks(m->g.nz, default_tend_kc(m->g), E->tab[i].tend_kc, E->tab[i].tend_nchunk);
                set_chunks(m->g.nz, default_fric_kc(m->g), E->tab[i].fric_kc, E->tab[i].fric_nchunk);
            }
            bool reg = true, split = true;
            for (ps3d_model *m : E->mem)
                if (m->prm.enable_vertical_boundaries) {
                    const uint64_t fsb = (uint64_t)m->g.S * (E->mixed ? sizeof(float2) : sizeof(double2));
                    reg = reg && m->thomas_reg && thomas_reg_ok(m->g.nz, fsb);
                    split = split && m->thomas_reg && thomas_split_ok(m->g.nz, fsb);
                }
            E->thomas_reg = reg;
            E->thomas_split = !reg && split;
            build_maps(*E);
            PS3D_HIP_CHECK(hipMalloc(&E->dev_tab, sizeof(EnsMember) * n));
            PS3D_HIP_CHECK(hipMalloc(&E->dev_state, sizeof(EnsState)));
            PS3D_HIP_CHECK(hipMemcpyAsync(E->dev_state, &E->state, sizeof(EnsState), hipMemcpyHostToDevice,
                                          E->stream));
            upload_table(*E);
        } catch (...) {
            destroy(E);
            throw;
        }
        *out = E;
    });
}

int ps3d_ensemble_destroy(ps3d_ensemble *e)
{
    return guard([&] { destroy(e); });
}

int ps3d_ensemble_set_field(ps3d_ensemble *e, int member, int field, const double *host)
{
    return guard([&] {
        check_member(e, member);
        PS3D_REQUIRE(host, "null host pointer");
        ps3d_model *m = e->mem[member];
        if (e->mixed) {
            PS3D_HIP_CHECK(hipStreamSynchronize(e->stream));
            with_parity(m, e->state.par, [&] {
                const int rc = ps3d_set_field(m, field, host);
                if (rc != PS3D_OK) throw Error(rc, error_slot());
            });
            return;
        }
        if (field >= PS3D_U && field <= PS3D_B) {
            PS3D_HIP_CHECK(hipStreamSynchronize(e->stream));
            PS3D_HIP_CHECK(hipMemcpyAsync(e->tab[member].f[field][e->state.par], host, sizeof(double) * m->g.H,
                                          hipMemcpyHostToDevice, e->stream));
            PS3D_HIP_CHECK(hipStreamSynchronize(e->stream));
            return;
        }
        const int rc = ps3d_set_field(m, field, host);
        if (rc != PS3D_OK) throw Error(rc, error_slot());
    });
}

int ps3d_ensemble_get_field(ps3d_ensemble *e, int member, int field, double *host)
{
    return guard([&] {
        check_member(e, member);
        PS3D_REQUIRE(host, "null host pointer");
        ps3d_model *m = e->mem[member];
        if (e->mixed) {
            with_parity(m, e->state.par, [&] {
                const int rc = ps3d_get_field(m, field, host);
                if (rc != PS3D_OK) throw Error(rc, error_slot());
            });
            return;
        }
        if (field >= PS3D_U && field <= PS3D_B) {
            PS3D_HIP_CHECK(hipMemcpyAsync(host, e->tab[member].f[field][e->state.par], sizeof(double) * m->g.H,
                                          hipMemcpyDeviceToHost, e->stream));
            PS3D_HIP_CHECK(hipStreamSynchronize(e->stream));
            return;
        }
        const int rc = ps3d_get_field(m, field, host);
        if (rc != PS3D_OK) throw Error(rc, error_slot());
    });
}

int ps3d_ensemble_run(ps3d_ensemble *e, int nsteps, int32_t *st, int use_graph)
{
    return guard([&] {
        PS3D_REQUIRE(e && st && nsteps >= 0, "bad arguments");
        auto ok = [](int t) { return t >= 1 && t <= 3; };
        PS3D_REQUIRE(ok(st[1]) && ok(st[2]) && ok(st[3]) && st[1] != st[2] && st[1] != st[3] && st[2] != st[3],
                     "tau slots must be 1..3 and differ");
        PS3D_REQUIRE(st[0] >= 0, "itt must be >= 0");
        if (e->state.itt != st[0] || e->state.tau != st[1] || e->state.taum1 != st[2] || e->state.taum2 != st[3]) {
            e->state.itt = st[0];
            e->state.tau = st[1];
            e->state.taum1 = st[2];
            e->state.taum2 = st[3];
            PS3D_HIP_CHECK(hipMemcpyAsync(e->dev_state, &e->state, sizeof(EnsState), hipMemcpyHostToDevice,
                                          e->stream));
            PS3D_HIP_CHECK(hipStreamSynchronize(e->stream));
        }
        if (use_graph && nsteps > 0 && !e->exec) {
            PS3D_HIP_CHECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_step(*e);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(e->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            PS3D_HIP_CHECK(hipStreamEndCapture(e->stream, &e->graph));
            PS3D_HIP_CHECK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        }
        if (e->mixed && e->cold && nsteps > 0) {
            enqueue_cold_mixed(*e);
            e->cold = false;
            e->refine = 2;
        }
        for (int s = 0; s < nsteps; ++s) {
            if (e->mixed && e->refine > 0) {   // not in the graph: the steady step is
                enqueue_step_mixed(*e, true);
                e->refine -= 1;
            } else if (use_graph) {
                PS3D_HIP_CHECK(hipGraphLaunch(e->exec, e->stream));
            } else {
                enqueue_step(*e);
            }
            e->state.tau = e->state.tau % 3 + 1;
            e->state.taum1 = e->state.taum1 % 3 + 1;
            e->state.taum2 = e->state.taum2 % 3 + 1;
            e->state.itt += 1;
            e->state.par ^= 1;
            e->state.dp_cur = (e->state.dp_cur + 1) % 3;
        }
        if (e->mixed && nsteps > 0) {
            // as after a single model's step: the joined fields' halos are
            // rebuilt on the way out and a pressure set by the host is
            // superseded
            for (ps3d_model *m : e->mem) {
                for (int f = 0; f < 4; ++f) m->lo_stale[f] = true;
                m->p_host.clear();
            }
        }
        st[0] = e->state.itt;
        st[1] = e->state.tau;
        st[2] = e->state.taum1;
        st[3] = e->state.taum2;
    });
}

int ps3d_ensemble_synchronize(ps3d_ensemble *e)
{
    return guard([&] {
        PS3D_REQUIRE(e, "null ensemble");
        PS3D_HIP_CHECK(hipStreamSynchronize(e->stream));
    });
}

int ps3d_ensemble_tune(ps3d_ensemble *e, const char *cache_path, int reps)
{
    return guard([&] {
        PS3D_REQUIRE(e, "null ensemble");
        reps = std::max(1, reps);
        // the cache file: one "stage shape levels" line per entry; the
        // stages are named ens_* (a batched launch, timed over all members,
        // is a different measurement from a single model's)
        std::map<std::string, int> cache = tune_cache_read(cache_path);
        std::vector<std::string> shapes;
        for (ps3d_model *m : e->mem) {
            const std::string k = shape_key(m->prm);
            if (std::find(shapes.begin(), shapes.end(), k) == shapes.end()) shapes.push_back(k);
        }
        hipEvent_t t0, t1;
        PS3D_HIP_CHECK(hipEventCreate(&t0));
        PS3D_HIP_CHECK(hipEventCreate(&t1));
        const char *names[2] = {"ens_tendency", "ens_friction_ab"};
        const int stages[2] = {S_TEND, S_FRIC};
        try {
            for (int si = 0; si < 2; ++si) {
                const int stage = stages[si];
                for (const std::string &sh : shapes) {
                    const std::string key = std::string(names[si]) + " " + sh;
                    auto hit = cache.find(key);
                    if (hit != cache.end()) {
                        apply_tiles(*e, stage, sh, hit->second);
                        e->tuned[key] = hit->second;
                        continue;
                    }
                    int nz = 0;
                    for (ps3d_model *m : e->mem)
                        if (shape_key(m->prm) == sh) nz = m->g.nz;
                    // levels per work-group: powers of two up to nz, and nz
                    std::vector<int> cand;
                    for (int kc = 1; kc < nz; kc *= 2) cand.push_back(kc);
                    cand.push_back(nz);
                    // three interleaved rounds, the median of each
                    // candidate's (as the single-model tuner)
                    constexpr int kRounds = 3;
                    std::vector<std::vector<float>> tm(cand.size());
                    for (int round = 0; round < kRounds; ++round)
                        for (size_t q = 0; q < cand.size(); ++q) {
                            apply_tiles(*e, stage, sh, cand[q]);
                            build_maps(*e);
                            upload_table(*e);
                            launch_stage(*e, stage);   // warm-up
                            PS3D_HIP_CHECK(hipEventRecord(t0, e->stream));
                            for (int r = 0; r < reps; ++r) launch_stage(*e, stage);
                            PS3D_HIP_CHECK(hipEventRecord(t1, e->stream));
                            PS3D_HIP_CHECK(hipEventSynchronize(t1));
                            float ms = 0.f;
                            PS3D_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
                            tm[q].push_back(ms);
                        }
                    int best = cand[0];
                    float best_ms = 1e30f;
                    for (size_t q = 0; q < cand.size(); ++q) {
                        std::sort(tm[q].begin(), tm[q].end());
                        if (tm[q][kRounds / 2] < best_ms) {
                            best_ms = tm[q][kRounds / 2];
                            best = cand[q];
                        }
                    }
                    apply_tiles(*e, stage, sh, best);
                    e->tuned[key] = best;
                    cache[key] = best;
                }
            }
            build_maps(*e);
            upload_table(*e);
            if (e->mixed) tune_fft_plans(*e, cache, reps, t0, t1);
        } catch (...) {
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            throw;
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        build_maps(*e);
        upload_table(*e);
        tune_cache_write(cache_path, cache);
    });
}

int ps3d_ensemble_describe(ps3d_ensemble *e, char *buf, int len)
{
    int n = 0;
    const int rc = guard([&] {
        PS3D_REQUIRE(e, "null ensemble");
        std::map<std::string, std::pair<int, int>> seen;
        for (size_t i = 0; i < e->mem.size(); ++i)
            seen[shape_key(e->mem[i]->prm)] = {e->tab[i].tend_kc, e->tab[i].fric_kc};
        std::ostringstream o;
        for (auto &kv : seen) {
            o << "ens_tendency " << kv.first << " " << kv.second.first << "\n";
            o << "ens_friction_ab " << kv.first << " " << kv.second.second << "\n";
        }
        if (e->mixed) {
            // radix plans by axis and transform length
            std::map<std::string, int> plans;
            const char *names[3] = {"ens_fft_plan_x", "ens_fft_plan_y", "ens_fft_plan_z"};
            for (size_t i = 0; i < e->mem.size(); ++i) {
                const Grid &g = e->mem[i]->g;
                const int lens[3] = {g.nx / 2, g.ny, g.nz};
                for (int ax = 0; ax < 3; ++ax)
                    if (ax < 2 || !e->tab[i].c.lid)
                        plans[std::string(names[ax]) + " " + std::to_string(lens[ax]) + ":mixed"] =
                            e->tab[i].fft_plan[ax];
            }
            for (auto &kv : plans) o << kv.first << " " << kv.second << "\n";
        }
        const std::string s = o.str();
        n = (int)s.size();
        if (buf && len > 0) {
            const int c = std::min(n, len - 1);
            std::memcpy(buf, s.data(), (size_t)c);
            buf[c] = '\0';
        }
    });
    return rc == PS3D_OK ? n : -rc;
}

}  // extern "C"
