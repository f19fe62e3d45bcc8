// monitor.hip -- device-side monitors of diagnose.F90:2-61 (kinetic and
// potential energy, max |u_h|, max |w|, CFL and Peclet numbers, NaN check)
// and the restart records of restart.f90:2-82.
//
// The reductions run on the device and return a handful of numbers, so a
// driver can monitor every step without copying fields to the host.  Sums
// are formed per work-group in a fixed tree order and finished on the host;
// they differ from the reference's sequential loop only by the summation
// order (relative 1e-15 level).
#include "ps3d_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#pragma clang fp contract(off)

namespace ps3d {

namespace {

constexpr int MT = 256;

struct MonOut {
    double ke, pe, umax, wmax;
    int nan;
};

// one value of field f at interior point o: FP64, or (hi + lo) of the mixed form
struct FieldRef {
    const double *d;
    const float *hi, *lo;
    __device__ __forceinline__ double at(int64_t o) const
    {
        return d ? d[o] : (double)hi[o] + (double)lo[o];
    }
};

__global__ __launch_bounds__(MT) void k_monitor(StepConsts c, FieldRef u, FieldRef v, FieldRef w, FieldRef b,
                                                 double N0, MonOut *__restrict__ part)
{
    __shared__ double ske[MT], spe[MT], sum[MT], swm[MT];
    __shared__ int snan[MT];
    const int64_t N = (int64_t)c.nx * c.ny * c.nz;
    double ke = 0.0, pe = 0.0, um = 0.0, wm = 0.0;
    int nan = 0;
    for (int64_t t = (int64_t)blockIdx.x * MT + threadIdx.x; t < N; t += (int64_t)gridDim.x * MT) {
        const int i = (int)(t % c.nx) + 1;
        const int j = (int)((t / c.nx) % c.ny) + 1;
        const int k = (int)(t / ((int64_t)c.nx * c.ny)) + 1;
        const int64_t o = (int64_t)(i + 1) + c.X * (int64_t)(j + 1) + c.XY * (int64_t)(k + 1);
        const double uu = u.at(o), vv = v.at(o), ww = w.at(o), bb = b.at(o);
        const double sp = sqrt(uu * uu + vv * vv);
        if (sp != sp) nan = 1;
        um = fmax(um, sp);
        wm = fmax(wm, fabs(ww));
        ke = ke + 0.5 * (uu * uu + vv * vv + c.dsqr * (ww * ww));
        if (N0 > 0.0) pe = pe + 0.5 * (bb * bb) / (N0 * N0);
        else pe = pe + bb * ((double)k - 0.5) * c.dz;
    }
    const int t = threadIdx.x;
    ske[t] = ke;
    spe[t] = pe;
    sum[t] = um;
    swm[t] = wm;
    snan[t] = nan;
    __syncthreads();
    for (int s = MT / 2; s > 0; s >>= 1) {
        if (t < s) {
            ske[t] = ske[t] + ske[t + s];
            spe[t] = spe[t] + spe[t + s];
            sum[t] = fmax(sum[t], sum[t + s]);
            swm[t] = fmax(swm[t], swm[t + s]);
            snan[t] = snan[t] | snan[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        MonOut r;
        r.ke = ske[0];
        r.pe = spe[0];
        r.umax = sum[0];
        r.wmax = swm[0];
        r.nan = snan[0];
        part[blockIdx.x] = r;
    }
}

FieldRef field_ref(Model &m, int f)
{
    FieldRef r{};
    if (m.mixed) {
        r.hi = m.fh[f];
        r.lo = m.fl[f];
    } else {
        r.d = m.field_buf(f);
    }
    return r;
}

// Fortran unformatted sequential records: 4-byte length markers before and
// after each record; records longer than 2147483639 bytes are written as
// subrecords whose leading marker is negated when more follow and whose
// trailing marker is negated when one precedes (the gfortran convention)
constexpr int64_t kSubMax = 2147483639;

void put_record(FILE *f, const std::vector<std::pair<const void *, size_t>> &parts)
{
    size_t total = 0;
    for (auto &p : parts) total += p.second;
    size_t done = 0, pi = 0, off = 0;
    bool first = true;
    do {
        const size_t len = std::min<size_t>(total - done, (size_t)kSubMax);
        const bool more = done + len < total;
        int32_t lead = (int32_t)len, trail = (int32_t)len;
        if (more) lead = -lead;
        if (!first) trail = -trail;
        PS3D_REQUIRE(fwrite(&lead, 4, 1, f) == 1, "restart: write failed");
        size_t left = len;
        while (left > 0) {
            const size_t n = std::min(left, parts[pi].second - off);
            PS3D_REQUIRE(fwrite((const char *)parts[pi].first + off, 1, n, f) == n, "restart: write failed");
            left -= n;
            off += n;
            if (off == parts[pi].second) { ++pi; off = 0; }
        }
        PS3D_REQUIRE(fwrite(&trail, 4, 1, f) == 1, "restart: write failed");
        done += len;
        first = false;
    } while (done < total);
}

void get_record(FILE *f, const std::vector<std::pair<void *, size_t>> &parts)
{
    size_t total = 0;
    for (auto &p : parts) total += p.second;
    size_t done = 0, pi = 0, off = 0;
    bool more = true;
    while (more) {
        int32_t lead = 0, trail = 0;
        PS3D_REQUIRE(fread(&lead, 4, 1, f) == 1, "restart: truncated file");
        more = lead < 0;
        const size_t len = (size_t)(lead < 0 ? -(int64_t)lead : lead);
        PS3D_REQUIRE(done + len <= total, "restart: record longer than expected (grid mismatch?)");
        size_t left = len;
        while (left > 0) {
            const size_t n = std::min(left, parts[pi].second - off);
            PS3D_REQUIRE(fread((char *)parts[pi].first + off, 1, n, f) == n, "restart: truncated file");
            left -= n;
            off += n;
            if (off == parts[pi].second) { ++pi; off = 0; }
        }
        PS3D_REQUIRE(fread(&trail, 4, 1, f) == 1, "restart: truncated file");
        PS3D_REQUIRE((size_t)(trail < 0 ? -(int64_t)trail : trail) == len, "restart: inconsistent record markers");
        done += len;
    }
    PS3D_REQUIRE(done == total, "restart: record shorter than expected (grid mismatch?)");
}

// the monitors from the sums and maxima: diagnose.F90:14-41
void monitor_finish(const MonOut &tot, const ps3d_params &p, double *out)
{
    const double n = (double)p.nx * p.ny * p.nz;
    const double ke = tot.ke / n;
    const double pe = tot.pe / n;
    const double umax = tot.umax, wmax = tot.wmax;
    const double cfl_h = umax * p.dt / p.dx;
    const double cfl_v = wmax * p.dt / p.dz;
    const double uw = std::fmax(umax, wmax);
    const double pecl_h = std::fmin(uw * p.dx / (1e-32 + p.Ah), uw * p.dx * p.dx * p.dx / (1e-32 + p.Ahbi));
    const double pecl_v = std::fmin(uw * p.dz / (1e-32 + p.Av), uw * p.dz * p.dz * p.dz / (1e-32 + p.Avbi));
    const double v[8] = {ke, pe, umax, wmax, cfl_h, cfl_v, pecl_h, pecl_v};
    std::memcpy(out, v, sizeof(v));
}

constexpr int kMaxBlocks = 2048;   // work-groups of k_monitor (grid-stride beyond)

int monitor_blocks(int64_t n) { return (int)std::min<int64_t>((n + MT - 1) / MT, kMaxBlocks); }

// The partial records of one k_monitor launch summed in block order into
// part[nb]: the same sequence of additions and maxima as the host loop of
// monitor(), so both give the same bits.
__global__ void k_monitor_total(MonOut *part, int nb)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    MonOut t{0.0, 0.0, 0.0, 0.0, 0};
    for (int q = 0; q < nb; ++q) {
        const MonOut r = part[q];
        t.ke = t.ke + r.ke;
        t.pe = t.pe + r.pe;
        t.umax = fmax(t.umax, r.umax);
        t.wmax = fmax(t.wmax, r.wmax);
        t.nan = t.nan | r.nan;
    }
    part[nb] = t;
}

// Scratch of ps3d_diagnose_noalloc: the partial record of each work-group of
// k_monitor and, after them, the total.  One copy per GPU, held until the
// process ends; its device address is looked up once per device (the name is
// only a host-side stand-in for it).  Calls that overlap on one GPU race on it.
constexpr int kScratchRecords = kMaxBlocks + 1;
__device__ __align__(256) MonOut scratch[kScratchRecords];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
MonOut *scratch_ptr[kMaxDevices] = {};

MonOut *scratch_on_current_device()
{
    int dev = -1;
    PS3D_HIP_CHECK(hipGetDevice(&dev));
    PS3D_REQUIRE(dev >= 0 && dev < kMaxDevices, "diagnose_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *ptr = nullptr;
        PS3D_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        PS3D_REQUIRE(ptr, "diagnose_noalloc: no device address for the scratch records");
        scratch_ptr[dev] = static_cast<MonOut *>(ptr);
    }
    return scratch_ptr[dev];
}

}  // namespace

// diagnose.F90:11-53 on the device; out = ke, pe, max|u_h|, max|w|, cfl_h,
// cfl_v, pecl_h, pecl_v; returns whether a NaN was seen (diagnose.F90:42-45)
bool monitor(Model &m, double *out)
{
    const Grid &g = m.g;
    const int nb = monitor_blocks(g.N);
    if (!m.mon) {   // partial results, kept for the model's lifetime
        PS3D_HIP_CHECK(hipMalloc(&m.mon, sizeof(MonOut) * kMaxBlocks));
        m.allocations.push_back(m.mon);
    }
    MonOut *part = static_cast<MonOut *>(m.mon);
    std::vector<MonOut> h(nb);
    hipLaunchKernelGGL(k_monitor, dim3(nb), dim3(MT), 0, m.stream, m.c, field_ref(m, 0), field_ref(m, 1),
                       field_ref(m, 2), field_ref(m, 3), m.prm.N0, part);
    PS3D_HIP_CHECK(hipGetLastError());
    PS3D_HIP_CHECK(hipMemcpyAsync(h.data(), part, sizeof(MonOut) * nb, hipMemcpyDeviceToHost, m.stream));
    PS3D_HIP_CHECK(hipStreamSynchronize(m.stream));
    MonOut tot{0.0, 0.0, 0.0, 0.0, 0};
    for (const MonOut &r : h) {
        tot.ke = tot.ke + r.ke;
        tot.pe = tot.pe + r.pe;
        tot.umax = std::fmax(tot.umax, r.umax);
        tot.wmax = std::fmax(tot.wmax, r.wmax);
        tot.nan = tot.nan | r.nan;
    }
    monitor_finish(tot, m.prm, out);
    return tot.nan != 0;
}

// the same monitors of caller-owned FP64 device fields, with the partial
// records in the static scratch instead of model memory: no allocation
bool monitor_noalloc(const ps3d_params &p, const double *u, const double *v, const double *w, const double *b,
                     hipStream_t stream, double *out)
{
    PS3D_REQUIRE(p.nx >= 1 && p.ny >= 1 && p.nz >= 1, "diagnose_noalloc: nx, ny, nz must be >= 1");
    PS3D_REQUIRE(p.nx <= (1 << 20) && p.ny <= (1 << 20) && p.nz <= (1 << 20),
                 "diagnose_noalloc: nx, ny, nz must be <= 2^20");
    StepConsts c{};
    c.nx = p.nx;
    c.ny = p.ny;
    c.nz = p.nz;
    c.X = p.nx + 2 * ONX;
    c.XY = c.X * (p.ny + 2 * ONX);
    c.dsqr = p.dsqr;
    c.dz = p.dz;
    const int64_t n = (int64_t)p.nx * p.ny * p.nz;
    const int nb = monitor_blocks(n);
    // the scratch holds one record per work-group plus the total; k_monitor
    // strides over any grid with at most kMaxBlocks work-groups
    PS3D_REQUIRE(nb + 1 <= kScratchRecords, "diagnose_noalloc: more work-groups than scratch records");
    MonOut *part = scratch_on_current_device();
    FieldRef fu{u, nullptr, nullptr}, fv{v, nullptr, nullptr}, fw{w, nullptr, nullptr}, fb{b, nullptr, nullptr};
    hipLaunchKernelGGL(k_monitor, dim3(nb), dim3(MT), 0, stream, c, fu, fv, fw, fb, p.N0, part);
    PS3D_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_monitor_total, dim3(1), dim3(64), 0, stream, part, nb);
    PS3D_HIP_CHECK(hipGetLastError());
    MonOut tot;
    PS3D_HIP_CHECK(hipMemcpyAsync(&tot, part + nb, sizeof(MonOut), hipMemcpyDeviceToHost, stream));
    PS3D_HIP_CHECK(hipStreamSynchronize(stream));
    monitor_finish(tot, p, out);
    return tot.nan != 0;
}

// restart.f90:2-22 / :27-82 on host arrays: state = u, v, w, b, p, du, dv,
// dw, db (full arrays with halo, 17 H doubles) in the reference's record
// layout, one rank (is_pe = 1 .. ke_pe = nz)
void restart_write_host(const char *path, int nx, int ny, int nz, int32_t itt, double time, int32_t tau,
                        int32_t taum1, int32_t taum2, const double *state)
{
    const int64_t H = (int64_t)(nx + 2 * ONX) * (ny + 2 * ONX) * (nz + 2 * ONX);
    const int32_t hdr_i[4] = {nx, ny, nz, itt};
    const int32_t hdr_t[3] = {taum2, taum1, tau};
    const int32_t rng[6] = {1, nx, 1, ny, 1, nz};
    FILE *f = fopen(path, "wb");
    PS3D_REQUIRE(f, std::string("restart: cannot open ") + path + " for writing");
    try {
        put_record(f, {{hdr_i, sizeof(hdr_i)}, {&time, sizeof(double)}, {hdr_t, sizeof(hdr_t)}});
        put_record(f, {{rng, sizeof(rng)}});
        put_record(f, {{state, sizeof(double) * (size_t)(17 * H)}});
    } catch (...) {
        fclose(f);
        throw;
    }
    PS3D_REQUIRE(fclose(f) == 0, "restart: write failed");
}

void restart_read_host(const char *path, int nx, int ny, int nz, int32_t *itt, double *time, int32_t *tau,
                       int32_t *taum1, int32_t *taum2, double *state)
{
    const int64_t H = (int64_t)(nx + 2 * ONX) * (ny + 2 * ONX) * (nz + 2 * ONX);
    const int32_t rng[6] = {1, nx, 1, ny, 1, nz};
    FILE *f = fopen(path, "rb");
    PS3D_REQUIRE(f, std::string("restart: cannot open ") + path);
    int32_t rd_i[4], rd_t[3], rd_r[6];
    double t = 0.0;
    try {
        get_record(f, {{rd_i, sizeof(rd_i)}, {&t, sizeof(double)}, {rd_t, sizeof(rd_t)}});
        PS3D_REQUIRE(rd_i[0] == nx && rd_i[1] == ny && rd_i[2] == nz,
                     "restart: grid size in the file differs from the model's (restart.f90:47-52)");
        get_record(f, {{rd_r, sizeof(rd_r)}});
        PS3D_REQUIRE(std::memcmp(rd_r, rng, sizeof(rng)) == 0,
                     "restart: the file was written by a decomposed run (index ranges differ; restart.f90:54-60)");
        get_record(f, {{state, sizeof(double) * (size_t)(17 * H)}});
    } catch (...) {
        fclose(f);
        throw;
    }
    fclose(f);
    *itt = rd_i[3];
    *time = t;
    *taum2 = rd_t[0];
    *taum1 = rd_t[1];
    *tau = rd_t[2];
}

// the model's state through the host buffer
void restart_io(Model &m, const char *path, bool write, int32_t *itt, double *time, int32_t *tau, int32_t *taum1,
                int32_t *taum2)
{
    PS3D_REQUIRE(!m.mixed, "restart: the record layout holds FP64 tendencies; PS3D_FP64 models only");
    for (int f = 0; f < 4; ++f)
        PS3D_REQUIRE(m.cur[f] == 0, "restart: friction results not yet folded in by ab_update");
    const Grid &g = m.g;
    const int64_t H = g.H;
    std::vector<double> buf((size_t)(17 * H));
    double *dev[9] = {m.u, m.v, m.w, m.b, m.p, m.du, m.dv, m.dw, m.db};
    const int64_t len[9] = {H, H, H, H, H, 3 * H, 3 * H, 3 * H, 3 * H};
    if (write) {
        int64_t off = 0;
        for (int q = 0; q < 9; ++q) {
            PS3D_HIP_CHECK(hipMemcpyAsync(buf.data() + off, dev[q], sizeof(double) * len[q], hipMemcpyDeviceToHost,
                                          m.stream));
            off += len[q];
        }
        PS3D_HIP_CHECK(hipStreamSynchronize(m.stream));
        restart_write_host(path, g.nx, g.ny, g.nz, *itt, *time, *tau, *taum1, *taum2, buf.data());
        return;
    }
    restart_read_host(path, g.nx, g.ny, g.nz, itt, time, tau, taum1, taum2, buf.data());
    int64_t off = 0;
    for (int q = 0; q < 9; ++q) {
        PS3D_HIP_CHECK(hipMemcpyAsync(dev[q], buf.data() + off, sizeof(double) * len[q], hipMemcpyHostToDevice,
                                      m.stream));
        off += len[q];
    }
    PS3D_HIP_CHECK(hipStreamSynchronize(m.stream));
}

}  // namespace ps3d
