// transforms.hip -- the distributed FFT library of the reference (P3DFFT:
// p3dfft_ftran_r2c / p3dfft_btran_c2r, called by my_fft.f90:159-252) on one
// GPU.  The whole domain sits on the device, so P3DFFT's MPI pencil
// transposes become the LDS transposes inside the FFT kernels of fft.hip:
// the x pass reads rows, the y and z passes read column segments, all
// coalesced.  Unnormalised, FFTW's sign convention (e^{-ikx} forward), the
// same factorisation as the model's own transforms, so a transform here and
// inside the step agree bit for bit.
#include "ps3d_internal.h"

#include <cstring>

using namespace ps3d;

struct ps3d_fft {
    Model m;
    bool has_z = false;
};

namespace {

template <class T>
T *falloc(Model &m, int64_t n)
{
    void *p = nullptr;
    PS3D_HIP_CHECK(hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    m.allocations.push_back(p);
    return static_cast<T *>(p);
}

bool sizes_235(int n)
{
    if (n < 1) return false;
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

void release(ps3d_fft *f)
{
    if (!f) return;
    Model &m = f->m;
    if (m.stream) (void)hipStreamSynchronize(m.stream);
    for (void *p : m.allocations) (void)hipFree(p);
    if (m.own_stream && m.stream) (void)hipStreamDestroy(m.stream);
    delete f;
}

// P3DFFT's op string: three characters, one per transformed dimension in
// execution order (forward x, y, z; backward z, y, x), 'f' = FFT, 'n' = none.
// The forms my_fft.f90 uses are accepted: forward "fft" and "ffn", backward
// "fft" and "nff".
int parse_op(const char *op, bool forward)
{
    PS3D_REQUIRE(op, "null op string");
    char s[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        PS3D_REQUIRE(op[i] != '\0', "op string must have three characters");
        s[i] = (char)(op[i] | 0x20);   // P3DFFT accepts either case
    }
    if (std::strcmp(s, "fft") == 0) return 3;
    if (forward && std::strcmp(s, "ffn") == 0) return 2;
    if (!forward && std::strcmp(s, "nff") == 0) return 2;
    throw Error(PS3D_ERR_ARG, std::string("unsupported op string '") + s + "' (forward: fft, ffn; backward: fft, nff)");
}

void check(const ps3d_fft *f)
{
    PS3D_REQUIRE(f, "null transform plan");
}

}  // namespace

extern "C" {

int ps3d_fft_create(int nx, int ny, int nz, void *stream, ps3d_fft **out)
{
    return guard([&] {
        PS3D_REQUIRE(out, "null argument");
        *out = nullptr;
        PS3D_REQUIRE(nx >= 2 && nx % 2 == 0 && ny >= 2 && nz >= 1, "nx even >= 2, ny >= 2, nz >= 1");
        PS3D_REQUIRE(sizes_235(nx / 2) && sizes_235(ny), "nx/2 and ny must be products of 2, 3 and 5");
        PS3D_REQUIRE(ny <= 1024 && nz <= 1024 && nx <= 8192, "grid too large for the FFT kernels");
        auto *f = new ps3d_fft();
        try {
            Model &m = f->m;
            ps3d_params &p = m.prm;
            p.nx = nx; p.ny = ny; p.nz = nz;
            p.dx = p.dy = p.dz = p.dt = 1.0;
            p.Lx = nx; p.Ly = ny; p.Lz = nz;
            p.dsqr = 1.0;
            Grid &g = m.g;
            g.nx = nx; g.ny = ny; g.nz = nz;
            g.X = nx + 2 * ONX; g.Y = ny + 2 * ONX; g.Z = nz + 2 * ONX;
            g.XY = g.X * g.Y;
            g.H = g.XY * g.Z;
            g.N = (int64_t)nx * ny * nz;
            g.nkx = nx / 2 + 1;
            g.S = (int64_t)g.nkx * ny * nz;
            StepConsts &c = m.c;
            c.nx = nx; c.ny = ny; c.nz = nz;
            c.X = g.X; c.XY = g.XY;
            c.dx = c.dy = c.dz = c.dt = c.dsqr = c.Ro = 1.0;
            c.rdx = c.rdy = c.rdz = c.rdt = 1.0;
            if (stream) {
                m.stream = static_cast<hipStream_t>(stream);
            } else {
                PS3D_HIP_CHECK(hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking));
                m.own_stream = true;
            }
            m.FS = falloc<double2>(m, g.S);
            m.aloc = falloc<double>(m, g.N);
            fft_plan_init(m, m.px, nx / 2);
            fft_plan_init(m, m.py, ny);
            f->has_z = nz >= 2 && sizes_235(nz);
            if (f->has_z) fft_plan_init(m, m.pz, nz);
            std::vector<double2> wk(g.nkx);
            for (int k = 0; k < g.nkx; ++k) wk[k] = host_twiddle(k, nx);
            m.tw_r2c = falloc<double2>(m, g.nkx);
            PS3D_HIP_CHECK(hipMemcpyAsync(m.tw_r2c, wk.data(), sizeof(double2) * g.nkx, hipMemcpyHostToDevice,
                                          m.stream));
            PS3D_HIP_CHECK(hipStreamSynchronize(m.stream));
        } catch (...) {
            release(f);
            throw;
        }
        *out = f;
    });
}

int ps3d_fft_destroy(ps3d_fft *f)
{
    return guard([&] { release(f); });
}

int ps3d_fft_forward(ps3d_fft *f, const char *op, const double *qp, double *qs)
{
    return guard([&] {
        check(f);
        PS3D_REQUIRE(qp && qs, "null buffer");
        const int dims = parse_op(op, true);
        PS3D_REQUIRE(dims < 3 || f->has_z, "op 'fft' needs nz to be a product of 2, 3 and 5");
        Model &m = f->m;
        PS3D_HIP_CHECK(hipMemcpyAsync(m.aloc, qp, sizeof(double) * m.g.N, hipMemcpyHostToDevice, m.stream));
        fft_r2c_x(m, m.aloc, 0, 0, m.FS);
        fft_c2c_y(m, m.FS, -1, 0.0, 0.0);
        if (dims == 3) fft_c2c_z(m, m.FS, -1, 0.0, 0.0, 0.0);
        PS3D_HIP_CHECK(hipMemcpyAsync(qs, m.FS, sizeof(double2) * m.g.S, hipMemcpyDeviceToHost, m.stream));
        PS3D_HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}

int ps3d_fft_backward(ps3d_fft *f, const char *op, const double *qs, double *qp)
{
    return guard([&] {
        check(f);
        PS3D_REQUIRE(qp && qs, "null buffer");
        const int dims = parse_op(op, false);
        PS3D_REQUIRE(dims < 3 || f->has_z, "op 'fft' needs nz to be a product of 2, 3 and 5");
        Model &m = f->m;
        PS3D_HIP_CHECK(hipMemcpyAsync(m.FS, qs, sizeof(double2) * m.g.S, hipMemcpyHostToDevice, m.stream));
        if (dims == 3) fft_c2c_z(m, m.FS, +1, 0.0, 0.0, 0.0);
        fft_c2c_y(m, m.FS, +1, 0.0, 0.0);
        fft_c2r_x(m, m.FS, m.aloc, 0, 1.0);
        PS3D_HIP_CHECK(hipMemcpyAsync(qp, m.aloc, sizeof(double) * m.g.N, hipMemcpyDeviceToHost, m.stream));
        PS3D_HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}

}  // extern "C"
