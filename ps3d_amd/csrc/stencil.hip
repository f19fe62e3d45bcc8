// stencil.hip -- finite-volume stencils of the ps3D step on gfx950, FP64.
//
// Every kernel evaluates the reference's expressions with the same operands
// and the same association (contraction off), so results are bit-identical
// to the reference loops: advection.f90, mixing.f90, integrate.f90:21-54 and
// :87-102, pressure_solver.f90:16-24 and :98-110, parallel_mpi.f90:70-211.
//
// Fluxes are recomputed per point rather than staged in the flux_east /
// flux_north / flux_top scratch arrays: one pass over HBM per stencil
// instead of one per flux array.
#include "ps3d_internal.h"

#include <type_traits>

#pragma clang fp contract(off)

namespace ps3d {

namespace {

constexpr int BX = 64;   // one wavefront along i (unit stride)
constexpr int BY = 4;

__device__ __forceinline__ int64_t idx(const StepConsts &c, int i, int j, int k)
{
    return (int64_t)(i + 1) + c.X * (int64_t)(j + 1) + c.XY * (int64_t)(k + 1);
}

__device__ __forceinline__ int wrap(int i, int n)
{
    return i < 1 ? i + n : (i > n ? i - n : i);
}


// ---------------------------------------------------------------------------
// tendencies: advection.f90:2-184 + integrate.f90:21-54 (interior points)
// ---------------------------------------------------------------------------
template <bool DO_B, bool DO_M, bool DO_INLINE>
__global__ __launch_bounds__(BX * BY) void k_tendency(StepConsts c, const double *__restrict__ u,
                                                       const double *__restrict__ v,
                                                       const double *__restrict__ w,
                                                       const double *__restrict__ b,
                                                       double *__restrict__ du, double *__restrict__ dv,
                                                       double *__restrict__ dw, double *__restrict__ db, int kc)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    if (i > c.nx || j > c.ny) return;
    // each block marches up kc levels: planes k-1, k, k+1 are re-read from
    // this XCD's L2 instead of HBM
    const int kb = blockIdx.z * kc + 1;
    const int ke = min(c.nz, kb + kc - 1);
    for (int k = kb; k <= ke; ++k)
    {
    const int64_t X = c.X, XY = c.XY;
    const int64_t o = idx(c, i, j, k);
    const bool lid = c.lid != 0;
    const bool top0 = lid && (k >= c.nz);       // flux_top(k) zeroed at the lid
    const bool bot0 = lid && (k - 1 <= 0);      // flux_top(k-1) zeroed at the bottom
    const bool top0w = lid && (k >= c.nz - 1);  // w flux_top zeroed from nz-1 (advection.f90:169)
    const double Ro = c.Ro, dx = c.dx, dy = c.dy, dz = c.dz;

    if (DO_B) {
        double adv = 0.0;
        if (c.nonlinear) {
            const double fe1 = 0.5 * (b[o] + b[o + 1]) * u[o];
            const double fe0 = 0.5 * (b[o - 1] + b[o]) * u[o - 1];
            const double fn1 = 0.5 * (b[o] + b[o + X]) * v[o];
            const double fn0 = 0.5 * (b[o - X] + b[o]) * v[o - X];
            const double ft1 = top0 ? 0.0 : 0.5 * (b[o] + b[o + XY]) * w[o];
            const double ft0 = bot0 ? 0.0 : 0.5 * (b[o - XY] + b[o]) * w[o - XY];
            adv = -(Ro * (div_rn(fe1 - fe0, dx, c.rdx) + div_rn(fn1 - fn0, dy, c.rdy) + div_rn(ft1 - ft0, dz, c.rdz)));
        }
        if (DO_INLINE) adv = adv - 0.5 * (w[o] + w[o - XY]) * (c.N0 * c.N0);
        db[o] = adv;
    }
    if (DO_M) {
        double au = 0.0, av = 0.0, aw = 0.0;
        if (c.nonlinear) {
            {   // zonal momentum, advection.f90:65-100
                const double fe1 = 0.25 * (u[o] + u[o + 1]) * (u[o + 1] + u[o]);
                const double fe0 = 0.25 * (u[o - 1] + u[o]) * (u[o] + u[o - 1]);
                const double fn1 = 0.25 * (u[o] + u[o + X]) * (v[o + 1] + v[o]);
                const double fn0 = 0.25 * (u[o - X] + u[o]) * (v[o + 1 - X] + v[o - X]);
                const double ft1 = top0 ? 0.0 : 0.25 * (u[o + XY] + u[o]) * (w[o] + w[o + 1]);
                const double ft0 = bot0 ? 0.0 : 0.25 * (u[o] + u[o - XY]) * (w[o - XY] + w[o + 1 - XY]);
                au = -(Ro * (div_rn(fe1 - fe0, dx, c.rdx) + div_rn(fn1 - fn0, dy, c.rdy) + div_rn(ft1 - ft0, dz, c.rdz)));
            }
            {   // meridional momentum, :104-140
                const double fe1 = 0.25 * (v[o] + v[o + 1]) * (u[o + X] + u[o]);
                const double fe0 = 0.25 * (v[o - 1] + v[o]) * (u[o - 1 + X] + u[o - 1]);
                const double fn1 = 0.25 * (v[o] + v[o + X]) * (v[o + X] + v[o]);
                const double fn0 = 0.25 * (v[o - X] + v[o]) * (v[o] + v[o - X]);
                const double ft1 = top0 ? 0.0 : 0.25 * (v[o + XY] + v[o]) * (w[o] + w[o + X]);
                const double ft0 = bot0 ? 0.0 : 0.25 * (v[o] + v[o - XY]) * (w[o - XY] + w[o + X - XY]);
                av = -(Ro * (div_rn(fe1 - fe0, dx, c.rdx) + div_rn(fn1 - fn0, dy, c.rdy) + div_rn(ft1 - ft0, dz, c.rdz)));
            }
            {   // vertical momentum, :145-181
                const double fe1 = 0.25 * (w[o] + w[o + 1]) * (u[o] + u[o + XY]);
                const double fe0 = 0.25 * (w[o - 1] + w[o]) * (u[o - 1] + u[o - 1 + XY]);
                const double fn1 = 0.25 * (w[o] + w[o + X]) * (v[o] + v[o + XY]);
                const double fn0 = 0.25 * (w[o - X] + w[o]) * (v[o - X] + v[o - X + XY]);
                const double ft1 = top0w ? 0.0 : 0.25 * (w[o + XY] + w[o]) * (w[o] + w[o + XY]);
                const double ft0 = (bot0 || (lid && k - 1 >= c.nz - 1))
                                       ? 0.0
                                       : 0.25 * (w[o] + w[o - XY]) * (w[o - XY] + w[o]);
                aw = -((c.dsqr * Ro) * (div_rn(fe1 - fe0, dx, c.rdx) + div_rn(fn1 - fn0, dy, c.rdy) + div_rn(ft1 - ft0, dz, c.rdz)));
            }
        }
        if (DO_INLINE) {
            // Coriolis, integrate.f90:37-42
            au = au + (v[o] + v[o - X] + v[o + 1] + v[o + 1 - X]) * 0.25 * c.f0;
            av = av - (u[o - 1] + u[o] + u[o - 1 + X] + u[o + X]) * 0.25 * c.f0;
            // buoyancy force, :47-49, and the lid, :51-54
            aw = aw + 0.5 * (b[o] + b[o + XY]);
            if (lid && k >= c.nz) aw = 0.0;
        }
        du[o] = au;
        dv[o] = av;
        dw[o] = aw;
    }
    }
}

// ---------------------------------------------------------------------------
// the same tendencies, LDS-tiled: a work-group owns a TTX x TTY column tile
// and marches up a chunk of levels with planes k-1, k, k+1 of u, v, w, b
// (tile plus a one-point apron) in an LDS ring; the next plane is fetched
// into registers while the current level is computed, so every value is
// read from memory once per work-group.
// ---------------------------------------------------------------------------
constexpr int TTX = TEND_TX, TTY = TEND_TY;
constexpr int TPX = TTX + 2, TPY = TTY + 2;
constexpr int TTHR = TTX * TTY;
constexpr int TPER = (TPX * TPY + TTHR - 1) / TTHR;   // plane values per thread

template <class T>
struct TendArgs {
    const T *u, *v, *w, *b;
    T *du, *dv, *dw, *db;
    // mixed form: lo parts of u, v, b and the FP64 reference pressure (may be null)
    const float *ul, *vl, *bl;
    const double *pint;
    int kc, nchunk, tiles_x, tiles_y, nblocks;
};

// The flux-form advection terms at interior point (i, j, k) of level k,
// advection.f90:2-184: F(f, dk, dj, di) is field f (0..3: u, v, w, b) at
// (i+di, j+dj, k+dk).  Each returns -Ro * (flux divergence) (times dsqr for
// w), the expressions and their association those of the reference, so
// every kernel that evaluates them through these helpers gets the same bits.
template <class T, class A>
__device__ __forceinline__ T adv_b(const StepConsts &c, int k, A F)
{
    const bool lid = c.lid != 0;
    const bool top0 = lid && (k >= c.nz);
    const bool bot0 = lid && (k - 1 <= 0);
    const T h = T(0.5), zero = T(0), Ro = (T)c.Ro;
    const T fe1 = h * (F(3, 0, 0, 0) + F(3, 0, 0, 1)) * F(0, 0, 0, 0);
    const T fe0 = h * (F(3, 0, 0, -1) + F(3, 0, 0, 0)) * F(0, 0, 0, -1);
    const T fn1 = h * (F(3, 0, 0, 0) + F(3, 0, 1, 0)) * F(1, 0, 0, 0);
    const T fn0 = h * (F(3, 0, -1, 0) + F(3, 0, 0, 0)) * F(1, 0, -1, 0);
    const T ft1 = top0 ? zero : h * (F(3, 0, 0, 0) + F(3, 1, 0, 0)) * F(2, 0, 0, 0);
    const T ft0 = bot0 ? zero : h * (F(3, -1, 0, 0) + F(3, 0, 0, 0)) * F(2, -1, 0, 0);
    return -(Ro * (sdiv<T>(fe1 - fe0, c.dx, c.rdx) + sdiv<T>(fn1 - fn0, c.dy, c.rdy) +
                   sdiv<T>(ft1 - ft0, c.dz, c.rdz)));
}

template <class T, class A>
__device__ __forceinline__ T adv_u(const StepConsts &c, int k, A F)
{
    const bool lid = c.lid != 0;
    const bool top0 = lid && (k >= c.nz);
    const bool bot0 = lid && (k - 1 <= 0);
    const T q4 = T(0.25), zero = T(0), Ro = (T)c.Ro;
    const T fe1 = q4 * (F(0, 0, 0, 0) + F(0, 0, 0, 1)) * (F(0, 0, 0, 1) + F(0, 0, 0, 0));
    const T fe0 = q4 * (F(0, 0, 0, -1) + F(0, 0, 0, 0)) * (F(0, 0, 0, 0) + F(0, 0, 0, -1));
    const T fn1 = q4 * (F(0, 0, 0, 0) + F(0, 0, 1, 0)) * (F(1, 0, 0, 1) + F(1, 0, 0, 0));
    const T fn0 = q4 * (F(0, 0, -1, 0) + F(0, 0, 0, 0)) * (F(1, 0, -1, 1) + F(1, 0, -1, 0));
    const T ft1 = top0 ? zero : q4 * (F(0, 1, 0, 0) + F(0, 0, 0, 0)) * (F(2, 0, 0, 0) + F(2, 0, 0, 1));
    const T ft0 = bot0 ? zero : q4 * (F(0, 0, 0, 0) + F(0, -1, 0, 0)) * (F(2, -1, 0, 0) + F(2, -1, 0, 1));
    return -(Ro * (sdiv<T>(fe1 - fe0, c.dx, c.rdx) + sdiv<T>(fn1 - fn0, c.dy, c.rdy) +
                   sdiv<T>(ft1 - ft0, c.dz, c.rdz)));
}

template <class T, class A>
__device__ __forceinline__ T adv_v(const StepConsts &c, int k, A F)
{
    const bool lid = c.lid != 0;
    const bool top0 = lid && (k >= c.nz);
    const bool bot0 = lid && (k - 1 <= 0);
    const T q4 = T(0.25), zero = T(0), Ro = (T)c.Ro;
    const T fe1 = q4 * (F(1, 0, 0, 0) + F(1, 0, 0, 1)) * (F(0, 0, 1, 0) + F(0, 0, 0, 0));
    const T fe0 = q4 * (F(1, 0, 0, -1) + F(1, 0, 0, 0)) * (F(0, 0, 1, -1) + F(0, 0, 0, -1));
    const T fn1 = q4 * (F(1, 0, 0, 0) + F(1, 0, 1, 0)) * (F(1, 0, 1, 0) + F(1, 0, 0, 0));
    const T fn0 = q4 * (F(1, 0, -1, 0) + F(1, 0, 0, 0)) * (F(1, 0, 0, 0) + F(1, 0, -1, 0));
    const T ft1 = top0 ? zero : q4 * (F(1, 1, 0, 0) + F(1, 0, 0, 0)) * (F(2, 0, 0, 0) + F(2, 0, 1, 0));
    const T ft0 = bot0 ? zero : q4 * (F(1, 0, 0, 0) + F(1, -1, 0, 0)) * (F(2, -1, 0, 0) + F(2, -1, 1, 0));
    return -(Ro * (sdiv<T>(fe1 - fe0, c.dx, c.rdx) + sdiv<T>(fn1 - fn0, c.dy, c.rdy) +
                   sdiv<T>(ft1 - ft0, c.dz, c.rdz)));
}

template <class T, class A>
__device__ __forceinline__ T adv_w(const StepConsts &c, int k, A F)
{
    const bool lid = c.lid != 0;
    const bool bot0 = lid && (k - 1 <= 0);
    const bool top0w = lid && (k >= c.nz - 1);
    const T q4 = T(0.25), zero = T(0), Ro = (T)c.Ro;
    const T fe1 = q4 * (F(2, 0, 0, 0) + F(2, 0, 0, 1)) * (F(0, 0, 0, 0) + F(0, 1, 0, 0));
    const T fe0 = q4 * (F(2, 0, 0, -1) + F(2, 0, 0, 0)) * (F(0, 0, 0, -1) + F(0, 1, 0, -1));
    const T fn1 = q4 * (F(2, 0, 0, 0) + F(2, 0, 1, 0)) * (F(1, 0, 0, 0) + F(1, 1, 0, 0));
    const T fn0 = q4 * (F(2, 0, -1, 0) + F(2, 0, 0, 0)) * (F(1, 0, -1, 0) + F(1, 1, -1, 0));
    const T ft1 = top0w ? zero : q4 * (F(2, 1, 0, 0) + F(2, 0, 0, 0)) * (F(2, 0, 0, 0) + F(2, 1, 0, 0));
    const T ft0 = (bot0 || (lid && k - 1 >= c.nz - 1))
                      ? zero
                      : q4 * (F(2, 0, 0, 0) + F(2, -1, 0, 0)) * (F(2, -1, 0, 0) + F(2, 0, 0, 0));
    return -(((T)c.dsqr * Ro) * (sdiv<T>(fe1 - fe0, c.dx, c.rdx) + sdiv<T>(fn1 - fn0, c.dy, c.rdy) +
                                 sdiv<T>(ft1 - ft0, c.dz, c.rdz)));
}

// The FP64 tendencies of integrate.f90:12-54 at interior point (i, j, k):
// advection (when enable_nonlinear), then the inline terms -- N0^2 w for b,
// Coriolis for u and v, buoyancy and the lid for w -- as the reference adds
// them.  F as for adv_b.
template <int FLD, class A>
__device__ __forceinline__ double tendency_fp64(const StepConsts &c, int k, A F)
{
    const double h = 0.5, q4 = 0.25;
    double t = 0.0;
    if constexpr (FLD == 0) {
        if (c.nonlinear) t = adv_u<double>(c, k, F);
        t = t + (F(1, 0, 0, 0) + F(1, 0, -1, 0) + F(1, 0, 0, 1) + F(1, 0, -1, 1)) * q4 * c.f0;
    } else if constexpr (FLD == 1) {
        if (c.nonlinear) t = adv_v<double>(c, k, F);
        t = t - (F(0, 0, 0, -1) + F(0, 0, 0, 0) + F(0, 0, 1, -1) + F(0, 0, 1, 0)) * q4 * c.f0;
    } else if constexpr (FLD == 2) {
        if (c.nonlinear) t = adv_w<double>(c, k, F);
        t = t + h * (F(3, 0, 0, 0) + F(3, 1, 0, 0));
        if (c.lid != 0 && k >= c.nz) t = 0.0;
    } else {
        if (c.nonlinear) t = adv_b<double>(c, k, F);
        t = t - h * (F(2, 0, 0, 0) + F(2, -1, 0, 0)) * (c.N0 * c.N0);
    }
    return t;
}

// T = double: the drop-in (bit-exact); T = float: the FP32 stencils of the
// mixed form (same expressions, reciprocal multiplies instead of divisions)
template <class T>
__device__ __forceinline__ void tendency_body(const StepConsts &c, const TendArgs<T> &a, int bid)
{
    __shared__ T sp[4][3][TPY][TPX];
    // tiles fastest: the work-groups resident on an XCD at once hold many
    // rows of tiles of the same chunk, which march through the same levels
    // together, so the apron rows a tile shares with its y-neighbours are
    // read from HBM once and then found in L2
    const int ntile = a.tiles_x * a.tiles_y;
    const int tile = bid % ntile;
    const int ch = bid / ntile;
    const int i0 = (tile % a.tiles_x) * TTX + 1;
    const int j0 = (tile / a.tiles_x) * TTY + 1;
    const int kb = 1 + ch * a.kc;
    const int ke = min(c.nz, kb + a.kc - 1);
    const int t = threadIdx.x;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const T *fld[4] = {a.u, a.v, a.w, a.b};

    T pre[4][TPER];
    auto fetch = [&](int k) {
#pragma unroll
        for (int s = 0; s < TPER; ++s) {
            const int q = t + s * TTHR;
            const int rx = q % TPX, ry = q / TPX;
            const int ii = i0 - 1 + rx, jj = j0 - 1 + ry;
            const bool ok = q < TPX * TPY && ii <= nx + 2 && jj <= ny + 2;
            const int64_t o = idx(c, ii, jj, k);
#pragma unroll
            for (int f = 0; f < 4; ++f) pre[f][s] = ok ? fld[f][o] : T(0);
        }
    };
    auto store = [&](int k) {
        const int sl = (k - kb + 1) % 3;
#pragma unroll
        for (int s = 0; s < TPER; ++s) {
            const int q = t + s * TTHR;
            if (q < TPX * TPY) {
                const int rx = q % TPX, ry = q / TPX;
#pragma unroll
                for (int f = 0; f < 4; ++f) sp[f][sl][ry][rx] = pre[f][s];
            }
        }
    };

    fetch(kb - 1);
    store(kb - 1);
    fetch(kb);
    store(kb);
    fetch(kb + 1);
    const int px = t % TTX, py = t / TTX;
    const int i = i0 + px, j = j0 + py;
    const int rx = px + 1, ry = py + 1;
    const bool lid = c.lid != 0;
    const T h = T(0.5), zero = T(0);
    constexpr bool MIXED = std::is_same<T, float>::value;
    for (int k = kb; k <= ke; ++k) {
        store(k + 1);
        if (k + 2 <= nz + 2 && k + 1 <= ke) fetch(k + 2);
        // mixed form: the lo parts and the reference pressure this point
        // needs, issued before the barrier
        float lu[4] = {}, lv[4] = {}, lb[2] = {};
        double pp[4] = {};
        if constexpr (MIXED) {
            if (i <= nx && j <= ny) {
                const int64_t o = idx(c, i, j, k);
                const int64_t X = c.X, XY = c.XY;
                // the lo parts at the neighbours' periodic partners inside
                // the domain (the kernels keep only the hi parts' halos
                // current; under a lid b's level nz + 1 is not used)
                const int im = i == 1 ? nx : i - 1, ip = i == nx ? 1 : i + 1;
                const int jm = j == 1 ? ny : j - 1, jp = j == ny ? 1 : j + 1;
                const int kp = k == nz ? 1 : k + 1;
                lu[0] = a.ul[idx(c, im, j, k)]; lu[1] = a.ul[o]; lu[2] = a.ul[idx(c, im, jp, k)];
                lu[3] = a.ul[idx(c, i, jp, k)];
                lv[0] = a.vl[o]; lv[1] = a.vl[idx(c, i, jm, k)]; lv[2] = a.vl[idx(c, ip, j, k)];
                lv[3] = a.vl[idx(c, ip, jm, k)];
                lb[0] = a.bl[o]; lb[1] = a.bl[idx(c, i, j, kp)];
                if (a.pint) {
                    pp[0] = a.pint[o]; pp[1] = a.pint[o + 1]; pp[2] = a.pint[o + X]; pp[3] = a.pint[o + XY];
                }
            }
        }
        __syncthreads();
        const int sm = (k - kb) % 3, sc = (k - kb + 1) % 3, sx = (k - kb + 2) % 3;
        if (i <= nx && j <= ny) {
            // value of field f at (i+di, j+dj, k+dk)
            auto F = [&](int f, int dk, int dj, int di) -> T {
                const int sl = dk < 0 ? sm : (dk > 0 ? sx : sc);
                return sp[f][sl][ry + dj][rx + di];
            };
            auto U = [&](int dk, int dj, int di) { return F(0, dk, dj, di); };
            auto V = [&](int dk, int dj, int di) { return F(1, dk, dj, di); };
            auto W = [&](int dk, int dj, int di) { return F(2, dk, dj, di); };
            auto B = [&](int dk, int dj, int di) { return F(3, dk, dj, di); };
            const int64_t o = idx(c, i, j, k);
            if constexpr (!MIXED) {
                a.db[o] = tendency_fp64<3>(c, k, F);
                a.du[o] = tendency_fp64<0>(c, k, F);
                a.dv[o] = tendency_fp64<1>(c, k, F);
                a.dw[o] = tendency_fp64<2>(c, k, F);
            } else {
                T adv = zero, au = zero, av = zero, aw = zero;
                if (c.nonlinear) {
                    adv = adv_b<T>(c, k, F);
                    au = adv_u<T>(c, k, F);
                    av = adv_v<T>(c, k, F);
                    aw = adv_w<T>(c, k, F);
                }
                adv = adv - h * (W(0, 0, 0) + W(-1, 0, 0)) * (T)(c.N0 * c.N0);
                a.db[o] = adv;
                // mixed form: the Coriolis and buoyancy terms, which the
                // pressure gradient nearly balances, and that gradient in FP64
                // from (hi, lo); the slots keep tendency - grad(pint)
                auto u64 = [&](int dj, int di, int q) { return (double)U(0, dj, di) + (double)lu[q]; };
                auto v64 = [&](int dj, int di, int q) { return (double)V(0, dj, di) + (double)lv[q]; };
                const double cu = (v64(0, 0, 0) + v64(-1, 0, 1) + v64(0, 1, 2) + v64(-1, 1, 3)) * 0.25 * c.f0;
                const double cv = (u64(0, -1, 0) + u64(0, 0, 1) + u64(1, -1, 2) + u64(1, 0, 3)) * 0.25 * c.f0;
                const double bb = 0.5 * (((double)B(0, 0, 0) + (double)lb[0]) + ((double)B(1, 0, 0) + (double)lb[1]));
                double gx = 0.0, gy = 0.0, gz = 0.0;
                if (a.pint) {
                    gx = (pp[1] - pp[0]) * c.rdx;
                    gy = (pp[2] - pp[0]) * c.rdy;
                    gz = (pp[3] - pp[0]) * c.rdz;
                }
                const double su = ((double)au + cu) - gx;
                const double sv = ((double)av - cv) - gy;
                const double sw = (lid && k >= nz) ? 0.0 : ((double)aw + bb) - gz;
                a.du[o] = (float)su;
                a.dv[o] = (float)sv;
                a.dw[o] = (float)sw;
            }
        }
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(TTHR) void k_tendency_lds(StepConsts c, TendArgs<T> a)
{
    tendency_body<T>(c, a, xcd_remap((int)blockIdx.x, a.nblocks));
}

// ---------------------------------------------------------------------------
// friction, mixing.f90 -- out of place: interior from `in`, halo copied,
// so later passes see the stale halo exactly as the in-place reference does
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BX * BY) void k_harmonic(StepConsts c, const double *__restrict__ in,
                                                       double *__restrict__ out, double Hm, double Vm,
                                                       int zero_top)
{
    const int i = blockIdx.x * BX + threadIdx.x - 1;
    const int j = blockIdx.y * BY + threadIdx.y - 1;
    const int k = blockIdx.z - 1;
    if (i > c.nx + 2 || j > c.ny + 2) return;
    const int64_t o = idx(c, i, j, k);
    double r;
    if (i >= 1 && i <= c.nx && j >= 1 && j <= c.ny && k >= 1 && k <= c.nz) {
        const int64_t X = c.X, XY = c.XY;
        const bool lid = c.lid != 0;
        const double fe1 = div_rn(Hm * (in[o + 1] - in[o]), c.dx, c.rdx);
        const double fe0 = div_rn(Hm * (in[o] - in[o - 1]), c.dx, c.rdx);
        const double fn1 = div_rn(Hm * (in[o + X] - in[o]), c.dy, c.rdy);
        const double fn0 = div_rn(Hm * (in[o] - in[o - X]), c.dy, c.rdy);
        const double ft1 = (lid && k >= c.nz) ? 0.0 : div_rn(Vm * (in[o + XY] - in[o]), c.dz, c.rdz);
        const double ft0 = (lid && k - 1 <= 0) ? 0.0 : div_rn(Vm * (in[o] - in[o - XY]), c.dz, c.rdz);
        r = in[o] + c.dt * (div_rn(fe1 - fe0, c.dx, c.rdx) + div_rn(fn1 - fn0, c.dy, c.rdy) +
                            div_rn(ft1 - ft0, c.dz, c.rdz));
    } else {
        r = in[o];
    }
    if (zero_top && k >= c.nz) r = 0.0;
    out[o] = r;
}

__global__ __launch_bounds__(BX * BY) void k_biharmonic_h(StepConsts c, const double *__restrict__ in,
                                                           double *__restrict__ out, double Hm)
{
    const int i = blockIdx.x * BX + threadIdx.x - 1;
    const int j = blockIdx.y * BY + threadIdx.y - 1;
    const int k = blockIdx.z - 1;
    if (i > c.nx + 2 || j > c.ny + 2) return;
    const int64_t o = idx(c, i, j, k);
    double r;
    if (i >= 1 && i <= c.nx && j >= 1 && j <= c.ny && k >= 1 && k <= c.nz) {
        const int64_t X = c.X;
        const double dx = c.dx, dy = c.dy, rdx = c.rdx, rdy = c.rdy;
        auto f1e = [&](int64_t q) { return -div_rn(Hm * (in[q + 1] - in[q]), dx, rdx); };
        auto f1n = [&](int64_t q) { return -div_rn(Hm * (in[q + X] - in[q]), dy, rdy); };
        auto d2 = [&](int64_t q) {
            return div_rn(f1e(q) - f1e(q - 1), dx, rdx) + div_rn(f1n(q) - f1n(q - X), dy, rdy);
        };
        const double dc = d2(o), de = d2(o + 1), dwv = d2(o - 1), dn = d2(o + X), ds = d2(o - X);
        const double fe1 = div_rn(de - dc, dx, rdx), fe0 = div_rn(dc - dwv, dx, rdx);
        const double fn1 = div_rn(dn - dc, dy, rdy), fn0 = div_rn(dc - ds, dy, rdy);
        r = in[o] + c.dt * (div_rn(fe1 - fe0, dx, rdx) + div_rn(fn1 - fn0, dy, rdy));
    } else {
        r = in[o];
    }
    out[o] = r;
}

__global__ __launch_bounds__(BX * BY) void k_biharmonic_v(StepConsts c, const double *__restrict__ in,
                                                           double *__restrict__ out, double Vm,
                                                           int zero_top)
{
    const int i = blockIdx.x * BX + threadIdx.x - 1;
    const int j = blockIdx.y * BY + threadIdx.y - 1;
    const int k = blockIdx.z - 1;
    if (i > c.nx + 2 || j > c.ny + 2) return;
    const int64_t o = idx(c, i, j, k);
    double r;
    if (i >= 1 && i <= c.nx && j >= 1 && j <= c.ny && k >= 1 && k <= c.nz) {
        const int64_t XY = c.XY;
        const bool lid = c.lid != 0;
        const int nz = c.nz;
        const double dz = c.dz, rdz = c.rdz;
        auto f1 = [&](int kk, int64_t q) {
            return (lid && (kk <= 0 || kk >= nz)) ? 0.0 : -div_rn(Vm * (in[q + XY] - in[q]), dz, rdz);
        };
        auto d2 = [&](int kk, int64_t q) { return div_rn(f1(kk, q) - f1(kk - 1, q - XY), dz, rdz); };
        const double dm = d2(k - 1, o - XY), dc = d2(k, o), dp = d2(k + 1, o + XY);
        const double f21 = (lid && k >= nz) ? 0.0 : div_rn(dp - dc, dz, rdz);
        const double f20 = (lid && k - 1 <= 0) ? 0.0 : div_rn(dc - dm, dz, rdz);
        r = in[o] + c.dt * div_rn(f21 - f20, dz, rdz);
    } else {
        r = in[o];
    }
    if (zero_top && k >= c.nz) r = 0.0;
    out[o] = r;
}

// ---------------------------------------------------------------------------
// Adams-Bashforth, integrate.f90:87-102 (interior; the halo is refilled by
// the exchange that follows at :106-111)
// ---------------------------------------------------------------------------
struct AbArgs {
    const double *src[4];
    double *dst[4];
    const double *t0[4], *t1[4], *t2[4];   // tau, taum1, taum2 slots
    double scale[4];                        // dt, dt, dt/dsqr, dt
    double ca, cb, cc;                      // coefficients
    int order;                              // 1, 2, 3
    int halo;                               // also write the halo images (the exchange of integrate.f90:106-111)
};

// wavemaker forcing, for_src/config.f90:224-237 (lee_wave1.f90:320-323):
// q = q + dt*(cs*q_r + sn*q_i), cs = cos(omega t) and sn = sin(omega t)
// from the host.  Interior points only: the Adams-Bashforth update that
// follows reads the interior, and the exchange after it rebuilds the halo,
// which leaves the state the reference's whole-array update plus exchange
// leaves.  blockIdx.z runs over levels x forced fields.
struct WaveArgs {
    double *f[4];
    const double *re[4], *im[4];
};

__global__ __launch_bounds__(BX * BY) void k_wavemaker(StepConsts c, WaveArgs a, double cs, double sn)
{
    const int i = blockIdx.x * (int)blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * (int)blockDim.y + threadIdx.y + 1;
    if (i > c.nx || j > c.ny) return;
    const int k = (int)blockIdx.z % c.nz + 1, f = (int)blockIdx.z / c.nz;
    const int64_t o = idx(c, i, j, k);
    a.f[f][o] = a.f[f][o] + c.dt * (cs * a.re[f][o] + sn * a.im[f][o]);
}

// Advection by a background flow (u_back, 0, 0) and of a background
// buoyancy b_back, lee_wave1.f90:239-314, and a steady body force, added to
// the tendency slot tau at interior points.  The fluxes are the reference's
// flux_east/north/top values, formed per face where needed (the lid zeroes
// flux_top at k <= 0 and k >= nz); u, v, w, b are the post-friction fields,
// whose halos still hold the pre-friction values, as in the reference.
// The divisions and their association follow the reference's statements:
// du and db as -Ro*(dfe/dx + dfn/dy + dft/dz), dv as -(Ro*dfe)/dx, dw as
// -(dsqr*Ro*dfe)/dx.  T = float: the mixed form's version on the FP32 hi
// parts (applied before its fused friction + AB, see the step).
template <class T>
struct BgArgs {
    const T *u, *v, *w, *b;         // prognostic fields (post-friction, FP64 form)
    T *du, *dv, *dw, *db;           // tendency slot tau
    const double *ub, *bb;          // u_back, b_back (null: no background advection)
    const double *f[4];             // body force per field (null: none)
};

// the forced tendencies of interior point (i, j, k) = o: slot tau plus the
// background-flow terms and the body force
template <class T>
__device__ __forceinline__ void bg_point(const StepConsts &c, const BgArgs<T> &a, int k, int64_t o, double &du,
                                         double &dv, double &dw, double &db)
{
    const int64_t X = c.X, XY = c.XY;
    const double Ro = c.Ro;
    du = a.du[o];
    dv = a.dv[o];
    dw = a.dw[o];
    db = a.db[o];
    if (a.ub) {
        const double *UB = a.ub, *BB = a.bb;
        auto U = [&](int64_t q) { return (double)a.u[q]; };
        auto V = [&](int64_t q) { return (double)a.v[q]; };
        auto W = [&](int64_t q) { return (double)a.w[q]; };
        auto B = [&](int64_t q) { return (double)a.b[q]; };
        const bool lid = c.lid != 0;
        // u (:245-268); flux_east at face i (q = point i)
        auto fe_u = [&](int64_t q) { return 0.5 * (UB[q] + UB[q + 1]) * (U(q + 1) + U(q)); };
        auto fn_u = [&](int64_t q) { return 0.25 * (UB[q] + UB[q + X]) * (V(q + 1) + V(q)); };
        auto ft_u = [&](int64_t q, int kk) {
            return (lid && (kk <= 0 || kk >= c.nz)) ? 0.0 : 0.25 * (UB[q + XY] + UB[q]) * (W(q) + W(q + 1));
        };
        du = du - Ro * (div_rn(fe_u(o) - fe_u(o - 1), c.dx, c.rdx) + div_rn(fn_u(o) - fn_u(o - X), c.dy, c.rdy) +
                        div_rn(ft_u(o, k) - ft_u(o - XY, k - 1), c.dz, c.rdz));
        // v (:270-277)
        auto fe_v = [&](int64_t q) { return 0.25 * (V(q) + V(q + 1)) * (UB[q + X] + UB[q]); };
        dv = dv - div_rn(Ro * (fe_v(o) - fe_v(o - 1)), c.dx, c.rdx);
        // w (:279-286)
        auto fe_w = [&](int64_t q) { return 0.25 * (W(q) + W(q + 1)) * (UB[q] + UB[q + XY]); };
        dw = dw - div_rn(c.dsqr * Ro * (fe_w(o) - fe_w(o - 1)), c.dx, c.rdx);
        // b (:288-311)
        // (b_back = 0 when BB is null)
        auto fe_b = [&](int64_t q) {
            const double t = 0.5 * (B(q) + B(q + 1)) * UB[q];
            return BB ? t + 0.5 * (BB[q] + BB[q + 1]) * U(q) : t;
        };
        auto fn_b = [&](int64_t q) { return BB ? 0.5 * (BB[q] + BB[q + X]) * V(q) : 0.0; };
        auto ft_b = [&](int64_t q, int kk) {
            if (lid && (kk <= 0 || kk >= c.nz)) return 0.0;
            return BB ? 0.5 * (BB[q] + BB[q + XY]) * W(q) : 0.0;
        };
        db = db - Ro * (div_rn(fe_b(o) - fe_b(o - 1), c.dx, c.rdx) + div_rn(fn_b(o) - fn_b(o - X), c.dy, c.rdy) +
                        div_rn(ft_b(o, k) - ft_b(o - XY, k - 1), c.dz, c.rdz));
    }
    if (a.f[0]) du = du + a.f[0][o];
    if (a.f[1]) dv = dv + a.f[1][o];
    if (a.f[2]) dw = dw + a.f[2][o];
    if (a.f[3]) db = db + a.f[3][o];
}

template <class T>
__global__ __launch_bounds__(BX * BY) void k_background(StepConsts c, BgArgs<T> a)
{
    const int i = blockIdx.x * (int)blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * (int)blockDim.y + threadIdx.y + 1;
    if (i > c.nx || j > c.ny) return;
    const int k = (int)blockIdx.z + 1;
    const int64_t o = idx(c, i, j, k);
    double du, dv, dw, db;
    bg_point<T>(c, a, k, o, du, dv, dw, db);
    a.du[o] = (T)du;
    a.dv[o] = (T)dv;
    a.dw[o] = (T)dw;
    a.db[o] = (T)db;
}

// Adams-Bashforth of the mixed form after a separate friction pass (models
// with background forcing, whose terms need the post-friction state): from
// the post-friction (hi, lo) pair P, x = P + (AB increment - dt grad(Q)) in
// FP64, split back into (hi, lo), with the halo images of hi.  The same
// terms as the fused kernel's final stage (fric_body<float>).
struct AbMixedArgs {
    const float *ph[4], *pl[4];          // post-friction state
    float *oh[4], *ol[4];                // new state
    const float *t0[4], *t1[4], *t2[4];  // slots tau, taum1, taum2
    double scale[4];
    int64_t gstride[4];
    double grd[4];
    double ca, cb, cc;
    int order;
    const float *d1, *d2;
    double qc1, qc2;
    int halo;
};

__global__ __launch_bounds__(BX * BY) void k_ab_mixed(StepConsts c, AbMixedArgs a)
{
    const int i = blockIdx.x * (int)blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * (int)blockDim.y + threadIdx.y + 1;
    if (i > c.nx || j > c.ny) return;
    const int k = (int)blockIdx.z % c.nz + 1, f = (int)blockIdx.z / c.nz;
    const int64_t o = idx(c, i, j, k);
    const double s = a.scale[f];
    double ab;
    if (a.order == 1) ab = s * (double)a.t0[f][o];
    else if (a.order == 2) ab = s * (a.ca * (double)a.t0[f][o] - a.cb * (double)a.t1[f][o]);
    else ab = s * (a.ca * (double)a.t0[f][o] + a.cb * (double)a.t1[f][o] + a.cc * (double)a.t2[f][o]);
    double gq = 0.0;
    const int64_t gs = a.gstride[f];
    if (gs && a.d1) {
        double qo = a.qc1 * (double)a.d1[o], qn = a.qc1 * (double)a.d1[o + gs];
        if (a.d2) {
            qo += a.qc2 * (double)a.d2[o];
            qn += a.qc2 * (double)a.d2[o + gs];
        }
        gq = s * ((qn - qo) * a.grd[f]);
    }
    const double x = ((double)a.ph[f][o] + (double)a.pl[f][o]) + (ab - gq);
    const float hi = (float)x;
    a.oh[f][o] = hi;
    a.ol[f][o] = (float)(x - (double)hi);
    if (a.halo) write_images<float>(c, a.oh[f], i, j, k, hi, 0);
}

__global__ __launch_bounds__(BX * BY) void k_ab(StepConsts c, AbArgs a)
{
    const int i = blockIdx.x * (int)blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * (int)blockDim.y + threadIdx.y + 1;
    const int k = (int)(blockIdx.z % (unsigned)c.nz) + 1;
    const int f = (int)(blockIdx.z / (unsigned)c.nz);
    if (i > c.nx || j > c.ny) return;
    const int64_t o = idx(c, i, j, k);
    const double s = a.scale[f];
    double val;
    if (a.order == 1) {
        val = a.src[f][o] + s * a.t0[f][o];
    } else if (a.order == 2) {
        val = a.src[f][o] + s * (a.ca * a.t0[f][o] - a.cb * a.t1[f][o]);
    } else {
        val = a.src[f][o] + s * (a.ca * a.t0[f][o] + a.cb * a.t1[f][o] + a.cc * a.t2[f][o]);
    }
    a.dst[f][o] = val;
    if (a.halo) write_images<double>(c, a.dst[f], i, j, k, val, 0);
}

// the background forcing and the Adams-Bashforth update in one pass (a
// forced step whose only forcing is the background flow / body force, after
// the friction-only pass: the fields are read from the second buffers and
// the new values written to the canonical ones, so no point is read after it
// is written).  The slot tau keeps the forced tendency for later steps; the
// same expressions as k_background then k_ab.
__device__ __forceinline__ void background_ab_point(const StepConsts &c, const BgArgs<double> &bg, const AbArgs &a,
                                                    int i, int j, int k)
{
    const int64_t o = idx(c, i, j, k);
    double d[4];
    bg_point<double>(c, bg, k, o, d[0], d[1], d[2], d[3]);
    bg.du[o] = d[0];
    bg.dv[o] = d[1];
    bg.dw[o] = d[2];
    bg.db[o] = d[3];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const double s = a.scale[f];
        double val;
        if (a.order == 1) {
            val = a.src[f][o] + s * d[f];
        } else if (a.order == 2) {
            val = a.src[f][o] + s * (a.ca * d[f] - a.cb * a.t1[f][o]);
        } else {
            val = a.src[f][o] + s * (a.ca * d[f] + a.cb * a.t1[f][o] + a.cc * a.t2[f][o]);
        }
        a.dst[f][o] = val;
        if (a.halo) write_images<double>(c, a.dst[f], i, j, k, val, 0);
    }
}

__global__ __launch_bounds__(BX * BY) void k_background_ab(StepConsts c, BgArgs<double> bg, AbArgs a)
{
    const int i = blockIdx.x * (int)blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * (int)blockDim.y + threadIdx.y + 1;
    if (i > c.nx || j > c.ny) return;
    background_ab_point(c, bg, a, i, j, (int)blockIdx.z + 1);
}

// ---------------------------------------------------------------------------
// friction chain and Adams-Bashforth of one field in one pass:
// integrate.f90:59-70, :77 (harmonic, then biharmonic_h, each over the stale
// halo that the reference's in-place passes leave) and :87-102.
// A work-group owns an FTX x FTY column tile and marches up a chunk of
// levels.  Planes of f0 over the tile plus a 3-point apron rotate through
// LDS; f1 = harmonic(f0) is formed on the tile plus a 2-point apron (halo
// points keep f0, as in the reference); each tile point then applies
// biharmonic_h to f1 and the AB update, writing a second buffer (neighbouring
// tiles still read f0).
// ---------------------------------------------------------------------------
constexpr int FTX = FRIC_TX, FTY = FRIC_TY;
constexpr int R0X = FTX + 6, R0Y = FTY + 6;
constexpr int R1X = FTX + 4, R1Y = FTY + 4;
constexpr int FTHREADS = FTX * FTY;

template <class T>
struct FricField {
    const T *in;                  // f0: the FP64 field, or the FP32 hi part (mixed form)
    const T *in_lo;               // FP32 lo part (mixed form), else null
    T *out, *out_lo;
    const T *t0, *t1, *t2;        // tendency slots tau, taum1, taum2
    double Hm, Vm, Hbi;           // harmonic (H, V) and biharmonic_h coefficients
    double scale;                 // dt, or dt/dsqr for w
    int harm, bih, zero_top;
    int64_t gstride;              // mixed form: grad(Q) direction (1, X, XY; 0 for b)
    double grd;                   // 1/dx, 1/dy or 1/dz
};

template <class T>
struct FricArgs {
    FricField<T> f[4];
    double ca, cb, cc;
    int order;
    int nchunk, kc, tiles_x, tiles_y, nblocks;
    // mixed form: Q = qc1 d1 + qc2 d2 from the last two pressure increments;
    // u, v, w subtract scale * grad(Q) (see the step in api.hip)
    const float *d1, *d2;
    double qc1, qc2;
    int halo;   // also write the new values' halo images (integrate.f90:106-111)
};

// T = double: the drop-in, bit-exact.  T = float: the mixed form -- stencils
// in FP32, and the new state (hi + lo) + harmonic increment + biharmonic
// increment + AB increment summed in FP64 and split back into (hi, lo).
// field of work-group bid: blocks run over (tile, chunk, field)
template <class T>
__device__ __forceinline__ int fric_field(const FricArgs<T> &a, int bid)
{
    return bid / (a.tiles_x * a.tiles_y) / a.nchunk;
}

// The tendency fused in (TF = 0..3, the field of the work-group, FP64 only):
// the work-group forms field TF's tendency at its tile points (the
// expressions of tendency_fp64) from its own ring of f0 and, for the other
// fields, values loaded from global memory at the few points the stencil
// reads -- the work-groups of the same tile and chunk that own those fields
// run next to it on the same XCD and bring them into L2.  The tendency goes
// straight into the AB sum and into slot tau, which is never read back.
template <int TF, class A>
__device__ __forceinline__ double fused_tendency(const StepConsts &c, int k, A ring, const double *e0,
                                                 const double *em, const double *ep)
{
    auto F = [&](int f, int dk, int dj, int di) -> double {
        if (f == TF) return ring(dk, dj, di);
        if constexpr (TF == 0) {   // V at k: (j, i+1), (j, i), (j-1, i+1), (j-1, i); W at k, k-1: (j, i), (j, i+1)
            if (f == 1) return e0[(dj == 0 ? 0 : 2) + (di == 1 ? 0 : 1)];
            return dk == 0 ? e0[4 + di] : em[di];
        } else if constexpr (TF == 1) {   // U at k: (j+1, i), (j, i), (j+1, i-1), (j, i-1); W at k, k-1: (j, i), (j+1, i)
            if (f == 0) return e0[(di == 0 ? 0 : 2) + (dj == 1 ? 0 : 1)];
            return dk == 0 ? e0[4 + dj] : em[dj];
        } else if constexpr (TF == 2) {   // at k and k+1: U (j, i), (j, i-1); V (j, i), (j-1, i); B (j, i)
            const double *e = dk == 0 ? e0 : ep;
            if (f == 0) return e[di == 0 ? 0 : 1];
            if (f == 1) return e[2 + (dj == 0 ? 0 : 1)];
            return e[4];
        } else {   // U at k: (j, i), (j, i-1); V at k: (j, i), (j-1, i); W at k, k-1: (j, i)
            if (f == 0) return e0[di == 0 ? 0 : 1];
            if (f == 1) return e0[2 + (dj == 0 ? 0 : 1)];
            return dk == 0 ? e0[4] : em[0];
        }
    };
    return tendency_fp64<TF>(c, k, F);
}

// the other fields' values of fused_tendency<TF> at level kk (see there)
template <int TF>
__device__ __forceinline__ void fused_tendency_load(const StepConsts &c, const FricArgs<double> &a, int i, int j,
                                                    int kk, double *e)
{
    const int64_t o = (int64_t)(i + 1) + c.X * (int64_t)(j + 1) + c.XY * (int64_t)(kk + 1);
    const int64_t X = c.X;
    const double *U = a.f[0].in, *V = a.f[1].in, *W = a.f[2].in, *B = a.f[3].in;
    if constexpr (TF == 0) {
        e[0] = V[o + 1]; e[1] = V[o]; e[2] = V[o + 1 - X]; e[3] = V[o - X]; e[4] = W[o]; e[5] = W[o + 1];
    } else if constexpr (TF == 1) {
        e[0] = U[o + X]; e[1] = U[o]; e[2] = U[o + X - 1]; e[3] = U[o - 1]; e[4] = W[o]; e[5] = W[o + X];
    } else if constexpr (TF == 2) {
        e[0] = U[o]; e[1] = U[o - 1]; e[2] = V[o]; e[3] = V[o - X]; e[4] = B[o];
    } else {
        e[0] = U[o]; e[1] = U[o - 1]; e[2] = V[o]; e[3] = V[o - X]; e[4] = W[o];
    }
}

template <class T, int TF = -1>
__device__ __forceinline__ void fric_body(const StepConsts &c, const FricArgs<T> &a, const FricField<T> &F, int bid)
{
    static_assert(TF < 0 || std::is_same<T, double>::value, "the fused tendency is FP64 only");
    constexpr bool TEND = TF >= 0;
    constexpr bool MIXED = std::is_same<T, float>::value;
    __shared__ T f0p[3][R0Y][R0X];
    __shared__ T f1p[R1Y][R1X];
    __shared__ T d2p[FTY + 2][FTX + 2];   // biharmonic_h's del2 on the tile + 1
    __shared__ T hinc[MIXED ? FTY : 1][MIXED ? FTX : 1];   // harmonic increment at tile points
    const int ntile = a.tiles_x * a.tiles_y;
    const int tile = bid % ntile;
    const int rest = bid / ntile;
    const int ch = rest % a.nchunk;
    const int i0 = (tile % a.tiles_x) * FTX + 1;   // x-adjacent tiles run together
    const int j0 = (tile / a.tiles_x) * FTY + 1;
    const int kb = 1 + ch * a.kc;
    const int ke = min(c.nz, kb + a.kc - 1);
    const int t = threadIdx.x;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const bool lid = c.lid != 0;
    const double dx = c.dx, dy = c.dy, dz = c.dz, rdx = c.rdx, rdy = c.rdy, rdz = c.rdz;
    const T dt = (T)c.dt;
    const T zero = T(0);

    auto slot = [&](int k) { return (k - kb + 1) % 3; };
    constexpr int PER = (R0X * R0Y + FTHREADS - 1) / FTHREADS;
    T pre[PER];
    auto fetch = [&](int k) {
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            const int q = t + s * FTHREADS;
            const int rx = q % R0X, ry = q / R0X;
            const int ii = i0 - 3 + rx, jj = j0 - 3 + ry;
            const bool ok = q < R0X * R0Y && ii >= -1 && ii <= nx + 2 && jj >= -1 && jj <= ny + 2;
            pre[s] = ok ? F.in[idx(c, ii, jj, k)] : zero;
        }
    };
    auto store = [&](int k) {
        const int sl = slot(k);
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            const int q = t + s * FTHREADS;
            if (q < R0X * R0Y) f0p[sl][q / R0X][q % R0X] = pre[s];
        }
    };
    const int tpx = t % FTX, tpy = t / FTX;
    const bool mine = (i0 + tpx <= nx) && (j0 + tpy <= ny);

    fetch(kb - 1);
    store(kb - 1);
    fetch(kb);
    store(kb);
    fetch(kb + 1);
    // the tendency slots (and lo part) of a level are loaded one level ahead,
    // so their latency overlaps a whole level's LDS phases
    struct Pt {
        T s0, s1, s2, lo0;
        float q1o, q1n, q2o, q2n;   // mixed form: pressure-history increments
    };
    auto load_pt = [&](int k) {
        Pt r{zero, zero, zero, zero, 0.f, 0.f, 0.f, 0.f};
        if (mine) {
            const int64_t o = idx(c, i0 + tpx, j0 + tpy, k);
            // read once: streamed past the L2, which keeps the f0 planes
            // that neighbouring tiles share
            if constexpr (!TEND) r.s0 = __builtin_nontemporal_load(F.t0 + o);
            if (a.order >= 2) r.s1 = __builtin_nontemporal_load(F.t1 + o);
            if (a.order >= 3) r.s2 = __builtin_nontemporal_load(F.t2 + o);
            if constexpr (MIXED) {
                r.lo0 = F.in_lo[o];
                if (F.gstride && a.d1) {
                    r.q1o = a.d1[o];
                    r.q1n = a.d1[o + F.gstride];
                    if (a.d2) {
                        r.q2o = a.d2[o];
                        r.q2n = a.d2[o + F.gstride];
                    }
                }
            }
        }
        return r;
    };
    Pt nxt = load_pt(kb);
    // harmonic: the vertical flux through each point's upper face, for the
    // next level's lower face
    constexpr int NSQ1 = (R1X * R1Y + FTHREADS - 1) / FTHREADS;
    T ftc[NSQ1];
#pragma unroll
    for (int sq = 0; sq < NSQ1; ++sq) ftc[sq] = zero;
    // fused tendency: the other fields' values at level k (e0), k - 1 (em:
    // W, carried from the level before) and k + 1 (ep, field w only)
    double e0[6] = {}, em[2] = {}, ep[6] = {};
    if constexpr (TEND) {
        if (mine) {
            if constexpr (TF == 2) fused_tendency_load<TF>(c, a, i0 + tpx, j0 + tpy, kb, ep);
            else fused_tendency_load<TF>(c, a, i0 + tpx, j0 + tpy, kb - 1, e0);
        }
    }
    for (int k = kb; k <= ke; ++k) {
        store(k + 1);
        if (k + 1 <= ke) fetch(k + 2);
        const Pt cur = nxt;
        if (k + 1 <= ke) nxt = load_pt(k + 1);
        if constexpr (TEND) {
            if (mine) {
                if constexpr (TF == 2) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) e0[q] = ep[q];
                    fused_tendency_load<TF>(c, a, i0 + tpx, j0 + tpy, k + 1, ep);
                } else {
                    if constexpr (TF == 3) em[0] = e0[4];
                    else { em[0] = e0[4]; em[1] = e0[5]; }
                    fused_tendency_load<TF>(c, a, i0 + tpx, j0 + tpy, k, e0);
                }
            }
        }
        const T s0 = cur.s0, s1 = cur.s1, s2 = cur.s2, lo0 = cur.lo0;
        const float q1o = cur.q1o, q1n = cur.q1n, q2o = cur.q2o, q2n = cur.q2n;
        __syncthreads();
        const int sm = slot(k - 1), sc = slot(k), sp = slot(k + 1);
#pragma unroll
        for (int sq = 0; sq < NSQ1; ++sq) {
            const int q = t + sq * FTHREADS;
            if (q >= R1X * R1Y) break;
            const int rx = q % R1X, ry = q / R1X;
            const int ii = i0 - 2 + rx, jj = j0 - 2 + ry;
            const int ox = rx + 1, oy = ry + 1;
            const T v0 = f0p[sc][oy][ox];
            T f1 = v0, inc = zero;
            if (F.harm && ii >= 1 && ii <= nx && jj >= 1 && jj <= ny) {
                const T Hm = (T)F.Hm, Vm = (T)F.Vm;
                const T fe1 = sdiv<T>(Hm * (f0p[sc][oy][ox + 1] - v0), dx, rdx);
                const T fe0 = sdiv<T>(Hm * (v0 - f0p[sc][oy][ox - 1]), dx, rdx);
                const T fn1 = sdiv<T>(Hm * (f0p[sc][oy + 1][ox] - v0), dy, rdy);
                const T fn0 = sdiv<T>(Hm * (v0 - f0p[sc][oy - 1][ox]), dy, rdy);
                const T ft1 = (lid && k >= nz) ? zero : sdiv<T>(Vm * (f0p[sp][oy][ox] - v0), dz, rdz);
                // the flux through the face below is the face above of the
                // level before (the same operands, so the same bits), carried
                // by the thread, which keeps its point from level to level;
                // formed here at the chunk's first level
                const T ft0 = k > kb ? ftc[sq]
                                     : ((lid && k - 1 <= 0) ? zero : sdiv<T>(Vm * (v0 - f0p[sm][oy][ox]), dz, rdz));
                ftc[sq] = ft1;
                inc = dt * (sdiv<T>(fe1 - fe0, dx, rdx) + sdiv<T>(fn1 - fn0, dy, rdy) + sdiv<T>(ft1 - ft0, dz, rdz));
                f1 = v0 + inc;
            }
            if (F.zero_top && k >= nz) f1 = zero;
            f1p[ry][rx] = f1;
            if (MIXED && rx >= 2 && rx < FTX + 2 && ry >= 2 && ry < FTY + 2) hinc[ry - 2][rx - 2] = inc;
        }
        __syncthreads();
        if (F.bih) {
            // del2 = (f1e(i) - f1e(i-1))/dx + (f1n(j) - f1n(j-1))/dy, each point once
            const T Hb = (T)F.Hbi;
#pragma unroll
            for (int sq = 0; sq < ((FTX + 2) * (FTY + 2) + FTHREADS - 1) / FTHREADS; ++sq) {
                const int q = t + sq * FTHREADS;
                if (q >= (FTX + 2) * (FTY + 2)) break;
                const int xx = q % (FTX + 2) + 1, yy = q / (FTX + 2) + 1;   // f1p coordinates
                const T f1e1 = -sdiv<T>(Hb * (f1p[yy][xx + 1] - f1p[yy][xx]), dx, rdx);
                const T f1e0 = -sdiv<T>(Hb * (f1p[yy][xx] - f1p[yy][xx - 1]), dx, rdx);
                const T f1n1 = -sdiv<T>(Hb * (f1p[yy + 1][xx] - f1p[yy][xx]), dy, rdy);
                const T f1n0 = -sdiv<T>(Hb * (f1p[yy][xx] - f1p[yy - 1][xx]), dy, rdy);
                d2p[yy - 1][xx - 1] = sdiv<T>(f1e1 - f1e0, dx, rdx) + sdiv<T>(f1n1 - f1n0, dy, rdy);
            }
            __syncthreads();
        }
        if (mine) {
            const int px = tpx, py = tpy;
            const int i = i0 + px, j = j0 + py;
            const int cx = px + 2, cy = py + 2;
            T f2 = f1p[cy][cx], binc = zero;
            if (F.bih) {
                const int ex = px + 1, ey = py + 1;   // d2p coordinates
                const T dc = d2p[ey][ex], de = d2p[ey][ex + 1], dwv = d2p[ey][ex - 1];
                const T dn = d2p[ey + 1][ex], ds = d2p[ey - 1][ex];
                const T fe1 = sdiv<T>(de - dc, dx, rdx), fe0 = sdiv<T>(dc - dwv, dx, rdx);
                const T fn1 = sdiv<T>(dn - dc, dy, rdy), fn0 = sdiv<T>(dc - ds, dy, rdy);
                binc = dt * (sdiv<T>(fe1 - fe0, dx, rdx) + sdiv<T>(fn1 - fn0, dy, rdy));
                f2 = f2 + binc;
            }
            const int64_t o = idx(c, i, j, k);
            if constexpr (!MIXED) {
                double s0t = s0;
                if constexpr (TEND) {
                    const int sm2 = slot(k - 1), sc2 = slot(k), sp2 = slot(k + 1);
                    const int ry0 = py + 3, rx0 = px + 3;
                    auto ring = [&](int dk, int dj, int di) -> double {
                        const int sl = dk < 0 ? sm2 : (dk > 0 ? sp2 : sc2);
                        return f0p[sl][ry0 + dj][rx0 + di];
                    };
                    s0t = fused_tendency<TF>(c, k, ring, e0, em, ep);
                    // slot tau (read-only elsewhere in this kernel), for the next steps' AB sums
                    const_cast<double *>(F.t0)[o] = s0t;
                }
                const double s = F.scale;
                double val;
                if (a.order == 0) val = f2;   // friction only (the split step of a forced model)
                else if (a.order == 1) val = f2 + s * s0t;
                else if (a.order == 2) val = f2 + s * (a.ca * s0t - a.cb * s1);
                else val = f2 + s * (a.ca * s0t + a.cb * s1 + a.cc * s2);
                F.out[o] = val;
                if (a.halo) write_images<double>(c, F.out, i, j, k, val, 0);
            } else {
                const double s = F.scale;
                double ab;
                if (a.order == 0) ab = 0.0;   // friction only (the split step of a forced mixed model)
                else if (a.order == 1) ab = s * (double)s0;
                else if (a.order == 2) ab = s * (a.ca * (double)s0 - a.cb * (double)s1);
                else ab = s * (a.ca * (double)s0 + a.cb * (double)s1 + a.cc * (double)s2);
                // the harmonic pass zeroes w(nz) under a rigid lid before the
                // biharmonic and AB increments are added
                const bool zt = F.zero_top && k >= nz;
                const double base = zt ? 0.0 : (double)f0p[sc][cy + 1][cx + 1] + (double)lo0;
                double gq = 0.0;
                if (F.gstride && a.d1) {
                    double qo = a.qc1 * (double)q1o, qn = a.qc1 * (double)q1n;
                    if (a.d2) {
                        qo += a.qc2 * (double)q2o;
                        qn += a.qc2 * (double)q2n;
                    }
                    gq = F.scale * ((qn - qo) * F.grd);
                }
                const double x = base + ((zt ? 0.0 : (double)hinc[py][px]) + (double)binc + ab - gq);
                const float hi = (float)x;
                F.out[o] = hi;
                F.out_lo[o] = (float)(x - (double)hi);
                if (a.halo) write_images<float>(c, F.out, i, j, k, hi, 0);
            }
        }
        // no barrier here: the next level's first barrier orders this level's
        // reads of f1p, d2p and hinc before their next writes, and the f0p
        // slot refilled next held level k-2, last read before this level's
        // second barrier
    }
}

template <class T>
__global__ __launch_bounds__(FTHREADS) void k_fric_ab(StepConsts c, FricArgs<T> a)
{
    const int bid = xcd_remap((int)blockIdx.x, a.nblocks);
    fric_body<T>(c, a, a.f[fric_field(a, bid)], bid);
}

// tendency (integrate.f90:12-54), friction chain and Adams-Bashforth of the
// four fields in one launch (FP64): blocks run over (field, tile, chunk) with
// the field fastest, so the four work-groups of a tile and chunk sit together
// on one XCD and share the state's planes in its L2
__global__ __launch_bounds__(FTHREADS) void k_tend_fric_ab(StepConsts c, FricArgs<double> a)
{
    const int bid = xcd_remap((int)blockIdx.x, a.nblocks);
    const int f = bid & 3, rest = bid >> 2;
    if (f == 0) fric_body<double, 0>(c, a, a.f[0], rest);
    else if (f == 1) fric_body<double, 1>(c, a, a.f[1], rest);
    else if (f == 2) fric_body<double, 2>(c, a, a.f[2], rest);
    else fric_body<double, 3>(c, a, a.f[3], rest);
}

// ---------------------------------------------------------------------------
// halo exchange, parallel_mpi.f90:70-211 on one rank.  z, then y, then x
// periodic copies of whole planes leave every halo point holding the value
// at its fully wrapped interior index, which is what each thread copies.
// Only halo points are enumerated: 4 full z-planes, 4 y-rows per interior
// level, 4 x-points per interior row.  neumann_z: pressure_solver.f90:82-93
// (copies of levels 1 and nz into the z-halo instead of periodic images).
// ---------------------------------------------------------------------------
template <class T>
struct HaloArgsT {
    T *f[8];
    int neumann[8];
};
using HaloArgs = HaloArgsT<double>;

// coordinates of halo point t (0 <= t < number of halo points)
__device__ __forceinline__ void halo_ijk(const StepConsts &c, int64_t t, int &i, int &j, int &k)
{
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int64_t X = c.X, Y = ny + 4;
    const int64_t nzp = 4 * X * Y, nyr = 4 * X * (int64_t)nz;
    if (t < nzp) {
        const int64_t q = t / (X * Y), r = t % (X * Y);
        k = q < 2 ? (int)q - 1 : nz + (int)q - 1;   // -1, 0, nz+1, nz+2
        j = (int)(r / X) - 1;
        i = (int)(r % X) - 1;
    } else if (t < nzp + nyr) {
        const int64_t s = t - nzp;
        const int64_t q = s / (X * nz), r = s % (X * nz);
        j = q < 2 ? (int)q - 1 : ny + (int)q - 1;
        k = (int)(r / X) + 1;
        i = (int)(r % X) - 1;
    } else {
        const int64_t s = t - nzp - nyr;
        const int64_t q = s % 4, r = s / 4;
        i = q < 2 ? (int)q - 1 : nx + (int)q - 1;
        j = (int)(r % ny) + 1;
        k = (int)(r / ny) + 1;
    }
}

// halo point t (0 <= t < number of halo points) of array a
template <class T>
__device__ __forceinline__ void halo_point(const StepConsts &c, T *__restrict__ a, int neu, int64_t t)
{
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    int i, j, k;
    halo_ijk(c, t, i, j, k);
    const int si = wrap(i, nx), sj = wrap(j, ny);
    const int sk = neu ? (k < 1 ? 1 : (k > nz ? nz : k)) : wrap(k, nz);
    a[idx(c, i, j, k)] = a[idx(c, si, sj, sk)];
}

template <class T>
__global__ __launch_bounds__(256) void k_halo(StepConsts c, HaloArgsT<T> h, int64_t nhalo)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nhalo) return;
    halo_point<T>(c, h.f[blockIdx.y], h.neumann[blockIdx.y], t);
}

// dst's halo points take src's values (the same points): the halos of an
// out-of-place friction result keep the pre-friction values, as the
// reference's in-place friction leaves them
template <class T>
struct HaloCopyArgs {
    const T *src[4];
    T *dst[4];
};

template <class T>
__global__ __launch_bounds__(256) void k_halo_copy(StepConsts c, HaloCopyArgs<T> h, int64_t nhalo)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nhalo) return;
    int i, j, k;
    halo_ijk(c, t, i, j, k);
    const int64_t o = idx(c, i, j, k);
    h.dst[blockIdx.y][o] = h.src[blockIdx.y][o];
}

// ---------------------------------------------------------------------------
// projection, pressure_solver.f90:98-110 (interior).  p is read at wrapped
// indices (periodic in x, y; periodic or Neumann in z), which is what its
// exchanged halo holds, so p's own halo can be filled afterwards.
// ---------------------------------------------------------------------------
// With halo set, the halo images of the new u, v, w and of p (Neumann in z
// under a lid) are written too: the exchange of pressure_solver.f90:112-118
// and of p itself (pressure_solver.f90:82-93).  Only interior points of p
// are read, so its halo can be written in the same pass.
__device__ __forceinline__ void project_point(const StepConsts &c, double *__restrict__ u, double *__restrict__ v,
                                              double *__restrict__ w, double *__restrict__ p, int i, int j, int k,
                                              int halo)
{
    const int64_t o = idx(c, i, j, k);
    const double pc = p[o];
    const double pe = p[idx(c, i == c.nx ? 1 : i + 1, j, k)];
    const double pn = p[idx(c, i, j == c.ny ? 1 : j + 1, k)];
    const double pt = p[idx(c, i, j, k == c.nz ? (c.lid ? c.nz : 1) : k + 1)];
    const double un = u[o] - div_rn(c.dt * (pe - pc), c.dx, c.rdx);
    const double vn = v[o] - div_rn(c.dt * (pn - pc), c.dy, c.rdy);
    const double wn = (c.lid && k >= c.nz) ? 0.0 : w[o] - div_rn((c.dt / c.dsqr) * (pt - pc), c.dz, c.rdz);
    u[o] = un;
    v[o] = vn;
    w[o] = wn;
    if (halo) {
        write_images<double>(c, u, i, j, k, un, 0);
        write_images<double>(c, v, i, j, k, vn, 0);
        write_images<double>(c, w, i, j, k, wn, 0);
        write_images<double>(c, p, i, j, k, pc, c.lid);
    }
}

__global__ __launch_bounds__(256) void k_project(StepConsts c, double *__restrict__ u,
                                                  double *__restrict__ v, double *__restrict__ w,
                                                  double *__restrict__ p, int halo)
{
    // blocks renumbered per XCD: the rows of p a block reads past its own
    // tile belong to the next blocks, which then share its L2.  Work-groups
    // of 256 x 1 points where rows are that long, else BX x BY.
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int r = xcd_remap((int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z), gx * gy * (int)gridDim.z);
    const int i = (r % gx) * (int)blockDim.x + threadIdx.x + 1;
    const int j = ((r / gx) % gy) * (int)blockDim.y + threadIdx.y + 1;
    if (i > c.nx || j > c.ny) return;
    project_point(c, u, v, w, p, i, j, r / (gx * gy) + 1, halo);
}

// projection of the mixed form with the FP32 pressure increment dp: the
// state update accumulated in FP64 on (hi + lo) and split back, and the FP64
// pressure advanced, pint += dp
__device__ __forceinline__ void project_f_point(const StepConsts &c, float *__restrict__ uh, float *__restrict__ ul,
                                                float *__restrict__ vh, float *__restrict__ vl,
                                                float *__restrict__ wh, float *__restrict__ wl,
                                                float *__restrict__ dp, double *__restrict__ pint, int i, int j,
                                                int k, int halo)
{
    const int64_t o = idx(c, i, j, k);
    const double pc = dp[o];
    const double pe = dp[idx(c, i == c.nx ? 1 : i + 1, j, k)];
    const double pn = dp[idx(c, i, j == c.ny ? 1 : j + 1, k)];
    const double pt = dp[idx(c, i, j, k == c.nz ? (c.lid ? c.nz : 1) : k + 1)];
    // with halo set, the halo images of the hi parts, of the increment and of
    // pint are written too (Neumann in z under a lid for the last two)
    auto put = [&](float *h, float *l, double x) {
        const float hi = (float)x;
        h[o] = hi;
        l[o] = (float)(x - (double)hi);
        if (halo) write_images<float>(c, h, i, j, k, hi, 0);
    };
    put(uh, ul, ((double)uh[o] + (double)ul[o]) - c.dt * ((pe - pc) * c.rdx));
    put(vh, vl, ((double)vh[o] + (double)vl[o]) - c.dt * ((pn - pc) * c.rdy));
    if (c.lid && k >= c.nz) put(wh, wl, 0.0);
    else put(wh, wl, ((double)wh[o] + (double)wl[o]) - (c.dt / c.dsqr) * ((pt - pc) * c.rdz));
    const double pn1 = pint[o] + pc;
    pint[o] = pn1;
    if (halo) {
        write_images<float>(c, dp, i, j, k, (float)pc, c.lid);
        write_images<double>(c, pint, i, j, k, pn1, c.lid);
    }
}

// work-groups of 256 x 1 points where rows are that long (each of the
// fifteen FP32/FP64 streams is then read and written in whole rows), else
// BX x BY
__global__ __launch_bounds__(256) void k_project_f(StepConsts c, float *__restrict__ uh, float *__restrict__ ul,
                                                    float *__restrict__ vh, float *__restrict__ vl,
                                                    float *__restrict__ wh, float *__restrict__ wl,
                                                    float *__restrict__ dp, double *__restrict__ pint, int halo)
{
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int r = xcd_remap((int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z), gx * gy * (int)gridDim.z);
    const int i = (r % gx) * (int)blockDim.x + threadIdx.x + 1;
    const int j = ((r / gx) % gy) * (int)blockDim.y + threadIdx.y + 1;
    const int k = r / (gx * gy) + 1;
    if (i > c.nx || j > c.ny) return;
    project_f_point(c, uh, ul, vh, vl, wh, wl, dp, pint, i, j, k, halo);
}

// host-layout conversions of the mixed form: 0 split FP64 -> (hi, lo),
// 1 join (hi, lo) -> FP64, 2 FP64 -> FP32, 3 FP32 -> FP64, 4 zero FP32,
// 5 FP32 accumulate (df += sf)
__global__ __launch_bounds__(256) void k_convert(int dir, int64_t n, const double *__restrict__ sd,
                                                 double *__restrict__ dd, const float *__restrict__ sf,
                                                 float *__restrict__ df, const float *__restrict__ slo,
                                                 float *__restrict__ dlo)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    if (dir == 0) {
        const double x = sd[t];
        const float hi = (float)x;
        df[t] = hi;
        dlo[t] = (float)(x - (double)hi);
    } else if (dir == 1) {
        dd[t] = (double)sf[t] + (double)slo[t];
    } else if (dir == 2) {
        df[t] = (float)sd[t];
    } else if (dir == 3) {
        dd[t] = (double)sf[t];
    } else if (dir == 4) {
        df[t] = 0.f;
    } else {
        df[t] = df[t] + sf[t];
    }
}

inline dim3 interior_grid(const Grid &g, int zmult = 1)
{
    return dim3((unsigned)((g.nx + BX - 1) / BX), (unsigned)((g.ny + BY - 1) / BY), (unsigned)(g.nz * zmult));
}

// launch shape of the per-point kernels over the interior (blockIdx.z =
// level x zmult): rows of 256 points where nx allows it (each array is then
// read and written in whole rows), else BX x BY tiles
struct PointShape {
    dim3 grid, block;
};

inline PointShape point_shape(const Grid &g, int zmult = 1)
{
    if (g.nx >= 256)
        return {dim3((unsigned)((g.nx + 255) / 256), (unsigned)g.ny, (unsigned)(g.nz * zmult)), dim3(256, 1)};
    return {interior_grid(g, zmult), dim3(BX, BY)};
}

inline dim3 full_grid(const Grid &g)
{
    return dim3((unsigned)((g.X + BX - 1) / BX), (unsigned)((g.Y + BY - 1) / BY), (unsigned)g.Z);
}

// ---------------------------------------------------------------------------
// batched ensemble stages: one launch over all members (see EnsMember)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TTHR) void k_ens_tendency(const EnsMember *__restrict__ mem, EnsMap map,
                                                        const EnsState *__restrict__ st)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int par = st->par;
    const int64_t off = (int64_t)(st->tau - 1) * e.H;
    TendArgs<double> a{};
    a.u = e.f[0][par]; a.v = e.f[1][par]; a.w = e.f[2][par]; a.b = e.f[3][par];
    a.du = e.tend[0] + off; a.dv = e.tend[1] + off; a.dw = e.tend[2] + off; a.db = e.tend[3] + off;
    a.kc = e.tend_kc; a.nchunk = e.tend_nchunk; a.tiles_x = e.tend_tx; a.tiles_y = e.tend_ty;
    tendency_body<double>(c, a, g - map.start[mi]);
}

__global__ __launch_bounds__(FTHREADS) void k_ens_fric(const EnsMember *__restrict__ mem, EnsMap map,
                                                        const EnsState *__restrict__ st)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int par = st->par, itt = st->itt;
    const int lb = g - map.start[mi];
    FricArgs<double> a{};   // tiling and AB coefficients; the field record is built alone
    a.kc = e.fric_kc; a.nchunk = e.fric_nchunk; a.tiles_x = e.fric_tx; a.tiles_y = e.fric_ty;
    if (itt == 0) {
        a.order = 1;
    } else if (itt == 1 || !e.ab3) {
        a.order = 2;
        a.ca = 1.5 + e.eps_ab;
        a.cb = 0.5 + e.eps_ab;
    } else {
        a.order = 3;
        a.ca = 23.0 / 12.0;
        a.cb = -16.0 / 12.0;
        a.cc = 5.0 / 12.0;
    }
    const int f = fric_field(a, lb);
    const int q = f < 3 ? 0 : 1;
    FricField<double> F{};
    F.in = e.f[f][par];
    F.out = e.f[f][par ^ 1];
    F.t0 = e.tend[f] + (int64_t)(st->tau - 1) * e.H;
    F.t1 = e.tend[f] + (int64_t)(st->taum1 - 1) * e.H;
    F.t2 = e.tend[f] + (int64_t)(st->taum2 - 1) * e.H;
    F.scale = (f == 2) ? c.dt / c.dsqr : c.dt;
    F.harm = e.harm[q];
    F.Hm = e.Hm[q];
    F.Vm = e.Vm[q];
    F.bih = e.bih[q];
    F.Hbi = e.Hbi[q];
    F.zero_top = (f == 2 && F.harm && c.lid) ? 1 : 0;
    a.halo = 1;
    if (e.forced) {
        // set_forcing sees the post-friction state (integrate.f90:82): the
        // friction chain alone into the third buffers; the halo copy and the
        // forcing + AB stages follow (launch_friction_fused_only, then
        // launch_background_ab, of a single model)
        a.order = 0;
        F.out = e.tmp[f];
        a.halo = 0;
    }
    fric_body<double>(c, a, F, lb);
}

// forced members: the third buffers' halos take the pre-friction values (the
// halo the reference's in-place friction leaves), k_halo_copy per member
__global__ __launch_bounds__(256) void k_ens_halo_copy(const EnsMember *__restrict__ mem, EnsMap map,
                                                        const EnsState *__restrict__ st, int mixed)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int64_t nhalo = e.H - (int64_t)c.nx * c.ny * c.nz;
    const int64_t per = (nhalo + 255) / 256;
    const int64_t lb = g - map.start[mi];
    const int f = (int)(lb / per);
    const int64_t t = (lb % per) * 256 + threadIdx.x;
    if (f > 3 || t >= nhalo) return;
    int i, j, k;
    halo_ijk(c, t, i, j, k);
    const int64_t o = idx(c, i, j, k);
    if (mixed) e.tmph[f][o] = e.fh[f][st->par][o];   // the hi parts (halo_copy<float>)
    else e.tmp[f][o] = e.f[f][st->par][o];
}

// forced members: background flow / body force and the Adams-Bashforth
// update in one pass (k_background_ab of a single model), from the
// post-friction third buffers into the state's next buffers, halo images
// included
__global__ __launch_bounds__(BX * BY) void k_ens_background_ab(const EnsMember *__restrict__ mem, EnsMap map,
                                                                const EnsState *__restrict__ st)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int bx = (c.nx + BX - 1) / BX, by = (c.ny + BY - 1) / BY;
    const int k = lb / (bx * by) + 1, r = lb % (bx * by);
    const int i = (r % bx) * BX + (int)threadIdx.x + 1;
    const int j = (r / bx) * BY + (int)threadIdx.y + 1;
    if (i > c.nx || j > c.ny) return;
    const int par = st->par, itt = st->itt;
    const int64_t H = e.H;
    BgArgs<double> bg{};
    bg.u = e.tmp[0]; bg.v = e.tmp[1]; bg.w = e.tmp[2]; bg.b = e.tmp[3];
    const int64_t o0 = (int64_t)(st->tau - 1) * H;
    bg.du = e.tend[0] + o0; bg.dv = e.tend[1] + o0; bg.dw = e.tend[2] + o0; bg.db = e.tend[3] + o0;
    bg.ub = e.ub;
    bg.bb = e.bb;
    AbArgs a{};
    for (int q = 0; q < 4; ++q) {
        bg.f[q] = e.bf[q];
        a.src[q] = e.tmp[q];
        a.dst[q] = e.f[q][par ^ 1];
        a.t0[q] = e.tend[q] + o0;
        a.t1[q] = e.tend[q] + (int64_t)(st->taum1 - 1) * H;
        a.t2[q] = e.tend[q] + (int64_t)(st->taum2 - 1) * H;
        a.scale[q] = c.dt;
    }
    a.scale[2] = c.dt / c.dsqr;
    if (itt == 0) {
        a.order = 1;
    } else if (itt == 1 || !e.ab3) {
        a.order = 2;
        a.ca = 1.5 + e.eps_ab;
        a.cb = 0.5 + e.eps_ab;
    } else {
        a.order = 3;
        a.ca = 23.0 / 12.0;
        a.cb = -16.0 / 12.0;
        a.cc = 5.0 / 12.0;
    }
    a.halo = 1;
    background_ab_point(c, bg, a, i, j, k);
}

// mixed forced members: background flow / body force on the pre-friction hi
// parts into the FP32 slot tau (k_background<float> of a single model: the
// cold start's provisional step)
__global__ __launch_bounds__(BX * BY) void k_ens_background_f(const EnsMember *__restrict__ mem, EnsMap map,
                                                               const EnsState *__restrict__ st)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int bx = (c.nx + BX - 1) / BX, by = (c.ny + BY - 1) / BY;
    const int k = lb / (bx * by) + 1, r = lb % (bx * by);
    const int i = (r % bx) * BX + (int)threadIdx.x + 1;
    const int j = (r / bx) * BY + (int)threadIdx.y + 1;
    if (i > c.nx || j > c.ny) return;
    const int par = st->par;
    const int64_t o0 = (int64_t)(st->tau - 1) * e.H;
    BgArgs<float> bg{};
    bg.u = e.fh[0][par]; bg.v = e.fh[1][par]; bg.w = e.fh[2][par]; bg.b = e.fh[3][par];
    bg.du = e.tf[0] + o0; bg.dv = e.tf[1] + o0; bg.dw = e.tf[2] + o0; bg.db = e.tf[3] + o0;
    bg.ub = e.ub;
    bg.bb = e.bb;
    for (int q = 0; q < 4; ++q) bg.f[q] = e.bf[q];
    const int64_t o = idx(c, i, j, k);
    double du, dv, dw, db;
    bg_point<float>(c, bg, k, o, du, dv, dw, db);
    bg.du[o] = (float)du;
    bg.dv[o] = (float)dv;
    bg.dw[o] = (float)dw;
    bg.db[o] = (float)db;
}

// mixed forced members, the steady step: the forcing on the post-friction
// hi parts of the third buffers into the FP32 slot tau (k_background<float>),
// then the Adams-Bashforth update with the pressure-history term from the
// third (hi, lo) buffers into the state's next buffers, halo images of the
// hi parts included (k_ab_mixed) -- per point, with the slot's FP32 values,
// so the same bits as the two kernels of a single model
__global__ __launch_bounds__(BX * BY) void k_ens_background_ab_f(const EnsMember *__restrict__ mem, EnsMap map,
                                                                  const EnsState *__restrict__ st)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int bx = (c.nx + BX - 1) / BX, by = (c.ny + BY - 1) / BY;
    const int k = lb / (bx * by) + 1, r = lb % (bx * by);
    const int i = (r % bx) * BX + (int)threadIdx.x + 1;
    const int j = (r / bx) * BY + (int)threadIdx.y + 1;
    if (i > c.nx || j > c.ny) return;
    const int par = st->par, itt = st->itt, cur = st->dp_cur;
    const int64_t H = e.H;
    const int64_t o0 = (int64_t)(st->tau - 1) * H;
    BgArgs<float> bg{};
    bg.u = e.tmph[0]; bg.v = e.tmph[1]; bg.w = e.tmph[2]; bg.b = e.tmph[3];
    bg.du = e.tf[0] + o0; bg.dv = e.tf[1] + o0; bg.dw = e.tf[2] + o0; bg.db = e.tf[3] + o0;
    bg.ub = e.ub;
    bg.bb = e.bb;
    for (int q = 0; q < 4; ++q) bg.f[q] = e.bf[q];
    const int64_t o = idx(c, i, j, k);
    double fd[4];
    bg_point<float>(c, bg, k, o, fd[0], fd[1], fd[2], fd[3]);
    float sl[4];
    for (int q = 0; q < 4; ++q) sl[q] = (float)fd[q];
    bg.du[o] = sl[0];
    bg.dv[o] = sl[1];
    bg.dw[o] = sl[2];
    bg.db[o] = sl[3];
    // the AB order and pressure-history coefficients of the step
    // (integrate_mixed of a single model)
    int order;
    double ca = 0.0, cb = 0.0, cc = 0.0, qc1 = 0.0, qc2 = 0.0;
    if (itt == 0) {
        order = 1;
    } else if (itt == 1 || !e.ab3) {
        order = 2;
        ca = 1.5 + e.eps_ab;
        cb = 0.5 + e.eps_ab;
        qc1 = -(0.5 + e.eps_ab);
    } else {
        order = 3;
        ca = 23.0 / 12.0;
        cb = -16.0 / 12.0;
        cc = 5.0 / 12.0;
        qc1 = -16.0 / 12.0 + 5.0 / 12.0;
        qc2 = 5.0 / 12.0;
    }
    const float *d1 = qc1 != 0.0 ? e.dpr[cur] : nullptr;
    const float *d2 = qc2 != 0.0 ? e.dpr[(cur + 2) % 3] : nullptr;
    for (int f = 0; f < 4; ++f) {
        const double s = (f == 2) ? c.dt / c.dsqr : c.dt;
        const float *t1 = e.tf[f] + (int64_t)(st->taum1 - 1) * H;
        const float *t2 = e.tf[f] + (int64_t)(st->taum2 - 1) * H;
        double ab;
        if (order == 1) ab = s * (double)sl[f];
        else if (order == 2) ab = s * (ca * (double)sl[f] - cb * (double)t1[o]);
        else ab = s * (ca * (double)sl[f] + cb * (double)t1[o] + cc * (double)t2[o]);
        double gq = 0.0;
        const int64_t gs = f == 0 ? 1 : (f == 1 ? c.X : (f == 2 ? c.XY : 0));
        if (gs && d1) {
            const double grd = f == 0 ? c.rdx : (f == 1 ? c.rdy : c.rdz);
            double qo = qc1 * (double)d1[o], qn = qc1 * (double)d1[o + gs];
            if (d2) {
                qo += qc2 * (double)d2[o];
                qn += qc2 * (double)d2[o + gs];
            }
            gq = s * ((qn - qo) * grd);
        }
        const double x = ((double)e.tmph[f][o] + (double)e.tmpl[f][o]) + (ab - gq);
        const float hi = (float)x;
        e.fh[f][par ^ 1][o] = hi;
        e.fl[f][par ^ 1][o] = (float)(x - (double)hi);
        write_images<float>(c, e.fh[f][par ^ 1], i, j, k, hi, 0);
    }
}

__global__ __launch_bounds__(BX * BY) void k_ens_project(const EnsMember *__restrict__ mem, EnsMap map,
                                                          const EnsState *__restrict__ st)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int bx = (c.nx + BX - 1) / BX, by = (c.ny + BY - 1) / BY;
    const int i = (lb % bx) * BX + threadIdx.x + 1;
    const int j = ((lb / bx) % by) * BY + threadIdx.y + 1;
    const int k = lb / (bx * by) + 1;
    if (i > c.nx || j > c.ny) return;
    const int nb = st->par ^ 1;
    project_point(c, e.f[0][nb], e.f[1][nb], e.f[2][nb], e.p, i, j, k, 1);
}

// end of the step: main.f90:156-160 (tau rotation, itt) and the buffer swap
__global__ void k_ens_advance(EnsState *st)
{
    if (threadIdx.x == 0) {
        EnsState s = *st;
        s.tau = s.tau % 3 + 1;
        s.taum1 = s.taum1 % 3 + 1;
        s.taum2 = s.taum2 % 3 + 1;
        s.itt += 1;
        s.par ^= 1;
        s.dp_cur = (s.dp_cur + 1) % 3;
        *st = s;
    }
}

// ---- the same stages for members of the mixed form -------------------------
// (the bodies and arguments of tendency_impl<float>, friction_ab_impl<float>
// and k_project_f, with the member's buffers of the current parity)
__global__ __launch_bounds__(TTHR) void k_ens_tendency_f(const EnsMember *__restrict__ mem, EnsMap map,
                                                          const EnsState *__restrict__ st, int cold)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int par = st->par;
    const int64_t off = (int64_t)(st->tau - 1) * e.H;
    TendArgs<float> a{};
    a.u = e.fh[0][par]; a.v = e.fh[1][par]; a.w = e.fh[2][par]; a.b = e.fh[3][par];
    a.ul = e.fl[0][par]; a.vl = e.fl[1][par]; a.bl = e.fl[3][par];
    a.pint = cold ? nullptr : e.pint;
    a.du = e.tf[0] + off; a.dv = e.tf[1] + off; a.dw = e.tf[2] + off; a.db = e.tf[3] + off;
    a.kc = e.tend_kc; a.nchunk = e.tend_nchunk; a.tiles_x = e.tend_tx; a.tiles_y = e.tend_ty;
    tendency_body<float>(c, a, g - map.start[mi]);
}

__global__ __launch_bounds__(FTHREADS) void k_ens_fric_f(const EnsMember *__restrict__ mem, EnsMap map,
                                                          const EnsState *__restrict__ st, int cold)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int par = st->par, itt = cold ? 0 : st->itt, cur = st->dp_cur;
    const int lb = g - map.start[mi];
    FricArgs<float> a{};
    a.kc = e.fric_kc; a.nchunk = e.fric_nchunk; a.tiles_x = e.fric_tx; a.tiles_y = e.fric_ty;
    double qc1 = 0.0, qc2 = 0.0;
    if (itt == 0) {
        a.order = 1;
    } else if (itt == 1 || !e.ab3) {
        a.order = 2;
        a.ca = 1.5 + e.eps_ab;
        a.cb = 0.5 + e.eps_ab;
        qc1 = -(0.5 + e.eps_ab);
    } else {
        a.order = 3;
        a.ca = 23.0 / 12.0;
        a.cb = -16.0 / 12.0;
        a.cc = 5.0 / 12.0;
        qc1 = -16.0 / 12.0 + 5.0 / 12.0;
        qc2 = 5.0 / 12.0;
    }
    // Q = qc1 dp_{n-1} + qc2 dp_{n-2}: the pressure-history term of the step
    a.qc1 = qc1;
    a.qc2 = qc2;
    a.d1 = qc1 != 0.0 ? e.dpr[cur] : nullptr;
    a.d2 = qc2 != 0.0 ? e.dpr[(cur + 2) % 3] : nullptr;
    const int f = fric_field(a, lb);
    const int q = f < 3 ? 0 : 1;
    FricField<float> F{};
    F.in = e.fh[f][par];
    F.in_lo = e.fl[f][par];
    F.out = e.fh[f][par ^ 1];
    F.out_lo = e.fl[f][par ^ 1];
    F.t0 = e.tf[f] + (int64_t)(st->tau - 1) * e.H;
    F.t1 = e.tf[f] + (int64_t)(st->taum1 - 1) * e.H;
    F.t2 = e.tf[f] + (int64_t)(st->taum2 - 1) * e.H;
    F.scale = (f == 2) ? c.dt / c.dsqr : c.dt;
    F.harm = e.harm[q];
    F.Hm = e.Hm[q];
    F.Vm = e.Vm[q];
    F.bih = e.bih[q];
    F.Hbi = e.Hbi[q];
    F.zero_top = (f == 2 && F.harm && c.lid) ? 1 : 0;
    F.gstride = f == 0 ? 1 : (f == 1 ? c.X : (f == 2 ? c.XY : 0));
    F.grd = f == 0 ? c.rdx : (f == 1 ? c.rdy : c.rdz);
    a.halo = 1;
    if (e.forced && !cold) {
        // the forcing sees the post-friction state: friction alone into the
        // third (hi, lo) buffers (launch_friction_forcing_ab_mixed of a
        // single model); the halo copy and forcing + AB stages follow
        a.order = 0;
        a.qc1 = a.qc2 = 0.0;
        a.d1 = a.d2 = nullptr;
        F.out = e.tmph[f];
        F.out_lo = e.tmpl[f];
        a.halo = 0;
    }
    fric_body<float>(c, a, F, lb);
}

__device__ __forceinline__ float *ens_increment(const EnsMember &e, int cur, int which)
{
    return which == 0 ? e.dpr[0] : e.dpr[(cur + which) % 3];
}

__global__ __launch_bounds__(BX * BY) void k_ens_project_f(const EnsMember *__restrict__ mem, EnsMap map,
                                                            const EnsState *__restrict__ st, int which)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int bx = (c.nx + BX - 1) / BX, by = (c.ny + BY - 1) / BY;
    const int i = (lb % bx) * BX + threadIdx.x + 1;
    const int j = ((lb / bx) % by) * BY + threadIdx.y + 1;
    const int k = lb / (bx * by) + 1;
    if (i > c.nx || j > c.ny) return;
    const int nb = st->par ^ 1;
    project_f_point(c, e.fh[0][nb], e.fl[0][nb], e.fh[1][nb], e.fl[1][nb], e.fh[2][nb], e.fl[2][nb],
                    ens_increment(e, st->dp_cur, which), e.pint, i, j, k, 1);
}

// cold start: pint = the FP32 estimate in dpr[0], with its halo images
// (Neumann in z under a lid), as the halo exchange of the estimate followed
// by the conversion to FP64
__global__ __launch_bounds__(BX * BY) void k_ens_cold_pint(const EnsMember *__restrict__ mem, EnsMap map)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int bx = (c.nx + BX - 1) / BX, by = (c.ny + BY - 1) / BY;
    const int i = (lb % bx) * BX + threadIdx.x + 1;
    const int j = ((lb / bx) % by) * BY + threadIdx.y + 1;
    const int k = lb / (bx * by) + 1;
    if (i > c.nx || j > c.ny) return;
    const int64_t o = idx(c, i, j, k);
    const double x = (double)e.dpr[0][o];
    e.pint[o] = x;
    write_images<double>(c, e.pint, i, j, k, x, c.lid);
}

// over the H points of each member: what 0 zeroes the three increments,
// what 1 adds the correction to the new increment
__global__ __launch_bounds__(256) void k_ens_points_f(const EnsMember *__restrict__ mem, EnsMap map,
                                                       const EnsState *__restrict__ st, int what)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const int64_t t = (int64_t)(g - map.start[mi]) * 256 + threadIdx.x;
    if (t >= e.H) return;
    if (what == 0) {
        e.dpr[0][t] = 0.f;
        e.dpr[1][t] = 0.f;
        e.dpr[2][t] = 0.f;
    } else {
        const int cur = st->dp_cur;
        float *dn = e.dpr[(cur + 1) % 3];
        dn[t] = dn[t] + e.dpr[(cur + 2) % 3][t];
    }
}

}  // namespace

double *Model::field_buf(int id) const
{
    switch (id) {
    case PS3D_U: return u;
    case PS3D_V: return v;
    case PS3D_W: return w;
    case PS3D_B: return b;
    case PS3D_P: return p;
    default: return nullptr;
    }
}

// grid of column tiles x level chunks for the k-marching stencils: enough
// blocks to fill the device (~2048), chunks of at least 4 levels
inline dim3 march_grid(const Grid &g, int &kc)
{
    const int tiles = ((g.nx + BX - 1) / BX) * ((g.ny + BY - 1) / BY);
    int nch = (2048 + tiles - 1) / tiles;
    nch = nch < 1 ? 1 : nch;
    kc = (g.nz + nch - 1) / nch;
    kc = kc < 4 ? (g.nz < 4 ? g.nz : 4) : kc;
    nch = (g.nz + kc - 1) / kc;
    return dim3((unsigned)((g.nx + BX - 1) / BX), (unsigned)((g.ny + BY - 1) / BY), (unsigned)nch);
}

template <class T>
static void tendency_impl(Model &m, const T *u, const T *v, const T *w, const T *b, T *const *tend, int tau,
                          const float *ul = nullptr, const float *vl = nullptr, const float *bl = nullptr,
                          const double *pint = nullptr)
{
    ProfScope ps(m, "tendency");
    const Grid &g = m.g;
    const int64_t off = (int64_t)(tau - 1) * g.H;
    TendArgs<T> a{};
    a.u = u; a.v = v; a.w = w; a.b = b;
    a.ul = ul; a.vl = vl; a.bl = bl; a.pint = pint;
    a.du = tend[0] + off; a.dv = tend[1] + off; a.dw = tend[2] + off; a.db = tend[3] + off;
    a.tiles_x = (g.nx + TTX - 1) / TTX;
    a.tiles_y = (g.ny + TTY - 1) / TTY;
    const int tiles = a.tiles_x * a.tiles_y;
    // about 2 work-groups per CU over 256 CUs, chunks of at least 8 levels
    // (or the tuned number of levels): longer chunks re-read fewer apron
    // planes, which at 256x256x64 outweighs the extra work-groups
    int nchunk = (512 + tiles - 1) / tiles;
    const int maxch = g.nz / 8 > 1 ? g.nz / 8 : 1;
    nchunk = nchunk > maxch ? maxch : nchunk;
    if (m.tend_kc > 0) nchunk = (g.nz + m.tend_kc - 1) / m.tend_kc;
    a.kc = (g.nz + nchunk - 1) / nchunk;
    a.nchunk = (g.nz + a.kc - 1) / a.kc;
    a.nblocks = tiles * a.nchunk;
    hipLaunchKernelGGL(k_tendency_lds<T>, dim3((unsigned)a.nblocks), dim3(TTHR), 0, m.stream, m.c, a);
    PS3D_HIP_CHECK(hipGetLastError());
}

void launch_tendency(Model &m, int tau)
{
    PS3D_REQUIRE(!m.mixed, "tendency: FP64 models only");
    double *tend[4] = {m.du, m.dv, m.dw, m.db};
    tendency_impl<double>(m, m.u, m.v, m.w, m.b, tend, tau);
}

void launch_tendency_mixed(Model &m, int tau, const double *pint)
{
    tendency_impl<float>(m, m.fh[0], m.fh[1], m.fh[2], m.fh[3], m.tf, tau, m.fl[0], m.fl[1], m.fl[3], pint);
}

void launch_buoyancy_advection(Model &m, int tau)
{
    const int64_t off = (int64_t)(tau - 1) * m.g.H;
    int kc = 1;
    const dim3 grid = march_grid(m.g, kc);
    hipLaunchKernelGGL((k_tendency<true, false, false>), grid, dim3(BX, BY), 0, m.stream, m.c, m.u, m.v, m.w,
                       m.b, nullptr, nullptr, nullptr, m.db + off, kc);
    PS3D_HIP_CHECK(hipGetLastError());
}

// advection.f90:2-184 of arbitrary fields (diag_balance.f90:500-535): the
// advective tendencies only, no Coriolis / buoyancy terms
void launch_advection_fields(Model &m, const double *u, const double *v, const double *w, const double *b,
                             double *du, double *dv, double *dw, double *db)
{
    int kc = 1;
    const dim3 grid = march_grid(m.g, kc);
    hipLaunchKernelGGL((k_tendency<true, true, false>), grid, dim3(BX, BY), 0, m.stream, m.c, u, v, w, b, du, dv,
                       dw, db, kc);
    PS3D_HIP_CHECK(hipGetLastError());
}

void launch_momentum_advection(Model &m, int tau)
{
    const int64_t off = (int64_t)(tau - 1) * m.g.H;
    int kc = 1;
    const dim3 grid = march_grid(m.g, kc);
    hipLaunchKernelGGL((k_tendency<false, true, false>), grid, dim3(BX, BY), 0, m.stream, m.c, m.u, m.v, m.w,
                       m.b, m.du + off, m.dv + off, m.dw + off, nullptr, kc);
    PS3D_HIP_CHECK(hipGetLastError());
}

void launch_mixing(Model &m, int op, int f, double hmix, double vmix, int zero_top)
{
    const double *in = m.cur_buf(f);
    double *out = m.cur[f] ? m.field_buf(f) : m.alt[f];
    switch (op) {
    case PS3D_HARMONIC: {
        ProfScope ps(m, "harmonic");
        hipLaunchKernelGGL(k_harmonic, full_grid(m.g), dim3(BX, BY), 0, m.stream, m.c, in, out, hmix, vmix,
                           zero_top);
        break;
    }
    case PS3D_BIHARMONIC_H: {
        ProfScope ps(m, "biharmonic_h");
        hipLaunchKernelGGL(k_biharmonic_h, full_grid(m.g), dim3(BX, BY), 0, m.stream, m.c, in, out, hmix);
        break;
    }
    case PS3D_BIHARMONIC_V: {
        ProfScope ps(m, "biharmonic_v");
        hipLaunchKernelGGL(k_biharmonic_v, full_grid(m.g), dim3(BX, BY), 0, m.stream, m.c, in, out, vmix,
                           zero_top);
        break;
    }
    default: throw Error(PS3D_ERR_ARG, "unknown mixing operator");
    }
    PS3D_HIP_CHECK(hipGetLastError());
    m.cur[f] ^= 1;
}

// integrate.f90:59-78
void launch_friction_block(Model &m)
{
    const ps3d_params &p = m.prm;
    const int lid = p.enable_vertical_boundaries;
    if (p.Ah > 0. || p.Av > 0.) {
        launch_mixing(m, PS3D_HARMONIC, PS3D_U, p.Ah, p.Av, 0);
        launch_mixing(m, PS3D_HARMONIC, PS3D_V, p.Ah, p.Av, 0);
        launch_mixing(m, PS3D_HARMONIC, PS3D_W, p.Ah, p.Av, lid);
    }
    if (p.Kh > 0. || p.Kv > 0.) launch_mixing(m, PS3D_HARMONIC, PS3D_B, p.Kh, p.Kv, 0);
    if (p.Ahbi > 0.) {
        launch_mixing(m, PS3D_BIHARMONIC_H, PS3D_U, p.Ahbi, 0.0, 0);
        launch_mixing(m, PS3D_BIHARMONIC_H, PS3D_V, p.Ahbi, 0.0, 0);
        launch_mixing(m, PS3D_BIHARMONIC_H, PS3D_W, p.Ahbi, 0.0, 0);
    }
    if (p.Avbi > 0.) {
        launch_mixing(m, PS3D_BIHARMONIC_V, PS3D_U, 0.0, p.Avbi, 0);
        launch_mixing(m, PS3D_BIHARMONIC_V, PS3D_V, 0.0, p.Avbi, 0);
        launch_mixing(m, PS3D_BIHARMONIC_V, PS3D_W, 0.0, p.Avbi, lid);
    }
    if (p.Khbi > 0.) launch_mixing(m, PS3D_BIHARMONIC_H, PS3D_B, p.Khbi, 0.0, 0);
    if (p.Kvbi > 0.) launch_mixing(m, PS3D_BIHARMONIC_V, PS3D_B, 0.0, p.Kvbi, 0);
}

static BgArgs<double> bg_args(Model &m, int tau)
{
    const int64_t off = (int64_t)(tau - 1) * m.g.H;
    BgArgs<double> a{};
    a.u = m.cur_buf(0); a.v = m.cur_buf(1); a.w = m.cur_buf(2); a.b = m.cur_buf(3);
    a.du = m.du + off; a.dv = m.dv + off; a.dw = m.dw + off; a.db = m.db + off;
    a.ub = m.bg_on ? m.bg_u : nullptr;
    a.bb = (m.bg_on && m.bg_b_on) ? m.bg_b : nullptr;
    for (int f = 0; f < 4; ++f) a.f[f] = m.bf_on[f] ? m.bf[f] : nullptr;
    return a;
}

static BgArgs<float> bg_args_mixed(Model &m, int tau, float *const *fields)
{
    const int64_t off = (int64_t)(tau - 1) * m.g.H;
    BgArgs<float> a{};
    a.u = fields[0]; a.v = fields[1]; a.w = fields[2]; a.b = fields[3];
    a.du = m.tf[0] + off; a.dv = m.tf[1] + off; a.dw = m.tf[2] + off; a.db = m.tf[3] + off;
    a.ub = m.bg_on ? m.bg_u : nullptr;
    a.bb = (m.bg_on && m.bg_b_on) ? m.bg_b : nullptr;
    for (int f = 0; f < 4; ++f) a.f[f] = m.bf_on[f] ? m.bf[f] : nullptr;
    return a;
}

void launch_background(Model &m, int tau, float *const *mixed_fields)
{
    if (!m.mixed) {
        const PointShape sh = point_shape(m.g);
        hipLaunchKernelGGL(k_background<double>, sh.grid, sh.block, 0, m.stream, m.c,
                           bg_args(m, tau));
    } else {
        const PointShape sh = point_shape(m.g);
        hipLaunchKernelGGL(k_background<float>, sh.grid, sh.block, 0, m.stream, m.c,
                           bg_args_mixed(m, tau, mixed_fields ? mixed_fields : m.fh));
    }
    PS3D_HIP_CHECK(hipGetLastError());
}

void launch_wavemaker(Model &m, double cs, double sn)
{
    WaveArgs a{};
    int n = 0;
    for (int f = 0; f < 4; ++f) {
        if (!m.wm_field[f]) continue;
        a.f[n] = m.cur_buf(f);
        a.re[n] = m.wm_re[f];
        a.im[n] = m.wm_im[f];
        ++n;
    }
    if (!n) return;
    const PointShape sh = point_shape(m.g, n);
    hipLaunchKernelGGL(k_wavemaker, sh.grid, sh.block, 0, m.stream, m.c, a, cs, sn);
    PS3D_HIP_CHECK(hipGetLastError());
}

// integrate.f90:87-101: the AB update from the current buffers into the
// canonical ones
static AbArgs ab_args(Model &m, int itt, int tau, int taum1, int taum2)
{
    AbArgs a{};
    const int64_t H = m.g.H;
    double *tend[4] = {m.du, m.dv, m.dw, m.db};
    for (int f = 0; f < 4; ++f) {
        a.src[f] = m.cur_buf(f);
        a.dst[f] = m.field_buf(f);
        a.t0[f] = tend[f] + (int64_t)(tau - 1) * H;
        a.t1[f] = tend[f] + (int64_t)(taum1 - 1) * H;
        a.t2[f] = tend[f] + (int64_t)(taum2 - 1) * H;
        a.scale[f] = m.prm.dt;
    }
    a.scale[2] = m.prm.dt / m.prm.dsqr;
    if (itt == 0) {
        a.order = 1;
    } else if (itt == 1 || !m.prm.enable_AB_3_order) {
        a.order = 2;
        a.ca = 1.5 + m.prm.eps_ab;
        a.cb = 0.5 + m.prm.eps_ab;
    } else {
        a.order = 3;
        a.ca = 23.0 / 12.0;   // main_module.f90:57
        a.cb = -16.0 / 12.0;
        a.cc = 5.0 / 12.0;
    }
    a.halo = m.fuse_halo ? 1 : 0;
    return a;
}

// after the AB update: the canonical buffers are current; their halos
// (integrate.f90:106-111) were written as images or are exchanged now
static void ab_finish(Model &m)
{
    for (int f = 0; f < 4; ++f) m.cur[f] = 0;
    if (m.fuse_halo) return;
    HaloArgs h{};
    for (int f = 0; f < 4; ++f) { h.f[f] = m.field_buf(f); h.neumann[f] = 0; }
    const Grid &g = m.g;
    const int64_t nhalo = g.H - g.N;
    ProfScope ps(m, "halo");
    hipLaunchKernelGGL(k_halo<double>, dim3((unsigned)((nhalo + 255) / 256), 4), dim3(256), 0, m.stream, m.c, h, nhalo);
    PS3D_HIP_CHECK(hipGetLastError());
}

// integrate.f90:87-111
void launch_ab_exchange(Model &m, int itt, int tau, int taum1, int taum2)
{
    const AbArgs a = ab_args(m, itt, tau, taum1, taum2);
    {
        ProfScope ps(m, "ab_update");
        const PointShape sh = point_shape(m.g, 4);
        hipLaunchKernelGGL(k_ab, sh.grid, sh.block, 0, m.stream, m.c, a);
        PS3D_HIP_CHECK(hipGetLastError());
    }
    ab_finish(m);
}

// set_forcing (background flow and body force only) and integrate.f90:87-111
// in one pass, after launch_friction_fused_only
void launch_background_ab(Model &m, int itt, int tau, int taum1, int taum2)
{
    for (int f = 0; f < 4; ++f)
        PS3D_REQUIRE(m.cur[f] == 1, "forcing + AB in one pass: the friction results must be in the second buffers");
    const AbArgs a = ab_args(m, itt, tau, taum1, taum2);
    {
        ProfScope ps(m, "ab_update");
        const PointShape sh = point_shape(m.g);
        hipLaunchKernelGGL(k_background_ab, sh.grid, sh.block, 0, m.stream, m.c, bg_args(m, tau), a);
        PS3D_HIP_CHECK(hipGetLastError());
    }
    ab_finish(m);
}

bool fused_friction_ok(const Model &m)
{
    return !(m.prm.Avbi > 0.) && !(m.prm.Kvbi > 0.);
}

bool fused_tendency_active(const Model &m)
{
    const bool forced = m.forcing || m.wm_on || m.bg_on ||
                        (m.bf_on[0] || m.bf_on[1] || m.bf_on[2] || m.bf_on[3]);
    return !m.mixed && fused_friction_ok(m) && m.fuse_friction && m.fuse_tendency && !forced;
}

// integrate.f90:59-111 for configs without vertical biharmonic friction:
// the friction chain and AB of all four fields in one launch, which also
// writes the new values' halo images (the exchange of integrate.f90:106-111).
// The new fields are written to the second buffers, which then become the
// canonical ones.
// itt < 0: friction only (no AB).  halo: write the new values' halo images
// (-1: where the model does, fuse_halo)
template <class T>
static void friction_ab_impl(Model &m, const T *const *in, const T *const *in_lo, T *const *out, T *const *out_lo,
                             T *const *tend, int itt, int tau, int taum1, int taum2, double qc1 = 0.0,
                             const float *d1 = nullptr, double qc2 = 0.0, const float *d2 = nullptr, int halo = -1,
                             bool fuse_tend = false)
{
    const ps3d_params &p = m.prm;
    const Grid &g = m.g;
    FricArgs<T> a{};
    a.d1 = d1;
    a.d2 = d2;
    a.qc1 = qc1;
    a.qc2 = qc2;
    for (int f = 0; f < 4; ++f) {
        FricField<T> &F = a.f[f];
        F.in = in[f];
        F.in_lo = in_lo ? in_lo[f] : nullptr;
        F.out = out[f];
        F.out_lo = out_lo ? out_lo[f] : nullptr;
        F.t0 = tend[f] + (int64_t)(tau - 1) * g.H;
        F.t1 = tend[f] + (int64_t)(taum1 - 1) * g.H;
        F.t2 = tend[f] + (int64_t)(taum2 - 1) * g.H;
        F.scale = (f == 2) ? p.dt / p.dsqr : p.dt;
        if (f < 3) {
            F.harm = (p.Ah > 0. || p.Av > 0.);
            F.Hm = p.Ah;
            F.Vm = p.Av;
            F.bih = p.Ahbi > 0.;
            F.Hbi = p.Ahbi;
        } else {
            F.harm = (p.Kh > 0. || p.Kv > 0.);
            F.Hm = p.Kh;
            F.Vm = p.Kv;
            F.bih = p.Khbi > 0.;
            F.Hbi = p.Khbi;
        }
        F.zero_top = (f == 2 && F.harm && p.enable_vertical_boundaries) ? 1 : 0;
        F.gstride = f == 0 ? 1 : (f == 1 ? g.X : (f == 2 ? g.XY : 0));
        F.grd = f == 0 ? m.c.rdx : (f == 1 ? m.c.rdy : m.c.rdz);
    }
    if (itt < 0) {
        a.order = 0;   // friction only (mixed form with background forcing)
    } else if (itt == 0) {
        a.order = 1;
    } else if (itt == 1 || !p.enable_AB_3_order) {
        a.order = 2;
        a.ca = 1.5 + p.eps_ab;
        a.cb = 0.5 + p.eps_ab;
    } else {
        a.order = 3;
        a.ca = 23.0 / 12.0;
        a.cb = -16.0 / 12.0;
        a.cc = 5.0 / 12.0;
    }
    a.tiles_x = (g.nx + FTX - 1) / FTX;
    a.tiles_y = (g.ny + FTY - 1) / FTY;
    const int tiles = a.tiles_x * a.tiles_y;
    const int target = 2048 * 512 / FTHREADS;   // ~8 work-groups of 512 threads per CU
    int nchunk = (target + tiles * 4 - 1) / (tiles * 4);
    nchunk = nchunk < 1 ? 1 : (nchunk > g.nz ? g.nz : nchunk);
    const int kc_set = fuse_tend ? m.tfric_kc : m.fric_kc;
    if (kc_set > 0) nchunk = (g.nz + kc_set - 1) / kc_set;
    a.kc = (g.nz + nchunk - 1) / nchunk;
    a.nchunk = (g.nz + a.kc - 1) / a.kc;
    a.nblocks = tiles * a.nchunk * 4;
    a.halo = halo < 0 ? (m.fuse_halo ? 1 : 0) : halo;
    if constexpr (std::is_same<T, double>::value) {
        if (fuse_tend) {
            PS3D_REQUIRE(itt >= 0, "tendency + friction + AB: itt >= 0");
            ProfScope ps(m, "tend_friction_ab");
            hipLaunchKernelGGL(k_tend_fric_ab, dim3((unsigned)a.nblocks), dim3(FTHREADS), 0, m.stream, m.c, a);
            PS3D_HIP_CHECK(hipGetLastError());
            return;
        }
    }
    PS3D_REQUIRE(!fuse_tend, "tendency + friction + AB: FP64 models only");
    ProfScope ps(m, "friction_ab");
    hipLaunchKernelGGL(k_fric_ab<T>, dim3((unsigned)a.nblocks), dim3(FTHREADS), 0, m.stream, m.c, a);
    PS3D_HIP_CHECK(hipGetLastError());
}

// the halo points of four arrays copied into four others
template <class T>
static void halo_copy(Model &m, const T *const *src, T *const *dst)
{
    HaloCopyArgs<T> h{};
    for (int f = 0; f < 4; ++f) {
        h.src[f] = src[f];
        h.dst[f] = dst[f];
    }
    const int64_t nhalo = m.g.H - m.g.N;
    ProfScope ps(m, "halo");
    hipLaunchKernelGGL(k_halo_copy<T>, dim3((unsigned)((nhalo + 255) / 256), 4), dim3(256), 0, m.stream, m.c, h,
                       nhalo);
    PS3D_HIP_CHECK(hipGetLastError());
}

template <class T>
static void halo_launch(Model &m, const HaloArgsT<T> &h, int n)
{
    const int64_t nhalo = m.g.H - m.g.N;
    ProfScope ps(m, "halo");
    hipLaunchKernelGGL(k_halo<T>, dim3((unsigned)((nhalo + 255) / 256), (unsigned)n), dim3(256), 0, m.stream, m.c,
                       h, nhalo);
    PS3D_HIP_CHECK(hipGetLastError());
}

// integrate.f90:59-111 for configs without vertical biharmonic friction:
// the friction chain and AB of all four fields in one launch, which also
// writes the new values' halo images (the exchange of integrate.f90:106-111).
// The new fields are written to the second buffers, which then become the
// canonical ones.
void launch_friction_ab_mixed(Model &m, int itt, int tau, int taum1, int taum2, double qc1, const float *d1,
                              double qc2, const float *d2, bool swap)
{
    friction_ab_impl<float>(m, m.fh, m.fl, m.ah, m.al, m.tf, itt, tau, taum1, taum2, qc1, d1, qc2, d2);
    if (!swap) return;
    for (int f = 0; f < 4; ++f) {
        std::swap(m.fh[f], m.ah[f]);
        std::swap(m.fl[f], m.al[f]);
    }
    // the stencils and the divergence read the hi parts' halos (written by the
    // kernel itself when fuse_halo); the lo parts are read at interior points
    // only (the mixed tendency wraps its neighbours' indices)
    if (m.fuse_halo) return;
    HaloArgsT<float> h{};
    for (int f = 0; f < 4; ++f) h.f[f] = m.fh[f];
    halo_launch(m, h, 4);
}

// the mixed step with background forcing: friction alone into the second
// buffers (whose halos keep the pre-friction values, as the reference's
// in-place friction leaves them), the forcing on that post-friction state
// into the tau slot, then the Adams-Bashforth pass back into the canonical
// buffers
void launch_friction_forcing_ab_mixed(Model &m, int itt, int tau, int taum1, int taum2, double qc1, const float *d1,
                                      double qc2, const float *d2)
{
    // interior only (halo 0); the hi halos take the pre-friction values, which
    // the forcing reads (the lo parts and the AB pass use interior points only)
    friction_ab_impl<float>(m, m.fh, m.fl, m.ah, m.al, m.tf, -1, tau, taum1, taum2, 0.0, nullptr, 0.0, nullptr, 0);
    halo_copy<float>(m, m.fh, m.ah);
    launch_background(m, tau, m.ah);
    AbMixedArgs a{};
    const ps3d_params &p = m.prm;
    for (int f = 0; f < 4; ++f) {
        a.ph[f] = m.ah[f];
        a.pl[f] = m.al[f];
        a.oh[f] = m.fh[f];
        a.ol[f] = m.fl[f];
        a.t0[f] = m.tf[f] + (int64_t)(tau - 1) * m.g.H;
        a.t1[f] = m.tf[f] + (int64_t)(taum1 - 1) * m.g.H;
        a.t2[f] = m.tf[f] + (int64_t)(taum2 - 1) * m.g.H;
        a.scale[f] = (f == 2) ? p.dt / p.dsqr : p.dt;
        a.gstride[f] = f == 0 ? 1 : (f == 1 ? m.g.X : (f == 2 ? m.g.XY : 0));
        a.grd[f] = f == 0 ? m.c.rdx : (f == 1 ? m.c.rdy : m.c.rdz);
    }
    if (itt == 0) {
        a.order = 1;
    } else if (itt == 1 || !p.enable_AB_3_order) {
        a.order = 2;
        a.ca = 1.5 + p.eps_ab;
        a.cb = 0.5 + p.eps_ab;
    } else {
        a.order = 3;
        a.ca = 23.0 / 12.0;
        a.cb = -16.0 / 12.0;
        a.cc = 5.0 / 12.0;
    }
    a.d1 = d1;
    a.d2 = d2;
    a.qc1 = qc1;
    a.qc2 = qc2;
    a.halo = m.fuse_halo ? 1 : 0;
    const PointShape sh = point_shape(m.g, 4);
    hipLaunchKernelGGL(k_ab_mixed, sh.grid, sh.block, 0, m.stream, m.c, a);
    PS3D_HIP_CHECK(hipGetLastError());
    if (!m.fuse_halo) {
        HaloArgsT<float> h{};
        for (int f = 0; f < 4; ++f) h.f[f] = m.fh[f];
        halo_launch(m, h, 4);
    }
}

// the fused kernel alone, into the second buffers (no swap, no halo): the
// autotuner's timing launch
// integrate.f90:59-78 for a forced FP64 step without vertical biharmonic
// friction: the fused kernel's friction chain alone (no AB) into the second
// buffers, whose halos then take the pre-friction values; the forcing and
// the Adams-Bashforth pass (launch_ab_exchange) read the second buffers
void launch_friction_fused_only(Model &m)
{
    PS3D_REQUIRE(!m.mixed && fused_friction_ok(m), "fused friction: FP64 without vertical biharmonic friction");
    double *in[4] = {m.field_buf(0), m.field_buf(1), m.field_buf(2), m.field_buf(3)};
    double *out[4] = {m.alt[0], m.alt[1], m.alt[2], m.alt[3]};
    double *tend[4] = {m.du, m.dv, m.dw, m.db};
    friction_ab_impl<double>(m, in, nullptr, out, nullptr, tend, -1, 1, 2, 3, 0.0, nullptr, 0.0, nullptr, 0);
    halo_copy<double>(m, in, out);
    for (int f = 0; f < 4; ++f) m.cur[f] = 1;
}

void launch_friction_ab_raw(Model &m, int itt, int tau, int taum1, int taum2, bool with_tendency)
{
    if (m.mixed) {
        launch_friction_ab_mixed(m, itt, tau, taum1, taum2, 0.0, nullptr, 0.0, nullptr, false);
        return;
    }
    double *in[4] = {m.u, m.v, m.w, m.b};
    double *tend[4] = {m.du, m.dv, m.dw, m.db};
    friction_ab_impl<double>(m, in, nullptr, m.alt, nullptr, tend, itt, tau, taum1, taum2, 0.0, nullptr, 0.0,
                             nullptr, -1, with_tendency);
}

void launch_friction_ab(Model &m, int itt, int tau, int taum1, int taum2, bool with_tendency)
{
    PS3D_REQUIRE(!m.mixed, "friction_ab: FP64 models only");
    double *in[4] = {m.u, m.v, m.w, m.b};
    double *tend[4] = {m.du, m.dv, m.dw, m.db};
    friction_ab_impl<double>(m, in, nullptr, m.alt, nullptr, tend, itt, tau, taum1, taum2, 0.0, nullptr, 0.0,
                             nullptr, -1, with_tendency);
    std::swap(m.u, m.alt[0]);
    std::swap(m.v, m.alt[1]);
    std::swap(m.w, m.alt[2]);
    std::swap(m.b, m.alt[3]);
    if (m.fuse_halo) return;
    HaloArgs h{};
    for (int f = 0; f < 4; ++f) { h.f[f] = m.field_buf(f); h.neumann[f] = 0; }
    halo_launch(m, h, 4);
}

void launch_border_exchg(Model &m, double *a, int neumann_z)
{
    HaloArgs h{};
    h.f[0] = a;
    h.neumann[0] = neumann_z;
    const int64_t nhalo = m.g.H - m.g.N;
    ProfScope ps(m, "halo");
    hipLaunchKernelGGL(k_halo<double>, dim3((unsigned)((nhalo + 255) / 256), 1), dim3(256), 0, m.stream, m.c, h, nhalo);
    PS3D_HIP_CHECK(hipGetLastError());
}

// pressure_solver.f90:75-118 after p's interior is in place: the projection
// writes the halo images of u, v, w and p as well (fuse_halo), or an
// exchange pass follows
void launch_project_exchange(Model &m)
{
    {
        ProfScope ps(m, "project");
        const bool rows = m.g.nx >= 256;
        const dim3 grid = rows ? dim3((unsigned)((m.g.nx + 255) / 256), (unsigned)m.g.ny, (unsigned)m.g.nz)
                               : interior_grid(m.g);
        hipLaunchKernelGGL(k_project, grid, rows ? dim3(256, 1) : dim3(BX, BY), 0, m.stream, m.c, m.u, m.v, m.w, m.p,
                           m.fuse_halo ? 1 : 0);
        PS3D_HIP_CHECK(hipGetLastError());
    }
    if (m.fuse_halo) return;
    HaloArgs h{};
    h.f[0] = m.u; h.f[1] = m.v; h.f[2] = m.w; h.f[3] = m.p;
    h.neumann[3] = m.prm.enable_vertical_boundaries;
    halo_launch(m, h, 4);
}

// pressure_solver.f90:98-118 of the mixed form
void launch_project_exchange_f(Model &m, float *dp)
{
    // the increment's halo (Neumann in z under a lid) serves the gradient of
    // the next steps' pressure-history term
    {
        ProfScope ps(m, "project");
        const bool rows = m.g.nx >= 256;
        const dim3 grid = rows ? dim3((unsigned)((m.g.nx + 255) / 256), (unsigned)m.g.ny, (unsigned)m.g.nz)
                               : interior_grid(m.g);
        hipLaunchKernelGGL(k_project_f, grid, rows ? dim3(256, 1) : dim3(BX, BY), 0, m.stream, m.c, m.fh[0], m.fl[0],
                           m.fh[1], m.fl[1], m.fh[2], m.fl[2], dp, m.pint, m.fuse_halo ? 1 : 0);
        PS3D_HIP_CHECK(hipGetLastError());
    }
    if (m.fuse_halo) return;
    HaloArgsT<float> h{};
    h.f[0] = m.fh[0]; h.f[1] = m.fh[1]; h.f[2] = m.fh[2]; h.f[3] = dp;
    h.neumann[3] = m.prm.enable_vertical_boundaries;
    halo_launch(m, h, 4);
    launch_border_exchg(m, m.pint, m.prm.enable_vertical_boundaries);
}

void launch_halo_f(Model &m, float *const *arrays, const int *neumann, int n)
{
    PS3D_REQUIRE(n >= 1 && n <= 8, "halo exchange of 1..8 arrays");
    HaloArgsT<float> h{};
    for (int q = 0; q < n; ++q) { h.f[q] = arrays[q]; h.neumann[q] = neumann ? neumann[q] : 0; }
    halo_launch(m, h, n);
}

void launch_convert(Model &m, int dir, int64_t n, const double *src_d, double *dst_d, const float *src_f,
                    float *dst_f, const float *src_lo, float *dst_lo)
{
    hipLaunchKernelGGL(k_convert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, dir, n, src_d, dst_d,
                       src_f, dst_f, src_lo, dst_lo);
    PS3D_HIP_CHECK(hipGetLastError());
}

// ---- batched ensemble stages ------------------------------------------------
void ens_tendency(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_tendency, dim3((unsigned)map.start[map.n]), dim3(TTHR), 0, s, mem, map, st);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_friction_ab(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_fric, dim3((unsigned)map.start[map.n]), dim3(FTHREADS), 0, s, mem, map, st);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_halo_copy(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, bool mixed)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_halo_copy, dim3((unsigned)map.start[map.n]), dim3(256), 0, s, mem, map, st,
                       mixed ? 1 : 0);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_background_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_background_f, dim3((unsigned)map.start[map.n]), dim3(BX, BY), 0, s, mem, map, st);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_background_ab_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_background_ab_f, dim3((unsigned)map.start[map.n]), dim3(BX, BY), 0, s, mem, map, st);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_background_ab(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_background_ab, dim3((unsigned)map.start[map.n]), dim3(BX, BY), 0, s, mem, map, st);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_project(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_project, dim3((unsigned)map.start[map.n]), dim3(BX, BY), 0, s, mem, map, st);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_advance(hipStream_t s, EnsState *st)
{
    hipLaunchKernelGGL(k_ens_advance, dim3(1), dim3(64), 0, s, st);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_tendency_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int cold)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_tendency_f, dim3((unsigned)map.start[map.n]), dim3(TTHR), 0, s, mem, map, st, cold);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_friction_ab_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int cold)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_fric_f, dim3((unsigned)map.start[map.n]), dim3(FTHREADS), 0, s, mem, map, st, cold);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_project_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int which)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_project_f, dim3((unsigned)map.start[map.n]), dim3(BX, BY), 0, s, mem, map, st, which);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_cold_pint(hipStream_t s, const EnsMember *mem, const EnsMap &map)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_cold_pint, dim3((unsigned)map.start[map.n]), dim3(BX, BY), 0, s, mem, map);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_points_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int what)
{
    if (!map.start[map.n]) return;
    hipLaunchKernelGGL(k_ens_points_f, dim3((unsigned)map.start[map.n]), dim3(256), 0, s, mem, map, st, what);
    PS3D_HIP_CHECK(hipGetLastError());
}

}  // namespace ps3d
