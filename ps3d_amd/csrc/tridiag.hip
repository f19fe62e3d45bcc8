// tridiag.hip -- rigid-lid vertical solve of pressure_solver.f90:30-54.
//
// For every horizontal wavenumber column (kx, ky) with ksqr > 0 the system
// is tridiagonal with a = c = 1/dz^2/dsqr and b = -ksqr - 2a (-ksqr - a at
// k = 1 and k = nz); solve_tridiag (pressure_solver.f90:126-162) eliminates
// it with the Thomas algorithm.  The elimination factors c'(k) and 1/m(k)
// depend only on ksqr, so they are computed once per model, with the same
// operations in the same order as the reference, and each step only runs
// the two sweeps over the complex right-hand side.  The (0,0) column takes
// the reference's double-cumulative-sum recurrence (:44-50).
//
// One thread per column; a wavefront covers 64 consecutive kx, so every
// level of the sweep is one coalesced access.  FS is updated in place: the
// forward sweep's d'(k) are kept in LDS when a work-group's columns fit
// (else left in FS), and the backward sweep writes the solution.
#include "ps3d_internal.h"

#include <type_traits>

#pragma clang fp contract(off)

namespace ps3d {

namespace {

__global__ __launch_bounds__(256) void k_tri_factor(const double *__restrict__ ksqr, double *__restrict__ cp,
                                                    double *__restrict__ rm, int64_t ncols, int nz, double fxa)
{
    const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= ncols) return;
    const double ks = ksqr[col];
    if (!(ks > 0)) return;
    // b(1) = -ksqr - fxa ; cp(1) = c(1)/b(1)
    const double b1 = -ks - fxa;
    double cprev = fxa / b1;
    cp[col] = cprev;
    rm[col] = b1;                 // level 1 keeps b(1): d'(1) = d(1)/b(1) is a division
    for (int k = 2; k <= nz; ++k) {
        const double bk = (k == nz) ? -ks - fxa : -ks - 2 * fxa;
        const double mm = bk - cprev * fxa;
        const double r = 1.0 / mm;
        const double ck = (k == nz) ? 0.0 : fxa * r;   // c(nz) is never used
        cp[col + (int64_t)(k - 1) * ncols] = ck;
        rm[col + (int64_t)(k - 1) * ncols] = r;
        cprev = ck;
    }
}

// one wavefront per work-group: the (nx/2+1)*ny columns are few (33K at
// 256x256), and 64-lane groups spread them over every CU.  IN_LDS: the
// forward sweep's d'(k) go to LDS (dpl[k * ds]) instead of back to FS, so FS
// is read once and written once.
template <bool IN_LDS>
__device__ __forceinline__ void thomas_col(double2 *__restrict__ fs, const double *__restrict__ ksqr,
                                           const double *__restrict__ cp, const double *__restrict__ rm,
                                           int64_t ncols, int nz, double fxa, double dz2, double dsqr, int64_t col,
                                           double2 *dpl = nullptr, int ds = 0)
{
    double2 *d = fs + col;
    const double ks = ksqr[col];
    if (ks > 0) {
        // the sweeps are sequential in k, but their loads are not: fetch
        // LOOK levels at a time so each lane keeps that many requests in flight
        constexpr int LOOK = 16;
        double2 dp = d[0];
        const double b1 = rm[col];
        dp = make_double2(dp.x / b1, dp.y / b1);
        if constexpr (IN_LDS) dpl[0] = dp;
        else d[0] = dp;
        for (int k0 = 1; k0 < nz; k0 += LOOK) {
            double2 dd[LOOK];
            double rr[LOOK];
#pragma unroll
            for (int q = 0; q < LOOK; ++q) {
                if (k0 + q < nz) {
                    const int64_t o = (int64_t)(k0 + q) * ncols;
                    dd[q] = d[o];
                    rr[q] = rm[col + o];
                }
            }
#pragma unroll
            for (int q = 0; q < LOOK; ++q) {
                if (k0 + q < nz) {
                    dp = make_double2((dd[q].x - dp.x * fxa) * rr[q], (dd[q].y - dp.y * fxa) * rr[q]);
                    if constexpr (IN_LDS) dpl[(k0 + q) * ds] = dp;
                    else d[(int64_t)(k0 + q) * ncols] = dp;
                }
            }
        }
        double2 x = dp;
        if constexpr (IN_LDS) d[(int64_t)(nz - 1) * ncols] = x;
        for (int k0 = nz - 2; k0 >= 0; k0 -= LOOK) {
            double2 dd[LOOK];
            double cc[LOOK];
#pragma unroll
            for (int q = 0; q < LOOK; ++q) {
                if (k0 - q >= 0) {
                    const int64_t o = (int64_t)(k0 - q) * ncols;
                    if constexpr (IN_LDS) dd[q] = dpl[(k0 - q) * ds];
                    else dd[q] = d[o];
                    cc[q] = cp[col + o];
                }
            }
#pragma unroll
            for (int q = 0; q < LOOK; ++q) {
                if (k0 - q >= 0) {
                    x = make_double2(dd[q].x - cc[q] * x.x, dd[q].y - cc[q] * x.y);
                    d[(int64_t)(k0 - q) * ncols] = x;
                }
            }
        }
    } else {
        // zero mode: p(1) = 0, p(2) = f(1) dz^2 dsqr, p(k+1) = 2p(k) - p(k-1) + f(k) dz^2 dsqr
        double2 bk = d[0];
        double2 pm1 = make_double2(0.0, 0.0);
        d[0] = pm1;
        if (nz < 2) return;
        double2 nxt = d[ncols];
        double2 pk = make_double2(bk.x * dz2 * dsqr, bk.y * dz2 * dsqr);
        d[ncols] = pk;
        bk = nxt;
        for (int k = 2; k < nz; ++k) {
            const int64_t o = (int64_t)k * ncols;
            nxt = d[o];
            const double2 pn = make_double2(2 * pk.x - pm1.x + bk.x * dz2 * dsqr,
                                            2 * pk.y - pm1.y + bk.y * dz2 * dsqr);
            d[o] = pn;
            pm1 = pk;
            pk = pn;
            bk = nxt;
        }
    }
}

// cpw columns per 64-lane work-group: with few columns (33K at 256x256) a
// full wavefront per 64 columns leaves SIMDs idle and the sequential sweeps'
// latency exposed; fewer columns per wave give more waves to interleave
template <bool IN_LDS>
__global__ __launch_bounds__(64) void k_thomas(double2 *__restrict__ fs, const double *__restrict__ ksqr,
                                               const double *__restrict__ cp, const double *__restrict__ rm,
                                               int64_t ncols, int nz, double fxa, double dz2, double dsqr, int cpw)
{
    extern __shared__ double2 thomas_lds[];   // cpw x nz when IN_LDS
    if ((int)threadIdx.x >= cpw) return;
    const int64_t col = (int64_t)xcd_remap((int)blockIdx.x, (int)gridDim.x) * cpw + threadIdx.x;
    if (col >= ncols) return;
    thomas_col<IN_LDS>(fs, ksqr, cp, rm, ncols, nz, fxa, dz2, dsqr, col, thomas_lds + threadIdx.x, cpw);
}

// The sweeps with the elimination factors formed on the fly, by the same
// operations as k_tri_factor (so the same bits), d'(k) and c'(k) kept in
// LDS (dpl, cpl: stride ds): FS is read once and written once and no factor
// table is read.  The factor recurrence (one division per level) is the
// sweep's critical path; the FS loads are independent of it and go LOOK
// levels ahead.  C = double2: the FP64 path; C = float2: the mixed form (FP32
// right-hand side and solution, FP64 elimination).
__device__ __forceinline__ double2 to_d2(double2 a) { return a; }
__device__ __forceinline__ double2 to_d2(float2 a) { return make_double2((double)a.x, (double)a.y); }
template <class C> __device__ __forceinline__ C from_d2(double2 a);
template <> __device__ __forceinline__ double2 from_d2<double2>(double2 a) { return a; }
template <> __device__ __forceinline__ float2 from_d2<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }

template <class C>
__device__ __forceinline__ void thomas_col_fly(C *__restrict__ fs, const double *__restrict__ ksqr, int64_t ncols,
                                               int nz, double fxa, double dz2, double dsqr, int64_t col,
                                               double2 *dpl, double *cpl, int ds)
{
    C *d = fs + col;
    const double ks = ksqr[col];
    if (ks > 0) {
        constexpr int LOOK = 16;
        const double b1 = -ks - fxa;
        double cprev = fxa / b1;
        cpl[0] = cprev;
        double2 dp = to_d2(d[0]);
        dp = make_double2(dp.x / b1, dp.y / b1);
        dpl[0] = dp;
        for (int k0 = 1; k0 < nz; k0 += LOOK) {
            C dd[LOOK];
#pragma unroll
            for (int q = 0; q < LOOK; ++q)
                if (k0 + q < nz) dd[q] = d[(int64_t)(k0 + q) * ncols];
#pragma unroll
            for (int q = 0; q < LOOK; ++q) {
                if (k0 + q < nz) {
                    const int k = k0 + q + 1;   // level, 1-based
                    const double bk = (k == nz) ? -ks - fxa : -ks - 2 * fxa;
                    const double mm = bk - cprev * fxa;
                    const double r = 1.0 / mm;
                    const double ck = (k == nz) ? 0.0 : fxa * r;
                    cpl[(k0 + q) * ds] = ck;
                    cprev = ck;
                    const double2 e = to_d2(dd[q]);
                    dp = make_double2((e.x - dp.x * fxa) * r, (e.y - dp.y * fxa) * r);
                    dpl[(k0 + q) * ds] = dp;
                }
            }
        }
        double2 x = dp;
        d[(int64_t)(nz - 1) * ncols] = from_d2<C>(x);
        for (int k = nz - 2; k >= 0; --k) {
            const double2 e = dpl[k * ds];
            const double c = cpl[k * ds];
            x = make_double2(e.x - c * x.x, e.y - c * x.y);
            d[(int64_t)k * ncols] = from_d2<C>(x);
        }
    } else {
        // zero mode, pressure_solver.f90:44-50, as in thomas_col (FP64 sums)
        double2 bk = to_d2(d[0]);
        double2 pm1 = make_double2(0.0, 0.0);
        d[0] = from_d2<C>(pm1);
        double2 nxt = to_d2(d[ncols]);
        double2 pk = make_double2(bk.x * dz2 * dsqr, bk.y * dz2 * dsqr);
        d[ncols] = from_d2<C>(pk);
        bk = nxt;
        for (int k = 2; k < nz; ++k) {
            const int64_t o = (int64_t)k * ncols;
            nxt = to_d2(d[o]);
            const double2 pn = make_double2(2 * pk.x - pm1.x + bk.x * dz2 * dsqr,
                                            2 * pk.y - pm1.y + bk.y * dz2 * dsqr);
            d[o] = from_d2<C>(pn);
            pm1 = pk;
            pk = pn;
            bk = nxt;
        }
    }
}

template <class C>
__global__ __launch_bounds__(64) void k_thomas_fly(C *__restrict__ fs, const double *__restrict__ ksqr,
                                                   int64_t ncols, int nz, double fxa, double dz2, double dsqr,
                                                   int cpw)
{
    extern __shared__ double thomas_fly_lds[];   // cpw x nz double2 (d'), then cpw x nz double (c')
    if ((int)threadIdx.x >= cpw) return;
    const int64_t col = (int64_t)xcd_remap((int)blockIdx.x, (int)gridDim.x) * cpw + threadIdx.x;
    if (col >= ncols) return;
    double2 *dpl = reinterpret_cast<double2 *>(thomas_fly_lds) + threadIdx.x;
    double *cpl = thomas_fly_lds + 2 * (int64_t)cpw * nz + threadIdx.x;
    thomas_col_fly<C>(fs, ksqr, ncols, nz, fxa, dz2, dsqr, col, dpl, cpl, cpw);
}

// The sweeps with whole columns in registers, NZ levels known at compile
// time: one lane per column, 64 columns per wave.  Each lane loads its
// column's NZ values at once (all loads in flight before the first is used),
// keeps d'(k) in registers and c'(k) in LDS (NZ x 64 doubles), and stores
// the solution.  With one wave per work-group and little LDS, every wave of
// the solve is resident at once, so the sequential sweeps of all columns
// run concurrently.  The arithmetic is that of thomas_col_fly, operation for
// operation: the same bits.  C = double2: the FP64 path; C = float2: the
// mixed form (FP32 right-hand side and solution, FP64 elimination).
// cpl: NZ x 64 doubles of LDS; columns col (this lane's) of ncols.
template <int NZ, class C>
__device__ __forceinline__ void thomas_reg_col(C *__restrict__ fs, const double *__restrict__ ksqr, int64_t ncols,
                                               int64_t col, double fxa, double dz2, double dsqr, double *cpl)
{
    double *cpk = cpl + threadIdx.x;   // c'(level q) at cpk[64 q]
    // 32-bit byte offsets from the array base (the launchers check that FS
    // is below 4 GiB), so each access is one base register pair plus one
    // offset register
    char *base = reinterpret_cast<char *>(fs);
    constexpr uint32_t E = sizeof(C);
    const uint32_t o0 = (uint32_t)col * E, st = (uint32_t)ncols * E;
    double2 d[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) d[k] = to_d2(*reinterpret_cast<const C *>(base + (o0 + (uint32_t)k * st)));
    const double ks = ksqr[col];
    if (ks > 0) {
        const double b1 = -ks - fxa;
        double cprev = fxa / b1;
        cpk[0] = cprev;
        double2 dp = make_double2(d[0].x / b1, d[0].y / b1);
        d[0] = dp;
#pragma unroll
        for (int q = 1; q < NZ; ++q) {
            const int k = q + 1;   // level, 1-based
            const double bk = (k == NZ) ? -ks - fxa : -ks - 2 * fxa;
            const double mm = bk - cprev * fxa;
            const double r = 1.0 / mm;
            const double ck = (k == NZ) ? 0.0 : fxa * r;
            cpk[64 * q] = ck;
            cprev = ck;
            const double2 e = d[q];
            dp = make_double2((e.x - dp.x * fxa) * r, (e.y - dp.y * fxa) * r);
            d[q] = dp;
        }
        double2 x = dp;
        // c'(k) is read back from LDS (not kept in registers across the
        // sweep); one wave per work-group, so the barrier holds nobody up
        __syncthreads();
#pragma unroll
        for (int q = NZ - 2; q >= 0; --q) {
            const double2 e = d[q];
            const double c = cpk[64 * q];
            x = make_double2(e.x - c * x.x, e.y - c * x.y);
            d[q] = x;
        }
    } else {
        // zero mode, pressure_solver.f90:44-50, as in thomas_col_fly
        double2 pm1 = make_double2(0.0, 0.0);
        double2 pk = make_double2(d[0].x * dz2 * dsqr, d[0].y * dz2 * dsqr);
        double2 bk = d[1];
        d[0] = pm1;
        d[1] = pk;
#pragma unroll
        for (int k = 2; k < NZ; ++k) {
            const double2 nxt = d[k];
            const double2 pn = make_double2(2 * pk.x - pm1.x + bk.x * dz2 * dsqr,
                                            2 * pk.y - pm1.y + bk.y * dz2 * dsqr);
            d[k] = pn;
            pm1 = pk;
            pk = pn;
            bk = nxt;
        }
    }
    // the store offsets are formed here, one after the other (the empty asm
    // makes each one opaque), rather than computed with the loads' and held
    // in registers through both sweeps
    uint32_t off = o0;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        asm volatile("" : "+v"(off));
        *reinterpret_cast<C *>(base + off) = from_d2<C>(d[k]);
        off += st;
    }
}

template <int NZ, class C>
__global__ __launch_bounds__(64, 1) void k_thomas_reg(C *__restrict__ fs, const double *__restrict__ ksqr,
                                                      int64_t ncols, double fxa, double dz2, double dsqr)
{
    __shared__ double cpl[NZ * 64];
    const int64_t col = (int64_t)xcd_remap((int)blockIdx.x, (int)gridDim.x) * 64 + threadIdx.x;
    if (col >= ncols) return;
    thomas_reg_col<NZ, C>(fs, ksqr, ncols, col, fxa, dz2, dsqr, cpl);
}

// Taller columns in registers: the real and imaginary parts of a column on
// two neighbouring lanes (32 columns per wave), so a lane keeps NZ values of
// d' instead of 2 NZ and columns of 128 levels still fit its registers.  Both
// lanes of a column run its c' recurrence (the same operations on the same
// operands, so the same bits; the even lane writes c' to LDS) and each runs
// the d' sweeps of its part with the operations thomas_reg_col applies to
// that component: the same bits as every other form.  R = double: the FP64
// path; R = float: the mixed form (FP32 right-hand side and solution, FP64
// elimination).  cpl: NZ x 32 doubles of LDS.
template <int NZ, class R>
__device__ __forceinline__ void thomas_split_col(R *__restrict__ fs, const double *__restrict__ ksqr, int64_t ncols,
                                                 int64_t col, double fxa, double dz2, double dsqr, double *cpl)
{
    const int part = threadIdx.x & 1;
    double *cpk = cpl + (threadIdx.x >> 1);   // c'(level q) at cpk[32 q]
    // component `part` of column col at level k is element 2 (k ncols + col)
    // + part; 32-bit byte offsets as in thomas_reg_col
    char *base = reinterpret_cast<char *>(fs);
    constexpr uint32_t E = sizeof(R);
    const uint32_t o0 = ((uint32_t)col * 2 + (uint32_t)part) * E, st = (uint32_t)ncols * 2 * E;
    double d[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) d[k] = (double)*reinterpret_cast<const R *>(base + (o0 + (uint32_t)k * st));
    const double ks = ksqr[col];
    const bool solve = ks > 0;
    double x = 0.0;
    if (solve) {
        const double b1 = -ks - fxa;
        double cprev = fxa / b1;
        if (part == 0) cpk[0] = cprev;
        double dp = d[0] / b1;
        d[0] = dp;
#pragma unroll
        for (int q = 1; q < NZ; ++q) {
            const int k = q + 1;   // level, 1-based
            const double bk = (k == NZ) ? -ks - fxa : -ks - 2 * fxa;
            const double mm = bk - cprev * fxa;
            const double r = 1.0 / mm;
            const double ck = (k == NZ) ? 0.0 : fxa * r;
            if (part == 0) cpk[32 * q] = ck;
            cprev = ck;
            dp = (d[q] - dp * fxa) * r;
            d[q] = dp;
        }
        x = dp;
    }
    // the odd lanes read back the c' their even neighbours wrote.  The
    // work-group is one wave: the barrier orders that wave's LDS accesses,
    // and lanes of columns past the last one, which returned before the
    // call, take no part (predicating their loads instead, so that every
    // lane reaches it, measured 3.6 times slower with FP32 data)
    __syncthreads();
    if (solve) {
#pragma unroll
        for (int q = NZ - 2; q >= 0; --q) {
            x = d[q] - cpk[32 * q] * x;
            d[q] = x;
        }
    } else {
        // zero mode, pressure_solver.f90:44-50, as in thomas_reg_col
        double pm1 = 0.0;
        double pk = d[0] * dz2 * dsqr;
        double bk = d[1];
        d[0] = pm1;
        d[1] = pk;
#pragma unroll
        for (int k = 2; k < NZ; ++k) {
            const double nxt = d[k];
            const double pn = 2 * pk - pm1 + bk * dz2 * dsqr;
            d[k] = pn;
            pm1 = pk;
            pk = pn;
            bk = nxt;
        }
    }
    uint32_t off = o0;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        asm volatile("" : "+v"(off));
        *reinterpret_cast<R *>(base + off) = (R)d[k];
        off += st;
    }
}

template <int NZ, class R>
__global__ __launch_bounds__(64, 1) void k_thomas_split(R *__restrict__ fs, const double *__restrict__ ksqr,
                                                        int64_t ncols, double fxa, double dz2, double dsqr)
{
    __shared__ double cpl[NZ * 32];
    const int64_t col = (int64_t)xcd_remap((int)blockIdx.x, (int)gridDim.x) * 32 + (threadIdx.x >> 1);
    if (col >= ncols) return;
    thomas_split_col<NZ, R>(fs, ksqr, ncols, col, fxa, dz2, dsqr, cpl);
}

// the same for the rigid-lid members of a batched ensemble, 64 columns per
// work-group; every member's nz is one of the compiled-in lengths
// (thomas_reg_ok), selected at run time
template <class C>
__global__ __launch_bounds__(64, 1) void k_ens_thomas_reg(const EnsMember *__restrict__ mem, EnsMap map)
{
    __shared__ double cpl[64 * 64];
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const int64_t ncols = (int64_t)e.nkx * e.c.ny;
    const int64_t col = (int64_t)(g - map.start[mi]) * 64 + threadIdx.x;
    if (col >= ncols) return;
    C *fs;
    if constexpr (std::is_same<C, float2>::value) fs = e.FSf;
    else fs = e.FS;
    const double fxa = e.fxa, dz2 = e.c.dz * e.c.dz, dsqr = e.c.dsqr;
    switch (e.c.nz) {
    case 16: thomas_reg_col<16, C>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    case 32: thomas_reg_col<32, C>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    case 48: thomas_reg_col<48, C>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    case 64: thomas_reg_col<64, C>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    default: break;
    }
}

// the split form for the rigid-lid members of a batched ensemble, 32
// columns per work-group, where every member's nz is one of its compiled-in
// lengths (thomas_split_ok) and some is too tall for k_ens_thomas_reg
template <class C>
__global__ __launch_bounds__(64, 1) void k_ens_thomas_split(const EnsMember *__restrict__ mem, EnsMap map)
{
    __shared__ double cpl[128 * 32];
    using R = typename std::conditional<std::is_same<C, float2>::value, float, double>::type;
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const int64_t ncols = (int64_t)e.nkx * e.c.ny;
    const int64_t col = (int64_t)(g - map.start[mi]) * 32 + (threadIdx.x >> 1);
    if (col >= ncols) return;
    R *fs;
    if constexpr (std::is_same<C, float2>::value) fs = reinterpret_cast<R *>(e.FSf);
    else fs = reinterpret_cast<R *>(e.FS);
    const double fxa = e.fxa, dz2 = e.c.dz * e.c.dz, dsqr = e.c.dsqr;
    switch (e.c.nz) {
    case 16: thomas_split_col<16, R>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    case 32: thomas_split_col<32, R>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    case 48: thomas_split_col<48, R>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    case 64: thomas_split_col<64, R>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    case 96: thomas_split_col<96, R>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    case 128: thomas_split_col<128, R>(fs, e.ksqr, ncols, col, fxa, dz2, dsqr, cpl); break;
    default: break;
    }
}

// the Thomas stage of a batched ensemble (rigid-lid members), ENS_TC columns
// per 64-lane work-group; FLY: the on-the-fly form with LDS for ENS_TC
// columns of nz_max levels, else the factor tables
constexpr int ENS_TC = 32;

template <bool FLY>
__global__ __launch_bounds__(64) void k_ens_thomas(const EnsMember *__restrict__ mem, EnsMap map, int nz_max)
{
    extern __shared__ double ens_thomas_lds[];   // FLY: ENS_TC x nz_max double2, then ENS_TC x nz_max double
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    if ((int)threadIdx.x >= ENS_TC) return;
    const int64_t ncols = (int64_t)e.nkx * e.c.ny;
    const int64_t col = (int64_t)(g - map.start[mi]) * ENS_TC + threadIdx.x;
    if (col >= ncols) return;
    if constexpr (FLY) {
        double2 *dpl = reinterpret_cast<double2 *>(ens_thomas_lds) + threadIdx.x;
        double *cpl = ens_thomas_lds + 2 * (int64_t)ENS_TC * nz_max + threadIdx.x;
        thomas_col_fly<double2>(e.FS, e.ksqr, ncols, e.c.nz, e.fxa, e.c.dz * e.c.dz, e.c.dsqr, col, dpl, cpl,
                                ENS_TC);
    } else {
        thomas_col<false>(e.FS, e.ksqr, e.cp, e.rm, ncols, e.c.nz, e.fxa, e.c.dz * e.c.dz, e.c.dsqr, col);
    }
}

// the Thomas stage of a batched ensemble of mixed-form members: FP32
// right-hand side and solution, FP64 elimination on the fly (the body of
// k_thomas_fly<float2>)
__global__ __launch_bounds__(64) void k_ens_thomas_f(const EnsMember *__restrict__ mem, EnsMap map, int nz_max)
{
    extern __shared__ double ens_thomas_lds_f[];   // ENS_TC x nz_max double2, then ENS_TC x nz_max double
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    if ((int)threadIdx.x >= ENS_TC) return;
    const int64_t ncols = (int64_t)e.nkx * e.c.ny;
    const int64_t col = (int64_t)(g - map.start[mi]) * ENS_TC + threadIdx.x;
    if (col >= ncols) return;
    double2 *dpl = reinterpret_cast<double2 *>(ens_thomas_lds_f) + threadIdx.x;
    double *cpl = ens_thomas_lds_f + 2 * (int64_t)ENS_TC * nz_max + threadIdx.x;
    thomas_col_fly<float2>(e.FSf, e.ksqr, ncols, e.c.nz, e.fxa, e.c.dz * e.c.dz, e.c.dsqr, col, dpl, cpl, ENS_TC);
}

// periodic spectral step without the z transforms (building-block API)
__global__ __launch_bounds__(256) void k_divide_ksqr(double2 *__restrict__ fs, const double *__restrict__ ksqr,
                                                     int64_t S)
{
    const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= S) return;
    const double ks = ksqr[n];
    double2 x = fs[n];
    if (ks > 0.) x = make_double2(-(x.x / ks), -(x.y / ks));
    else x = make_double2(0.0, 0.0);
    fs[n] = x;
}

// Mixed form for columns too tall for the LDS of k_thomas_fly<float2>: the
// right-hand side and the solution are FP32, the elimination runs in FP64
// with the factor tables, and the forward-sweep values d'(k) stay in FP64 in
// the dp64 scratch array.
__global__ __launch_bounds__(64) void k_thomas_f(float2 *__restrict__ fs, double2 *__restrict__ dpg,
                                                 const double *__restrict__ ksqr, const double *__restrict__ cp,
                                                 const double *__restrict__ rm, int64_t ncols, int nz, double fxa,
                                                 double dz2, double dsqr)
{
    const int64_t col = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (col >= ncols) return;
    float2 *d = fs + col;
    double2 *dp_ = dpg + col;
    const int64_t dstr = ncols;
    const double ks = ksqr[col];
    if (ks > 0) {
        constexpr int LOOK = 16;
        const float2 d0 = d[0];
        const double b1 = rm[col];
        double2 dp = make_double2((double)d0.x / b1, (double)d0.y / b1);
        dp_[0] = dp;
        for (int k0 = 1; k0 < nz; k0 += LOOK) {
            float2 dd[LOOK];
            double rr[LOOK];
#pragma unroll
            for (int q = 0; q < LOOK; ++q) {
                if (k0 + q < nz) {
                    const int64_t o = (int64_t)(k0 + q) * ncols;
                    dd[q] = d[o];
                    rr[q] = rm[col + o];
                }
            }
#pragma unroll
            for (int q = 0; q < LOOK; ++q) {
                if (k0 + q < nz) {
                    dp = make_double2(((double)dd[q].x - dp.x * fxa) * rr[q], ((double)dd[q].y - dp.y * fxa) * rr[q]);
                    dp_[(int64_t)(k0 + q) * dstr] = dp;
                }
            }
        }
        double2 x = dp;
        d[(int64_t)(nz - 1) * ncols] = make_float2((float)x.x, (float)x.y);
        for (int k0 = nz - 2; k0 >= 0; k0 -= LOOK) {
            double cc[LOOK];
#pragma unroll
            for (int q = 0; q < LOOK; ++q)
                if (k0 - q >= 0) cc[q] = cp[col + (int64_t)(k0 - q) * ncols];
#pragma unroll
            for (int q = 0; q < LOOK; ++q) {
                if (k0 - q >= 0) {
                    const double2 e = dp_[(int64_t)(k0 - q) * dstr];
                    x = make_double2(e.x - cc[q] * x.x, e.y - cc[q] * x.y);
                    d[(int64_t)(k0 - q) * ncols] = make_float2((float)x.x, (float)x.y);
                }
            }
        }
    } else {
        // zero mode, pressure_solver.f90:44-50, accumulated in FP64
        float2 bk = d[0];
        double2 pm1 = make_double2(0.0, 0.0);
        d[0] = make_float2(0.f, 0.f);
        float2 nxt = d[ncols];
        double2 pk = make_double2((double)bk.x * dz2 * dsqr, (double)bk.y * dz2 * dsqr);
        d[ncols] = make_float2((float)pk.x, (float)pk.y);
        bk = nxt;
        for (int k = 2; k < nz; ++k) {
            const int64_t o = (int64_t)k * ncols;
            nxt = d[o];
            const double2 pn = make_double2(2 * pk.x - pm1.x + (double)bk.x * dz2 * dsqr,
                                            2 * pk.y - pm1.y + (double)bk.y * dz2 * dsqr);
            d[o] = make_float2((float)pn.x, (float)pn.y);
            pm1 = pk;
            pk = pn;
            bk = nxt;
        }
    }
}

}  // namespace

// LDS of the on-the-fly form for cpw columns of nz levels, and the most
// one work-group may take (gfx950: 160 KiB per CU; above 64 KiB the kernel
// must ask for it)
static size_t thomas_fly_lds(int cpw, int nz) { return (sizeof(double2) + sizeof(double)) * (size_t)cpw * nz; }
constexpr size_t kLdsMax = 160 * 1024;

template <class K>
static void allow_lds(K kernel, size_t bytes)
{
    if (bytes > 64 * 1024)
        PS3D_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

bool thomas_reg_ok(int nz, uint64_t fs_bytes)
{
    return (nz == 16 || nz == 32 || nz == 48 || nz == 64) && fs_bytes < (uint64_t(1) << 32);
}

bool thomas_split_ok(int nz, uint64_t fs_bytes)
{
    return (nz == 16 || nz == 32 || nz == 48 || nz == 64 || nz == 96 || nz == 128) &&
           fs_bytes < (uint64_t(1) << 32);
}

// the split register form (k_thomas_split) over data of element type R
template <class R>
static void launch_thomas_split(Model &m, R *data)
{
    const int64_t ncols = (int64_t)m.g.nkx * m.g.ny;
    const double fxa = 1.0 / (m.prm.dz * m.prm.dz) / m.prm.dsqr;
    const double dz2 = m.prm.dz * m.prm.dz, dsqr = m.prm.dsqr;
    const dim3 grid((unsigned)((ncols + 31) / 32));
    switch (m.g.nz) {
#define CASE(n)                                                                                                 \
    case n:                                                                                                     \
        hipLaunchKernelGGL((k_thomas_split<n, R>), grid, dim3(64), 0, m.stream, data, m.ksqr, ncols, fxa, dz2, dsqr); \
        break;
        CASE(16) CASE(32) CASE(48) CASE(64) CASE(96) CASE(128)
#undef CASE
    default: throw Error(PS3D_ERR_ARG, "split Thomas form: no compiled-in column length " + std::to_string(m.g.nz));
    }
    PS3D_HIP_CHECK(hipGetLastError());
}

void tridiag_factor(Model &m)
{
    if (!m.prm.enable_vertical_boundaries) return;
    const int64_t ncols = (int64_t)m.g.nkx * m.g.ny;
    const double fxa = 1.0 / (m.prm.dz * m.prm.dz) / m.prm.dsqr;
    hipLaunchKernelGGL(k_tri_factor, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, m.stream, m.ksqr,
                       m.tri_cp, m.tri_rm, ncols, m.g.nz, fxa);
    PS3D_HIP_CHECK(hipGetLastError());
}

void launch_thomas(Model &m, double2 *data)
{
    const int64_t ncols = (int64_t)m.g.nkx * m.g.ny;
    const double fxa = 1.0 / (m.prm.dz * m.prm.dz) / m.prm.dsqr;
    const int form = m.thomas_cpw;
    const uint64_t fsb = (uint64_t)m.g.S * sizeof(double2);
    const bool reg = thomas_reg_ok(m.g.nz, fsb) && (form == 1 || (form == 0 && m.thomas_reg));
    // form 2, or untuned where the whole-column form has no length: the
    // real and imaginary parts on two lanes
    if (!reg && thomas_split_ok(m.g.nz, fsb) && (form == 2 || (form == 0 && m.thomas_reg))) {
        ProfScope ps(m, "thomas");
        launch_thomas_split<double>(m, reinterpret_cast<double *>(data));
        return;
    }
    if (reg) {
        const dim3 grid((unsigned)((ncols + 63) / 64));
        const double dz2 = m.prm.dz * m.prm.dz, dsqr = m.prm.dsqr;
        ProfScope ps(m, "thomas");
        switch (m.g.nz) {
#define CASE(n)                                                                                             \
    case n:                                                                                                 \
        hipLaunchKernelGGL((k_thomas_reg<n, double2>), grid, dim3(64), 0, m.stream, data, m.ksqr, ncols, fxa, dz2, \
                           dsqr);                                                                           \
        break;
            CASE(16) CASE(32) CASE(48) CASE(64)
#undef CASE
        }
        PS3D_HIP_CHECK(hipGetLastError());
        return;
    }
    // form: 16, 32, 64 columns per wave on the fly (c' and d' in LDS);
    // -16, -32, -64 the factor tables (d' in LDS where it fits 64 KiB, else
    // in FS); 0 (untuned): on the fly with 64 or 32 columns where that
    // keeps several work-groups on a CU (LDS <= 64 KiB), else the tables
    // with 64 columns (measured faster at nz = 128 than 16 columns on the
    // fly: a quarter of the lanes would idle)
    int cpw;
    bool table;
    if (form == 16 || form == 32 || form == 64) {
        cpw = form;
        table = thomas_fly_lds(cpw, m.g.nz) > kLdsMax;
    } else if (form == -16 || form == -32 || form == -64) {
        cpw = -form;
        table = true;
    } else {
        cpw = 64;
        while (cpw > 32 && thomas_fly_lds(cpw, m.g.nz) > 64 * 1024) cpw /= 2;
        table = thomas_fly_lds(cpw, m.g.nz) > 64 * 1024;
        if (table) cpw = 64;
    }
    const size_t lds = sizeof(double2) * (size_t)cpw * m.g.nz;
    const size_t lds_fly = thomas_fly_lds(cpw, m.g.nz);
    const dim3 grid((unsigned)((ncols + cpw - 1) / cpw));
    ProfScope ps(m, "thomas");
    if (!table) {
        allow_lds(k_thomas_fly<double2>, lds_fly);
        hipLaunchKernelGGL(k_thomas_fly<double2>, grid, dim3(64), lds_fly, m.stream, data, m.ksqr, ncols, m.g.nz,
                           fxa, m.prm.dz * m.prm.dz, m.prm.dsqr, cpw);
    } else if (lds <= 64 * 1024)
        hipLaunchKernelGGL(k_thomas<true>, grid, dim3(64), lds, m.stream, data, m.ksqr, m.tri_cp, m.tri_rm, ncols,
                           m.g.nz, fxa, m.prm.dz * m.prm.dz, m.prm.dsqr, cpw);
    else
        hipLaunchKernelGGL(k_thomas<false>, grid, dim3(64), 0, m.stream, data, m.ksqr, m.tri_cp, m.tri_rm, ncols,
                           m.g.nz, fxa, m.prm.dz * m.prm.dz, m.prm.dsqr, cpw);
    PS3D_HIP_CHECK(hipGetLastError());
}

void launch_divide_ksqr(Model &m, double2 *data)
{
    hipLaunchKernelGGL(k_divide_ksqr, dim3((unsigned)((m.g.S + 255) / 256)), dim3(256), 0, m.stream, data,
                       m.ksqr, m.g.S);
    PS3D_HIP_CHECK(hipGetLastError());
}

void launch_thomas_f(Model &m, float2 *data)
{
    const int64_t ncols = (int64_t)m.g.nkx * m.g.ny;
    const double fxa = 1.0 / (m.prm.dz * m.prm.dz) / m.prm.dsqr;
    const unsigned nb = (unsigned)((ncols + 63) / 64);
    ProfScope ps(m, "thomas");
    // forms as in launch_thomas: 1 whole columns in registers, 2 the parts
    // on two lanes, 32 on the fly, -64 the factor tables; 0 (untuned): the
    // first of these that applies
    const int form = m.thomas_cpw;
    const uint64_t fsb = (uint64_t)m.g.S * sizeof(float2);
    const bool reg = thomas_reg_ok(m.g.nz, fsb) && (form == 1 || (form == 0 && m.thomas_reg));
    if (!reg && thomas_split_ok(m.g.nz, fsb) && (form == 2 || (form == 0 && m.thomas_reg))) {
        launch_thomas_split<float>(m, reinterpret_cast<float *>(data));
        return;
    }
    if (reg) {
        const dim3 grid((unsigned)((ncols + 63) / 64));
        const double dz2 = m.prm.dz * m.prm.dz, dsqr = m.prm.dsqr;
        switch (m.g.nz) {
#define CASE(n)                                                                                             \
    case n:                                                                                                 \
        hipLaunchKernelGGL((k_thomas_reg<n, float2>), grid, dim3(64), 0, m.stream, data, m.ksqr, ncols, fxa, dz2,  \
                           dsqr);                                                                           \
        break;
            CASE(16) CASE(32) CASE(48) CASE(64)
#undef CASE
        }
        PS3D_HIP_CHECK(hipGetLastError());
        return;
    }
    // on the fly with 32 columns per wave where its LDS keeps several
    // work-groups on a CU, else the factor tables with FP64 d' in a scratch
    // array (faster than 16 columns on the fly for tall columns)
    constexpr int cpw = 32;
    const size_t lds_fly = thomas_fly_lds(cpw, m.g.nz);
    bool fly = lds_fly <= 64 * 1024;
    if (form == 32) fly = lds_fly <= kLdsMax;
    else if (form == -64 && m.dp64) fly = false;
    if (fly) {
        allow_lds(k_thomas_fly<float2>, lds_fly);
        hipLaunchKernelGGL(k_thomas_fly<float2>, dim3((unsigned)((ncols + cpw - 1) / cpw)), dim3(64), lds_fly,
                           m.stream, data, m.ksqr, ncols, m.g.nz, fxa, m.prm.dz * m.prm.dz, m.prm.dsqr, cpw);
    } else {
        PS3D_REQUIRE(m.dp64, "Thomas scratch not allocated");
        hipLaunchKernelGGL(k_thomas_f, dim3(nb), dim3(64), 0, m.stream, data, m.dp64, m.ksqr, m.tri_cp, m.tri_rm,
                           ncols, m.g.nz, fxa, m.prm.dz * m.prm.dz, m.prm.dsqr);
    }
    PS3D_HIP_CHECK(hipGetLastError());
}

int ens_thomas_cols(bool reg, bool split) { return reg ? 64 : (split ? 32 : ENS_TC); }

void ens_thomas_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nz_max, bool reg, bool split)
{
    if (!map.start[map.n]) return;
    if (reg || split) {
        if (reg) hipLaunchKernelGGL(k_ens_thomas_reg<float2>, dim3((unsigned)map.start[map.n]), dim3(64), 0, s, mem, map);
        else hipLaunchKernelGGL(k_ens_thomas_split<float2>, dim3((unsigned)map.start[map.n]), dim3(64), 0, s, mem, map);
        PS3D_HIP_CHECK(hipGetLastError());
        return;
    }
    const size_t lds = thomas_fly_lds(ENS_TC, nz_max);
    PS3D_REQUIRE(lds <= kLdsMax, "mixed ensemble: rigid-lid members of at most 213 levels");
    allow_lds(k_ens_thomas_f, lds);
    hipLaunchKernelGGL(k_ens_thomas_f, dim3((unsigned)map.start[map.n]), dim3(64), lds, s, mem, map, nz_max);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_thomas(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nz_max, bool reg, bool split)
{
    if (!map.start[map.n]) return;
    if (reg || split) {
        if (reg) hipLaunchKernelGGL(k_ens_thomas_reg<double2>, dim3((unsigned)map.start[map.n]), dim3(64), 0, s, mem, map);
        else hipLaunchKernelGGL(k_ens_thomas_split<double2>, dim3((unsigned)map.start[map.n]), dim3(64), 0, s, mem, map);
        PS3D_HIP_CHECK(hipGetLastError());
        return;
    }
    const size_t lds = thomas_fly_lds(ENS_TC, nz_max);
    if (lds <= kLdsMax) {
        allow_lds(k_ens_thomas<true>, lds);
        hipLaunchKernelGGL(k_ens_thomas<true>, dim3((unsigned)map.start[map.n]), dim3(64), lds, s, mem, map, nz_max);
    } else {
        hipLaunchKernelGGL(k_ens_thomas<false>, dim3((unsigned)map.start[map.n]), dim3(64), 0, s, mem, map, nz_max);
    }
    PS3D_HIP_CHECK(hipGetLastError());
}

}  // namespace ps3d
