// balance.hip -- normal-mode balancing diagnostic (Warn et al., to third
// order) of the model's state, in FP64 on the device: diag_balance.f90:89-362
// with the discrete normal modes of eigenvectors.f90:6-218 and the
// vertical-mode transforms of my_fft.f90:112-127, :211-252.
//
// The spectral work is point-wise: each thread evaluates the eigenvectors
// (vec_q, vec_p, omega2) of its own (kx, ky, kz) on the fly and applies one
// step of the diag_balance algebra, so no per-point eigen tables are stored.
// The vertical-mode transform under a rigid lid is a dense nz x nz matrix
// applied to each (kx, ky) column; one thread forms one output level of one
// column, summing in the reference's order (wavefronts run along kx, so the
// column loads are coalesced and the matrix entry is uniform).
#include "ps3d_internal.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#pragma clang fp contract(off)

namespace ps3d {

// the arrays of module_diag_balance (diag_balance.f90:1-20) and scratch
struct BalanceWork {
    enum S {   // spectral arrays (complex, nx/2+1 x ny x nz)
        US, VS, WS, BS, T0, I_00, I_0p, I_0m, f1p, f1m, I_f1p, I_f1m, I_0f1p, I_0f1m, I_pT0p, I_pT0m,
        f2p, f2m, pTf2p, pTf2m, I_f2p, I_f2m, I_0f2p, I_0f2m, X1, X2, NS
    };
    enum R {   // real arrays with halo
        u0, v0, w0, b0, u1, v1, w1, b1, u2, v2, w2, b2, u3, v3, w3, b3, pTu0, pTv0, pTw0, pTb0,
        t0, t1, t2, t3, d0, d1, d2, d3, NR
    };
    double2 *s[NS] = {};
    double *r[NR] = {};
    double *packed = nullptr;          // (nx, ny, nz) input of the x transform
    double *mode[3] = {};              // phi, psi, psib (nz x nz, [z + nz * mode])
    double *save[9] = {};              // model state during model_run
    double *k = nullptr;               // kx, ky, kz
};

namespace {

constexpr int BT = 256;

struct BalConsts {
    int nx, ny, nz, nkx, vbc;
    int64_t S;
    double dx, dy, dz, f0, N0, dsqr, Ro;
    const double *kx, *ky, *kz;
};

__device__ __forceinline__ double2 cm(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 ca(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cs(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cr(double2 a, double r) { return make_double2(a.x * r, a.y * r); }
__device__ __forceinline__ double2 cdr(double2 a, double r) { return make_double2(a.x / r, a.y / r); }
__device__ __forceinline__ double2 ci(double2 a) { return make_double2(-a.y, a.x); }    // i a
__device__ __forceinline__ double2 cmi(double2 a) { return make_double2(a.y, -a.x); }  // -i a
__device__ __forceinline__ double2 cj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cd(double2 a, double2 b)
{
    const double den = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__device__ __forceinline__ double2 cz() { return make_double2(0.0, 0.0); }

// eigenvectors.f90:224-270
__device__ __forceinline__ double hksqr(double k, double d) { return 2.0 * (1.0 - cos(k * d)) / (d * d); }
__device__ __forceinline__ double hepm(double k, double d) { return 0.5 * (1 + cos(k * d)); }
__device__ __forceinline__ double2 hkp(double k, double d)   // -i (e^{ikd} - 1) / d
{
    double s, c;
    sincos(k * d, &s, &c);
    return make_double2(s / d, (1.0 - c) / d);
}
__device__ __forceinline__ double2 hep(double k, double d)   // (e^{ikd} + 1) / 2
{
    double s, c;
    sincos(k * d, &s, &c);
    return make_double2((c + 1.0) * 0.5, s * 0.5);
}
__device__ __forceinline__ double2 hem(double k, double d)   // (e^{-ikd} + 1) / 2
{
    double s, c;
    sincos(k * d, &s, &c);
    return make_double2((c + 1.0) * 0.5, -s * 0.5);
}

struct V4 { double2 x, y, z, b; };

// the discrete normal modes at one wavenumber (eigenvectors.f90:6-218)
struct Eig {
    const BalConsts &P;
    double kx, ky, kz;
    double k2, f2, m2, N2, arg, om2;
    bool hor, ver;

    __device__ Eig(const BalConsts &p, int i, int j, int k) : P(p)
    {
        kx = P.kx[i];
        ky = P.ky[j];
        kz = P.kz[k];
        k2 = hksqr(kx, P.dx) + hksqr(ky, P.dy);
        f2 = P.f0 * P.f0 * hepm(kx, P.dx) * hepm(ky, P.dy);
        double K;
        if (P.vbc) {
            arg = kz / 2 * P.dz;
            const double sn = sin(arg), cn = cos(arg);
            m2 = 4 * (sn * sn) / (P.dz * P.dz);
            N2 = P.N0 * P.N0 * (cn * cn);
            K = P.dsqr * k2 + m2;
            om2 = K > 0.0 ? (N2 * k2 + f2 * m2) / K : 0.0;
        } else {
            arg = 0.0;
            m2 = hksqr(kz, P.dz);
            N2 = P.N0 * P.N0 * hepm(kz, P.dz);
            K = P.dsqr * k2 + hksqr(kz, P.dz);
            om2 = K > 0.0 ? (hepm(kz, P.dz) * (P.N0 * P.N0) * k2 + P.f0 * P.f0 * hksqr(kz, P.dz) * hepm(kx, P.dx) * hepm(ky, P.dy)) / K
                          : 0.0;
        }
        hor = kx * kx + ky * ky > 0.0;
        ver = kz * kz > 0.0;
    }

    __device__ double coeff_p(double s) const
    {
        const double d = k2 * N2 + f2 * m2;
        if (s == 0.0 && d > 0.0) return f2 * N2 / d;
        if (d > 0.0) return (N2 - P.dsqr * om2) * (om2 - f2) / (2 * d);
        return 0.0;
    }

    __device__ V4 vec_q(double s) const
    {
        V4 q{cz(), cz(), cz(), cz()};
        const double om = sqrt(om2);
        if (hor) {
            const double fxa = f2 - s * s * om2;
            const double2 mif0 = make_double2(0.0, -P.f0);   // -i f0
            const double2 pif0 = make_double2(0.0, P.f0);    // +i f0
            q.x = cdr(ca(cm(cm(cm(mif0, hep(kx, P.dx)), hem(ky, P.dy)), hkp(ky, P.dy)), cr(hkp(kx, P.dx), s * om)), fxa);
            q.y = cdr(ca(cm(cm(cm(pif0, hem(kx, P.dx)), hep(ky, P.dy)), hkp(kx, P.dx)), cr(hkp(ky, P.dy), s * om)), fxa);
            if (ver) {
                const double fxb = N2 - P.dsqr * s * s * om2;
                if (P.vbc) {
                    const double lam = 2 * sin(arg) / P.dz;
                    q.z = make_double2(0.0, lam * s * om / fxb);
                    q.b = make_double2(-lam * cos(arg) * (P.N0 * P.N0) / fxb, 0.0);
                } else {
                    q.z = cdr(cr(cr(hkp(kz, P.dz), s), om), fxb);
                    q.b = cdr(cr(ci(cm(hem(kz, P.dz), hkp(kz, P.dz))), P.N0 * P.N0), fxb);
                }
            }
        } else {
            q.x = make_double2(0.0, -s);
            q.y = make_double2(s * s, 0.0);
        }
        if (s == 0.0 && (f2 == 0.0 || N2 == 0.0)) q = V4{cz(), cz(), cz(), cz()};
        if (s != 0.0 && om2 == 0.0) q = V4{cz(), cz(), cz(), cz()};
        if (!P.vbc && hor && !ver && s != 0.0) {   // eigenvectors.f90:87-92
            q.x = cz();
            q.y = cz();
            q.z = make_double2(0.0, s);
            q.b = make_double2(s * s * P.N0 * sqrt(P.dsqr), 0.0);
        }
        return q;
    }

    __device__ V4 vec_p(double s) const
    {
        const double c = coeff_p(s);
        V4 q = vec_q(s);
        V4 p;
        p.x = cr(cj(q.x), c);
        p.y = cr(cj(q.y), c);
        p.z = cr(cr(cj(q.z), c), P.dsqr);
        p.b = cdr(cr(cj(q.b), c), P.N0 * P.N0);
        if (!P.vbc) {   // eigenvectors.f90:106-117
            if (!hor && ver) {
                p.x = make_double2(0.0, s / 2.0);
                p.y = make_double2(s * s / 2.0, 0.0);
            }
            if (hor && !ver && s != 0.0) {
                p.x = cz();
                p.y = cz();
                p.z = make_double2(0.0, -s / 2.0);
                p.b = make_double2(s * s / (2.0 * P.N0 * sqrt(P.dsqr)), 0.0);
            }
        }
        return p;
    }
};

// p . S = px*US + py*VS + pz*WS + pb*BS (diag_balance.f90:123)
__device__ __forceinline__ double2 pdot(const V4 &p, double2 a, double2 b, double2 c, double2 d)
{
    return ca(ca(ca(cm(p.x, a), cm(p.y, b)), cm(p.z, c)), cm(p.b, d));
}

enum Op {
    OP_ZERO,      // g0 = p0.S; out = g0 q0                              (:122-127)
    OP_DOT3,      // o0 = i p0.S, o1 = i p+.S, o2 = i p-.S               (:158-160)
    OP_DOT2,      // o0 = i p+.S, o1 = i p-.S                            (:200-201)
    OP_FIRST,     // f1p, f1m from I_0p, I_0m; out = Ro (f1p q+ + f1m q-) (:161-171)
    OP_PTG0,      // out = (-i I_00) q0                                  (:213-217)
    OP_SECOND,    // f2p, f2m; out = Ro^2 (f2p q+ + f2m q-)               (:250-264)
    OP_THIRD,     // f3p, f3m; out = Ro^3 (f3p q+ + f3m q-)               (:345-356)
    OP_FILTER,    // zero where kz^2 >= cut^2                            (:520-533)
    OP_GEO_BASE,  // OP_ZERO, g0 kept in aux0         (diag_opt_balance.f90:262-272)
    OP_GEO_EXCH   // g0 = p0.S; out = S + (g0b - g0) q0, g0b = in 4     (:303-313)
};

struct OpArgs {
    int op;
    const double2 *in[10];
    double2 *out[4];
    double2 *aux[2];   // f1p, f1m (OP_FIRST) or f2p, f2m (OP_SECOND)
    double cut2;
};

__global__ __launch_bounds__(BT) void k_bal_op(BalConsts P, OpArgs A)
{
    const int64_t n = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (n >= P.S) return;
    const int i = (int)(n % P.nkx);
    const int j = (int)((n / P.nkx) % P.ny);
    const int k = (int)(n / ((int64_t)P.nkx * P.ny));
    if (A.op == OP_FILTER) {
        const double kz = P.kz[k];
        if (kz * kz >= A.cut2)
            for (int c = 0; c < 4; ++c) A.out[c][n] = cz();
        return;
    }
    const Eig E(P, i, j, k);
    auto in = [&](int q) { return A.in[q][n]; };
    const double om = sqrt(E.om2);
    switch (A.op) {
    case OP_ZERO: {
        const double2 g0 = pdot(E.vec_p(0.0), in(0), in(1), in(2), in(3));
        const V4 q = E.vec_q(0.0);
        A.out[0][n] = cm(g0, q.x);
        A.out[1][n] = cm(g0, q.y);
        A.out[2][n] = cm(g0, q.z);
        A.out[3][n] = cm(g0, q.b);
        break;
    }
    case OP_GEO_BASE: {
        const double2 g0 = pdot(E.vec_p(0.0), in(0), in(1), in(2), in(3));
        const V4 q = E.vec_q(0.0);
        A.aux[0][n] = g0;
        A.out[0][n] = cm(g0, q.x);
        A.out[1][n] = cm(g0, q.y);
        A.out[2][n] = cm(g0, q.z);
        A.out[3][n] = cm(g0, q.b);
        break;
    }
    case OP_GEO_EXCH: {
        const double2 g0 = pdot(E.vec_p(0.0), in(0), in(1), in(2), in(3));
        const V4 q = E.vec_q(0.0);
        const double2 dg = cs(in(4), g0);
        A.out[0][n] = ca(in(0), cm(dg, q.x));
        A.out[1][n] = ca(in(1), cm(dg, q.y));
        A.out[2][n] = ca(in(2), cm(dg, q.z));
        A.out[3][n] = ca(in(3), cm(dg, q.b));
        break;
    }
    case OP_DOT3:
        A.out[0][n] = ci(pdot(E.vec_p(0.0), in(0), in(1), in(2), in(3)));
        A.out[1][n] = ci(pdot(E.vec_p(1.0), in(0), in(1), in(2), in(3)));
        A.out[2][n] = ci(pdot(E.vec_p(-1.0), in(0), in(1), in(2), in(3)));
        break;
    case OP_DOT2:
        A.out[0][n] = ci(pdot(E.vec_p(1.0), in(0), in(1), in(2), in(3)));
        A.out[1][n] = ci(pdot(E.vec_p(-1.0), in(0), in(1), in(2), in(3)));
        break;
    case OP_FIRST: {   // in: I_0p, I_0m
        double2 f1p = cz(), f1m = cz();
        if (om > 0.0) {
            f1p = cdr(in(0), om);
            f1m = cdr(make_double2(-in(1).x, -in(1).y), om);
        }
        const V4 qp = E.vec_q(1.0), qm = E.vec_q(-1.0);
        const double Ro = P.Ro;
        A.out[0][n] = cr(ca(cm(f1p, qp.x), cm(f1m, qm.x)), Ro);
        A.out[1][n] = cr(ca(cm(f1p, qp.y), cm(f1m, qm.y)), Ro);
        A.out[2][n] = cr(ca(cm(f1p, qp.z), cm(f1m, qm.z)), Ro);
        A.out[3][n] = cr(ca(cm(f1p, qp.b), cm(f1m, qm.b)), Ro);
        A.aux[0][n] = f1p;
        A.aux[1][n] = f1m;
        break;
    }
    case OP_PTG0: {
        const double2 g = cmi(in(0));
        const V4 q = E.vec_q(0.0);
        A.out[0][n] = cm(g, q.x);
        A.out[1][n] = cm(g, q.y);
        A.out[2][n] = cm(g, q.z);
        A.out[3][n] = cm(g, q.b);
        break;
    }
    case OP_SECOND: {
        // in: 0 I_0pT0p, 1 I_0pT0m, 2 I_0p, 3 I_0m, 4 I_pT0p, 5 I_pT0m, 6 I_0f1p, 7 I_0f1m, 8 I_f1p, 9 I_f1m
        double2 f2p = cz(), f2m = cz();
        if (om > 0.0) {
            const double2 pTf1p = cdr(cs(cs(in(0), in(2)), in(4)), om);
            const double2 t = cdr(cs(cs(in(1), in(3)), in(5)), om);
            const double2 pTf1m = make_double2(-t.x, -t.y);
            f2p = cdr(cs(cs(cs(in(6), in(2)), in(8)), ci(pTf1p)), om);
            const double2 u = cdr(cs(cs(cs(in(7), in(3)), in(9)), ci(pTf1m)), om);
            f2m = make_double2(-u.x, -u.y);
        }
        const V4 qp = E.vec_q(1.0), qm = E.vec_q(-1.0);
        const double R2 = P.Ro * P.Ro;
        A.out[0][n] = cr(ca(cm(f2p, qp.x), cm(f2m, qm.x)), R2);
        A.out[1][n] = cr(ca(cm(f2p, qp.y), cm(f2m, qm.y)), R2);
        A.out[2][n] = cr(ca(cm(f2p, qp.z), cm(f2m, qm.z)), R2);
        A.out[3][n] = cr(ca(cm(f2p, qp.b), cm(f2m, qm.b)), R2);
        A.aux[0][n] = f2p;
        A.aux[1][n] = f2m;
        break;
    }
    case OP_THIRD: {
        // in: 0 I_0f2p, 1 I_0f2m, 2 I_f1p, 3 I_f1m, 4 I_f2p, 5 I_f2m, 6 I_0p, 7 I_0m, 8 pTf2p, 9 pTf2m
        double2 f3p = cz(), f3m = cz();
        if (om > 0.0) {
            f3p = cdr(cs(cs(cs(ca(in(0), in(2)), in(4)), in(6)), ci(in(8))), om);
            const double2 t = cdr(cs(cs(cs(ca(in(1), in(3)), in(5)), in(7)), ci(in(9))), om);
            f3m = make_double2(-t.x, -t.y);
        }
        const V4 qp = E.vec_q(1.0), qm = E.vec_q(-1.0);
        const double R3 = P.Ro * P.Ro * P.Ro;
        A.out[0][n] = cr(ca(cm(f3p, qp.x), cm(f3m, qm.x)), R3);
        A.out[1][n] = cr(ca(cm(f3p, qp.y), cm(f3m, qm.y)), R3);
        A.out[2][n] = cr(ca(cm(f3p, qp.z), cm(f3m, qm.z)), R3);
        A.out[3][n] = cr(ca(cm(f3p, qp.b), cm(f3m, qm.b)), R3);
        break;
    }
    default: break;
    }
}

// pTf2 = (f2 - pTf2) / (10 dt) / Ro (diag_balance.f90:334-335)
__global__ __launch_bounds__(BT) void k_bal_dt(int64_t S, const double2 *__restrict__ f, double2 *__restrict__ pt,
                                               double tdt, double Ro)
{
    const int64_t n = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (n >= S) return;
    pt[n] = cdr(cdr(cs(f[n], pt[n]), tdt), Ro);
}

// QS(:,:,k1) = sum_k2 Qloc(:,:,k2) mode(k2,k1) dz   (my_fft.f90:246-250)
__global__ __launch_bounds__(BT) void k_mode_fwd(const double2 *__restrict__ q, double2 *__restrict__ out,
                                                 const double *__restrict__ mode, int64_t ncols, int nz, double dz)
{
    const int64_t t = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (t >= ncols * nz) return;
    const int64_t col = t % ncols;
    const int k1 = (int)(t / ncols);
    double2 acc = cz();
    for (int k2 = 0; k2 < nz; ++k2)
        acc = ca(acc, cr(cr(q[col + ncols * k2], mode[k2 + (int64_t)nz * k1]), dz));
    out[t] = acc;
}

// Qloc(:,:,k2) = sum_k1 QS(:,:,k1) mode(k2,k1)   (my_fft.f90:222-226)
__global__ __launch_bounds__(BT) void k_mode_bwd(const double2 *__restrict__ qs, double2 *__restrict__ out,
                                                 const double *__restrict__ mode, int64_t ncols, int nz)
{
    const int64_t t = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (t >= ncols * nz) return;
    const int64_t col = t % ncols;
    const int k2 = (int)(t / ncols);
    double2 acc = cz();
    for (int k1 = 0; k1 < nz; ++k1) acc = ca(acc, cr(qs[col + ncols * k1], mode[k2 + (int64_t)nz * k1]));
    out[t] = acc;
}

// interior of a field -> packed (nx, ny, nz), divided by d1 then d2 (0: not)
__global__ __launch_bounds__(BT) void k_pack(StepConsts c, const double *__restrict__ src, double *__restrict__ dst,
                                             double d1, double d2)
{
    const int64_t t = (int64_t)blockIdx.x * BT + threadIdx.x;
    const int64_t N = (int64_t)c.nx * c.ny * c.nz;
    if (t >= N) return;
    const int i = (int)(t % c.nx) + 1;
    const int j = (int)((t / c.nx) % c.ny) + 1;
    const int k = (int)(t / ((int64_t)c.nx * c.ny)) + 1;
    double x = src[(int64_t)(i + 1) + c.X * (int64_t)(j + 1) + c.XY * (int64_t)(k + 1)];
    if (d1 != 0.0) x = x / d1;
    if (d2 != 0.0) x = x / d2;
    dst[t] = x;
}

// out = a + b/db + c over whole arrays (halo included, as the reference's
// whole-array expressions); null terms are left out, db = 0: no division
__global__ __launch_bounds__(BT) void k_comb(int64_t n, double *__restrict__ out, const double *__restrict__ a,
                                             const double *__restrict__ b, double db, const double *__restrict__ c)
{
    const int64_t t = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (t >= n) return;
    double x = a ? a[t] : 0.0;
    if (b) x = a ? x + (db != 0.0 ? b[t] / db : b[t]) : (db != 0.0 ? b[t] / db : b[t]);
    if (c) x = x + c[t];
    out[t] = x;
}

unsigned blocks(int64_t n) { return (unsigned)((n + BT - 1) / BT); }

// ---- optimal balance (diag_opt_balance.f90) ---------------------------------
// whole-array updates, as the reference's array expressions
__global__ __launch_bounds__(BT) void k_ob_acc(int64_t n, double *__restrict__ acc, const double *__restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (t < n) acc[t] = acc[t] + x[t];   // :227-230
}

__global__ __launch_bounds__(BT) void k_ob_div(int64_t n, double *__restrict__ a, double d)
{
    const int64_t t = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (t < n) a[t] = a[t] / d;   // :232-233
}

__global__ __launch_bounds__(BT) void k_ob_bc(int64_t n, double *__restrict__ out, const double *__restrict__ save,
                                              const double *__restrict__ ave, const double *__restrict__ base)
{
    const int64_t t = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (t < n) out[t] = save[t] - ave[t] + base[t];   // :153-156
}

// opt_balance_norm (:338-363): per-block partial sums of e1, e2 and d over
// the interior, tree-reduced in LDS; the host adds the blocks in order
__global__ __launch_bounds__(BT) void k_ob_norm(StepConsts c, const double *__restrict__ u,
                                                const double *__restrict__ v, const double *__restrict__ w,
                                                const double *__restrict__ b, const double *__restrict__ ub,
                                                const double *__restrict__ vb, const double *__restrict__ wb,
                                                const double *__restrict__ bb, double dsqr, double N02,
                                                double *__restrict__ part)
{
    __shared__ double sh[3][BT];
    const int64_t t = (int64_t)blockIdx.x * BT + threadIdx.x;
    const int64_t N = (int64_t)c.nx * c.ny * c.nz;
    double e1 = 0.0, e2 = 0.0, d = 0.0;
    if (t < N) {
        const int i = (int)(t % c.nx) + 1;
        const int j = (int)((t / c.nx) % c.ny) + 1;
        const int k = (int)(t / ((int64_t)c.nx * c.ny)) + 1;
        const int64_t o = (int64_t)(i + 1) + c.X * (int64_t)(j + 1) + c.XY * (int64_t)(k + 1);
        e1 = u[o] * u[o] + v[o] * v[o] + dsqr * (w[o] * w[o]) + (b[o] * b[o]) / N02;
        e2 = ub[o] * ub[o] + vb[o] * vb[o] + dsqr * (wb[o] * wb[o]) + (bb[o] * bb[o]) / N02;
        const double du = u[o] - ub[o], dv = v[o] - vb[o], dw = w[o] - wb[o], db = b[o] - bb[o];
        d = du * du + dv * dv + dsqr * (dw * dw) + (db * db) / N02;
    }
    sh[0][threadIdx.x] = e1;
    sh[1][threadIdx.x] = e2;
    sh[2][threadIdx.x] = d;
    __syncthreads();
    for (int h = BT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = sh[q][threadIdx.x] + sh[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * (int64_t)blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// ramp of diag_opt_balance.f90:367-378
double ob_ramp(double theta)
{
    const double t1 = 1.0 / std::max(1e-32, theta);
    const double t2 = 1.0 / std::max(1e-32, 1.0 - theta);
    return std::exp(-t1) / (std::exp(-t1) + std::exp(-t2));
}

}  // namespace

void balance_release(Model &m)
{
    delete m.bal;
    m.bal = nullptr;
}

namespace {

struct Bal {
    Model &m;
    BalanceWork &W;
    BalConsts P{};
    int filter_width;
    std::vector<double> kzh;

    double2 *S(int i) { return W.s[i]; }
    double *R(int i) { return W.r[i]; }

    void op(int code, std::initializer_list<const double2 *> in, std::initializer_list<double2 *> out,
            double cut2 = 0.0, double2 *aux0 = nullptr, double2 *aux1 = nullptr)
    {
        OpArgs A{};
        A.aux[0] = aux0;
        A.aux[1] = aux1;
        A.op = code;
        int q = 0;
        for (auto p : in) A.in[q++] = p;
        q = 0;
        for (auto p : out) A.out[q++] = p;
        A.cut2 = cut2;
        hipLaunchKernelGGL(k_bal_op, dim3(blocks(P.S)), dim3(BT), 0, m.stream, P, A);
        PS3D_HIP_CHECK(hipGetLastError());
    }

    void comb(double *out, const double *a, const double *b, double db, const double *c)
    {
        hipLaunchKernelGGL(k_comb, dim3(blocks(m.g.H)), dim3(BT), 0, m.stream, m.g.H, out, a, b, db, c);
        PS3D_HIP_CHECK(hipGetLastError());
    }

    // diag_balance.f90:421-438: fields -> spectral (out[4]); d1, d2 divide
    void forward(const double *const *f, double2 *const *out, const double *d1, const double *d2)
    {
        const ps3d_params &p = m.prm;
        const int64_t ncols = (int64_t)m.g.nkx * m.g.ny;
        for (int c = 0; c < 4; ++c) {
            hipLaunchKernelGGL(k_pack, dim3(blocks(m.g.N)), dim3(BT), 0, m.stream, m.c, f[c], W.packed,
                               d1 ? d1[c] : 0.0, d2 ? d2[c] : 0.0);
            PS3D_HIP_CHECK(hipGetLastError());
            if (P.vbc) {
                fft_r2c_x(m, W.packed, 0, 0, S(BalanceWork::T0));
                fft_c2c_y(m, S(BalanceWork::T0), -1, p.dx, p.dy);
                const double *mode = W.mode[c < 2 ? 0 : (c == 2 ? 1 : 2)];
                hipLaunchKernelGGL(k_mode_fwd, dim3(blocks(m.g.S)), dim3(BT), 0, m.stream, S(BalanceWork::T0), out[c],
                                   mode, ncols, m.g.nz, p.dz);
                PS3D_HIP_CHECK(hipGetLastError());
            } else {
                fft_r2c_x(m, W.packed, 0, 0, out[c]);
                fft_c2c_y(m, out[c], -1, 0.0, 0.0);
                fft_c2c_z(m, out[c], -1, p.dx, p.dy, p.dz);
            }
        }
    }

    // diag_balance.f90:441-462 (the spectral inputs are used as scratch)
    void backward(double2 *const *in, double *const *out)
    {
        const ps3d_params &p = m.prm;
        const int64_t ncols = (int64_t)m.g.nkx * m.g.ny;
        for (int c = 0; c < 4; ++c) {
            if (P.vbc) {
                const double *mode = W.mode[c < 2 ? 0 : (c == 2 ? 1 : 2)];
                hipLaunchKernelGGL(k_mode_bwd, dim3(blocks(m.g.S)), dim3(BT), 0, m.stream, in[c], S(BalanceWork::T0),
                                   mode, ncols, m.g.nz);
                PS3D_HIP_CHECK(hipGetLastError());
                fft_c2c_y(m, S(BalanceWork::T0), +1, 0.0, 0.0);
                fft_c2r_x(m, S(BalanceWork::T0), out[c], 1, p.Lx * p.Ly);
            } else {
                fft_c2c_z(m, in[c], +1, 0.0, 0.0, 0.0);
                fft_c2c_y(m, in[c], +1, 0.0, 0.0);
                fft_c2r_x(m, in[c], out[c], 1, p.Lx * p.Ly * p.Lz);
            }
            launch_border_exchg(m, out[c], 0);
        }
    }

    // diag_balance.f90:500-535 -> US, VS, WS, BS
    void nonlin(const double *u, const double *v, const double *w, const double *b)
    {
        using B = BalanceWork;
        launch_advection_fields(m, u, v, w, b, R(B::d0), R(B::d1), R(B::d2), R(B::d3));
        const double *f[4] = {R(B::d0), R(B::d1), R(B::d2), R(B::d3)};
        const double Ro = m.prm.Ro;
        const double d1[4] = {Ro, Ro, Ro, Ro};
        const double d2[4] = {0.0, 0.0, m.prm.dsqr, 0.0};
        double2 *o[4] = {S(B::US), S(B::VS), S(B::WS), S(B::BS)};
        forward(f, o, d1, d2);
        if (filter_width >= 0) {
            const int idx = m.g.nz / 2 + filter_width;   // kz(nz/2+1+width), 1-based
            PS3D_REQUIRE(idx >= 0 && idx < m.g.nz, "diag_balance filter width outside the kz table");
            const double fxa = std::fabs(kzh[idx]);
            op(OP_FILTER, {}, {o[0], o[1], o[2], o[3]}, fxa * fxa);
        }
    }

    // diag_balance.f90:465-497 (itt stays 0: every step is Euler)
    void model_run(const double *u, const double *v, const double *w, const double *b, double *const *out, int n)
    {
        double *st[9] = {m.u, m.v, m.w, m.b, m.p, m.du, m.dv, m.dw, m.db};
        const int64_t len[9] = {m.g.H, m.g.H, m.g.H, m.g.H, m.g.H, 3 * m.g.H, 3 * m.g.H, 3 * m.g.H, 3 * m.g.H};
        for (int q = 0; q < 9; ++q)
            PS3D_HIP_CHECK(hipMemcpyAsync(W.save[q], st[q], sizeof(double) * len[q], hipMemcpyDeviceToDevice, m.stream));
        const double *in[4] = {u, v, w, b};
        for (int q = 0; q < 4; ++q)
            PS3D_HIP_CHECK(hipMemcpyAsync(st[q], in[q], sizeof(double) * m.g.H, hipMemcpyDeviceToDevice, m.stream));
        int tau = 3, taum1 = 2, taum2 = 1;
        for (int s = 0; s < n; ++s) {
            model_step(m, 0, tau, taum1, taum2);
            tau = tau % 3 + 1;
            taum1 = taum1 % 3 + 1;
            taum2 = taum2 % 3 + 1;
        }
        double *now[9] = {m.u, m.v, m.w, m.b, m.p, m.du, m.dv, m.dw, m.db};   // after buffer swaps
        for (int q = 0; q < 4; ++q)
            PS3D_HIP_CHECK(hipMemcpyAsync(out[q], now[q], sizeof(double) * m.g.H, hipMemcpyDeviceToDevice, m.stream));
        for (int q = 0; q < 9; ++q)
            PS3D_HIP_CHECK(hipMemcpyAsync(now[q], W.save[q], sizeof(double) * len[q], hipMemcpyDeviceToDevice, m.stream));
    }

    // diag_balance.f90:109-131 (the model's state -> u0..b0)
    void zero_order()
    {
        using B = BalanceWork;
        const double *f[4] = {m.u, m.v, m.w, m.b};
        double2 *o[4] = {S(B::US), S(B::VS), S(B::WS), S(B::BS)};
        forward(f, o, nullptr, nullptr);
        op(OP_ZERO, {o[0], o[1], o[2], o[3]}, {o[0], o[1], o[2], o[3]});
        double *r[4] = {R(B::u0), R(B::v0), R(B::w0), R(B::b0)};
        backward(o, r);
    }

    // diag_balance.f90:134-176
    void first_order()
    {
        using B = BalanceWork;
        nonlin(R(B::u0), R(B::v0), R(B::w0), R(B::b0));
        double2 *o[4] = {S(B::US), S(B::VS), S(B::WS), S(B::BS)};
        op(OP_DOT3, {o[0], o[1], o[2], o[3]}, {S(B::I_00), S(B::I_0p), S(B::I_0m)});
        op(OP_FIRST, {S(B::I_0p), S(B::I_0m)}, {o[0], o[1], o[2], o[3]}, 0.0, S(B::f1p), S(B::f1m));
        double *r[4] = {R(B::u1), R(B::v1), R(B::w1), R(B::b1)};
        backward(o, r);
    }

    // diag_balance.f90:181-268
    void second_order()
    {
        using B = BalanceWork;
        const double Ro = m.prm.Ro;
        double2 *o[4] = {S(B::US), S(B::VS), S(B::WS), S(B::BS)};
        comb(R(B::t0), nullptr, R(B::u1), Ro, nullptr);
        comb(R(B::t1), nullptr, R(B::v1), Ro, nullptr);
        comb(R(B::t2), nullptr, R(B::w1), Ro, nullptr);
        comb(R(B::t3), nullptr, R(B::b1), Ro, nullptr);
        nonlin(R(B::t0), R(B::t1), R(B::t2), R(B::t3));
        op(OP_DOT2, {o[0], o[1], o[2], o[3]}, {S(B::I_f1p), S(B::I_f1m)});
        comb(R(B::t0), R(B::u0), R(B::u1), Ro, nullptr);
        comb(R(B::t1), R(B::v0), R(B::v1), Ro, nullptr);
        comb(R(B::t2), R(B::w0), R(B::w1), Ro, nullptr);
        comb(R(B::t3), R(B::b0), R(B::b1), Ro, nullptr);
        nonlin(R(B::t0), R(B::t1), R(B::t2), R(B::t3));
        op(OP_DOT2, {o[0], o[1], o[2], o[3]}, {S(B::I_0f1p), S(B::I_0f1m)});
        op(OP_PTG0, {S(B::I_00)}, {o[0], o[1], o[2], o[3]});
        double *pt[4] = {R(B::pTu0), R(B::pTv0), R(B::pTw0), R(B::pTb0)};
        backward(o, pt);
        nonlin(pt[0], pt[1], pt[2], pt[3]);
        op(OP_DOT2, {o[0], o[1], o[2], o[3]}, {S(B::I_pT0p), S(B::I_pT0m)});
        comb(R(B::t0), R(B::u0), pt[0], 0.0, nullptr);
        comb(R(B::t1), R(B::v0), pt[1], 0.0, nullptr);
        comb(R(B::t2), R(B::w0), pt[2], 0.0, nullptr);
        comb(R(B::t3), R(B::b0), pt[3], 0.0, nullptr);
        nonlin(R(B::t0), R(B::t1), R(B::t2), R(B::t3));
        op(OP_DOT2, {o[0], o[1], o[2], o[3]}, {S(B::X1), S(B::X2)});   // I_0pT0p, I_0pT0m
        op(OP_SECOND,
           {S(B::X1), S(B::X2), S(B::I_0p), S(B::I_0m), S(B::I_pT0p), S(B::I_pT0m), S(B::I_0f1p), S(B::I_0f1m),
            S(B::I_f1p), S(B::I_f1m)},
           {o[0], o[1], o[2], o[3]}, 0.0, S(B::f2p), S(B::f2m));
        double *r[4] = {R(B::u2), R(B::v2), R(B::w2), R(B::b2)};
        backward(o, r);
    }

    // diag_balance.f90:273-362
    void third_order()
    {
        using B = BalanceWork;
        const double Ro = m.prm.Ro, R2 = Ro * Ro;
        double2 *o[4] = {S(B::US), S(B::VS), S(B::WS), S(B::BS)};
        comb(R(B::t0), nullptr, R(B::u2), R2, nullptr);
        comb(R(B::t1), nullptr, R(B::v2), R2, nullptr);
        comb(R(B::t2), nullptr, R(B::w2), R2, nullptr);
        comb(R(B::t3), nullptr, R(B::b2), R2, nullptr);
        nonlin(R(B::t0), R(B::t1), R(B::t2), R(B::t3));
        op(OP_DOT2, {o[0], o[1], o[2], o[3]}, {S(B::I_f2p), S(B::I_f2m)});
        comb(R(B::t0), R(B::u0), R(B::u2), R2, nullptr);
        comb(R(B::t1), R(B::v0), R(B::v2), R2, nullptr);
        comb(R(B::t2), R(B::w0), R(B::w2), R2, nullptr);
        comb(R(B::t3), R(B::b0), R(B::b2), R2, nullptr);
        nonlin(R(B::t0), R(B::t1), R(B::t2), R(B::t3));
        op(OP_DOT2, {o[0], o[1], o[2], o[3]}, {S(B::I_0f2p), S(B::I_0f2m)});
        comb(R(B::t0), R(B::u0), R(B::u1), 0.0, R(B::u2));
        comb(R(B::t1), R(B::v0), R(B::v1), 0.0, R(B::v2));
        comb(R(B::t2), R(B::w0), R(B::w1), 0.0, R(B::w2));
        comb(R(B::t3), R(B::b0), R(B::b1), 0.0, R(B::b2));
        double *u3[4] = {R(B::u3), R(B::v3), R(B::w3), R(B::b3)};
        model_run(R(B::t0), R(B::t1), R(B::t2), R(B::t3), u3, 10);
        const double *f[4] = {u3[0], u3[1], u3[2], u3[3]};
        forward(f, o, nullptr, nullptr);
        op(OP_ZERO, {o[0], o[1], o[2], o[3]}, {o[0], o[1], o[2], o[3]});
        double *r0[4] = {R(B::u0), R(B::v0), R(B::w0), R(B::b0)};
        backward(o, r0);
        PS3D_HIP_CHECK(hipMemcpyAsync(S(B::pTf2p), S(B::f2p), sizeof(double2) * P.S, hipMemcpyDeviceToDevice, m.stream));
        PS3D_HIP_CHECK(hipMemcpyAsync(S(B::pTf2m), S(B::f2m), sizeof(double2) * P.S, hipMemcpyDeviceToDevice, m.stream));
        first_order();
        second_order();
        hipLaunchKernelGGL(k_bal_dt, dim3(blocks(P.S)), dim3(BT), 0, m.stream, P.S, S(B::f2p), S(B::pTf2p),
                           10.0 * m.prm.dt, Ro);
        hipLaunchKernelGGL(k_bal_dt, dim3(blocks(P.S)), dim3(BT), 0, m.stream, P.S, S(B::f2m), S(B::pTf2m),
                           10.0 * m.prm.dt, Ro);
        PS3D_HIP_CHECK(hipGetLastError());
        zero_order();
        first_order();
        second_order();
        op(OP_THIRD,
           {S(B::I_0f2p), S(B::I_0f2m), S(B::I_f1p), S(B::I_f1m), S(B::I_f2p), S(B::I_f2m), S(B::I_0p), S(B::I_0m),
            S(B::pTf2p), S(B::pTf2m)},
           {o[0], o[1], o[2], o[3]});
        backward(o, u3);
    }
};

template <class T>
T *walloc(Model &m, int64_t n)
{
    void *p = nullptr;
    PS3D_HIP_CHECK(hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    m.allocations.push_back(p);
    return static_cast<T *>(p);
}

// my_fft.f90:112-127
void vertical_modes(int nz, double Lz, std::vector<double> &phi, std::vector<double> &psi, std::vector<double> &psib)
{
    const double pi = std::acos(0.0) * 2.0;
    phi.assign((size_t)nz * nz, 0.0);
    psi.assign((size_t)nz * nz, 0.0);
    psib.assign((size_t)nz * nz, 0.0);
    auto at = [nz](std::vector<double> &a, int k, int k2) -> double & { return a[(k - 1) + (size_t)nz * (k2 - 1)]; };
    for (int k = 1; k <= nz; ++k) at(phi, k, 1) = std::sqrt(1.0 / Lz);
    for (int k2 = 2; k2 <= nz; ++k2) {
        for (int k = 1; k <= nz; ++k) at(psi, k, k2) = std::sqrt(2.0 / Lz) * std::sin((k2 - 1) * k * pi / nz);
        const double sn = 2 * std::sin((k2 - 1) * pi / (2 * nz));
        const double cs = 2 * std::cos((k2 - 1) * pi / (2 * nz));
        at(phi, 1, k2) = at(psi, 1, k2) / sn;
        at(psib, 1, k2) = at(psi, 1, k2) / cs;
        for (int k = 2; k <= nz; ++k) {
            at(phi, k, k2) = (at(psi, k, k2) - at(psi, k - 1, k2)) / sn;
            at(psib, k, k2) = (at(psi, k, k2) + at(psi, k - 1, k2)) / cs;
        }
    }
}

// the work arrays, made on first use
BalanceWork &bal_work(Model &m)
{
    const Grid &g = m.g;
    if (!m.bal) {
        auto *W = new BalanceWork();
        m.bal = W;
        for (int i = 0; i < BalanceWork::NS; ++i) W->s[i] = walloc<double2>(m, g.S);
        for (int i = 0; i < BalanceWork::NR; ++i) W->r[i] = walloc<double>(m, g.H);
        for (int i = 0; i < 4; ++i) W->save[i] = walloc<double>(m, g.H);
        W->save[4] = walloc<double>(m, g.H);
        for (int i = 5; i < 9; ++i) W->save[i] = walloc<double>(m, 3 * g.H);
        W->packed = walloc<double>(m, g.N);
        W->k = walloc<double>(m, g.nkx + g.ny + g.nz);
        if (m.prm.enable_vertical_boundaries) {
            std::vector<double> tab[3];
            vertical_modes(g.nz, m.prm.Lz, tab[0], tab[1], tab[2]);
            for (int q = 0; q < 3; ++q) {
                W->mode[q] = walloc<double>(m, (int64_t)g.nz * g.nz);
                PS3D_HIP_CHECK(hipMemcpyAsync(W->mode[q], tab[q].data(), sizeof(double) * tab[q].size(),
                                              hipMemcpyHostToDevice, m.stream));
            }
        }
    }
    return *m.bal;
}

// the wavenumbers on the device and the constants of the spectral operators
void bal_consts(Model &m, Bal &b, int filter_width, const double *kx, const double *ky, const double *kz)
{
    const Grid &g = m.g;
    BalanceWork &W = b.W;
    b.filter_width = filter_width;
    b.kzh.assign(kz, kz + g.nz);
    double *dk = W.k;
    PS3D_HIP_CHECK(hipMemcpyAsync(dk, kx, sizeof(double) * g.nkx, hipMemcpyHostToDevice, m.stream));
    PS3D_HIP_CHECK(hipMemcpyAsync(dk + g.nkx, ky, sizeof(double) * g.ny, hipMemcpyHostToDevice, m.stream));
    PS3D_HIP_CHECK(hipMemcpyAsync(dk + g.nkx + g.ny, kz, sizeof(double) * g.nz, hipMemcpyHostToDevice, m.stream));
    BalConsts &P = b.P;
    P.nx = g.nx; P.ny = g.ny; P.nz = g.nz; P.nkx = g.nkx;
    P.vbc = m.prm.enable_vertical_boundaries;
    P.S = g.S;
    P.dx = m.prm.dx; P.dy = m.prm.dy; P.dz = m.prm.dz;
    P.f0 = m.prm.f0; P.N0 = m.prm.N0; P.dsqr = m.prm.dsqr; P.Ro = m.prm.Ro;
    P.kx = dk; P.ky = dk + g.nkx; P.kz = dk + g.nkx + g.ny;
}

}  // namespace

// diag_balance.f90:89-106 on the model's current state; out[16] = u0, v0,
// w0, b0, u1 .. b1, u2 .. b2, u3 .. b3 (host arrays with halo, nullable)
void diag_balance(Model &m, int filter_width, double *const *out, const double *kx, const double *ky,
                  const double *kz)
{
    PS3D_REQUIRE(!m.mixed, "diag_balance: FP64 models only");
    const Grid &g = m.g;
    for (int f = 0; f < 4; ++f)
        PS3D_REQUIRE(m.cur[f] == 0, "diag_balance: friction results not yet folded in by ab_update");
    BalanceWork &W = bal_work(m);
    Bal b{m, W};
    bal_consts(m, b, filter_width, kx, ky, kz);
    b.zero_order();
    b.first_order();
    b.second_order();
    b.third_order();
    using B = BalanceWork;
    const int ids[16] = {B::u0, B::v0, B::w0, B::b0, B::u1, B::v1, B::w1, B::b1,
                         B::u2, B::v2, B::w2, B::b2, B::u3, B::v3, B::w3, B::b3};
    for (int q = 0; q < 16; ++q)
        if (out && out[q])
            PS3D_HIP_CHECK(hipMemcpyAsync(out[q], W.r[ids[q]], sizeof(double) * g.H, hipMemcpyDeviceToHost, m.stream));
    PS3D_HIP_CHECK(hipStreamSynchronize(m.stream));
}

namespace {

// dt and Ro of the next steps, biharmonic friction on or off
void ob_set(Model &m, const ps3d_params &base, double dt, double Ro, bool bih)
{
    m.prm.dt = dt;
    m.prm.Ro = Ro;
    m.prm.Ahbi = bih ? base.Ahbi : 0.0;
    m.prm.Avbi = bih ? base.Avbi : 0.0;
    m.prm.Khbi = bih ? base.Khbi : 0.0;
    m.prm.Kvbi = bih ? base.Kvbi : 0.0;
    m.c.dt = dt;
    m.c.rdt = 1.0 / dt;
    m.c.Ro = Ro;
}

}  // namespace

// diag_opt_balance.f90:55-179 on the model's current state: out[4] =
// u_bal .. b_bal (host arrays with halo, nullable), norms[i] = norm_diff of
// iteration i+1 (nullable, max_iterations entries), *iterations = iterations
// run.  The model's state, dt, Ro and friction are restored at the end.
void opt_balance(Model &m, const ps3d_opt_balance_params &ob, double *const *out, double *norms, int *iterations,
                 const double *kx, const double *ky, const double *kz)
{
    PS3D_REQUIRE(!m.mixed, "opt_balance: FP64 models only");
    for (int f = 0; f < 4; ++f)
        PS3D_REQUIRE(m.cur[f] == 0, "opt_balance: friction results not yet folded in by ab_update");
    PS3D_REQUIRE(ob.period > 0 && ob.max_iterations >= 1, "opt_balance: period > 0 and max_iterations >= 1");
    PS3D_REQUIRE(!ob.time_average || (ob.average > 0 && ob.average_times >= 1),
                 "opt_balance: average > 0 and average_times >= 1 for the time-average variant");
    const Grid &g = m.g;
    BalanceWork &W = bal_work(m);
    Bal b{m, W};
    bal_consts(m, b, -1, kx, ky, kz);
    using B = BalanceWork;
    const int64_t H = g.H;
    hipStream_t st = m.stream;
    auto copy = [&](double *dst, const double *src, int64_t n) {
        PS3D_HIP_CHECK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    };
    auto fields = [&]() { return std::array<double *, 4>{m.u, m.v, m.w, m.b}; };   // after buffer swaps

    // save the model state (:80-83)
    const ps3d_params base = m.prm;
    {
        double *now[9] = {m.u, m.v, m.w, m.b, m.p, m.du, m.dv, m.dw, m.db};
        for (int q = 0; q < 9; ++q) copy(W.save[q], now[q], q < 5 ? H : 3 * H);
    }
    double *ubal[4] = {b.R(B::u0), b.R(B::v0), b.R(B::w0), b.R(B::b0)};
    double *ubase[4] = {b.R(B::u1), b.R(B::v1), b.R(B::w1), b.R(B::b1)};
    double *usave[4] = {b.R(B::u2), b.R(B::v2), b.R(B::w2), b.R(B::b2)};
    double *uave[4] = {b.R(B::u3), b.R(B::v3), b.R(B::w3), b.R(B::b3)};
    for (int q = 0; q < 4; ++q) copy(ubal[q], fields()[q], H);
    const double dt_loc = base.dt, Ro_loc = base.Ro;
    const int n_end = (int)(ob.period / dt_loc) + 1;      // :87-88
    const int n_end_ave = (int)(ob.average / dt_loc) + 1;

    // n Euler steps (itt stays 0) from tau = 3; acc: running sums of u..b
    auto steps = [&](int n, double dt, auto &&ro_of, bool bih, double *const *acc) {
        int tau = 3, taum1 = 2, taum2 = 1;
        for (int s = 0; s < n; ++s) {
            ob_set(m, base, dt, ro_of(s), bih);
            model_step(m, 0, tau, taum1, taum2);
            tau = tau % 3 + 1;
            taum1 = taum1 % 3 + 1;
            taum2 = taum2 % 3 + 1;
            if (acc) {
                const auto f = fields();
                for (int q = 0; q < 4; ++q) {
                    hipLaunchKernelGGL(k_ob_acc, dim3(blocks(H)), dim3(BT), 0, st, H, acc[q], f[q]);
                    PS3D_HIP_CHECK(hipGetLastError());
                }
            }
        }
    };
    // opt_time_ave (:212-235) into dst, then u = dst
    auto time_ave = [&](double *const *dst) {
        for (int q = 0; q < 4; ++q) PS3D_HIP_CHECK(hipMemsetAsync(dst[q], 0, sizeof(double) * H, st));
        steps(n_end_ave, std::fabs(dt_loc), [](int) { return 0.0; }, true, dst);
        for (int q = 0; q < 4; ++q) {
            hipLaunchKernelGGL(k_ob_div, dim3(blocks(H)), dim3(BT), 0, st, H, dst[q], (double)n_end_ave);
            PS3D_HIP_CHECK(hipGetLastError());
            copy(fields()[q], dst[q], H);
        }
    };
    // the geostrophic operators (:238-314) on the current fields
    auto geo = [&](int code) {
        const auto f = fields();
        const double *in[4] = {f[0], f[1], f[2], f[3]};
        double2 *o[4] = {b.S(B::US), b.S(B::VS), b.S(B::WS), b.S(B::BS)};
        b.forward(in, o, nullptr, nullptr);
        if (code == OP_GEO_EXCH)
            b.op(OP_GEO_EXCH, {o[0], o[1], o[2], o[3], b.S(B::I_00)}, {o[0], o[1], o[2], o[3]});
        else
            b.op(code, {o[0], o[1], o[2], o[3]}, {o[0], o[1], o[2], o[3]}, 0.0, b.S(B::I_00));
        double *r[4] = {f[0], f[1], f[2], f[3]};
        if (code == OP_GEO_BASE) {
            for (int q = 0; q < 4; ++q) r[q] = ubase[q];
        }
        b.backward(o, r);
    };
    const int64_t nblk = blocks(g.N);
    std::vector<double> part(3 * (size_t)nblk);
    double *dpart = b.R(B::t0);   // 3 * nblk <= H doubles of scratch
    auto norm = [&]() {
        const auto f = fields();
        hipLaunchKernelGGL(k_ob_norm, dim3((unsigned)nblk), dim3(BT), 0, st, m.c, f[0], f[1], f[2], f[3], ubal[0],
                           ubal[1], ubal[2], ubal[3], base.dsqr, base.N0 * base.N0, dpart);
        PS3D_HIP_CHECK(hipGetLastError());
        PS3D_HIP_CHECK(hipMemcpyAsync(part.data(), dpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
        PS3D_HIP_CHECK(hipStreamSynchronize(st));
        double e1 = 0.0, e2 = 0.0, d = 0.0;
        for (int64_t q = 0; q < nblk; ++q) {
            e1 += part[3 * q];
            e2 += part[3 * q + 1];
            d += part[3 * q + 2];
        }
        return 2.0 * std::sqrt(d) / (std::sqrt(e1) + std::sqrt(e2));   // :363
    };

    // base point (:96-110)
    if (ob.time_average) {
        for (int q = 0; q < ob.average_times; ++q) time_ave(ubase);
        for (int q = 0; q < 4; ++q) copy(fields()[q], ubal[q], H);
    } else {
        geo(OP_GEO_BASE);
    }
    int its = 0;
    for (int n = 1; n <= ob.max_iterations; ++n) {
        // :117-118, :183-201: backward, biharmonic friction off, Ro ramped down
        steps(n_end, -std::fabs(dt_loc), [&](int s) { return (1.0 - ob_ramp(s * dt_loc / ob.period)) * Ro_loc; },
              false, nullptr);
        if (ob.time_average)
            for (int q = 0; q < ob.average_times; ++q) time_ave(uave);
        else
            geo(OP_ZERO);
        // :137-139, :205-223: forward, Ro ramped up
        steps(n_end, std::fabs(dt_loc), [&](int s) { return ob_ramp(s * dt_loc / ob.period) * Ro_loc; }, true,
              nullptr);
        if (ob.time_average) {
            for (int q = 0; q < 4; ++q) copy(usave[q], fields()[q], H);
            for (int q = 0; q < ob.average_times; ++q) time_ave(uave);
            const auto f = fields();
            for (int q = 0; q < 4; ++q) {
                hipLaunchKernelGGL(k_ob_bc, dim3(blocks(H)), dim3(BT), 0, st, H, f[q], usave[q], uave[q], ubase[q]);
                PS3D_HIP_CHECK(hipGetLastError());
            }
        } else {
            geo(OP_GEO_EXCH);
        }
        const double nd = norm();
        if (norms) norms[n - 1] = nd;
        for (int q = 0; q < 4; ++q) copy(ubal[q], fields()[q], H);
        its = n;
        if (n > 1 && nd < ob.tol) break;   // :170-175
    }
    // results, then the model as it was (:180-184)
    for (int q = 0; q < 4; ++q)
        if (out && out[q])
            PS3D_HIP_CHECK(hipMemcpyAsync(out[q], ubal[q], sizeof(double) * H, hipMemcpyDeviceToHost, st));
    {
        double *now[9] = {m.u, m.v, m.w, m.b, m.p, m.du, m.dv, m.dw, m.db};
        for (int q = 0; q < 9; ++q) copy(now[q], W.save[q], q < 5 ? H : 3 * H);
    }
    ob_set(m, base, base.dt, base.Ro, true);
    PS3D_HIP_CHECK(hipStreamSynchronize(st));
    if (iterations) *iterations = its;
}

}  // namespace ps3d
