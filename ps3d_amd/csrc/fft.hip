// fft.hip -- mixed-radix (4, 2, 3, 5) FP64 FFTs for the spectral Poisson
// solve of pressure_solver.f90, replacing P3DFFT (my_fft.f90:159-205).
//
// Conventions of the transforms they replace: forward r2c unnormalised with
// e^{-i k x}; inverse c2r unnormalised with e^{+i k x}, the imaginary parts of
// the kx = 0 and kx = nx/2 modes ignored; the spectral array is (nx/2+1, ny, nz)
// with kx fastest.  The scalings of my_fft_2D/3D (QS*dx*dy(*dz)) and of
// my_ifft_2D/3D (QP/(Lx*Ly(*Lz))) are applied inside the kernels.
//
// Layout on the device: every 1-D transform of a pass is staged in LDS,
// batched so that global loads and stores are coalesced along the unit-stride
// axis.  x-pass: 1 row per transform, rows of one k-level side by side.
// y- and z-passes: TX consecutive kx columns per work-group, so each global
// access touches TX*16 contiguous bytes; the transposition needed to run the
// y (or z) transform on those columns happens in LDS, which is what the
// pencil transposes of P3DFFT become on one device.
//
// Each pass is a Stockham autosort step on one LDS buffer: a thread reads the
// R inputs of its butterflies into registers, the work-group synchronises,
// and the outputs are written back in natural order.
#include "ps3d_internal.h"

#include <cmath>
#include <type_traits>

// the factorisation (radix 4 passes, then 2, 3, 5), the twiddle table and the
// butterfly formulas fix the operation order; with contraction off the
// transforms are reproducible bit for bit by a CPU restatement of them
#pragma clang fp contract(off)

namespace ps3d {

namespace {

constexpr int MAXE = 16;   // max complex values one thread stages per pass

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// i*a for sign=+1, -i*a for sign=-1
__device__ __forceinline__ double2 mul_isign(double2 a, int sign)
{
    return sign > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

template <int R>
__device__ __forceinline__ void dft(double2 *v, int sign)
{
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const double c = -0.5, s = 0.86602540378443864676;   // cos, sin(2 pi/3)
        const double2 t = cadd(v[1], v[2]);
        const double2 d = csub(v[1], v[2]);
        const double2 m = make_double2(v[0].x + c * t.x, v[0].y + c * t.y);
        const double2 q = mul_isign(cscale(d, s), sign);
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, q);
        v[2] = csub(m, q);
    } else if constexpr (R == 4) {
        const double2 a = cadd(v[0], v[2]), bq = csub(v[0], v[2]);
        const double2 c = cadd(v[1], v[3]), d = mul_isign(csub(v[1], v[3]), sign);
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(bq, d);
        v[3] = csub(bq, d);
    } else if constexpr (R == 5) {
        const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
        const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
        const double2 a14 = cadd(v[1], v[4]), d14 = csub(v[1], v[4]);
        const double2 a23 = cadd(v[2], v[3]), d23 = csub(v[2], v[3]);
        const double2 m1 = make_double2(v[0].x + c1 * a14.x + c2 * a23.x, v[0].y + c1 * a14.y + c2 * a23.y);
        const double2 m2 = make_double2(v[0].x + c2 * a14.x + c1 * a23.x, v[0].y + c2 * a14.y + c1 * a23.y);
        const double2 q1 = mul_isign(make_double2(s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y), sign);
        const double2 q2 = mul_isign(make_double2(s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y), sign);
        v[0] = make_double2(v[0].x + a14.x + a23.x, v[0].y + a14.y + a23.y);
        v[1] = cadd(m1, q1);
        v[4] = csub(m1, q1);
        v[2] = cadd(m2, q2);
        v[3] = csub(m2, q2);
    }
}

// One Stockham pass of radix R on ncol sequences of length N in LDS.
// Element e of column col lives at buf[e*es + col*cs].  COLFAST maps
// consecutive threads to consecutive columns (for es > 1 layouts).
template <int R, bool COLFAST>
__device__ __forceinline__ void stockham_pass(double2 *buf, int N, int Ns, int ncol, int es, int cs,
                                              const double2 *__restrict__ tw, int sign)
{
    constexpr int MAXB = MAXE / R > 0 ? MAXE / R + (MAXE % R ? 1 : 0) : 1;
    const int nb = N / R;
    const int total = nb * ncol;
    const int T = blockDim.x;
    const int twstride = N / (Ns * R);
    double2 v[MAXB][R];
#pragma unroll
    for (int s = 0; s < MAXB; ++s) {
        const int t = threadIdx.x + s * T;
        if (t < total) {
            const int col = COLFAST ? t % ncol : t / nb;
            const int jb = COLFAST ? t / ncol : t % nb;
#pragma unroll
            for (int r = 0; r < R; ++r) v[s][r] = buf[(jb + r * nb) * es + col * cs];
            if (Ns > 1) {
                const int jm = jb % Ns;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    double2 w = tw[r * jm * twstride];
                    if (sign > 0) w.y = -w.y;
                    v[s][r] = cmul(v[s][r], w);
                }
            }
            dft<R>(v[s], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < MAXB; ++s) {
        const int t = threadIdx.x + s * T;
        if (t < total) {
            const int col = COLFAST ? t % ncol : t / nb;
            const int jb = COLFAST ? t / ncol : t % nb;
            const int jm = jb % Ns;
            const int base = (jb / Ns) * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[(base + r * Ns) * es + col * cs] = v[s][r];
        }
    }
    __syncthreads();
}

template <bool COLFAST>
__device__ void lds_fft(double2 *buf, const FftPlan &pl, int ncol, int es, int cs, int sign)
{
    int Ns = 1;
    for (int f = 0; f < pl.nf; ++f) {
        const int R = pl.fac[f];
        switch (R) {
        case 4: stockham_pass<4, COLFAST>(buf, pl.n, Ns, ncol, es, cs, pl.tw, sign); break;
        case 2: stockham_pass<2, COLFAST>(buf, pl.n, Ns, ncol, es, cs, pl.tw, sign); break;
        case 3: stockham_pass<3, COLFAST>(buf, pl.n, Ns, ncol, es, cs, pl.tw, sign); break;
        case 5: stockham_pass<5, COLFAST>(buf, pl.n, Ns, ncol, es, cs, pl.tw, sign); break;
        default: break;
        }
        Ns *= R;
    }
}

__device__ __forceinline__ int64_t fidx(const StepConsts &c, int i, int j, int k)
{
    return (int64_t)(i + 1) + c.X * (int64_t)(j + 1) + c.XY * (int64_t)(k + 1);
}

// ---------------------------------------------------------------------------
// x-pass, forward r2c.  Rows r = j + ny*k (0-based).  Source: a dense real
// (nx,ny,nz) array, or (FUSE_DIV) the divergence of u, v, w divided by dt,
// pressure_solver.f90:16-24, computed while loading.
// ---------------------------------------------------------------------------
template <bool FUSE_DIV>
__global__ __launch_bounds__(256) void k_r2c_x(StepConsts c, FftPlan pl, const double2 *__restrict__ wk,
                                               const double *__restrict__ src, const double *__restrict__ u,
                                               const double *__restrict__ v, const double *__restrict__ w,
                                               double2 *__restrict__ dst, int rows_per_block, int64_t nrows)
{
    extern __shared__ double2 lds[];
    const int n2 = pl.n;
    const int nkx = n2 + 1;
    const int cs = n2 + 1;   // padded row stride in LDS
    const int64_t row0 = (int64_t)blockIdx.x * rows_per_block;
    const int nrow = (int)min((int64_t)rows_per_block, nrows - row0);
    // load: z[m] = x[2m] + i x[2m+1]
    for (int t = threadIdx.x; t < nrow * n2; t += blockDim.x) {
        const int rr = t / n2, m = t % n2;
        const int64_t row = row0 + rr;
        double a0, a1;
        if (FUSE_DIV) {
            const int j = (int)(row % c.ny) + 1, k = (int)(row / c.ny) + 1;
            const int64_t X = c.X, XY = c.XY;
            int64_t o = fidx(c, 2 * m + 1, j, k);
            a0 = div_rn(u[o] - u[o - 1], c.dx, c.rdx) + div_rn(v[o] - v[o - X], c.dy, c.rdy) +
                 div_rn(w[o] - w[o - XY], c.dz, c.rdz);
            o += 1;
            a1 = div_rn(u[o] - u[o - 1], c.dx, c.rdx) + div_rn(v[o] - v[o - X], c.dy, c.rdy) +
                 div_rn(w[o] - w[o - XY], c.dz, c.rdz);
            a0 = div_rn(a0, c.dt, c.rdt);
            a1 = div_rn(a1, c.dt, c.rdt);
        } else {
            const double2 pr = reinterpret_cast<const double2 *>(src + row * (2 * (int64_t)n2))[m];
            a0 = pr.x;
            a1 = pr.y;
        }
        lds[rr * cs + m] = make_double2(a0, a1);
    }
    __syncthreads();
    lds_fft<false>(lds, pl, nrow, 1, cs, -1);
    // post-process: X[k] = E[k] + W^k O[k]
    for (int t = threadIdx.x; t < nrow * nkx; t += blockDim.x) {
        const int rr = t / nkx, kk = t % nkx;
        const double2 zk = lds[rr * cs + (kk == n2 ? 0 : kk)];
        const double2 zm = lds[rr * cs + (kk == 0 ? 0 : n2 - kk)];
        const double2 e = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
        const double2 bq = make_double2(zk.x - zm.x, zk.y + zm.y);   // zk - conj(zm)
        const double2 o = make_double2(0.5 * bq.y, -0.5 * bq.x);     // bq / (2i)
        const double2 xk = cadd(e, cmul(wk[kk], o));
        dst[(row0 + rr) * nkx + kk] = xk;
    }
}

// ---------------------------------------------------------------------------
// x-pass, inverse c2r, scaled by 1/inv_scale (my_fft.f90:166, :191).
// Destination: the interior of a field (halo layout) or a dense array.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_c2r_x(StepConsts c, FftPlan pl, const double2 *__restrict__ wk,
                                               const double2 *__restrict__ src, double *__restrict__ dst,
                                               int dst_is_field, double inv_scale, double rinv_scale,
                                               int rows_per_block, int64_t nrows)
{
    extern __shared__ double2 lds[];
    const int n2 = pl.n;
    const int nkx = n2 + 1;
    const int cs = n2 + 1;
    double2 *raw = lds + rows_per_block * cs;
    const int64_t row0 = (int64_t)blockIdx.x * rows_per_block;
    const int nrow = (int)min((int64_t)rows_per_block, nrows - row0);
    for (int t = threadIdx.x; t < nrow * nkx; t += blockDim.x) {
        const int rr = t / nkx, kk = t % nkx;
        double2 x = src[(row0 + rr) * nkx + kk];
        if (kk == 0 || kk == n2) x.y = 0.0;
        raw[rr * cs + kk] = x;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nrow * n2; t += blockDim.x) {
        const int rr = t / n2, kk = t % n2;
        const double2 xk = raw[rr * cs + kk];
        const double2 xm = raw[rr * cs + n2 - kk];
        const double2 a = make_double2(xk.x + xm.x, xk.y - xm.y);   // xk + conj(xm)
        const double2 d = make_double2(xk.x - xm.x, xk.y + xm.y);   // xk - conj(xm)
        const double2 wc = make_double2(wk[kk].x, -wk[kk].y);       // conj(W^k)
        const double2 q = cmul(wc, d);
        lds[rr * cs + kk] = make_double2(a.x - q.y, a.y + q.x);     // a + i q
    }
    __syncthreads();
    lds_fft<false>(lds, pl, nrow, 1, cs, +1);
    for (int t = threadIdx.x; t < nrow * n2; t += blockDim.x) {
        const int rr = t / n2, m = t % n2;
        const int64_t row = row0 + rr;
        const double2 z = lds[rr * cs + m];
        const double a0 = div_rn(z.x, inv_scale, rinv_scale), a1 = div_rn(z.y, inv_scale, rinv_scale);
        if (dst_is_field) {
            const int j = (int)(row % c.ny) + 1, k = (int)(row / c.ny) + 1;
            const int64_t o = fidx(c, 2 * m + 1, j, k);
            dst[o] = a0;
            dst[o + 1] = a1;
        } else {
            reinterpret_cast<double2 *>(dst + row * (2 * (int64_t)n2))[m] = make_double2(a0, a1);
        }
    }
}

// ---------------------------------------------------------------------------
// strided c2c pass over TX consecutive kx columns.  Element e of column col
// is at data[kx0+col + e*estride + lane_off]; optional scaling of the result
// by s1 then s2 (my_fft.f90:177: QS*dx*dy).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_c2c_cols(FftPlan pl, double2 *__restrict__ data, int nkx, int TX,
                                                  int64_t estride, int64_t lstride, int sign, double s1,
                                                  double s2, double s3, int nscale)
{
    extern __shared__ double2 lds[];
    const int N = pl.n;
    const int kx0 = blockIdx.x * TX;
    const int ncol = min(TX, nkx - kx0);
    double2 *base = data + kx0 + (int64_t)blockIdx.y * lstride;
    for (int t = threadIdx.x; t < N * TX; t += blockDim.x) {
        const int col = t % TX, e = t / TX;
        if (col < ncol) lds[e * TX + col] = base[col + e * estride];
    }
    __syncthreads();
    lds_fft<true>(lds, pl, TX, TX, 1, sign);
    for (int t = threadIdx.x; t < N * TX; t += blockDim.x) {
        const int col = t % TX, e = t / TX;
        if (col < ncol) {
            double2 x = lds[e * TX + col];
            if (nscale >= 1) x = cscale(x, s1);
            if (nscale >= 2) x = cscale(x, s2);
            if (nscale >= 3) x = cscale(x, s3);
            base[col + e * estride] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// periodic solve along z, pressure_solver.f90:58-70 with my_fft.f90's z
// transforms: forward z FFT, QS*dx*dy*dz, division by -ksqr (0 where
// ksqr <= 0), inverse z FFT -- one kernel, the column block stays in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_z_periodic_solve(FftPlan pl, double2 *__restrict__ data,
                                                          const double *__restrict__ ksqr, int nkx, int ny,
                                                          int TX, double dx, double dy, double dz)
{
    extern __shared__ double2 lds[];
    const int N = pl.n;
    const int kx0 = blockIdx.x * TX;
    const int j = blockIdx.y;
    const int ncol = min(TX, nkx - kx0);
    const int64_t estride = (int64_t)nkx * ny;
    double2 *base = data + kx0 + (int64_t)j * nkx;
    const double *kb = ksqr + kx0 + (int64_t)j * nkx;
    for (int t = threadIdx.x; t < N * TX; t += blockDim.x) {
        const int col = t % TX, e = t / TX;
        if (col < ncol) lds[e * TX + col] = base[col + e * estride];
    }
    __syncthreads();
    lds_fft<true>(lds, pl, TX, TX, 1, -1);
    for (int t = threadIdx.x; t < N * TX; t += blockDim.x) {
        const int col = t % TX, e = t / TX;
        if (col < ncol) {
            double2 x = lds[e * TX + col];
            x = cscale(cscale(cscale(x, dx), dy), dz);
            const double ks = kb[col + e * estride];
            if (ks > 0.0) x = make_double2(-(x.x / ks), -(x.y / ks));
            else x = make_double2(0.0, 0.0);
            lds[e * TX + col] = x;
        }
    }
    __syncthreads();
    lds_fft<true>(lds, pl, TX, TX, 1, +1);
    for (int t = threadIdx.x; t < N * TX; t += blockDim.x) {
        const int col = t % TX, e = t / TX;
        if (col < ncol) base[col + e * estride] = lds[e * TX + col];
    }
}

// ===========================================================================
// Compile-time-sized transforms.  Same factorisation, twiddles, butterflies
// and pre/post-processing as the generic kernels above (so the same bits),
// with the length, batch, strides and thread count known to the compiler:
// every index of every pass is a constant expression of the thread id.
// ===========================================================================
namespace ft {

// complex arithmetic over float2 / double2, same formulas as above
template <class C> struct real_of;
template <> struct real_of<double2> { using type = double; };
template <> struct real_of<float2> { using type = float; };

template <class C>
__device__ __forceinline__ C mk(typename real_of<C>::type x, typename real_of<C>::type y)
{
    C r;
    r.x = x;
    r.y = y;
    return r;
}
template <class C> __device__ __forceinline__ C mul(C a, C b) { return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <class C> __device__ __forceinline__ C add(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <class C> __device__ __forceinline__ C sub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <class C> __device__ __forceinline__ C scl(C a, typename real_of<C>::type s) { return mk<C>(a.x * s, a.y * s); }
template <class C, int SIGN> __device__ __forceinline__ C isg(C a) { return SIGN > 0 ? mk<C>(-a.y, a.x) : mk<C>(a.y, -a.x); }

template <int R, int SIGN, class C>
__device__ __forceinline__ void dftc(C *v)
{
    using T = typename real_of<C>::type;
    if constexpr (R == 2) {
        const C a = v[0], b = v[1];
        v[0] = add(a, b);
        v[1] = sub(a, b);
    } else if constexpr (R == 3) {
        const T c = T(-0.5), s = T(0.86602540378443864676);
        const C t = add(v[1], v[2]);
        const C d = sub(v[1], v[2]);
        const C m = mk<C>(v[0].x + c * t.x, v[0].y + c * t.y);
        const C q = isg<C, SIGN>(scl(d, s));
        v[0] = add(v[0], t);
        v[1] = add(m, q);
        v[2] = sub(m, q);
    } else if constexpr (R == 4) {
        const C a = add(v[0], v[2]), bq = sub(v[0], v[2]);
        const C c = add(v[1], v[3]), d = isg<C, SIGN>(sub(v[1], v[3]));
        v[0] = add(a, c);
        v[2] = sub(a, c);
        v[1] = add(bq, d);
        v[3] = sub(bq, d);
    } else if constexpr (R == 5) {
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
        const C a14 = add(v[1], v[4]), d14 = sub(v[1], v[4]);
        const C a23 = add(v[2], v[3]), d23 = sub(v[2], v[3]);
        const C m1 = mk<C>(v[0].x + c1 * a14.x + c2 * a23.x, v[0].y + c1 * a14.y + c2 * a23.y);
        const C m2 = mk<C>(v[0].x + c2 * a14.x + c1 * a23.x, v[0].y + c2 * a14.y + c1 * a23.y);
        const C q1 = isg<C, SIGN>(mk<C>(s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y));
        const C q2 = isg<C, SIGN>(mk<C>(s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y));
        v[0] = mk<C>(v[0].x + a14.x + a23.x, v[0].y + a14.y + a23.y);
        v[1] = add(m1, q1);
        v[4] = sub(m1, q1);
        v[2] = add(m2, q2);
        v[3] = sub(m2, q2);
    } else if constexpr (R == 8) {
        // two radix-4 butterflies on the even and odd inputs, joined by the
        // eighth roots w^q = e^(SIGN i pi q / 4): X(q) = A(q) + w^q B(q),
        // X(q + 4) = A(q) - w^q B(q)
        C a[4] = {v[0], v[2], v[4], v[6]}, b[4] = {v[1], v[3], v[5], v[7]};
        dftc<4, SIGN>(a);
        dftc<4, SIGN>(b);
        const T h = T(0.70710678118654752440);   // 1/sqrt(2)
        const C b1 = scl(add(b[1], isg<C, SIGN>(b[1])), h);
        const C b2 = isg<C, SIGN>(b[2]);
        const C b3 = scl(sub(isg<C, SIGN>(b[3]), b[3]), h);
        v[0] = add(a[0], b[0]);
        v[4] = sub(a[0], b[0]);
        v[1] = add(a[1], b1);
        v[5] = sub(a[1], b1);
        v[2] = add(a[2], b2);
        v[6] = sub(a[2], b2);
        v[3] = add(a[3], b3);
        v[7] = sub(a[3], b3);
    }
}

// Radix plans of the compiled-in bodies.  Plan 0: radix-4 passes, then 2,
// 3, 5 -- the FP64 plan, the order the oracle's transforms follow, so FP64
// results are reproducible bit for bit.  Plan 1: radix-8 passes first, then
// 4, 2, 3, 5 -- fewer passes through LDS; a different rounding, offered to
// the FP32 transforms of the mixed form, where the autotuner picks the plan
// per transform length.
constexpr bool plan_differs(int N) { return N % 8 == 0; }

struct Fac {
    int n;
    int f[MAX_FAC];
};

constexpr Fac factorize(int N, int PL)
{
    Fac F{0, {}};
    int r = N;
    if (PL == 1)
        while (r % 8 == 0) { F.f[F.n++] = 8; r /= 8; }
    while (r % 4 == 0) { F.f[F.n++] = 4; r /= 4; }
    while (r % 2 == 0) { F.f[F.n++] = 2; r /= 2; }
    while (r % 3 == 0) { F.f[F.n++] = 3; r /= 3; }
    while (r % 5 == 0) { F.f[F.n++] = 5; r /= 5; }
    return F;
}

// one Stockham pass; element e of column col at buf[e*ES + col*CS]
template <int R, int N, int NS, int NCOL, int T, int ES, int CS, bool COLFAST, int SIGN, class C>
__device__ __forceinline__ void pass(C *buf, const C *__restrict__ tw)
{
    constexpr int NB = N / R;
    constexpr int TOTAL = NB * NCOL;
    constexpr int PER = (TOTAL + T - 1) / T;
    constexpr int TWS = N / (NS * R);
    C v[PER][R];
#pragma unroll
    for (int s = 0; s < PER; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if (TOTAL % T == 0 || t < TOTAL) {
            const int col = COLFAST ? t % NCOL : t / NB;
            const int jb = COLFAST ? t / NCOL : t % NB;
#pragma unroll
            for (int r = 0; r < R; ++r) v[s][r] = buf[(jb + r * NB) * ES + col * CS];
            if constexpr (NS > 1) {
                const int jm = jb % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C w = tw[r * jm * TWS];
                    if constexpr (SIGN > 0) w.y = -w.y;
                    v[s][r] = mul(v[s][r], w);
                }
            }
            dftc<R, SIGN>(v[s]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < PER; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if (TOTAL % T == 0 || t < TOTAL) {
            const int col = COLFAST ? t % NCOL : t / NB;
            const int jb = COLFAST ? t / NCOL : t % NB;
            const int jm = jb % NS;
            const int base = (jb / NS) * NS * R + jm;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[(base + r * NS) * ES + col * CS] = v[s][r];
        }
    }
    __syncthreads();
}

template <int N, int P, int NS, int NCOL, int T, int ES, int CS, bool COLFAST, int SIGN, int PL, class C>
__device__ __forceinline__ void passes(C *buf, const C *__restrict__ tw)
{
    constexpr Fac F = factorize(N, PL);
    if constexpr (P < F.n) {
        constexpr int R = F.f[P];
        pass<R, N, NS, NCOL, T, ES, CS, COLFAST, SIGN>(buf, tw);
        passes<N, P + 1, NS * R, NCOL, T, ES, CS, COLFAST, SIGN, PL>(buf, tw);
    }
}

template <int N, int NCOL, int T, int ES, int CS, bool COLFAST, int SIGN, int PL = 0, class C>
__device__ __forceinline__ void fft(C *buf, const C *__restrict__ tw)
{
    passes<N, 0, 1, NCOL, T, ES, CS, COLFAST, SIGN, PL>(buf, tw);
}

// x-pass, forward r2c of ROWS rows (see k_r2c_x); the post-processing works
// on the pair (k, n2-k) in place.  RT = double (drop-in) or float (mixed).
template <int N2, int ROWS, int T, bool FUSE_DIV, class RT, int PL = 0>
__device__ __forceinline__ void r2c_body(typename cplx_of<RT>::type *lds, int blk, const StepConsts &c,
                                         const typename cplx_of<RT>::type *__restrict__ tw,
                                         const typename cplx_of<RT>::type *__restrict__ wk,
                                         const RT *__restrict__ src, const RT *__restrict__ u,
                                         const RT *__restrict__ v, const RT *__restrict__ w,
                                         typename cplx_of<RT>::type *__restrict__ dst, int64_t nrows, int dstride)
{
    using C = typename cplx_of<RT>::type;
    constexpr int CS = N2 + 1;
    const int64_t row0 = (int64_t)blk * ROWS;
    const RT h = RT(0.5);
#pragma unroll
    for (int s = 0; s < (ROWS * N2 + T - 1) / T; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if ((ROWS * N2) % T == 0 || t < ROWS * N2) {
            const int rr = t / N2, m = t % N2;
            const int64_t row = row0 + rr;
            RT a0 = RT(0), a1 = RT(0);
            if (row < nrows) {
                if (FUSE_DIV) {
                    const int j = (int)(row % c.ny) + 1, k = (int)(row / c.ny) + 1;
                    const int64_t X = c.X, XY = c.XY;
                    int64_t o = fidx(c, 2 * m + 1, j, k);
                    a0 = sdiv<RT>(u[o] - u[o - 1], c.dx, c.rdx) + sdiv<RT>(v[o] - v[o - X], c.dy, c.rdy) +
                         sdiv<RT>(w[o] - w[o - XY], c.dz, c.rdz);
                    o += 1;
                    a1 = sdiv<RT>(u[o] - u[o - 1], c.dx, c.rdx) + sdiv<RT>(v[o] - v[o - X], c.dy, c.rdy) +
                         sdiv<RT>(w[o] - w[o - XY], c.dz, c.rdz);
                    a0 = sdiv<RT>(a0, c.dt, c.rdt);
                    a1 = sdiv<RT>(a1, c.dt, c.rdt);
                } else {
                    a0 = src[row * (2 * (int64_t)N2) + 2 * m];
                    a1 = src[row * (2 * (int64_t)N2) + 2 * m + 1];
                }
            }
            lds[rr * CS + m] = mk<C>(a0, a1);
        }
    }
    __syncthreads();
    fft<N2, ROWS, T, 1, CS, false, -1, PL>(lds, tw);
    constexpr int HALF = N2 / 2 + 1;   // k = 0..N2/2 handles (k, N2-k)
#pragma unroll
    for (int s = 0; s < (ROWS * HALF + T - 1) / T; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if ((ROWS * HALF) % T == 0 || t < ROWS * HALF) {
            const int rr = t / HALF, kk = t % HALF;
            const int64_t row = row0 + rr;
            if (row < nrows) {
                const C z0 = lds[rr * CS + kk];
                const C z1 = lds[rr * CS + (kk == 0 ? 0 : N2 - kk)];
                C *d = dst + row * dstride;
                {   // X[kk] from zk = z0, zm = z1
                    const C e = mk<C>(h * (z0.x + z1.x), h * (z0.y - z1.y));
                    const C bq = mk<C>(z0.x - z1.x, z0.y + z1.y);
                    const C o = mk<C>(h * bq.y, -h * bq.x);
                    d[kk] = add(e, mul(wk[kk], o));
                }
                const int km = (kk == 0) ? N2 : N2 - kk;   // its partner
                if (km != kk) {                             // X[km] from zk = z1 (or Z[0]), zm = z0
                    const C zk = (kk == 0) ? z0 : z1;
                    const C zm = z0;
                    const C e = mk<C>(h * (zk.x + zm.x), h * (zk.y - zm.y));
                    const C bq = mk<C>(zk.x - zm.x, zk.y + zm.y);
                    const C o = mk<C>(h * bq.y, -h * bq.x);
                    d[km] = add(e, mul(wk[km], o));
                }
            }
        }
    }
}

template <int N2, int ROWS, int T, bool FUSE_DIV, class RT, int PL = 0>
__global__ __launch_bounds__(T) void r2c_x(StepConsts c, const typename cplx_of<RT>::type *__restrict__ tw,
                                           const typename cplx_of<RT>::type *__restrict__ wk,
                                           const RT *__restrict__ src, const RT *__restrict__ u,
                                           const RT *__restrict__ v, const RT *__restrict__ w,
                                           typename cplx_of<RT>::type *__restrict__ dst, int64_t nrows, int dstride)
{
    __shared__ typename cplx_of<RT>::type lds[ROWS * (N2 + 1)];
    // (measured: the XCD renumbering of the other passes makes this one slower)
    r2c_body<N2, ROWS, T, FUSE_DIV, RT, PL>(lds, (int)blockIdx.x, c, tw, wk, src, u, v, w, dst, nrows, dstride);
}

// x-pass, inverse c2r of ROWS rows (see k_c2r_x), pre-processing in place
template <int N2, int ROWS, int T, class RT, int PL = 0>
__device__ __forceinline__ void c2r_body(typename cplx_of<RT>::type *lds, int blk, const StepConsts &c,
                                         const typename cplx_of<RT>::type *__restrict__ tw,
                                         const typename cplx_of<RT>::type *__restrict__ wk,
                                         const typename cplx_of<RT>::type *__restrict__ src,
                                         RT *__restrict__ dst, int dst_is_field, double inv_scale,
                                         double rinv_scale, int64_t nrows, int sstride)
{
    using C = typename cplx_of<RT>::type;
    constexpr int CS = N2 + 1;
    constexpr int HALF = N2 / 2 + 1;
#pragma unroll
    for (int s = 0; s < (ROWS * HALF + T - 1) / T; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if ((ROWS * HALF) % T == 0 || t < ROWS * HALF) {
            const int rr = t / HALF, kk = t % HALF;
            const int64_t row = (int64_t)blk * ROWS + rr;
            if (row < nrows) {
                const C *sr = src + row * sstride;
                C xk = sr[kk];
                const int km = N2 - kk;
                C xm = sr[km];
                if (kk == 0) { xk.y = RT(0); xm.y = RT(0); }   // kx = 0 and kx = nx/2
                {   // Z'[kk] from xk, xm
                    const C a = mk<C>(xk.x + xm.x, xk.y - xm.y);
                    const C dd = mk<C>(xk.x - xm.x, xk.y + xm.y);
                    const C wc = mk<C>(wk[kk].x, -wk[kk].y);
                    const C q = mul(wc, dd);
                    lds[rr * CS + kk] = mk<C>(a.x - q.y, a.y + q.x);
                }
                if (kk != 0 && km != kk) {   // Z'[km] from xk' = xm, xm' = xk
                    const C a = mk<C>(xm.x + xk.x, xm.y - xk.y);
                    const C dd = mk<C>(xm.x - xk.x, xm.y + xk.y);
                    const C wc = mk<C>(wk[km].x, -wk[km].y);
                    const C q = mul(wc, dd);
                    lds[rr * CS + km] = mk<C>(a.x - q.y, a.y + q.x);
                }
            }
        }
    }
    __syncthreads();
    fft<N2, ROWS, T, 1, CS, false, +1, PL>(lds, tw);
#pragma unroll
    for (int s = 0; s < (ROWS * N2 + T - 1) / T; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if ((ROWS * N2) % T == 0 || t < ROWS * N2) {
            const int rr = t / N2, m = t % N2;
            const int64_t row = (int64_t)blk * ROWS + rr;
            if (row < nrows) {
                const C z = lds[rr * CS + m];
                const RT a0 = sdiv<RT>(z.x, inv_scale, rinv_scale), a1 = sdiv<RT>(z.y, inv_scale, rinv_scale);
                if (dst_is_field) {
                    const int j = (int)(row % c.ny) + 1, k = (int)(row / c.ny) + 1;
                    const int64_t o = fidx(c, 2 * m + 1, j, k);
                    dst[o] = a0;
                    dst[o + 1] = a1;
                } else {
                    dst[row * (2 * (int64_t)N2) + 2 * m] = a0;
                    dst[row * (2 * (int64_t)N2) + 2 * m + 1] = a1;
                }
            }
        }
    }
}

template <int N2, int ROWS, int T, class RT, int PL = 0>
__global__ __launch_bounds__(T) void c2r_x(StepConsts c, const typename cplx_of<RT>::type *__restrict__ tw,
                                           const typename cplx_of<RT>::type *__restrict__ wk,
                                           const typename cplx_of<RT>::type *__restrict__ src,
                                           RT *__restrict__ dst, int dst_is_field, double inv_scale,
                                           double rinv_scale, int64_t nrows, int sstride)
{
    __shared__ typename cplx_of<RT>::type lds[ROWS * (N2 + 1)];
    c2r_body<N2, ROWS, T, RT, PL>(lds, xcd_remap((int)blockIdx.x, (int)gridDim.x), c, tw, wk, src, dst, dst_is_field,
                              inv_scale, rinv_scale, nrows, sstride);
}

// strided c2c pass over TX consecutive kx columns (see k_c2c_cols)
template <int N, int TX, int T, int SIGN, class C, int PL = 0>
__device__ __forceinline__ void cols_body(C *lds, int bx, int by, const C *__restrict__ tw, C *__restrict__ data,
                                          int nkx, int64_t estride, int64_t lstride, double s1, double s2, double s3,
                                          int nscale)
{
    using RT = typename real_of<C>::type;
    const int kx0 = bx * TX;
    const int ncol = min(TX, nkx - kx0);
    C *base = data + kx0 + (int64_t)by * lstride;
#pragma unroll
    for (int s = 0; s < (N * TX + T - 1) / T; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if ((N * TX) % T == 0 || t < N * TX) {
            const int col = t % TX, e = t / TX;
            lds[e * TX + col] = (col < ncol) ? base[col + e * estride] : mk<C>(RT(0), RT(0));
        }
    }
    __syncthreads();
    fft<N, TX, T, TX, 1, true, SIGN, PL>(lds, tw);
#pragma unroll
    for (int s = 0; s < (N * TX + T - 1) / T; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if ((N * TX) % T == 0 || t < N * TX) {
            const int col = t % TX, e = t / TX;
            if (col < ncol) {
                C x = lds[e * TX + col];
                if (nscale >= 1) x = scl(x, (RT)s1);
                if (nscale >= 2) x = scl(x, (RT)s2);
                if (nscale >= 3) x = scl(x, (RT)s3);
                base[col + e * estride] = x;
            }
        }
    }
}

template <int N, int TX, int T, int SIGN, class C, int PL = 0>
__global__ __launch_bounds__(T) void c2c_cols(const C *__restrict__ tw, C *__restrict__ data, int nkx,
                                              int64_t estride, int64_t lstride, double s1, double s2,
                                              double s3, int nscale)
{
    __shared__ C lds[N * TX];
    // the column groups of one row of the grid share cache lines (a group's
    // rows are not line-aligned): keep them on one XCD's L2
    const int gx = (int)gridDim.x;
    const int r = xcd_remap((int)blockIdx.x + gx * (int)blockIdx.y, gx * (int)gridDim.y);
    cols_body<N, TX, T, SIGN, C, PL>(lds, r % gx, r / gx, tw, data, nkx, estride, lstride, s1, s2, s3, nscale);
}

// periodic z solve (see k_z_periodic_solve); the scaling and the division by
// ksqr are done in FP64 whatever the storage type
template <int N, int TX, int T, class C, int PL = 0>
__device__ __forceinline__ void zsolve_body(C *lds, int bx, int by, const C *__restrict__ tw, C *__restrict__ data,
                                            const double *__restrict__ ksqr, int nkx, int nkxs, int ny, double dx,
                                            double dy, double dz)
{
    using RT = typename real_of<C>::type;
    const int kx0 = bx * TX;
    const int j = by;
    const int ncol = min(TX, nkx - kx0);
    const int64_t estride = (int64_t)nkxs * ny;
    C *base = data + kx0 + (int64_t)j * nkxs;
    const double *kb = ksqr + kx0 + (int64_t)j * nkx;
#pragma unroll
    for (int s = 0; s < (N * TX + T - 1) / T; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if ((N * TX) % T == 0 || t < N * TX) {
            const int col = t % TX, e = t / TX;
            lds[e * TX + col] = (col < ncol) ? base[col + e * estride] : mk<C>(RT(0), RT(0));
        }
    }
    __syncthreads();
    fft<N, TX, T, TX, 1, true, -1, PL>(lds, tw);
#pragma unroll
    for (int s = 0; s < (N * TX + T - 1) / T; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if ((N * TX) % T == 0 || t < N * TX) {
            const int col = t % TX, e = t / TX;
            if (col < ncol) {
                const C xs = lds[e * TX + col];
                double xr = (double)xs.x * dx * dy * dz, xi = (double)xs.y * dx * dy * dz;
                const double ks = kb[col + e * (int64_t)nkx * ny];
                if (ks > 0.0) { xr = -(xr / ks); xi = -(xi / ks); }
                else { xr = 0.0; xi = 0.0; }
                lds[e * TX + col] = mk<C>((RT)xr, (RT)xi);
            }
        }
    }
    __syncthreads();
    fft<N, TX, T, TX, 1, true, +1, PL>(lds, tw);
#pragma unroll
    for (int s = 0; s < (N * TX + T - 1) / T; ++s) {
        const int t = (int)threadIdx.x + s * T;
        if ((N * TX) % T == 0 || t < N * TX) {
            const int col = t % TX, e = t / TX;
            if (col < ncol) base[col + e * estride] = lds[e * TX + col];
        }
    }
}

template <int N, int TX, int T, class C, int PL = 0>
__global__ __launch_bounds__(T) void z_periodic_solve(const C *__restrict__ tw, C *__restrict__ data,
                                                      const double *__restrict__ ksqr, int nkx, int nkxs, int ny,
                                                      double dx, double dy, double dz)
{
    __shared__ C lds[N * TX];
    const int gx = (int)gridDim.x;
    const int r = xcd_remap((int)blockIdx.x + gx * (int)blockIdx.y, gx * (int)gridDim.y);
    zsolve_body<N, TX, T, C, PL>(lds, r % gx, r / gx, tw, data, ksqr, nkx, nkxs, ny, dx, dy, dz);
}

// launch configurations: x-pass rows so that ROWS*N2 = 1024 values; 256
// threads (more for the longest columns of the y/z passes)
constexpr int XROWS(int n2) { return n2 >= 1024 ? 1 : 1024 / n2; }
// y/z passes: the columns of one work-group span one 64-byte segment of each
// row (4 FP64 or 8 FP32 complex values), which gives the most work-groups
// per pass at full-segment accesses
template <class C> constexpr int TCOL = 64 / (int)sizeof(C);
constexpr int TTHREADS = 256;
template <class C> constexpr int TCOLS(int n) { return n * TCOL<C> / 16 > 256 ? n * TCOL<C> / 16 : 256; }
// single-model passes: half the columns for the longest transforms (LDS per work-group)
template <class C, int N> constexpr int TCOLN = N >= 1024 ? TCOL<C> / 2 : TCOL<C>;
template <class C, int N> constexpr int TCOLSN = N * TCOLN<C, N> / 16 > 256 ? N * TCOLN<C, N> / 16 : 256;

// plan 1 bodies are compiled for the FP32 transforms of lengths it changes
template <int N, class RT> constexpr bool has_plan1 = std::is_same<RT, float>::value && plan_differs(N);

template <int N2, class RT, int PL>
void launch_r2c_pl(Model &m, const RT *src, const RT *u, const RT *v, const RT *w, int fuse_div,
                   typename cplx_of<RT>::type *dst, const typename cplx_of<RT>::type *tw,
                   const typename cplx_of<RT>::type *wk, int64_t nrows, int dstride)
{
    constexpr int ROWS = XROWS(N2);
    const dim3 grid((unsigned)((nrows + ROWS - 1) / ROWS));
    if (fuse_div)
        hipLaunchKernelGGL((r2c_x<N2, ROWS, TTHREADS, true, RT, PL>), grid, dim3(TTHREADS), 0, m.stream, m.c, tw, wk,
                           nullptr, u, v, w, dst, nrows, dstride);
    else
        hipLaunchKernelGGL((r2c_x<N2, ROWS, TTHREADS, false, RT, PL>), grid, dim3(TTHREADS), 0, m.stream, m.c, tw, wk,
                           src, nullptr, nullptr, nullptr, dst, nrows, dstride);
}

template <int N2, class RT>
bool launch_r2c(Model &m, const RT *src, const RT *u, const RT *v, const RT *w, int fuse_div,
                typename cplx_of<RT>::type *dst, const typename cplx_of<RT>::type *tw,
                const typename cplx_of<RT>::type *wk, int64_t nrows, int dstride, int plan)
{
    if constexpr (has_plan1<N2, RT>)
        if (plan == 1) {
            launch_r2c_pl<N2, RT, 1>(m, src, u, v, w, fuse_div, dst, tw, wk, nrows, dstride);
            return true;
        }
    launch_r2c_pl<N2, RT, 0>(m, src, u, v, w, fuse_div, dst, tw, wk, nrows, dstride);
    return true;
}

template <int N2, class RT>
bool launch_c2r(Model &m, const typename cplx_of<RT>::type *src, RT *dst, int dst_is_field, double inv_scale,
                const typename cplx_of<RT>::type *tw, const typename cplx_of<RT>::type *wk, int64_t nrows,
                int sstride, int plan)
{
    constexpr int ROWS = XROWS(N2);
    const dim3 grid((unsigned)((nrows + ROWS - 1) / ROWS));
    if constexpr (has_plan1<N2, RT>)
        if (plan == 1) {
            hipLaunchKernelGGL((c2r_x<N2, ROWS, TTHREADS, RT, 1>), grid, dim3(TTHREADS), 0, m.stream, m.c, tw, wk, src,
                               dst, dst_is_field, inv_scale, 1.0 / inv_scale, nrows, sstride);
            return true;
        }
    hipLaunchKernelGGL((c2r_x<N2, ROWS, TTHREADS, RT, 0>), grid, dim3(TTHREADS), 0, m.stream, m.c, tw, wk, src, dst,
                       dst_is_field, inv_scale, 1.0 / inv_scale, nrows, sstride);
    return true;
}

template <int N, class C, int PL>
void launch_cols_pl(const C *tw, hipStream_t st, C *data, int nkx, int64_t estride, int64_t lstride,
                    unsigned ny_or_nz, int sign, double s1, double s2, double s3, int nscale)
{
    const dim3 grid((unsigned)((nkx + TCOLN<C, N> - 1) / TCOLN<C, N>), ny_or_nz);
    constexpr int T = TCOLSN<C, N>;
    if (sign < 0)
        hipLaunchKernelGGL((c2c_cols<N, TCOLN<C, N>, T, -1, C, PL>), grid, dim3(T), 0, st, tw, data, nkx, estride,
                           lstride, s1, s2, s3, nscale);
    else
        hipLaunchKernelGGL((c2c_cols<N, TCOLN<C, N>, T, +1, C, PL>), grid, dim3(T), 0, st, tw, data, nkx, estride,
                           lstride, s1, s2, s3, nscale);
}

template <int N, class C>
bool launch_cols(const C *tw, hipStream_t st, C *data, int nkx, int64_t estride, int64_t lstride,
                 unsigned ny_or_nz, int sign, double s1, double s2, double s3, int nscale, int plan)
{
    if constexpr (has_plan1<N, typename real_of<C>::type>)
        if (plan == 1) {
            launch_cols_pl<N, C, 1>(tw, st, data, nkx, estride, lstride, ny_or_nz, sign, s1, s2, s3, nscale);
            return true;
        }
    launch_cols_pl<N, C, 0>(tw, st, data, nkx, estride, lstride, ny_or_nz, sign, s1, s2, s3, nscale);
    return true;
}

template <int N, class C>
bool launch_zsolve(Model &m, C *data, const C *tw, int nkxs, int plan)
{
    const dim3 grid((unsigned)((m.g.nkx + TCOLN<C, N> - 1) / TCOLN<C, N>), (unsigned)m.g.ny);
    constexpr int T = TCOLSN<C, N>;
    if constexpr (has_plan1<N, typename real_of<C>::type>)
        if (plan == 1) {
            hipLaunchKernelGGL((z_periodic_solve<N, TCOLN<C, N>, T, C, 1>), grid, dim3(T), 0, m.stream, tw, data, m.ksqr,
                               m.g.nkx, nkxs, m.g.ny, m.prm.dx, m.prm.dy, m.prm.dz);
            return true;
        }
    hipLaunchKernelGGL((z_periodic_solve<N, TCOLN<C, N>, T, C, 0>), grid, dim3(T), 0, m.stream, tw, data, m.ksqr,
                       m.g.nkx, nkxs, m.g.ny, m.prm.dx, m.prm.dy, m.prm.dz);
    return true;
}

// the sizes compiled in; others take the generic FP64 kernels
#define PS3D_FFT_SIZES(X) X(8) X(10) X(12) X(15) X(16) X(20) X(24) X(30) X(32) X(40) X(48) X(50) X(60) X(64) \
    X(80) X(96) X(100) X(120) X(128) X(160) X(240) X(256) X(512) X(1024)

template <class RT>
bool dispatch_r2c(Model &m, const RT *src, const RT *u, const RT *v, const RT *w, int fuse_div,
                  typename cplx_of<RT>::type *dst, const typename cplx_of<RT>::type *tw,
                  const typename cplx_of<RT>::type *wk, int64_t nrows, int dstride, int plan = 0)
{
    switch (m.px.n) {
#define CASE(n) case n: return launch_r2c<n, RT>(m, src, u, v, w, fuse_div, dst, tw, wk, nrows, dstride, plan);
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: return false;
    }
}

template <class RT>
bool dispatch_c2r(Model &m, const typename cplx_of<RT>::type *src, RT *dst, int dst_is_field, double inv_scale,
                  const typename cplx_of<RT>::type *tw, const typename cplx_of<RT>::type *wk, int64_t nrows,
                  int sstride, int plan = 0)
{
    switch (m.px.n) {
#define CASE(n) case n: return launch_c2r<n, RT>(m, src, dst, dst_is_field, inv_scale, tw, wk, nrows, sstride, plan);
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: return false;
    }
}

template <class C>
bool dispatch_cols(int n, const C *tw, hipStream_t st, C *data, int nkx, int64_t estride, int64_t lstride,
                   unsigned ny_or_nz, int sign, double s1, double s2, double s3, int nscale, int plan = 0)
{
    switch (n) {
#define CASE(n) case n: return launch_cols<n, C>(tw, st, data, nkx, estride, lstride, ny_or_nz, sign, s1, s2, s3, nscale, plan);
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: return false;
    }
}

template <class C>
bool dispatch_zsolve(Model &m, C *data, const C *tw, int nkxs, int plan = 0)
{
    switch (m.pz.n) {
#define CASE(n) case n: return launch_zsolve<n, C>(m, data, tw, nkxs, plan);
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: return false;
    }
}

// ---- batched ensemble stages (one launch over all members) -----------------
// The member's transform length selects a compiled-in body at run time;
// every body of a stage runs with 256 threads and the stage's dynamic LDS is
// sized for its largest member.
extern __shared__ double2 ens_lds[];

__global__ __launch_bounds__(TTHREADS) void k_ens_r2c(const EnsMember *__restrict__ mem, EnsMap map,
                                                       const EnsState *__restrict__ st)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int nb = st->par ^ 1;
    const int64_t nrows = (int64_t)c.ny * c.nz;
    switch (e.n2) {
#define CASE(n)                                                                                              \
    case n:                                                                                                  \
        r2c_body<n, XROWS(n), TTHREADS, true, double>(ens_lds, lb, c, e.twx, e.wk, nullptr, e.f[0][nb],      \
                                                      e.f[1][nb], e.f[2][nb], e.FS, nrows, e.nkx);            \
        break;
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: break;
    }
}

__global__ __launch_bounds__(TTHREADS) void k_ens_c2r(const EnsMember *__restrict__ mem, EnsMap map)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int64_t nrows = (int64_t)c.ny * c.nz;
    const double inv = c.lid ? e.lxy : e.lxyz, rinv = c.lid ? e.rlxy : e.rlxyz;
    switch (e.n2) {
#define CASE(n)                                                                                              \
    case n:                                                                                                  \
        c2r_body<n, XROWS(n), TTHREADS, double>(ens_lds, lb, c, e.twx, e.wk, e.FS, e.p, 1, inv, rinv, nrows,  \
                                                e.nkx);                                                      \
        break;
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: break;
    }
}

// y-pass over every member: forward scaled by dx dy under a rigid lid
// (my_fft_2D), unscaled otherwise (the 3-D scaling is in the z solve)
// NMAX bounds the lengths compiled into the kernel: the register budget of
// the longest body sets the occupancy of every member
template <int SIGN, int NMAX>
__global__ __launch_bounds__(TTHREADS) void k_ens_cols_y(const EnsMember *__restrict__ mem, EnsMap map)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const int lb = g - map.start[mi];
    const int nbx = (e.nkx + TCOL<double2> - 1) / TCOL<double2>;
    const int bx = lb % nbx, by = lb / nbx;
    const int nscale = (SIGN < 0 && e.c.lid) ? 2 : 0;
    const double s1 = nscale ? e.c.dx : 0.0, s2 = nscale ? e.c.dy : 0.0;
    switch (e.c.ny) {
#define CASE(n)                                                                                              \
    case n:                                                                                                  \
        if constexpr (n <= NMAX && TCOLS<double2>(n) == TTHREADS)                                            \
            cols_body<n, TCOL<double2>, TTHREADS, SIGN, double2>(ens_lds, bx, by, e.twy, e.FS, e.nkx, (int64_t)e.nkx,  \
                                                        (int64_t)e.nkx * e.c.ny, s1, s2, 0.0, nscale);       \
        break;
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: break;
    }
}

template <int NMAX>
__global__ __launch_bounds__(TTHREADS) void k_ens_zsolve(const EnsMember *__restrict__ mem, EnsMap map)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const int lb = g - map.start[mi];
    const int nbx = (e.nkx + TCOL<double2> - 1) / TCOL<double2>;
    const int bx = lb % nbx, by = lb / nbx;
    switch (e.c.nz) {
#define CASE(n)                                                                                              \
    case n:                                                                                                  \
        if constexpr (n <= NMAX && TCOLS<double2>(n) == TTHREADS)                                            \
            zsolve_body<n, TCOL<double2>, TTHREADS, double2>(ens_lds, bx, by, e.twz, e.FS, e.ksqr, e.nkx, e.nkx,       \
                                                    e.c.ny, e.c.dx, e.c.dy, e.c.dz);                         \
        break;
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: break;
    }
}

// the same stages for members of the mixed form: FP32 data and tables, the
// bodies and arguments of fft_r2c_x_f, fft_c2c_y_f, fft_z_periodic_solve_f and
// fft_c2r_x_f; the x-passes read the state of the new parity and the inverse
// writes the pressure increment selected by which (see ens_c2r_f)
__global__ __launch_bounds__(TTHREADS) void k_ens_r2c_f(const EnsMember *__restrict__ mem, EnsMap map,
                                                         const EnsState *__restrict__ st)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int nb = st->par ^ 1;
    const int64_t nrows = (int64_t)c.ny * c.nz;
    float2 *lds = reinterpret_cast<float2 *>(ens_lds);
    switch (e.n2) {
#define CASE(n)                                                                                              \
    case n:                                                                                                  \
        if constexpr (plan_differs(n))                                                                       \
            if (e.fft_plan[0] == 1) {                                                                        \
                r2c_body<n, XROWS(n), TTHREADS, true, float, 1>(lds, lb, c, e.twxf, e.wkf, nullptr, e.fh[0][nb], \
                                                                e.fh[1][nb], e.fh[2][nb], e.FSf, nrows, e.nkx); \
                break;                                                                                       \
            }                                                                                                \
        r2c_body<n, XROWS(n), TTHREADS, true, float>(lds, lb, c, e.twxf, e.wkf, nullptr, e.fh[0][nb],         \
                                                     e.fh[1][nb], e.fh[2][nb], e.FSf, nrows, e.nkx);          \
        break;
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: break;
    }
}

__global__ __launch_bounds__(TTHREADS) void k_ens_c2r_f(const EnsMember *__restrict__ mem, EnsMap map,
                                                         const EnsState *__restrict__ st, int which)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const StepConsts c = e.c;
    const int lb = g - map.start[mi];
    const int64_t nrows = (int64_t)c.ny * c.nz;
    const double inv = c.lid ? e.lxy : e.lxyz, rinv = c.lid ? e.rlxy : e.rlxyz;
    float *dst = which == 0 ? e.dpr[0] : e.dpr[(st->dp_cur + which) % 3];
    float2 *lds = reinterpret_cast<float2 *>(ens_lds);
    switch (e.n2) {
#define CASE(n)                                                                                              \
    case n:                                                                                                  \
        if constexpr (plan_differs(n))                                                                       \
            if (e.fft_plan[0] == 1) {                                                                        \
                c2r_body<n, XROWS(n), TTHREADS, float, 1>(lds, lb, c, e.twxf, e.wkf, e.FSf, dst, 1, inv, rinv, \
                                                          nrows, e.nkx);                                     \
                break;                                                                                       \
            }                                                                                                \
        c2r_body<n, XROWS(n), TTHREADS, float>(lds, lb, c, e.twxf, e.wkf, e.FSf, dst, 1, inv, rinv, nrows,    \
                                               e.nkx);                                                       \
        break;
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: break;
    }
}

template <int SIGN, int NMAX>
__global__ __launch_bounds__(TTHREADS) void k_ens_cols_y_f(const EnsMember *__restrict__ mem, EnsMap map)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const int lb = g - map.start[mi];
    const int nbx = (e.nkx + TCOL<float2> - 1) / TCOL<float2>;
    const int bx = lb % nbx, by = lb / nbx;
    const int nscale = (SIGN < 0 && e.c.lid) ? 2 : 0;
    const double s1 = nscale ? e.c.dx : 0.0, s2 = nscale ? e.c.dy : 0.0;
    float2 *lds = reinterpret_cast<float2 *>(ens_lds);
    switch (e.c.ny) {
#define CASE(n)                                                                                              \
    case n:                                                                                                  \
        if constexpr (n <= NMAX && TCOLS<float2>(n) == TTHREADS) {                                           \
            if constexpr (plan_differs(n))                                                                   \
                if (e.fft_plan[1] == 1) {                                                                    \
                    cols_body<n, TCOL<float2>, TTHREADS, SIGN, float2, 1>(lds, bx, by, e.twyf, e.FSf, e.nkx,  \
                                                                          (int64_t)e.nkx, (int64_t)e.nkx * e.c.ny, \
                                                                          s1, s2, 0.0, nscale);              \
                    break;                                                                                   \
                }                                                                                            \
            cols_body<n, TCOL<float2>, TTHREADS, SIGN, float2>(lds, bx, by, e.twyf, e.FSf, e.nkx, (int64_t)e.nkx,     \
                                                       (int64_t)e.nkx * e.c.ny, s1, s2, 0.0, nscale);        \
        }                                                                                                    \
        break;
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: break;
    }
}

template <int NMAX>
__global__ __launch_bounds__(TTHREADS) void k_ens_zsolve_f(const EnsMember *__restrict__ mem, EnsMap map)
{
    const int g = xcd_remap((int)blockIdx.x, map.start[map.n]);
    const int mi = ens_find(map, g);
    const EnsMember &e = mem[mi];
    const int lb = g - map.start[mi];
    const int nbx = (e.nkx + TCOL<float2> - 1) / TCOL<float2>;
    const int bx = lb % nbx, by = lb / nbx;
    float2 *lds = reinterpret_cast<float2 *>(ens_lds);
    switch (e.c.nz) {
#define CASE(n)                                                                                              \
    case n:                                                                                                  \
        if constexpr (n <= NMAX && TCOLS<float2>(n) == TTHREADS) {                                           \
            if constexpr (plan_differs(n))                                                                   \
                if (e.fft_plan[2] == 1) {                                                                    \
                    zsolve_body<n, TCOL<float2>, TTHREADS, float2, 1>(lds, bx, by, e.twzf, e.FSf, e.ksqr, e.nkx, \
                                                                      e.nkx, e.c.ny, e.c.dx, e.c.dy, e.c.dz); \
                    break;                                                                                   \
                }                                                                                            \
            zsolve_body<n, TCOL<float2>, TTHREADS, float2>(lds, bx, by, e.twzf, e.FSf, e.ksqr, e.nkx, e.nkx, e.c.ny,  \
                                                   e.c.dx, e.c.dy, e.c.dz);                                  \
        }                                                                                                    \
        break;
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: break;
    }
}

}  // namespace ft

// rows per block of an x-pass so that every radix's butterflies fit in
// MAXE staged values per thread (256 threads)
int rows_for(const FftPlan &pl, int threads)
{
    int rows = 64;
    for (; rows > 1; --rows) {
        bool ok = true;
        for (int f = 0; f < pl.nf; ++f) {
            const int R = pl.fac[f];
            const int maxb = MAXE / R + (MAXE % R ? 1 : 0);
            if ((int64_t)(pl.n / R) * rows > (int64_t)maxb * threads) ok = false;
        }
        if (ok && (int64_t)rows * (pl.n + 1) * 2 * 16 <= 64 * 1024) break;
    }
    return rows;
}

int threads_for_cols(const FftPlan &pl, int TX)
{
    for (int T = 64; T <= 512; T *= 2) {
        bool ok = true;
        for (int f = 0; f < pl.nf; ++f) {
            const int R = pl.fac[f];
            const int maxb = MAXE / R + (MAXE % R ? 1 : 0);
            if ((int64_t)(pl.n / R) * TX > (int64_t)maxb * T) ok = false;
        }
        if (ok) return T;
    }
    return -1;
}

constexpr int TXC = 8;   // kx columns per work-group in the y and z passes

// dynamic LDS above 64 KiB must be requested per kernel
template <class K>
void allow_lds(K kernel, size_t bytes)
{
    if (bytes > 64 * 1024)
        PS3D_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

}  // namespace

void fft_plan_init(Model &m, FftPlan &pl, int n)
{
    pl.n = n;
    pl.nf = 0;
    int r = n;
    while (r % 4 == 0) { pl.fac[pl.nf++] = 4; r /= 4; }
    while (r % 2 == 0) { pl.fac[pl.nf++] = 2; r /= 2; }
    while (r % 3 == 0) { pl.fac[pl.nf++] = 3; r /= 3; }
    while (r % 5 == 0) { pl.fac[pl.nf++] = 5; r /= 5; }
    PS3D_REQUIRE(r == 1, "FFT length " + std::to_string(n) + " has a prime factor other than 2, 3, 5");
    PS3D_REQUIRE(pl.nf <= MAX_FAC, "FFT length too long");
    std::vector<double2> tw(n);
    for (int q = 0; q < n; ++q) tw[q] = host_twiddle(q, n);
    double2 *d = nullptr;
    PS3D_HIP_CHECK(hipMalloc(&d, sizeof(double2) * n));
    m.allocations.push_back(d);
    PS3D_HIP_CHECK(hipMemcpy(d, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    pl.tw = d;
    std::vector<float2> twf(n);
    for (int q = 0; q < n; ++q) twf[q] = make_float2((float)tw[q].x, (float)tw[q].y);
    float2 *df = nullptr;
    PS3D_HIP_CHECK(hipMalloc(&df, sizeof(float2) * n));
    m.allocations.push_back(df);
    PS3D_HIP_CHECK(hipMemcpy(df, twf.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    pl.twf = df;
    PS3D_REQUIRE(threads_for_cols(pl, TXC) > 0, "FFT length too long for one work-group");
}

void fft_r2c_x(Model &m, const double *src, int /*src_is_field*/, int fuse_div, double2 *dst)
{
    const Grid &g = m.g;
    const int rows = rows_for(m.px, 256);
    const int64_t nrows = (int64_t)g.ny * g.nz;
    const size_t lds = (size_t)rows * (m.px.n + 1) * sizeof(double2);
    const dim3 grid((unsigned)((nrows + rows - 1) / rows));
    ProfScope ps(m, "fft_x_r2c");
    if (ft::dispatch_r2c<double>(m, src, m.u, m.v, m.w, fuse_div, dst, m.px.tw, m.tw_r2c, nrows, g.nkx)) {
        PS3D_HIP_CHECK(hipGetLastError());
        return;
    }
    if (fuse_div)
        hipLaunchKernelGGL(k_r2c_x<true>, grid, dim3(256), lds, m.stream, m.c, m.px, m.tw_r2c, nullptr, m.u,
                           m.v, m.w, dst, rows, nrows);
    else
        hipLaunchKernelGGL(k_r2c_x<false>, grid, dim3(256), lds, m.stream, m.c, m.px, m.tw_r2c, src, nullptr,
                           nullptr, nullptr, dst, rows, nrows);
    PS3D_HIP_CHECK(hipGetLastError());
}

void fft_c2r_x(Model &m, const double2 *src, double *dst, int dst_is_field, double inv_scale)
{
    const Grid &g = m.g;
    const int rows = rows_for(m.px, 256);
    const int64_t nrows = (int64_t)g.ny * g.nz;
    const size_t lds = (size_t)2 * rows * (m.px.n + 1) * sizeof(double2);
    const dim3 grid((unsigned)((nrows + rows - 1) / rows));
    ProfScope ps(m, "fft_x_c2r");
    if (ft::dispatch_c2r<double>(m, src, dst, dst_is_field, inv_scale, m.px.tw, m.tw_r2c, nrows, g.nkx)) {
        PS3D_HIP_CHECK(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_c2r_x, grid, dim3(256), lds, m.stream, m.c, m.px, m.tw_r2c, src, dst, dst_is_field,
                       inv_scale, 1.0 / inv_scale, rows, nrows);
    PS3D_HIP_CHECK(hipGetLastError());
}

void fft_c2c_y(Model &m, double2 *data, int sign, double scale1, double scale2)
{
    const Grid &g = m.g;
    const int T = threads_for_cols(m.py, TXC);
    const size_t lds = (size_t)m.py.n * TXC * sizeof(double2);
    const dim3 grid((unsigned)((g.nkx + TXC - 1) / TXC), (unsigned)g.nz);
    const int nscale = (scale1 != 0.0) ? 2 : 0;
    allow_lds(k_c2c_cols, lds);
    ProfScope ps(m, sign < 0 ? "fft_y_fwd" : "fft_y_inv");
    if (ft::dispatch_cols<double2>(m.py.n, m.py.tw, m.stream, data, g.nkx, (int64_t)g.nkx, (int64_t)g.nkx * g.ny,
                                   (unsigned)g.nz, sign, scale1, scale2, 0.0, nscale)) {
        PS3D_HIP_CHECK(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_c2c_cols, grid, dim3(T), lds, m.stream, m.py, data, g.nkx, TXC, (int64_t)g.nkx,
                       (int64_t)g.nkx * g.ny, sign, scale1, scale2, 0.0, nscale);
    PS3D_HIP_CHECK(hipGetLastError());
}

void fft_c2c_z(Model &m, double2 *data, int sign, double s1, double s2, double s3)
{
    const Grid &g = m.g;
    const int T = threads_for_cols(m.pz, TXC);
    const size_t lds = (size_t)m.pz.n * TXC * sizeof(double2);
    const dim3 grid((unsigned)((g.nkx + TXC - 1) / TXC), (unsigned)g.ny);
    const int nscale = (s1 != 0.0) ? 3 : 0;
    allow_lds(k_c2c_cols, lds);
    ProfScope ps(m, sign < 0 ? "fft_z_fwd" : "fft_z_inv");
    if (ft::dispatch_cols<double2>(m.pz.n, m.pz.tw, m.stream, data, g.nkx, (int64_t)g.nkx * g.ny, (int64_t)g.nkx,
                                   (unsigned)g.ny, sign, s1, s2, s3, nscale)) {
        PS3D_HIP_CHECK(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_c2c_cols, grid, dim3(T), lds, m.stream, m.pz, data, g.nkx, TXC,
                       (int64_t)g.nkx * g.ny, (int64_t)g.nkx, sign, s1, s2, s3, nscale);
    PS3D_HIP_CHECK(hipGetLastError());
}

void fft_z_periodic_solve(Model &m, double2 *data)
{
    const Grid &g = m.g;
    const int T = threads_for_cols(m.pz, TXC);
    const size_t lds = (size_t)m.pz.n * TXC * sizeof(double2);
    const dim3 grid((unsigned)((g.nkx + TXC - 1) / TXC), (unsigned)g.ny);
    allow_lds(k_z_periodic_solve, lds);
    ProfScope ps(m, "z_solve");
    if (ft::dispatch_zsolve<double2>(m, data, m.pz.tw, g.nkx)) {
        PS3D_HIP_CHECK(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_z_periodic_solve, grid, dim3(T), lds, m.stream, m.pz, data, m.ksqr, g.nkx, g.ny, TXC,
                       m.prm.dx, m.prm.dy, m.prm.dz);
    PS3D_HIP_CHECK(hipGetLastError());
}

// ---- FP32 transforms of the mixed form (compiled-in lengths only) ----------

void fft_r2c_x_f(Model &m, const float *u, const float *v, const float *w, float2 *dst)
{
    const Grid &g = m.g;
    ProfScope ps(m, "fft_x_r2c");
    PS3D_REQUIRE(ft::dispatch_r2c<float>(m, nullptr, u, v, w, 1, dst, m.px.twf, m.twf_r2c, (int64_t)g.ny * g.nz,
                                         g.nkx, m.fft_plan[0]),
                 "mixed precision: nx/2 must be one of the compiled-in FFT lengths");
    PS3D_HIP_CHECK(hipGetLastError());
}

void fft_c2r_x_f(Model &m, const float2 *src, float *dst, double inv_scale)
{
    const Grid &g = m.g;
    ProfScope ps(m, "fft_x_c2r");
    PS3D_REQUIRE(ft::dispatch_c2r<float>(m, src, dst, 1, inv_scale, m.px.twf, m.twf_r2c, (int64_t)g.ny * g.nz, g.nkx,
                                         m.fft_plan[0]),
                 "mixed precision: nx/2 must be one of the compiled-in FFT lengths");
    PS3D_HIP_CHECK(hipGetLastError());
}

void fft_c2c_y_f(Model &m, float2 *data, int sign, double scale1, double scale2)
{
    const Grid &g = m.g;
    const int nscale = (scale1 != 0.0) ? 2 : 0;
    ProfScope ps(m, sign < 0 ? "fft_y_fwd" : "fft_y_inv");
    PS3D_REQUIRE(ft::dispatch_cols<float2>(m.py.n, m.py.twf, m.stream, data, g.nkx, (int64_t)g.nkx,
                                           (int64_t)g.nkx * g.ny, (unsigned)g.nz, sign, scale1, scale2, 0.0, nscale,
                                           m.fft_plan[1]),
                 "mixed precision: ny must be one of the compiled-in FFT lengths");
    PS3D_HIP_CHECK(hipGetLastError());
}

void fft_z_periodic_solve_f(Model &m, float2 *data)
{
    ProfScope ps(m, "z_solve");
    PS3D_REQUIRE(ft::dispatch_zsolve<float2>(m, data, m.pz.twf, m.g.nkx, m.fft_plan[2]),
                 "mixed precision: nz must be one of the compiled-in FFT lengths");
    PS3D_HIP_CHECK(hipGetLastError());
}

bool fft_sizes_compiled(int n)
{
    switch (n) {
#define CASE(n) case n: return true;
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: return false;
    }
}

bool fft_plan_choice(int n) { return fft_sizes_compiled(n) && ft::plan_differs(n); }

std::string fft_sizes_text()
{
    std::string t;
#define ITEM(n) t += (t.empty() ? "" : ", ") + std::to_string(n);
    PS3D_FFT_SIZES(ITEM)
#undef ITEM
    return t;
}

// ---- batched ensemble stages ------------------------------------------------
bool ens_fft_size_ok(int n, bool column_pass)
{
    switch (n) {
#define CASE(n) case n: return !column_pass || (ft::TCOLS<double2>(n) == ft::TTHREADS && ft::TCOLS<float2>(n) == ft::TTHREADS);
        PS3D_FFT_SIZES(CASE)
#undef CASE
    default: return false;
    }
}

size_t ens_x_lds(int n2) { return (size_t)ft::XROWS(n2) * (n2 + 1) * sizeof(double2); }

// the same for both forms: TCOL<C> * sizeof(C) = 64 bytes per row
size_t ens_cols_lds(int n) { return (size_t)n * 64; }

int ens_r2c_blocks(const EnsMember &e)
{
    const int rows = ft::XROWS(e.n2);
    return (int)(((int64_t)e.c.ny * e.c.nz + rows - 1) / rows);
}

int ens_cols_blocks(const EnsMember &e, int /*n_len*/, int lines)
{
    const int tc = e.FSf ? ft::TCOL<float2> : ft::TCOL<double2>;
    return ((e.nkx + tc - 1) / tc) * lines;
}

template <class K>
static void ens_lds_attr(K kern, size_t lds)
{
    if (lds > 65536) PS3D_HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

void ens_r2c(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, size_t lds)
{
    if (!map.start[map.n]) return;
    ens_lds_attr(ft::k_ens_r2c, lds);
    hipLaunchKernelGGL(ft::k_ens_r2c, dim3((unsigned)map.start[map.n]), dim3(ft::TTHREADS), lds, s, mem, map, st);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_c2r(hipStream_t s, const EnsMember *mem, const EnsMap &map, size_t lds)
{
    if (!map.start[map.n]) return;
    ens_lds_attr(ft::k_ens_c2r, lds);
    hipLaunchKernelGGL(ft::k_ens_c2r, dim3((unsigned)map.start[map.n]), dim3(ft::TTHREADS), lds, s, mem, map);
    PS3D_HIP_CHECK(hipGetLastError());
}

template <int SIGN, int NMAX>
static void cols_y_launch(hipStream_t s, const EnsMember *mem, const EnsMap &map, size_t lds)
{
    ens_lds_attr(ft::k_ens_cols_y<SIGN, NMAX>, lds);
    hipLaunchKernelGGL((ft::k_ens_cols_y<SIGN, NMAX>), dim3((unsigned)map.start[map.n]), dim3(ft::TTHREADS), lds, s,
                       mem, map);
}

template <int SIGN>
static void cols_y_nmax(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nmax, size_t lds)
{
    if (nmax <= 64) cols_y_launch<SIGN, 64>(s, mem, map, lds);
    else if (nmax <= 128) cols_y_launch<SIGN, 128>(s, mem, map, lds);
    else if (nmax <= 256) cols_y_launch<SIGN, 256>(s, mem, map, lds);
    else cols_y_launch<SIGN, 512>(s, mem, map, lds);
}

void ens_cols_y(hipStream_t s, const EnsMember *mem, const EnsMap &map, int sign, int nmax, size_t lds)
{
    if (!map.start[map.n]) return;
    if (sign < 0) cols_y_nmax<-1>(s, mem, map, nmax, lds);
    else cols_y_nmax<+1>(s, mem, map, nmax, lds);
    PS3D_HIP_CHECK(hipGetLastError());
}

template <int NMAX>
static void zsolve_launch(hipStream_t s, const EnsMember *mem, const EnsMap &map, size_t lds)
{
    ens_lds_attr(ft::k_ens_zsolve<NMAX>, lds);
    hipLaunchKernelGGL(ft::k_ens_zsolve<NMAX>, dim3((unsigned)map.start[map.n]), dim3(ft::TTHREADS), lds, s, mem,
                       map);
}

void ens_zsolve(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nmax, size_t lds)
{
    if (!map.start[map.n]) return;
    if (nmax <= 64) zsolve_launch<64>(s, mem, map, lds);
    else if (nmax <= 128) zsolve_launch<128>(s, mem, map, lds);
    else if (nmax <= 256) zsolve_launch<256>(s, mem, map, lds);
    else zsolve_launch<512>(s, mem, map, lds);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_r2c_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, size_t lds)
{
    if (!map.start[map.n]) return;
    ens_lds_attr(ft::k_ens_r2c_f, lds);
    hipLaunchKernelGGL(ft::k_ens_r2c_f, dim3((unsigned)map.start[map.n]), dim3(ft::TTHREADS), lds, s, mem, map, st);
    PS3D_HIP_CHECK(hipGetLastError());
}

void ens_c2r_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, const EnsState *st, int which, size_t lds)
{
    if (!map.start[map.n]) return;
    ens_lds_attr(ft::k_ens_c2r_f, lds);
    hipLaunchKernelGGL(ft::k_ens_c2r_f, dim3((unsigned)map.start[map.n]), dim3(ft::TTHREADS), lds, s, mem, map, st,
                       which);
    PS3D_HIP_CHECK(hipGetLastError());
}

template <int SIGN, int NMAX>
static void cols_y_launch_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, size_t lds)
{
    ens_lds_attr(ft::k_ens_cols_y_f<SIGN, NMAX>, lds);
    hipLaunchKernelGGL((ft::k_ens_cols_y_f<SIGN, NMAX>), dim3((unsigned)map.start[map.n]), dim3(ft::TTHREADS), lds,
                       s, mem, map);
}

template <int SIGN>
static void cols_y_nmax_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nmax, size_t lds)
{
    if (nmax <= 64) cols_y_launch_f<SIGN, 64>(s, mem, map, lds);
    else if (nmax <= 128) cols_y_launch_f<SIGN, 128>(s, mem, map, lds);
    else if (nmax <= 256) cols_y_launch_f<SIGN, 256>(s, mem, map, lds);
    else cols_y_launch_f<SIGN, 512>(s, mem, map, lds);
}

void ens_cols_y_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, int sign, int nmax, size_t lds)
{
    if (!map.start[map.n]) return;
    if (sign < 0) cols_y_nmax_f<-1>(s, mem, map, nmax, lds);
    else cols_y_nmax_f<+1>(s, mem, map, nmax, lds);
    PS3D_HIP_CHECK(hipGetLastError());
}

template <int NMAX>
static void zsolve_launch_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, size_t lds)
{
    ens_lds_attr(ft::k_ens_zsolve_f<NMAX>, lds);
    hipLaunchKernelGGL(ft::k_ens_zsolve_f<NMAX>, dim3((unsigned)map.start[map.n]), dim3(ft::TTHREADS), lds, s, mem,
                       map);
}

void ens_zsolve_f(hipStream_t s, const EnsMember *mem, const EnsMap &map, int nmax, size_t lds)
{
    if (!map.start[map.n]) return;
    if (nmax <= 64) zsolve_launch_f<64>(s, mem, map, lds);
    else if (nmax <= 128) zsolve_launch_f<128>(s, mem, map, lds);
    else if (nmax <= 256) zsolve_launch_f<256>(s, mem, map, lds);
    else zsolve_launch_f<512>(s, mem, map, lds);
    PS3D_HIP_CHECK(hipGetLastError());
}

}  // namespace ps3d
