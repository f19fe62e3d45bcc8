/*
 * ps3d_oracle.c -- CPU restatement of the ps3D model time step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the thing
 * measured or shipped.  See ps3d_oracle.h for who may load it.
 *
 * Every routine follows one reference routine (cited per function) loop by
 * loop, with the same index ranges and the same association of floating-point
 * operations, so that on the same inputs it reproduces the reference to the
 * last bit where the reference is compiled without FMA contraction.  Build
 * with -ffp-contract=off.
 *
 * The FFT is not the reference's (P3DFFT over FFTW, absent here; FFTW picks
 * its factorisation at run time, so there is no fixed operation order to
 * follow).  It is a Stockham autosort mixed-radix DFT: radix-4 passes first,
 * then 2, 3, 5, with the twiddle table tw[q] = exp(-2 pi i q/n) and the
 * butterfly formulas written out below; real transforms pack x[2m] + i x[2m+1]
 * into a half-length complex DFT.  Conventions are those of my_fft.f90:159-205:
 * unnormalised, e^{-i} forward, c2r ignoring the imaginary parts of the kx=0
 * and kx=nx/2 modes.  The operation order is the one the GPU kernels use, so
 * the oracle and the GPU agree bit for bit; the transform itself is checked
 * against numpy.fft and the DFT definition.
 */
#include "ps3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

typedef struct {
    int n;
    int nfac;
    int fac[32];
    cplx *tw;      /* tw[t] = exp(-2 pi i t / n), t = 0..n-1 */
    cplx *scratch; /* 2n */
} plan1d;

struct orc_model {
    orc_params p;
    int64_t X, Y, Z, H;       /* padded extents and field size */
    int64_t nkx, S;           /* nx/2+1 and spectral size */
    double *f[ORC_NFIELDS];
    int64_t len[ORC_NFIELDS];
    double *aloc;             /* nx*ny*nz */
    double *del2;             /* H, the automatic array of mixing.f90 */
    plan1d px, py, pz;        /* lengths nx/2, ny, nz */
    cplx *wk;                 /* exp(-2 pi i k/nx), k = 0..nx/2 */
    cplx *row;                /* 2*max(nx,ny,nz) + 2 */
    orc_forcing_fn forcing;   /* set_forcing (integrate.f90:82), NULL: empty */
    void *forcing_user;
    double wm_omega;          /* orc_set_wavemaker: amplitudes owned by the caller */
    const double *wm_re[4], *wm_im[4];
    const double *bg_u, *bg_b; /* orc_set_background: owned by the caller */
    const double *bf[4];
};

#define ONX 2

static inline int64_t IX(const orc_model *m, int i, int j, int k)
{
    return (int64_t)(i + 1) + m->X * ((int64_t)(j + 1) + m->Y * (int64_t)(k + 1));
}
#define A(arr, i, j, k) (arr)[IX(m, (i), (j), (k))]

/* ------------------------------------------------------------------------ */
/* 1-D complex DFT                                                           */
/* ------------------------------------------------------------------------ */

/* exp(-2 pi i q/n).  sincos() is called explicitly: compilers fuse a cos()
   and sin() of one argument into sincos() or not as they please, and the two
   can differ in the last bit, so the GPU library calls the same routine. */
static void twiddle(int q, int n, cplx *w)
{
    const double ang = 2.0 * M_PI * (double)q / (double)n;
    double s, c;
    sincos(ang, &s, &c);
    w->re = c;
    w->im = -s;
}

static void plan_init(plan1d *pl, int n)
{
    pl->n = n;
    pl->nfac = 0;
    int r = n;
    while (r % 4 == 0) { pl->fac[pl->nfac++] = 4; r /= 4; }
    while (r % 2 == 0) { pl->fac[pl->nfac++] = 2; r /= 2; }
    while (r % 3 == 0) { pl->fac[pl->nfac++] = 3; r /= 3; }
    while (r % 5 == 0) { pl->fac[pl->nfac++] = 5; r /= 5; }
    for (int f = 7; r > 1; f += 2)           /* other primes: plain DFT radices */
        while (r % f == 0) { pl->fac[pl->nfac++] = f; r /= f; }
    pl->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int q = 0; q < n; ++q) twiddle(q, n, &pl->tw[q]);
    pl->scratch = (cplx *)malloc(sizeof(cplx) * (size_t)(2 * n + 2));
}

static void plan_free(plan1d *pl)
{
    free(pl->tw);
    free(pl->scratch);
}

static inline cplx c_mul(cplx a, cplx b)
{
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cplx c_add(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx c_sub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx c_scale(cplx a, double s) { cplx r = {a.re * s, a.im * s}; return r; }
/* i*a for sign = +1, -i*a for sign = -1 */
static inline cplx c_isign(cplx a, int sign)
{
    cplx r;
    if (sign > 0) { r.re = -a.im; r.im = a.re; }
    else { r.re = a.im; r.im = -a.re; }
    return r;
}

/* R-point DFT in place, v[k] = sum_r v[r] exp(sign 2 pi i r k / R) */
static void dft_small(cplx *v, int R, int sign, const plan1d *pl)
{
    if (R == 2) {
        const cplx a = v[0], b = v[1];
        v[0] = c_add(a, b);
        v[1] = c_sub(a, b);
    } else if (R == 3) {
        const double c = -0.5, s = 0.86602540378443864676;
        const cplx t = c_add(v[1], v[2]);
        const cplx d = c_sub(v[1], v[2]);
        const cplx mm = {v[0].re + c * t.re, v[0].im + c * t.im};
        const cplx q = c_isign(c_scale(d, s), sign);
        v[0] = c_add(v[0], t);
        v[1] = c_add(mm, q);
        v[2] = c_sub(mm, q);
    } else if (R == 4) {
        const cplx a = c_add(v[0], v[2]), bq = c_sub(v[0], v[2]);
        const cplx c = c_add(v[1], v[3]), d = c_isign(c_sub(v[1], v[3]), sign);
        v[0] = c_add(a, c);
        v[2] = c_sub(a, c);
        v[1] = c_add(bq, d);
        v[3] = c_sub(bq, d);
    } else if (R == 5) {
        const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
        const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
        const cplx a14 = c_add(v[1], v[4]), d14 = c_sub(v[1], v[4]);
        const cplx a23 = c_add(v[2], v[3]), d23 = c_sub(v[2], v[3]);
        const cplx m1 = {v[0].re + c1 * a14.re + c2 * a23.re, v[0].im + c1 * a14.im + c2 * a23.im};
        const cplx m2 = {v[0].re + c2 * a14.re + c1 * a23.re, v[0].im + c2 * a14.im + c1 * a23.im};
        const cplx t1 = {s1 * d14.re + s2 * d23.re, s1 * d14.im + s2 * d23.im};
        const cplx t2 = {s2 * d14.re - s1 * d23.re, s2 * d14.im - s1 * d23.im};
        const cplx q1 = c_isign(t1, sign), q2 = c_isign(t2, sign);
        v[0].re = v[0].re + a14.re + a23.re;
        v[0].im = v[0].im + a14.im + a23.im;
        v[1] = c_add(m1, q1);
        v[4] = c_sub(m1, q1);
        v[2] = c_add(m2, q2);
        v[3] = c_sub(m2, q2);
    } else {
        /* other prime radices: direct sum with table twiddles */
        cplx out[64];
        const int st = pl->n / R;
        for (int k = 0; k < R; ++k) {
            cplx acc = v[0];
            for (int r = 1; r < R; ++r) {
                cplx w = pl->tw[((r * k) % R) * st];
                if (sign > 0) w.im = -w.im;
                acc = c_add(acc, c_mul(v[r], w));
            }
            out[k] = acc;
        }
        for (int k = 0; k < R; ++k) v[k] = out[k];
    }
}

/* Stockham autosort DFT of n contiguous values: pass f with radix R and
   Ns = product of the earlier radices reads v[r] = a[j + r n/R] for butterfly j,
   multiplies v[r] by W^{r (j mod Ns) n/(Ns R)}, runs the R-point DFT and
   writes b[(j/Ns) Ns R + j mod Ns + r Ns]. */
static void stockham(const plan1d *pl, cplx *a, cplx *b, int sign)
{
    const int N = pl->n;
    int Ns = 1;
    cplx *src = a, *dst = b;
    for (int f = 0; f < pl->nfac; ++f) {
        const int R = pl->fac[f];
        const int nb = N / R;
        const int twstride = N / (Ns * R);
        cplx v[64];
        for (int j = 0; j < nb; ++j) {
            for (int r = 0; r < R; ++r) v[r] = src[j + r * nb];
            const int jm = j % Ns;
            if (Ns > 1)
                for (int r = 1; r < R; ++r) {
                    cplx w = pl->tw[r * jm * twstride];
                    if (sign > 0) w.im = -w.im;
                    v[r] = c_mul(v[r], w);
                }
            dft_small(v, R, sign, pl);
            const int base = (j / Ns) * Ns * R + jm;
            for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
        }
        cplx *t = src; src = dst; dst = t;
        Ns *= R;
    }
    if (src != a) memcpy(a, src, sizeof(cplx) * (size_t)N);
}

static void cfft_strided(const plan1d *pl, cplx *data, int64_t stride, int sign)
{
    const int n = pl->n;
    cplx *in = pl->scratch;
    cplx *tmp = pl->scratch + n;
    for (int t = 0; t < n; ++t) in[t] = data[t * stride];
    stockham(pl, in, tmp, sign);
    for (int t = 0; t < n; ++t) data[t * stride] = in[t];
}

void orc_cfft(int n, double *data, int sign)
{
    plan1d pl;
    plan_init(&pl, n);
    cfft_strided(&pl, (cplx *)data, 1, sign);
    plan_free(&pl);
}

/* ------------------------------------------------------------------------ */
/* model                                                                     */
/* ------------------------------------------------------------------------ */

orc_model *orc_create(const orc_params *prm)
{
    orc_model *m = (orc_model *)calloc(1, sizeof(orc_model));
    m->p = *prm;
    const int nx = prm->nx, ny = prm->ny, nz = prm->nz;
    m->X = nx + 2 * ONX;
    m->Y = ny + 2 * ONX;
    m->Z = nz + 2 * ONX;
    m->H = m->X * m->Y * m->Z;
    m->nkx = nx / 2 + 1;
    m->S = m->nkx * ny * nz;
    for (int id = ORC_U; id <= ORC_P; ++id) m->len[id] = m->H;
    for (int id = ORC_DU; id <= ORC_DB; ++id) m->len[id] = 3 * m->H;
    for (int id = ORC_FLUX_EAST; id <= ORC_FLUX_TOP; ++id) m->len[id] = m->H;
    m->len[ORC_FS] = 2 * m->S;
    m->len[ORC_KSQR] = prm->enable_vertical_boundaries ? m->nkx * ny : m->S;
    m->len[ORC_KX] = m->nkx;
    m->len[ORC_KY] = ny;
    m->len[ORC_KZ] = nz;
    for (int id = 0; id < ORC_NFIELDS; ++id)
        m->f[id] = (double *)calloc((size_t)m->len[id], sizeof(double));
    m->aloc = (double *)calloc((size_t)nx * ny * nz, sizeof(double));
    m->del2 = (double *)calloc((size_t)m->H, sizeof(double));
    plan_init(&m->px, nx / 2);
    plan_init(&m->py, ny);
    plan_init(&m->pz, nz);
    m->wk = (cplx *)malloc(sizeof(cplx) * (size_t)m->nkx);
    for (int k = 0; k < m->nkx; ++k) twiddle(k, nx, &m->wk[k]);
    int nmax = nx > ny ? nx : ny;
    nmax = nmax > nz ? nmax : nz;
    m->row = (cplx *)malloc(sizeof(cplx) * (size_t)(2 * nmax + 2));
    orc_wavenumbers(m);
    return m;
}

void orc_destroy(orc_model *m)
{
    if (!m) return;
    for (int id = 0; id < ORC_NFIELDS; ++id) free(m->f[id]);
    free(m->aloc);
    free(m->del2);
    plan_free(&m->px);
    plan_free(&m->py);
    plan_free(&m->pz);
    free(m->wk);
    free(m->row);
    free(m);
}

double *orc_ptr(orc_model *m, int field)
{
    if (field < 0 || field >= ORC_NFIELDS) return NULL;
    return m->f[field];
}

int64_t orc_field_len(const orc_model *m, int field)
{
    if (field < 0 || field >= ORC_NFIELDS) return -1;
    return m->len[field];
}

/* eigenvectors.f90:224-229 */
double orc_hat_ksqr(double k, double dx)
{
    return 2.0 * (1.0 - cos(k * dx)) / (dx * dx);
}

/* my_fft.f90:87-156 (wavenumber grid and ksqr; the vertical modes phi, psi,
   psib belong to the balancing diagnostic and are not needed here) */
void orc_wavenumbers(orc_model *m)
{
    const orc_params *p = &m->p;
    const double pi = acos(0.0) * 2.0;   /* main_module.f90:116 */
    double *kx = m->f[ORC_KX], *ky = m->f[ORC_KY], *kz = m->f[ORC_KZ];
    for (int i = 1; i <= p->nx / 2 + 1; ++i) kx[i - 1] = (double)(2 * (i - 1)) * pi / p->Lx;
    for (int j = 1; j <= (p->ny + 1) / 2; ++j) ky[j - 1] = (double)(2 * (j - 1)) * pi / p->Ly;
    for (int j = (p->ny + 1) / 2 + 1; j <= p->ny; ++j)
        ky[j - 1] = -((double)(2 * (p->ny - j + 1)) * pi / p->Ly);
    if (p->enable_vertical_boundaries) {
        for (int k = 1; k <= p->nz; ++k) kz[k - 1] = (double)(float)(k - 1.0f) * pi / p->Lz;
    } else {
        for (int k = 1; k <= (p->nz + 1) / 2; ++k) kz[k - 1] = (double)(2 * (k - 1)) * pi / p->Lz;
        for (int k = (p->nz + 1) / 2 + 1; k <= p->nz; ++k)
            kz[k - 1] = -((double)(2 * (p->nz - k + 1)) * pi / p->Lz);
    }
    double *ks = m->f[ORC_KSQR];
    const int64_t nkx = m->nkx;
    if (p->enable_vertical_boundaries) {
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < nkx; ++i)
                ks[i + nkx * j] = orc_hat_ksqr(kx[i], p->dx) + orc_hat_ksqr(ky[j], p->dy);
    } else {
        for (int k = 0; k < p->nz; ++k)
            for (int j = 0; j < p->ny; ++j)
                for (int i = 0; i < nkx; ++i)
                    ks[i + nkx * (j + (int64_t)p->ny * k)] =
                        orc_hat_ksqr(kx[i], p->dx) + orc_hat_ksqr(ky[j], p->dy) +
                        orc_hat_ksqr(kz[k], p->dz) / p->dsqr;
    }
}

/* ------------------------------------------------------------------------ */
/* advection.f90                                                             */
/* ------------------------------------------------------------------------ */

static void lid_zero_flux_top(orc_model *m, double *ft, int top_from)
{
    /* flux_top(:,:,1-onx:0) = 0 ; flux_top(:,:,top_from:nz+onx) = 0 */
    const orc_params *p = &m->p;
    if (!p->enable_vertical_boundaries) return;
    for (int k = 1 - ONX; k <= 0; ++k)
        memset(&A(ft, -1, -1, k), 0, sizeof(double) * (size_t)(m->X * m->Y));
    for (int k = top_from; k <= p->nz + ONX; ++k)
        memset(&A(ft, -1, -1, k), 0, sizeof(double) * (size_t)(m->X * m->Y));
}

/* advection.f90:2-48 */
void orc_buoyancy_advection_2nd(orc_model *m, double *db_loc)
{
    const orc_params *p = &m->p;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double *u = m->f[ORC_U], *v = m->f[ORC_V], *w = m->f[ORC_W], *b = m->f[ORC_B];
    double *fe = m->f[ORC_FLUX_EAST], *fn = m->f[ORC_FLUX_NORTH], *ft = m->f[ORC_FLUX_TOP];
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i)
                A(fe, i, j, k) = 0.5 * (A(b, i, j, k) + A(b, i + 1, j, k)) * A(u, i, j, k);
    for (int k = 1; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(fn, i, j, k) = 0.5 * (A(b, i, j, k) + A(b, i, j + 1, k)) * A(v, i, j, k);
    for (int k = 0; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(ft, i, j, k) = 0.5 * (A(b, i, j, k) + A(b, i, j, k + 1)) * A(w, i, j, k);
    lid_zero_flux_top(m, ft, nz);
    const double Ro = p->Ro, dx = p->dx, dy = p->dy, dz = p->dz;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(db_loc, i, j, k) = -(Ro * ((A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx +
                                             (A(fn, i, j, k) - A(fn, i, j - 1, k)) / dy +
                                             (A(ft, i, j, k) - A(ft, i, j, k - 1)) / dz));
}

/* advection.f90:53-184 */
void orc_momentum_advection_2nd(orc_model *m, double *du_loc, double *dv_loc, double *dw_loc)
{
    const orc_params *p = &m->p;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double *u = m->f[ORC_U], *v = m->f[ORC_V], *w = m->f[ORC_W];
    double *fe = m->f[ORC_FLUX_EAST], *fn = m->f[ORC_FLUX_NORTH], *ft = m->f[ORC_FLUX_TOP];
    const double Ro = p->Ro, dx = p->dx, dy = p->dy, dz = p->dz, dsqr = p->dsqr;

    /* zonal momentum (:62-100) */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i)
                A(fe, i, j, k) = 0.25 * (A(u, i, j, k) + A(u, i + 1, j, k)) *
                                 (A(u, i + 1, j, k) + A(u, i, j, k));
    for (int k = 1; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(fn, i, j, k) = 0.25 * (A(u, i, j, k) + A(u, i, j + 1, k)) *
                                 (A(v, i + 1, j, k) + A(v, i, j, k));
    for (int k = 0; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(ft, i, j, k) = 0.25 * (A(u, i, j, k + 1) + A(u, i, j, k)) *
                                 (A(w, i, j, k) + A(w, i + 1, j, k));
    lid_zero_flux_top(m, ft, nz);
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(du_loc, i, j, k) = -(Ro * ((A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx +
                                             (A(fn, i, j, k) - A(fn, i, j - 1, k)) / dy +
                                             (A(ft, i, j, k) - A(ft, i, j, k - 1)) / dz));

    /* meridional momentum (:104-140) */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i)
                A(fe, i, j, k) = 0.25 * (A(v, i, j, k) + A(v, i + 1, j, k)) *
                                 (A(u, i, j + 1, k) + A(u, i, j, k));
    for (int k = 1; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(fn, i, j, k) = 0.25 * (A(v, i, j, k) + A(v, i, j + 1, k)) *
                                 (A(v, i, j + 1, k) + A(v, i, j, k));
    for (int k = 0; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(ft, i, j, k) = 0.25 * (A(v, i, j, k + 1) + A(v, i, j, k)) *
                                 (A(w, i, j, k) + A(w, i, j + 1, k));
    lid_zero_flux_top(m, ft, nz);
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(dv_loc, i, j, k) = -(Ro * ((A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx +
                                             (A(fn, i, j, k) - A(fn, i, j - 1, k)) / dy +
                                             (A(ft, i, j, k) - A(ft, i, j, k - 1)) / dz));

    /* vertical momentum (:145-181); the lid zeroing starts at nz-1 (:169) */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i)
                A(fe, i, j, k) = 0.25 * (A(w, i, j, k) + A(w, i + 1, j, k)) *
                                 (A(u, i, j, k) + A(u, i, j, k + 1));
    for (int k = 1; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(fn, i, j, k) = 0.25 * (A(w, i, j, k) + A(w, i, j + 1, k)) *
                                 (A(v, i, j, k) + A(v, i, j, k + 1));
    for (int k = 0; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(ft, i, j, k) = 0.25 * (A(w, i, j, k + 1) + A(w, i, j, k)) *
                                 (A(w, i, j, k) + A(w, i, j, k + 1));
    lid_zero_flux_top(m, ft, nz - 1);
    const double dsro = dsqr * Ro;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(dw_loc, i, j, k) = -(dsro * ((A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx +
                                               (A(fn, i, j, k) - A(fn, i, j - 1, k)) / dy +
                                               (A(ft, i, j, k) - A(ft, i, j, k - 1)) / dz));
}

/* ------------------------------------------------------------------------ */
/* mixing.f90                                                                */
/* ------------------------------------------------------------------------ */

/* mixing.f90:225-274 */
void orc_harmonic(orc_model *m, double *p1, double Hmix, double Vmix)
{
    const orc_params *p = &m->p;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double *fe = m->f[ORC_FLUX_EAST], *fn = m->f[ORC_FLUX_NORTH], *ft = m->f[ORC_FLUX_TOP];
    const double dx = p->dx, dy = p->dy, dz = p->dz, dt = p->dt;
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = -1; i <= nx + 1; ++i)
                A(fe, i, j, k) = Hmix * (A(p1, i + 1, j, k) - A(p1, i, j, k)) / dx;
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = -1; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i)
                A(fn, i, j, k) = Hmix * (A(p1, i, j + 1, k) - A(p1, i, j, k)) / dy;
    for (int k = -1; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i)
                A(ft, i, j, k) = Vmix * (A(p1, i, j, k + 1) - A(p1, i, j, k)) / dz;
    lid_zero_flux_top(m, ft, nz);
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(p1, i, j, k) = A(p1, i, j, k) +
                                 dt * ((A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx +
                                       (A(fn, i, j, k) - A(fn, i, j - 1, k)) / dy +
                                       (A(ft, i, j, k) - A(ft, i, j, k - 1)) / dz);
}

/* mixing.f90:96-158 */
void orc_biharmonic_h(orc_model *m, double *p1, double Hmix)
{
    const orc_params *p = &m->p;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double *fe = m->f[ORC_FLUX_EAST], *fn = m->f[ORC_FLUX_NORTH];
    double *del2 = m->del2;
    const double dx = p->dx, dy = p->dy, dt = p->dt;
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = -1; i <= nx + 1; ++i)
                A(fe, i, j, k) = -(Hmix * (A(p1, i + 1, j, k) - A(p1, i, j, k)) / dx);
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = -1; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i)
                A(fn, i, j, k) = -(Hmix * (A(p1, i, j + 1, k) - A(p1, i, j, k)) / dy);
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i)
                A(del2, i, j, k) = (A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx +
                                   (A(fn, i, j, k) - A(fn, i, j - 1, k)) / dy;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i)
                A(fe, i, j, k) = (A(del2, i + 1, j, k) - A(del2, i, j, k)) / dx;
    for (int k = 1; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(fn, i, j, k) = (A(del2, i, j + 1, k) - A(del2, i, j, k)) / dy;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(p1, i, j, k) = A(p1, i, j, k) +
                                 dt * ((A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx +
                                       (A(fn, i, j, k) - A(fn, i, j - 1, k)) / dy);
}

/* mixing.f90:164-217 */
void orc_biharmonic_v(orc_model *m, double *p1, double Vmix)
{
    const orc_params *p = &m->p;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double *ft = m->f[ORC_FLUX_TOP];
    double *del2 = m->del2;
    const double dz = p->dz, dt = p->dt;
    for (int k = -1; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i)
                A(ft, i, j, k) = -(Vmix * (A(p1, i, j, k + 1) - A(p1, i, j, k)) / dz);
    lid_zero_flux_top(m, ft, nz);
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i)
                A(del2, i, j, k) = (A(ft, i, j, k) - A(ft, i, j, k - 1)) / dz;
    for (int k = 0; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(ft, i, j, k) = (A(del2, i, j, k + 1) - A(del2, i, j, k)) / dz;
    lid_zero_flux_top(m, ft, nz);
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(p1, i, j, k) = A(p1, i, j, k) + dt * ((A(ft, i, j, k) - A(ft, i, j, k - 1)) / dz);
}

/* ------------------------------------------------------------------------ */
/* parallel_mpi.f90:70-211, single rank: periodic copies z, y, x              */
/* ------------------------------------------------------------------------ */

void orc_border_exchg_3d(orc_model *m, double *a)
{
    const int nx = m->p.nx, ny = m->p.ny, nz = m->p.nz;
    const int64_t X = m->X, Y = m->Y;
    /* border_exchg_in_z (:205-210): whole x-y planes */
    for (int k = 1; k <= ONX; ++k) {
        memcpy(&A(a, -1, -1, nz + k), &A(a, -1, -1, k), sizeof(double) * (size_t)(X * Y));
        memcpy(&A(a, -1, -1, 1 - k), &A(a, -1, -1, nz - k + 1), sizeof(double) * (size_t)(X * Y));
    }
    /* border_exchg_in_y (:160-165): whole x-z planes */
    for (int j = 1; j <= ONX; ++j) {
        for (int k = -1; k <= nz + 2; ++k)
            memcpy(&A(a, -1, ny + j, k), &A(a, -1, j, k), sizeof(double) * (size_t)X);
        for (int k = -1; k <= nz + 2; ++k)
            memcpy(&A(a, -1, 1 - j, k), &A(a, -1, ny - j + 1, k), sizeof(double) * (size_t)X);
    }
    /* border_exchg_in_x (:117-122): whole y-z planes */
    for (int i = 1; i <= ONX; ++i) {
        for (int k = -1; k <= nz + 2; ++k)
            for (int j = -1; j <= ny + 2; ++j) A(a, nx + i, j, k) = A(a, i, j, k);
        for (int k = -1; k <= nz + 2; ++k)
            for (int j = -1; j <= ny + 2; ++j) A(a, 1 - i, j, k) = A(a, nx - i + 1, j, k);
    }
}

/* ------------------------------------------------------------------------ */
/* integrate.f90                                                             */
/* ------------------------------------------------------------------------ */

/* the parameters diag_opt_balance.f90 changes between steps (:90-91,
   :190-198): dt, Ro and the biharmonic coefficients */
void orc_set_dynamic(orc_model *m, double dt, double Ro, double Ahbi, double Avbi, double Khbi, double Kvbi)
{
    m->p.dt = dt;
    m->p.Ro = Ro;
    m->p.Ahbi = Ahbi;
    m->p.Avbi = Avbi;
    m->p.Khbi = Khbi;
    m->p.Kvbi = Kvbi;
}

void orc_set_forcing(orc_model *m, orc_forcing_fn fn, void *user)
{
    m->forcing = fn;
    m->forcing_user = user;
}

/* wavemaker forcing, for_src/config.f90:224-237 (lee_wave1.f90:320-323 with
   omega = -omega_e): q = q + dt*(cos(omega*t)*q_r + sin(omega*t)*q_i) over
   the whole array, t = dt*itt.  re[f], im[f] for f = u, v, w, b; NULL skips. */
void orc_wavemaker(orc_model *m, int itt, double omega, const double *const *re, const double *const *im)
{
    const double dt = m->p.dt;
    const double t = dt * itt;
    const double cs = cos(omega * t), sn = sin(omega * t);
    double *q[4] = {m->f[ORC_U], m->f[ORC_V], m->f[ORC_W], m->f[ORC_B]};
    for (int f = 0; f < 4; ++f) {
        if (!re[f]) continue;
        for (int64_t n = 0; n < m->H; ++n) q[f][n] = q[f][n] + dt * (cs * re[f][n] + sn * im[f][n]);
    }
}

/* Advection by the background flow (u_back, 0, 0) and of the background
   buoyancy b_back, lee_wave1.f90:239-314, into the tau slot, with the
   reference's loop ranges and flux scratch arrays (b_back NULL: zero). */
void orc_background_flow(orc_model *m, int tau, const double *ub, const double *bb)
{
    const orc_params *p = &m->p;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const int64_t H = m->H, X = m->X, Y = m->Y;
    const double Ro = p->Ro, dx = p->dx, dy = p->dy, dz = p->dz, dsqr = p->dsqr;
    const double *u = m->f[ORC_U], *v = m->f[ORC_V], *w = m->f[ORC_W], *b = m->f[ORC_B];
    double *du = m->f[ORC_DU] + (tau - 1) * H, *dv = m->f[ORC_DV] + (tau - 1) * H;
    double *dw = m->f[ORC_DW] + (tau - 1) * H, *db = m->f[ORC_DB] + (tau - 1) * H;
    double *fe = m->f[ORC_FLUX_EAST], *fn = m->f[ORC_FLUX_NORTH], *ft = m->f[ORC_FLUX_TOP];
    const int lid = p->enable_vertical_boundaries;
#define BBK(i, j, k) (bb ? A(bb, i, j, k) : 0.0)
    /* u (:245-268) */
    for (int i = 0; i <= nx; ++i)
        for (int k = -1; k <= nz + 2; ++k)
            for (int j = -1; j <= ny + 2; ++j)
                A(fe, i, j, k) = 2 * 0.25 * (A(ub, i, j, k) + A(ub, i + 1, j, k)) * (A(u, i + 1, j, k) + A(u, i, j, k));
    for (int j = 0; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
            for (int k = -1; k <= nz + 2; ++k)
                A(fn, i, j, k) = 0.25 * (A(ub, i, j, k) + A(ub, i, j + 1, k)) * (A(v, i + 1, j, k) + A(v, i, j, k));
    for (int k = 0; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i)
            for (int j = -1; j <= ny + 2; ++j)
                A(ft, i, j, k) = 0.25 * (A(ub, i, j, k + 1) + A(ub, i, j, k)) * (A(w, i, j, k) + A(w, i + 1, j, k));
    if (lid) {
        for (int k = 1 - ONX; k <= 0; ++k) memset(&A(ft, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
        for (int k = nz; k <= nz + ONX; ++k) memset(&A(ft, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
    }
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(du, i, j, k) = A(du, i, j, k) - Ro * ((A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx +
                                                        (A(fn, i, j, k) - A(fn, i, j - 1, k)) / dy +
                                                        (A(ft, i, j, k) - A(ft, i, j, k - 1)) / dz);
    /* v (:270-277) */
    for (int j = 1; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i)
            for (int k = -1; k <= nz + 2; ++k)
                A(fe, i, j, k) = 0.25 * (A(v, i, j, k) + A(v, i + 1, j, k)) * (A(ub, i, j + 1, k) + A(ub, i, j, k));
    for (int i = 1; i <= nx; ++i)
        for (int k = -1; k <= nz + 2; ++k)
            for (int j = -1; j <= ny + 2; ++j)
                A(dv, i, j, k) = A(dv, i, j, k) - Ro * (A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx;
    /* w (:279-286) */
    for (int k = 1; k <= nz; ++k)
        for (int i = 0; i <= nx; ++i)
            for (int j = -1; j <= ny + 2; ++j)
                A(fe, i, j, k) = 0.25 * (A(w, i, j, k) + A(w, i + 1, j, k)) * (A(ub, i, j, k) + A(ub, i, j, k + 1));
    for (int i = 1; i <= nx; ++i)
        for (int k = -1; k <= nz + 2; ++k)
            for (int j = -1; j <= ny + 2; ++j)
                A(dw, i, j, k) = A(dw, i, j, k) - dsqr * Ro * (A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx;
    /* b (:288-311) */
    for (int i = 0; i <= nx; ++i)
        for (int k = -1; k <= nz + 2; ++k)
            for (int j = -1; j <= ny + 2; ++j)
                A(fe, i, j, k) = 0.5 * (A(b, i, j, k) + A(b, i + 1, j, k)) * A(ub, i, j, k) +
                                 0.5 * (BBK(i, j, k) + BBK(i + 1, j, k)) * A(u, i, j, k);
    for (int j = 0; j <= ny; ++j)
        for (int k = -1; k <= nz + 2; ++k)
            for (int i = -1; i <= nx + 2; ++i)
                A(fn, i, j, k) = 0.5 * (BBK(i, j, k) + BBK(i, j + 1, k)) * A(v, i, j, k);
    for (int k = 0; k <= nz; ++k)
        for (int j = -1; j <= ny + 2; ++j)
            for (int i = -1; i <= nx + 2; ++i)
                A(ft, i, j, k) = 0.5 * (BBK(i, j, k) + BBK(i, j, k + 1)) * A(w, i, j, k);
    if (lid) {
        for (int k = 1 - ONX; k <= 0; ++k) memset(&A(ft, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
        for (int k = nz; k <= nz + ONX; ++k) memset(&A(ft, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
    }
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(db, i, j, k) = A(db, i, j, k) - Ro * ((A(fe, i, j, k) - A(fe, i - 1, j, k)) / dx +
                                                        (A(fn, i, j, k) - A(fn, i, j - 1, k)) / dy +
                                                        (A(ft, i, j, k) - A(ft, i, j, k - 1)) / dz);
#undef BBK
}

/* a steady body force added to the tau slot (whole arrays); NULL: none */
void orc_body_force(orc_model *m, int tau, const double *const *f)
{
    double *d[4] = {m->f[ORC_DU] + (tau - 1) * m->H, m->f[ORC_DV] + (tau - 1) * m->H,
                    m->f[ORC_DW] + (tau - 1) * m->H, m->f[ORC_DB] + (tau - 1) * m->H};
    for (int q = 0; q < 4; ++q) {
        if (!f[q]) continue;
        for (int64_t n = 0; n < m->H; ++n) d[q][n] = d[q][n] + f[q][n];
    }
}

static void background_hook(orc_model *m, void *user, int itt, int tau)
{
    (void)user;
    (void)itt;
    if (m->bg_u) orc_background_flow(m, tau, m->bg_u, m->bg_b);
    orc_body_force(m, tau, m->bf);
}

/* background flow (u_back NULL: none) and body force (entries NULL: none)
   as the model's set_forcing */
void orc_set_background(orc_model *m, const double *u_back, const double *b_back, const double *const *f)
{
    m->bg_u = u_back;
    m->bg_b = b_back;
    for (int q = 0; q < 4; ++q) m->bf[q] = f ? f[q] : NULL;
    orc_set_forcing(m, background_hook, NULL);
}

static void wavemaker_hook(orc_model *m, void *user, int itt, int tau)
{
    (void)user;
    (void)tau;
    orc_wavemaker(m, itt, m->wm_omega, m->wm_re, m->wm_im);
}

/* the wavemaker as the model's set_forcing */
void orc_set_wavemaker(orc_model *m, double omega, const double *const *re, const double *const *im)
{
    m->wm_omega = omega;
    for (int f = 0; f < 4; ++f) {
        m->wm_re[f] = re[f];
        m->wm_im[f] = im[f];
    }
    orc_set_forcing(m, wavemaker_hook, NULL);
}

void orc_integrate_pre(orc_model *m, int itt, int tau, int taum1, int taum2)
{
    const orc_params *p = &m->p;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const int64_t H = m->H, X = m->X, Y = m->Y;
    double *u = m->f[ORC_U], *v = m->f[ORC_V], *w = m->f[ORC_W], *b = m->f[ORC_B];
    double *du = m->f[ORC_DU] + (tau - 1) * H, *dv = m->f[ORC_DV] + (tau - 1) * H;
    double *dw = m->f[ORC_DW] + (tau - 1) * H, *db = m->f[ORC_DB] + (tau - 1) * H;

    /* buoyancy advection (:12-17) */
    if (p->enable_nonlinear) orc_buoyancy_advection_2nd(m, db);
    else memset(db, 0, sizeof(double) * (size_t)H);

    /* background stratification (:21-23), whole x-y planes */
    const double N02 = p->N0 * p->N0;
    for (int k = 1; k <= nz; ++k)
        for (int j = -1; j <= ny + 2; ++j)
            for (int i = -1; i <= nx + 2; ++i)
                A(db, i, j, k) = A(db, i, j, k) - 0.5 * (A(w, i, j, k) + A(w, i, j, k - 1)) * N02;

    /* momentum advection (:28-32) */
    if (p->enable_nonlinear) orc_momentum_advection_2nd(m, du, dv, dw);
    else {
        memset(du, 0, sizeof(double) * (size_t)H);
        memset(dv, 0, sizeof(double) * (size_t)H);
        memset(dw, 0, sizeof(double) * (size_t)H);
    }

    /* Coriolis (:37-42) over all k including the halo */
    const double f0 = p->f0;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
            for (int k = -1; k <= nz + 2; ++k) {
                A(du, i, j, k) = A(du, i, j, k) +
                    (A(v, i, j, k) + A(v, i, j - 1, k) + A(v, i + 1, j, k) + A(v, i + 1, j - 1, k)) * 0.25 * f0;
                A(dv, i, j, k) = A(dv, i, j, k) -
                    (A(u, i - 1, j, k) + A(u, i, j, k) + A(u, i - 1, j + 1, k) + A(u, i, j + 1, k)) * 0.25 * f0;
            }

    /* buoyancy force (:47-49), whole x-y planes */
    for (int k = 1; k <= nz; ++k)
        for (int j = -1; j <= ny + 2; ++j)
            for (int i = -1; i <= nx + 2; ++i)
                A(dw, i, j, k) = A(dw, i, j, k) + 0.5 * (A(b, i, j, k) + A(b, i, j, k + 1));

    /* (:51-54) */
    if (p->enable_vertical_boundaries) {
        for (int k = 1 - ONX; k <= 0; ++k) memset(&A(dw, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
        for (int k = nz; k <= nz + ONX; ++k) memset(&A(dw, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
    }

    /* friction (:59-78), in place, no halo exchange in between */
    if (p->Ah > 0. || p->Av > 0.) orc_harmonic(m, u, p->Ah, p->Av);
    if (p->Ah > 0. || p->Av > 0.) orc_harmonic(m, v, p->Ah, p->Av);
    if (p->Ah > 0. || p->Av > 0.) {
        orc_harmonic(m, w, p->Ah, p->Av);
        if (p->enable_vertical_boundaries)
            for (int k = nz; k <= nz + ONX; ++k) memset(&A(w, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
    }
    if (p->Kh > 0. || p->Kv > 0.) orc_harmonic(m, b, p->Kh, p->Kv);
    if (p->Ahbi > 0.) orc_biharmonic_h(m, u, p->Ahbi);
    if (p->Ahbi > 0.) orc_biharmonic_h(m, v, p->Ahbi);
    if (p->Ahbi > 0.) orc_biharmonic_h(m, w, p->Ahbi);
    if (p->Avbi > 0.) orc_biharmonic_v(m, u, p->Avbi);
    if (p->Avbi > 0.) orc_biharmonic_v(m, v, p->Avbi);
    if (p->Avbi > 0.) {
        orc_biharmonic_v(m, w, p->Avbi);
        if (p->enable_vertical_boundaries)
            for (int k = nz; k <= nz + ONX; ++k) memset(&A(w, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
    }
    if (p->Khbi > 0.) orc_biharmonic_h(m, b, p->Khbi);
    if (p->Kvbi > 0.) orc_biharmonic_v(m, b, p->Kvbi);

    /* set_forcing (:82): the caller's hook, else empty as in
       for_config/eddy1.f90:141-142 */
    if (m->forcing) m->forcing(m, m->forcing_user, itt, tau);

    /* Adams-Bashforth (:87-102), whole arrays */
    const double dt = p->dt, dtw = p->dt / p->dsqr;
    const double *du1 = m->f[ORC_DU] + (taum1 - 1) * H, *dv1 = m->f[ORC_DV] + (taum1 - 1) * H;
    const double *dw1 = m->f[ORC_DW] + (taum1 - 1) * H, *db1 = m->f[ORC_DB] + (taum1 - 1) * H;
    const double *du2 = m->f[ORC_DU] + (taum2 - 1) * H, *dv2 = m->f[ORC_DV] + (taum2 - 1) * H;
    const double *dw2 = m->f[ORC_DW] + (taum2 - 1) * H, *db2 = m->f[ORC_DB] + (taum2 - 1) * H;
    if (itt == 0) {
        for (int64_t n = 0; n < H; ++n) {
            u[n] = u[n] + dt * du[n];
            v[n] = v[n] + dt * dv[n];
            w[n] = w[n] + dtw * dw[n];
            b[n] = b[n] + dt * db[n];
        }
    } else if (itt == 1 || !p->enable_AB_3_order) {
        const double c1 = 1.5 + p->eps_ab, c2 = 0.5 + p->eps_ab;
        for (int64_t n = 0; n < H; ++n) {
            u[n] = u[n] + dt * (c1 * du[n] - c2 * du1[n]);
            v[n] = v[n] + dt * (c1 * dv[n] - c2 * dv1[n]);
            w[n] = w[n] + dtw * (c1 * dw[n] - c2 * dw1[n]);
            b[n] = b[n] + dt * (c1 * db[n] - c2 * db1[n]);
        }
    } else {
        const double a3 = 23.0 / 12.0, b3 = -16.0 / 12.0, c3 = 5.0 / 12.0;  /* main_module.f90:57 */
        for (int64_t n = 0; n < H; ++n) {
            u[n] = u[n] + dt * (a3 * du[n] + b3 * du1[n] + c3 * du2[n]);
            v[n] = v[n] + dt * (a3 * dv[n] + b3 * dv1[n] + c3 * dv2[n]);
            w[n] = w[n] + dtw * (a3 * dw[n] + b3 * dw1[n] + c3 * dw2[n]);
            b[n] = b[n] + dt * (a3 * db[n] + b3 * db1[n] + c3 * db2[n]);
        }
    }

    /* (:106-111) */
    orc_border_exchg_3d(m, u);
    orc_border_exchg_3d(m, v);
    orc_border_exchg_3d(m, w);
    orc_border_exchg_3d(m, b);
}

/* ------------------------------------------------------------------------ */
/* my_fft.f90 transforms                                                     */
/* ------------------------------------------------------------------------ */

/* real (nx,ny,nz) -> complex (nx/2+1,ny,nz) along x: z[m] = x[2m] + i x[2m+1],
   Z = DFT_{nx/2}(z), X[k] = E[k] + W^k O[k] with E = (Z[k] + conj Z[n2-k])/2,
   O = (Z[k] - conj Z[n2-k])/(2i), W = exp(-2 pi i/nx) */
static void r2c_x(orc_model *m, const double *qp, cplx *qs)
{
    const int nx = m->p.nx, ny = m->p.ny, nz = m->p.nz;
    const int n2 = nx / 2;
    const int64_t nkx = m->nkx;
    cplx *row = m->row;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            const double *src = qp + (int64_t)nx * (j + (int64_t)ny * k);
            for (int i = 0; i < n2; ++i) { row[i].re = src[2 * i]; row[i].im = src[2 * i + 1]; }
            cfft_strided(&m->px, row, 1, -1);
            cplx *dst = qs + nkx * (j + (int64_t)ny * k);
            for (int kk = 0; kk < nkx; ++kk) {
                const cplx zk = row[kk == n2 ? 0 : kk];
                const cplx zm = row[kk == 0 ? 0 : n2 - kk];
                const cplx e = {0.5 * (zk.re + zm.re), 0.5 * (zk.im - zm.im)};
                const cplx bq = {zk.re - zm.re, zk.im + zm.im};
                const cplx o = {0.5 * bq.im, -0.5 * bq.re};
                dst[kk] = c_add(e, c_mul(m->wk[kk], o));
            }
        }
}

/* inverse of r2c_x (unnormalised), imaginary parts of X[0], X[nx/2] ignored:
   Z'[k] = (X[k] + conj X[n2-k]) + i conj(W^k) (X[k] - conj X[n2-k]),
   z = IDFT_{nx/2}(Z'), x[2m] = Re z[m], x[2m+1] = Im z[m] */
static void c2r_x(orc_model *m, const cplx *qs, double *qp)
{
    const int nx = m->p.nx, ny = m->p.ny, nz = m->p.nz;
    const int n2 = nx / 2;
    const int64_t nkx = m->nkx;
    cplx *row = m->row;
    cplx *raw = m->row + nkx;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            const cplx *src = qs + nkx * (j + (int64_t)ny * k);
            for (int i = 0; i < nkx; ++i) raw[i] = src[i];
            raw[0].im = 0.0;
            raw[n2].im = 0.0;
            for (int kk = 0; kk < n2; ++kk) {
                const cplx xk = raw[kk], xm = raw[n2 - kk];
                const cplx a = {xk.re + xm.re, xk.im - xm.im};
                const cplx d = {xk.re - xm.re, xk.im + xm.im};
                const cplx wc = {m->wk[kk].re, -m->wk[kk].im};
                const cplx q = c_mul(wc, d);
                row[kk].re = a.re - q.im;
                row[kk].im = a.im + q.re;
            }
            cfft_strided(&m->px, row, 1, +1);
            double *dst = qp + (int64_t)nx * (j + (int64_t)ny * k);
            for (int i = 0; i < n2; ++i) { dst[2 * i] = row[i].re; dst[2 * i + 1] = row[i].im; }
        }
}

static void c2c_y(orc_model *m, cplx *qs, int sign)
{
    const int ny = m->p.ny, nz = m->p.nz;
    const int64_t nkx = m->nkx;
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nkx; ++i)
            cfft_strided(&m->py, qs + i + nkx * (int64_t)ny * k, nkx, sign);
}

static void c2c_z(orc_model *m, cplx *qs, int sign)
{
    const int ny = m->p.ny;
    const int64_t nkx = m->nkx;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nkx; ++i)
            cfft_strided(&m->pz, qs + i + nkx * (int64_t)j, nkx * ny, sign);
}

/* my_fft.f90:170-178 (p3dfft_ftran_r2c 'ffn', then QS = QS*dx*dy) */
void orc_my_fft_2d(orc_model *m, const double *qp, double *qs_)
{
    cplx *qs = (cplx *)qs_;
    r2c_x(m, qp, qs);
    c2c_y(m, qs, -1);
    const double dx = m->p.dx, dy = m->p.dy;
    for (int64_t n = 0; n < m->S; ++n) {
        qs[n].re = qs[n].re * dx * dy;
        qs[n].im = qs[n].im * dx * dy;
    }
}

/* my_fft.f90:159-167 (p3dfft_btran_c2r 'nff', then QP = QP/(Lx*Ly)) */
void orc_my_ifft_2d(orc_model *m, const double *qs_, double *qp)
{
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)m->S);
    memcpy(tmp, qs_, sizeof(cplx) * (size_t)m->S);
    c2c_y(m, tmp, +1);
    c2r_x(m, tmp, qp);
    free(tmp);
    const double s = m->p.Lx * m->p.Ly;
    const int64_t N = (int64_t)m->p.nx * m->p.ny * m->p.nz;
    for (int64_t n = 0; n < N; ++n) qp[n] = qp[n] / s;
}

/* my_fft.f90:197-205 ('fft', then QS = QS*dx*dy*dz) */
void orc_my_fft_3d(orc_model *m, const double *qp, double *qs_)
{
    cplx *qs = (cplx *)qs_;
    r2c_x(m, qp, qs);
    c2c_y(m, qs, -1);
    c2c_z(m, qs, -1);
    const double dx = m->p.dx, dy = m->p.dy, dz = m->p.dz;
    for (int64_t n = 0; n < m->S; ++n) {
        qs[n].re = qs[n].re * dx * dy * dz;
        qs[n].im = qs[n].im * dx * dy * dz;
    }
}

/* my_fft.f90:184-192 ('fft' backward, then QP = QP/(Lx*Ly*Lz)) */
void orc_my_ifft_3d(orc_model *m, const double *qs_, double *qp)
{
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)m->S);
    memcpy(tmp, qs_, sizeof(cplx) * (size_t)m->S);
    c2c_z(m, tmp, +1);
    c2c_y(m, tmp, +1);
    c2r_x(m, tmp, qp);
    free(tmp);
    const double s = m->p.Lx * m->p.Ly * m->p.Lz;
    const int64_t N = (int64_t)m->p.nx * m->p.ny * m->p.nz;
    for (int64_t n = 0; n < N; ++n) qp[n] = qp[n] / s;
}

/* ------------------------------------------------------------------------ */
/* pressure_solver.f90                                                       */
/* ------------------------------------------------------------------------ */

/* pressure_solver.f90:126-162 (Thomas algorithm; real a,b,c, complex d,x) */
void orc_solve_tridiag(const double *a, const double *b, const double *c,
                       const double *d_, double *x_, int n)
{
    const cplx *d = (const cplx *)d_;
    cplx *x = (cplx *)x_;
    double *cp = (double *)malloc(sizeof(double) * (size_t)n);
    cplx *dp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cp[0] = c[0] / b[0];
    dp[0].re = d[0].re / b[0];
    dp[0].im = d[0].im / b[0];
    for (int i = 1; i < n; ++i) {
        const double mm = b[i] - cp[i - 1] * a[i];
        const double fxa = 1.0 / mm;
        cp[i] = c[i] * fxa;
        dp[i].re = (d[i].re - dp[i - 1].re * a[i]) * fxa;
        dp[i].im = (d[i].im - dp[i - 1].im * a[i]) * fxa;
    }
    x[n - 1] = dp[n - 1];
    for (int i = n - 2; i >= 0; --i) {
        x[i].re = dp[i].re - cp[i] * x[i + 1].re;
        x[i].im = dp[i].im - cp[i] * x[i + 1].im;
    }
    free(cp);
    free(dp);
}

/* pressure_solver.f90:30-71, the part between the forward and inverse FFT */
void orc_spectral_solve(orc_model *m, double *fs_)
{
    const orc_params *p = &m->p;
    const int ny = p->ny, nz = p->nz;
    const int64_t nkx = m->nkx, plane = nkx * ny;
    cplx *FS = (cplx *)fs_;
    const double *ksqr = m->f[ORC_KSQR];
    if (p->enable_vertical_boundaries) {
        const double fxa = 1.0 / (p->dz * p->dz) / p->dsqr;
        double *a = (double *)malloc(sizeof(double) * (size_t)nz);
        double *bb = (double *)malloc(sizeof(double) * (size_t)nz);
        double *c = (double *)malloc(sizeof(double) * (size_t)nz);
        cplx *col = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
        cplx *sol = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
        a[0] = 0.0;                 /* a_tri(1) is never read */
        c[nz - 1] = 0.0;            /* c_tri(nz) is never used */
        for (int k = 1; k < nz; ++k) a[k] = fxa;
        for (int k = 0; k < nz - 1; ++k) c[k] = fxa;
        const double dz2 = p->dz * p->dz;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nkx; ++i) {
                const double ks = ksqr[i + nkx * j];
                for (int k = 0; k < nz; ++k) col[k] = FS[i + nkx * j + plane * k];
                if (ks > 0) {
                    bb[0] = -ks - fxa;
                    for (int k = 1; k < nz - 1; ++k) bb[k] = -ks - 2 * fxa;
                    bb[nz - 1] = -ks - fxa;
                    orc_solve_tridiag(a, bb, c, (const double *)col, (double *)sol, nz);
                } else {
                    sol[0].re = 0.0; sol[0].im = 0.0;
                    sol[1].re = col[0].re * dz2 * p->dsqr;
                    sol[1].im = col[0].im * dz2 * p->dsqr;
                    for (int k = 1; k < nz - 1; ++k) {
                        sol[k + 1].re = 2 * sol[k].re - sol[k - 1].re + col[k].re * dz2 * p->dsqr;
                        sol[k + 1].im = 2 * sol[k].im - sol[k - 1].im + col[k].im * dz2 * p->dsqr;
                    }
                }
                for (int k = 0; k < nz; ++k) FS[i + nkx * j + plane * k] = sol[k];
            }
        free(a); free(bb); free(c); free(col); free(sol);
    } else {
        for (int64_t n = 0; n < m->S; ++n) {
            const double ks = ksqr[n];
            if (ks > 0.) {
                FS[n].re = -(FS[n].re / ks);
                FS[n].im = -(FS[n].im / ks);
            } else {
                FS[n].re = 0.0;
                FS[n].im = 0.0;
            }
        }
    }
}

/* pressure_solver.f90:2-120 */
void orc_pressure_solver(orc_model *m)
{
    const orc_params *p = &m->p;
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const int64_t X = m->X, Y = m->Y;
    double *u = m->f[ORC_U], *v = m->f[ORC_V], *w = m->f[ORC_W], *pr = m->f[ORC_P];
    double *aloc = m->aloc;
    const double dx = p->dx, dy = p->dy, dz = p->dz, dt = p->dt;
#define AL(i, j, k) aloc[((i) - 1) + (int64_t)nx * (((j) - 1) + (int64_t)ny * ((k) - 1))]
    /* (:16-24) */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                AL(i, j, k) = (A(u, i, j, k) - A(u, i - 1, j, k)) / dx +
                              (A(v, i, j, k) - A(v, i, j - 1, k)) / dy +
                              (A(w, i, j, k) - A(w, i, j, k - 1)) / dz;
    const int64_t N = (int64_t)nx * ny * nz;
    for (int64_t n = 0; n < N; ++n) aloc[n] = aloc[n] / dt;

    /* (:30-72) */
    double *FS = m->f[ORC_FS];
    if (p->enable_vertical_boundaries) {
        orc_my_fft_2d(m, aloc, FS);
        orc_spectral_solve(m, FS);
        orc_my_ifft_2d(m, FS, aloc);
    } else {
        orc_my_fft_3d(m, aloc, FS);
        orc_spectral_solve(m, FS);
        orc_my_ifft_3d(m, FS, aloc);
    }

    /* (:75-93) */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) A(pr, i, j, k) = AL(i, j, k);
    orc_border_exchg_3d(m, pr);
    if (p->enable_vertical_boundaries) {
        for (int k = 1 - ONX; k <= 0; ++k)
            memcpy(&A(pr, -1, -1, k), &A(pr, -1, -1, 1), sizeof(double) * (size_t)(X * Y));
        for (int k = nz + 1; k <= nz + ONX; ++k)
            memcpy(&A(pr, -1, -1, k), &A(pr, -1, -1, nz), sizeof(double) * (size_t)(X * Y));
    }

    /* (:98-110) */
    const double dtw = dt / p->dsqr;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                A(u, i, j, k) = A(u, i, j, k) - dt * (A(pr, i + 1, j, k) - A(pr, i, j, k)) / dx;
                A(v, i, j, k) = A(v, i, j, k) - dt * (A(pr, i, j + 1, k) - A(pr, i, j, k)) / dy;
                A(w, i, j, k) = A(w, i, j, k) - dtw * (A(pr, i, j, k + 1) - A(pr, i, j, k)) / dz;
            }
    if (p->enable_vertical_boundaries) {
        for (int k = 1 - ONX; k <= 0; ++k) memset(&A(w, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
        for (int k = nz; k <= nz + ONX; ++k) memset(&A(w, -1, -1, k), 0, sizeof(double) * (size_t)(X * Y));
    }
    /* (:114-118) */
    orc_border_exchg_3d(m, u);
    orc_border_exchg_3d(m, v);
    orc_border_exchg_3d(m, w);
#undef AL
}

void orc_integrate(orc_model *m, int itt, int tau, int taum1, int taum2)
{
    orc_integrate_pre(m, itt, tau, taum1, taum2);
    orc_pressure_solver(m);
}

/* main.f90:139-161 without the diagnostics; tau rotation at :156-160 */
void orc_run(orc_model *m, int nsteps, int32_t *st)
{
    for (int s = 0; s < nsteps; ++s) {
        orc_integrate(m, st[0], st[1], st[2], st[3]);
        st[1] = st[1] % 3 + 1;
        st[2] = st[2] % 3 + 1;
        st[3] = st[3] % 3 + 1;
        st[0] += 1;
    }
}
