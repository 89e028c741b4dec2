/*
 * gev_oracle.c -- FP64 CPU restatement of gevolution 1.2's per-cycle PM loop.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
 * CPU baseline.  The product path never links or calls this file.
 *
 * Every function cites the reference lines it restates (paths relative to the
 * gevolution-1.2 source tree).  Expression order follows the reference
 * literally so that the HIP kernels, which follow the same order and are
 * compiled with -ffp-contract=off, can be compared bit for bit where the
 * inputs agree bit for bit (stencils, k-space solves, kicks, drifts, filing).
 *
 * Deviations, all documented where they occur:
 *   - `1.0l - w` (long double, gevolution.hpp:982,1102,1232) is evaluated in
 *     double: identical except for w < 2^-11, where the x87 path can double-
 *     round (<= 1 ulp).
 *   - LATfield2 is absent, so Particles::updateVel / moveParticles, the
 *     scalarProjectionCIC_project deposit and PlanFFT are restated from their
 *     call sites and the manual (sec. 4); the summation order inside them is
 *     this file's own.  These are "parity unpinned" against LATfield2 itself
 *     and pinned instead by closed forms / invariants / numpy (tests/).
 *   - particleHorizon uses GSL qng (background.hpp:200-213); here the same
 *     integral is evaluated by composite Gauss-Legendre to ~1e-14, so tau
 *     differs from the reference by at most qng's 1e-7 tolerance.  tau only
 *     enters outputs, never the dynamics.
 */
#include "gev_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { double re, im; } cplx;

/* ---- wall-clock timers of main.cpp's BENCHMARK categories (main.cpp:907-922) */
static double now_s(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return (double)clock() / CLOCKS_PER_SEC;
#endif
}

static double g_timer[ORC_NTIMERS];
static int64_t g_count[ORC_NTIMERS];

void orc_timers_reset(void)
{
    memset(g_timer, 0, sizeof g_timer);
    memset(g_count, 0, sizeof g_count);
}

void orc_timers_get(double *seconds, int64_t *counts)
{
    memcpy(seconds, g_timer, sizeof g_timer);
    memcpy(counts, g_count, sizeof g_count);
}

void orc_timer_add(int which, double seconds)
{
    g_timer[which] += seconds;
    g_count[which]++;
}

/* STREAM-style triad a = b + s c over three arrays of n doubles with the
   current OpenMP threads; the best of `reps` runs in GB/s (24 n bytes). */
double orc_stream_triad(int64_t n, int reps)
{
    double *a = malloc(sizeof(double) * n), *b = malloc(sizeof(double) * n), *c = malloc(sizeof(double) * n);
    if (!a || !b || !c) {
        free(a); free(b); free(c);
        return -1.;
    }
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) {
        a[i] = 0.;
        b[i] = 1.;
        c[i] = 2.;
    }
    double best = 0.;
    for (int r = 0; r < reps; r++) {
        const double t0 = now_s();
#pragma omp parallel for schedule(static)
        for (i = 0; i < n; i++) a[i] = b[i] + 3. * c[i];
        const double dt = now_s() - t0;
        const double gbs = 24. * (double)n / dt / 1e9;
        if (gbs > best) best = gbs;
    }
    volatile double sink = a[n / 2];
    (void)sink;
    free(a); free(b); free(c);
    return best;
}

static inline int wrapi(int i, int N) { return i < 0 ? i + N : (i >= N ? i - N : i); }
static inline int64_t lin(int N, int x, int y, int z)
{
    return (int64_t)x + (int64_t)N * ((int64_t)y + (int64_t)N * (int64_t)z);
}
/* field value at site (x,y,z) shifted by (dx,dy,dz) with periodic wrap */
static inline double at(const double *f, int N, int x, int y, int z, int dx, int dy, int dz)
{
    return f[lin(N, wrapi(x + dx, N), wrapi(y + dy, N), wrapi(z + dz, N))];
}

static inline cplx cmul(cplx a, cplx b)
{
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx rmul(double s, cplx a) { cplx r = {s * a.re, s * a.im}; return r; }
static inline cplx cmulr(cplx a, double s) { cplx r = {a.re * s, a.im * s}; return r; }
static inline cplx cdivr(cplx a, double s) { cplx r = {a.re / s, a.im / s}; return r; }
static inline cplx cconj(cplx a) { cplx r = {a.re, -a.im}; return r; }

/* ======================================================================== */
/* background (background.hpp:137-213)                                       */
/* ======================================================================== */

/* Hconf, background.hpp:137-140, with bg_ncdm = 0 (no ncdm species). */
double orc_hconf(double a, double fourpiG, const orc_cosmo *c)
{
    return sqrt((2. * fourpiG / 3.) *
                (((c->Omega_cdm + c->Omega_b + 0.0) / a) + (c->Omega_Lambda * a * a) +
                 (c->Omega_rad / a / a) +
                 (c->Omega_fld * exp(3. * c->wa_fld * (a - 1.)) /
                  pow(a, 1. + 3. * (c->w0_fld + c->wa_fld)))));
}

/* rungekutta4bg, background.hpp:167-177 */
void orc_rungekutta4bg(double *a_, double fourpiG, const orc_cosmo *c, double dtau)
{
    double a = *a_;
    double k1a, k2a, k3a, k4a;
    k1a = a * orc_hconf(a, fourpiG, c);
    k2a = (a + k1a * dtau / 2.) * orc_hconf(a + k1a * dtau / 2., fourpiG, c);
    k3a = (a + k2a * dtau / 2.) * orc_hconf(a + k2a * dtau / 2., fourpiG, c);
    k4a = (a + k3a * dtau) * orc_hconf(a + k3a * dtau, fourpiG, c);
    a += dtau * (k1a + 2. * k2a + 2. * k3a + k4a) / 6.;
    *a_ = a;
}

static double horizon_integrand(double s, const orc_cosmo *c)
{
    return 2. / (s * orc_hconf(s * s, 1., c));
}

/* particleHorizon, background.hpp:200-213: integral of 2/(s*H(s^2)) over
   s in [sqrt(a)*1e-7, sqrt(a)], divided by sqrt(fourpiG).  Composite 8-point
   Gauss-Legendre on 2048 panels (the integrand is smooth). */
double orc_particle_horizon(double a, double fourpiG, const orc_cosmo *c)
{
    static const double x8[4] = {0.1834346424956498, 0.5255324099163290,
                                 0.7966664774136267, 0.9602898564975363};
    static const double w8[4] = {0.3626837833783620, 0.3137066458778873,
                                 0.2223810344533745, 0.1012285362903763};
    double lo = sqrt(a) * 1.0e-7, hi = sqrt(a);
    int panels = 2048;
    double h = (hi - lo) / panels, sum = 0.;
    for (int p = 0; p < panels; p++) {
        double mid = lo + (p + 0.5) * h, half = 0.5 * h;
        for (int j = 0; j < 4; j++) {
            sum += w8[j] * half * (horizon_integrand(mid - half * x8[j], c) +
                                   horizon_integrand(mid + half * x8[j], c));
        }
    }
    return sum / sqrt(fourpiG);
}

/* dtau rule, main.cpp:292-295 and 869-872 */
double orc_next_dtau(const orc_sim *sim, double a)
{
    double dx = 1.0 / (double)sim->N;
    if (sim->Cf * dx < sim->steplimit / orc_hconf(a, sim->fourpiG, &sim->cosmo))
        return sim->Cf * dx;
    return sim->steplimit / orc_hconf(a, sim->fourpiG, &sim->cosmo);
}

/* ======================================================================== */
/* FFT: LATfield2 PlanFFT (FFTW3 r2c / c2r, unnormalised, sign -1 forward)   */
/* Restated as a recursive mixed-radix decimation-in-time DFT.               */
/* ======================================================================== */

static cplx *make_twiddles(int n)
{
    cplx *tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int t = 0; t < n; t++) {
        long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
        tw[t].re = (double)cosl(ang);
        tw[t].im = -(double)sinl(ang);
    }
    return tw;
}

/* out[k] = sum_j in[j*is] * W^(jk), W = exp(sign*2 pi i/n); tw holds the
   n0-point table exp(-2 pi i t/n0) and n = n0/ts. */
static void dft_rec(int n, const cplx *in, int is, cplx *out, const cplx *tw, int n0, int ts,
                    int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    int r = 2;
    while (n % r) r++;
    int m = n / r;
    for (int j = 0; j < r; j++) dft_rec(m, in + (size_t)j * is, is * r, out + (size_t)j * m, tw, n0, ts * r, sign);
    cplx t[16];
    for (int k = 0; k < m; k++) {
        for (int j = 0; j < r; j++) {
            cplx w = tw[((int64_t)j * k * ts) % n0];
            if (sign > 0) w.im = -w.im;
            t[j] = cmul(out[(size_t)j * m + k], w);
        }
        for (int qq = 0; qq < r; qq++) {
            cplx acc = t[0];
            for (int j = 1; j < r; j++) {
                cplx w = tw[((int64_t)j * qq * (n0 / r)) % n0];
                if (sign > 0) w.im = -w.im;
                acc = cadd(acc, cmul(t[j], w));
            }
            out[(size_t)qq * m + k] = acc;
        }
    }
}

void orc_fft1d(int n, const double *in, double *out, int sign)
{
    cplx *tw = make_twiddles(n);
    dft_rec(n, (const cplx *)in, 1, (cplx *)out, tw, n, 1, sign);
    free(tw);
}

/* complex FFT along y (dim=1) or z (dim=2) of a [N][N][H] complex array */
static void fft_axis(int N, cplx *a, int dim, int sign, const cplx *tw)
{
    const int H = N / 2 + 1;
    #pragma omp parallel
    {
        cplx *lin_in = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        cplx *lin_out = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        #pragma omp for schedule(static)
        for (int o = 0; o < N; o++) {
            for (int kx = 0; kx < H; kx++) {
                for (int i = 0; i < N; i++) {
                    size_t ix = dim == 1 ? (size_t)kx + (size_t)H * ((size_t)i + (size_t)N * o)
                                         : (size_t)kx + (size_t)H * ((size_t)o + (size_t)N * i);
                    lin_in[i] = a[ix];
                }
                dft_rec(N, lin_in, 1, lin_out, tw, N, 1, sign);
                for (int i = 0; i < N; i++) {
                    size_t ix = dim == 1 ? (size_t)kx + (size_t)H * ((size_t)i + (size_t)N * o)
                                         : (size_t)kx + (size_t)H * ((size_t)o + (size_t)N * i);
                    a[ix] = lin_out[i];
                }
            }
        }
        free(lin_in);
        free(lin_out);
    }
}

/* r2c forward: out[kx + H*(ky + N*kz)] = sum f(x,y,z) exp(-2 pi i k.x / N) */
void orc_fft_r2c(int N, const double *in, double *out_)
{
    const int H = N / 2 + 1;
    cplx *out = (cplx *)out_;
    cplx *tw = make_twiddles(N);
    #pragma omp parallel
    {
        cplx *lin_in = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        cplx *lin_out = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        #pragma omp for schedule(static)
        for (int64_t row = 0; row < (int64_t)N * N; row++) {
            for (int x = 0; x < N; x++) { lin_in[x].re = in[row * N + x]; lin_in[x].im = 0.; }
            dft_rec(N, lin_in, 1, lin_out, tw, N, 1, -1);
            for (int kx = 0; kx < H; kx++) out[row * H + kx] = lin_out[kx];
        }
        free(lin_in);
        free(lin_out);
    }
    fft_axis(N, out, 1, -1, tw);
    fft_axis(N, out, 2, -1, tw);
    free(tw);
}

/* c2r backward (input preserved, as main.cpp:586/593 require for BiFT).
   Along x the Hermitian line is completed from kx in [0,N/2]; the imaginary
   parts of the kx=0 and kx=N/2 entries are ignored (FFTW / numpy c2r). */
void orc_fft_c2r(int N, const double *in_, double *out)
{
    const int H = N / 2 + 1;
    size_t nk = (size_t)N * N * H;
    cplx *w = (cplx *)malloc(sizeof(cplx) * nk);
    memcpy(w, in_, sizeof(cplx) * nk);
    cplx *tw = make_twiddles(N);
    fft_axis(N, w, 2, +1, tw);
    fft_axis(N, w, 1, +1, tw);
    #pragma omp parallel
    {
        cplx *lin_in = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        cplx *lin_out = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        #pragma omp for schedule(static)
        for (int64_t row = 0; row < (int64_t)N * N; row++) {
            for (int k = 0; k < H; k++) lin_in[k] = w[row * H + k];
            lin_in[0].im = 0.;
            lin_in[N / 2].im = 0.;
            for (int k = H; k < N; k++) lin_in[k] = cconj(lin_in[N - k]);
            dft_rec(N, lin_in, 1, lin_out, tw, N, 1, +1);
            for (int x = 0; x < N; x++) out[row * N + x] = lin_out[x].re;
        }
        free(lin_in);
        free(lin_out);
    }
    free(tw);
    free(w);
}

/* ======================================================================== */
/* k-space solves                                                            */
/* ======================================================================== */

/* gridk2 / kshift tables, gevolution.hpp:222-227 */
static void ktables(int N, double *gridk2, cplx *kshift)
{
    for (int i = 0; i < N; i++) {
        gridk2[i] = 2. * (double)N * sin(M_PI * (double)i / (double)N);
        if (kshift) {
            cplx e = {cos(M_PI * (double)i / (double)N), -sin(M_PI * (double)i / (double)N)};
            kshift[i] = rmul(gridk2[i], e);
        }
        gridk2[i] *= gridk2[i];
    }
}

/* solveModifiedPoissonFT, gevolution.hpp:501-535 */
void orc_solve_modified_poisson_ft(int N, const double *src_, double *pot_, double coeff,
                                   double modif)
{
    const int H = N / 2 + 1;
    const cplx *src = (const cplx *)src_;
    cplx *pot = (cplx *)pot_;
    double *g = (double *)malloc(sizeof(double) * N);
    coeff /= -((long)N * (long)N * (long)N);
    ktables(N, g, NULL);
    #pragma omp parallel for schedule(static)
    for (int kz = 0; kz < N; kz++)
        for (int ky = 0; ky < N; ky++)
            for (int kx = 0; kx < H; kx++) {
                size_t i = (size_t)kx + (size_t)H * ((size_t)ky + (size_t)N * kz);
                if (kx == 0 && ky == 0 && kz == 0) {
                    if (modif == 0.) { pot[i].re = 0.; pot[i].im = 0.; }
                    else pot[i] = cdivr(cmulr(src[i], coeff), modif);
                } else {
                    pot[i] = cdivr(cmulr(src[i], coeff), g[kx] + g[ky] + g[kz] + modif);
                }
            }
    free(g);
}

/* projectFTscalar, gevolution.hpp:211-265 (components 00,01,02,11,12,22) */
void orc_project_ft_scalar(int N, const double *S_, double *chi_, int add)
{
    const int H = N / 2 + 1;
    const size_t nk = (size_t)N * N * H;
    const cplx *S = (const cplx *)S_;
    cplx *chi = (cplx *)chi_;
    double *g = (double *)malloc(sizeof(double) * N);
    cplx *ks = (cplx *)malloc(sizeof(cplx) * N);
    ktables(N, g, ks);
    #pragma omp parallel for schedule(static)
    for (int kz = 0; kz < N; kz++)
        for (int ky = 0; ky < N; ky++)
            for (int kx = 0; kx < H; kx++) {
                size_t i = (size_t)kx + (size_t)H * ((size_t)ky + (size_t)N * kz);
                if (kx == 0 && ky == 0 && kz == 0) {
                    chi[i].re = 0.; chi[i].im = 0.;
                    continue;
                }
                cplx S00 = S[i], S01 = S[nk + i], S02 = S[2 * nk + i];
                cplx S11 = S[3 * nk + i], S12 = S[4 * nk + i], S22 = S[5 * nk + i];
                double g0 = g[kx], g1 = g[ky], g2 = g[kz];
                cplx num = rmul(g1 + g2 - 2. * g0, S00);
                num = cadd(num, rmul(g0 + g2 - 2. * g1, S11));
                num = cadd(num, rmul(g0 + g1 - 2. * g2, S22));
                num = csub(num, cmul(cmul(rmul(6., ks[kx]), ks[ky]), S01));
                num = csub(num, cmul(cmul(rmul(6., ks[kx]), ks[kz]), S02));
                num = csub(num, cmul(cmul(rmul(6., ks[ky]), ks[kz]), S12));
                double den = 2. * (g0 + g1 + g2) * (g0 + g1 + g2) * (double)N;
                cplx r = cdivr(num, den);
                if (add) chi[i] = cadd(chi[i], r);
                else chi[i] = r;
            }
    free(g);
    free(ks);
}

/* evolveFTvector, gevolution.hpp:284-330 */
void orc_evolve_ft_vector(int N, const double *S_, double *B_, double a2dtau)
{
    const int H = N / 2 + 1;
    const size_t nk = (size_t)N * N * H;
    const cplx *S = (const cplx *)S_;
    cplx *B = (cplx *)B_;
    double *g = (double *)malloc(sizeof(double) * N);
    cplx *ks = (cplx *)malloc(sizeof(cplx) * N);
    ktables(N, g, ks);
    #pragma omp parallel for schedule(static)
    for (int kz = 0; kz < N; kz++)
        for (int ky = 0; ky < N; ky++)
            for (int kx = 0; kx < H; kx++) {
                size_t i = (size_t)kx + (size_t)H * ((size_t)ky + (size_t)N * kz);
                if (kx == 0 && ky == 0 && kz == 0) {
                    for (int c = 0; c < 3; c++) { B[c * nk + i].re = 0.; B[c * nk + i].im = 0.; }
                    continue;
                }
                cplx S00 = S[i], S01 = S[nk + i], S02 = S[2 * nk + i];
                cplx S11 = S[3 * nk + i], S12 = S[4 * nk + i], S22 = S[5 * nk + i];
                double g0 = g[kx], g1 = g[ky], g2 = g[kz];
                cplx k0 = ks[kx], k1 = ks[ky], k2 = ks[kz];
                double k4 = g0 + g1 + g2;
                k4 *= k4;
                cplx pre = {0., -2. * a2dtau / k4};
                cplx t, u;
                /* component 0 (gevolution.hpp:317-319) */
                t = rmul(g1 + g2, S00);
                t = csub(t, rmul(g1, S11));
                t = csub(t, rmul(g2, S22));
                t = csub(t, cmul(cmul(rmul(2., k1), k2), S12));
                t = cmul(cconj(k0), t);
                u = cadd(cmul(k1, S01), cmul(k2, S02));
                t = cadd(t, rmul(g1 + g2 - g0, u));
                B[i] = cadd(B[i], cmul(pre, t));
                /* component 1 (gevolution.hpp:320-322) */
                t = rmul(g0 + g2, S11);
                t = csub(t, rmul(g0, S00));
                t = csub(t, rmul(g2, S22));
                t = csub(t, cmul(cmul(rmul(2., k0), k2), S02));
                t = cmul(cconj(k1), t);
                u = cadd(cmul(k0, S01), cmul(k2, S12));
                t = cadd(t, rmul(g0 + g2 - g1, u));
                B[nk + i] = cadd(B[nk + i], cmul(pre, t));
                /* component 2 (gevolution.hpp:323-325) */
                t = rmul(g0 + g1, S22);
                t = csub(t, rmul(g0, S00));
                t = csub(t, rmul(g1, S11));
                t = csub(t, cmul(cmul(rmul(2., k0), k1), S01));
                t = cmul(cconj(k2), t);
                u = cadd(cmul(k0, S02), cmul(k1, S12));
                t = cadd(t, rmul(g0 + g1 - g2, u));
                B[2 * nk + i] = cadd(B[2 * nk + i], cmul(pre, t));
            }
    free(g);
    free(ks);
}

/* projectFTvector, gevolution.hpp:350-392 (SiFT may alias BiFT) */
void orc_project_ft_vector(int N, const double *S_, double *B_, double coeff, double modif)
{
    const int H = N / 2 + 1;
    const size_t nk = (size_t)N * N * H;
    const cplx *S = (const cplx *)S_;
    cplx *B = (cplx *)B_;
    double *g = (double *)malloc(sizeof(double) * N);
    cplx *ks = (cplx *)malloc(sizeof(cplx) * N);
    ktables(N, g, ks);
    #pragma omp parallel for schedule(static)
    for (int kz = 0; kz < N; kz++)
        for (int ky = 0; ky < N; ky++)
            for (int kx = 0; kx < H; kx++) {
                size_t i = (size_t)kx + (size_t)H * ((size_t)ky + (size_t)N * kz);
                if (kx == 0 && ky == 0 && kz == 0) {
                    for (int c = 0; c < 3; c++) { B[c * nk + i].re = 0.; B[c * nk + i].im = 0.; }
                    continue;
                }
                cplx S0 = S[i], S1 = S[nk + i], S2 = S[2 * nk + i];
                double k2 = g[kx] + g[ky] + g[kz];
                cplx tmp = cadd(cadd(cmul(ks[kx], S0), cmul(ks[ky], S1)), cmul(ks[kz], S2));
                tmp = cdivr(tmp, k2);
                B[i] = cdivr(cmulr(cmulr(csub(S0, cmul(cconj(ks[kx]), tmp)), 4.), coeff), k2 + modif);
                B[nk + i] = cdivr(cmulr(cmulr(csub(S1, cmul(cconj(ks[ky]), tmp)), 4.), coeff), k2 + modif);
                B[2 * nk + i] = cdivr(cmulr(cmulr(csub(S2, cmul(cconj(ks[kz]), tmp)), 4.), coeff), k2 + modif);
            }
    free(g);
    free(ks);
}

/* ======================================================================== */
/* stencils                                                                  */
/* ======================================================================== */

/* prepareFTsource (1), tensor form, gevolution.hpp:56-147 with PHINONLINEAR
   on and ORIGINALMETRIC off (makefile:21-23).  Sij may alias Tij. */
void orc_prepare_ft_source_tensor(int N, const double *phi, const double *Tij, double *Sij,
                                  double coeff)
{
    const size_t n3 = (size_t)N * N * N;
    #pragma omp parallel for schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                size_t i = (size_t)lin(N, x, y, z);
                double p = phi[i];
                double px = at(phi, N, x, y, z, 1, 0, 0), mx = at(phi, N, x, y, z, -1, 0, 0);
                double py = at(phi, N, x, y, z, 0, 1, 0), my = at(phi, N, x, y, z, 0, -1, 0);
                double pz = at(phi, N, x, y, z, 0, 0, 1), mz = at(phi, N, x, y, z, 0, 0, -1);
                double pxy = at(phi, N, x, y, z, 1, 1, 0);
                double pxz = at(phi, N, x, y, z, 1, 0, 1);
                double pyz = at(phi, N, x, y, z, 0, 1, 1);
                double s;
                /* 0-0 */
                s = coeff * Tij[i];
                s += 0.5 * (px - mx) * (px - mx);
                Sij[i] = s;
                /* 1-1 */
                s = coeff * Tij[3 * n3 + i];
                s += 0.5 * (py - my) * (py - my);
                Sij[3 * n3 + i] = s;
                /* 2-2 */
                s = coeff * Tij[5 * n3 + i];
                s += 0.5 * (pz - mz) * (pz - mz);
                Sij[5 * n3 + i] = s;
                /* 0-1 */
                s = coeff * Tij[n3 + i];
                s += px * py - p * pxy;
                s += 0.5 * p * p;
                s -= 0.5 * px * px;
                s -= 0.5 * py * py;
                s += 0.5 * pxy * pxy;
                Sij[n3 + i] = s;
                /* 0-2 */
                s = coeff * Tij[2 * n3 + i];
                s += px * pz - p * pxz;
                s += 0.5 * p * p;
                s -= 0.5 * px * px;
                s -= 0.5 * pz * pz;
                s += 0.5 * pxz * pxz;
                Sij[2 * n3 + i] = s;
                /* 1-2 */
                s = coeff * Tij[4 * n3 + i];
                s += py * pz - p * pyz;
                s += 0.5 * p * p;
                s -= 0.5 * py * py;
                s -= 0.5 * pz * pz;
                s += 0.5 * pyz * pyz;
                Sij[4 * n3 + i] = s;
            }
}

/* prepareFTsource (2), scalar form, gevolution.hpp:169-192 (PHINONLINEAR on,
   ORIGINALMETRIC off).  result may alias source. */
void orc_prepare_ft_source_scalar(int N, const double *phi, const double *chi,
                                  const double *source, double bgmodel, double *result,
                                  double coeff, double coeff2, double coeff3)
{
    #pragma omp parallel for schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                size_t i = (size_t)lin(N, x, y, z);
                double p = phi[i];
                double px = at(phi, N, x, y, z, 1, 0, 0), mx = at(phi, N, x, y, z, -1, 0, 0);
                double py = at(phi, N, x, y, z, 0, 1, 0), my = at(phi, N, x, y, z, 0, -1, 0);
                double pz = at(phi, N, x, y, z, 0, 0, 1), mz = at(phi, N, x, y, z, 0, 0, -1);
                double r = coeff2 * (source[i] - bgmodel);
                r *= 1. - 2. * p;
                r += 0.125 * (mx - px) * (mx - px);
                r += 0.125 * (my - py) * (my - py);
                r += 0.125 * (mz - pz) * (mz - pz);
                r += (coeff3 - coeff) * p - coeff3 * chi[i];
                result[i] = r;
            }
}

/* ======================================================================== */
/* particle filing                                                           */
/* ======================================================================== */

/* Cell of a coordinate: floor(pos/dx) with a true division (dx = 1/N), the
   LATfield2 re-filing rule (Particles_gevolution.hpp:170 idiom). */
static inline int cell_of(double p, double dx) { return (int)floor(p / dx); }

void orc_file_particles(int N, int64_t n, const double *pos, const int64_t *id, uint32_t *cell,
                        uint32_t *count, uint64_t *start, int64_t *perm)
{
    const double dx = 1.0 / (double)N;
    const size_t n3 = (size_t)N * N * N;
    memset(count, 0, sizeof(uint32_t) * n3);
    for (int64_t p = 0; p < n; p++) {
        int cx = cell_of(pos[3 * p], dx), cy = cell_of(pos[3 * p + 1], dx), cz = cell_of(pos[3 * p + 2], dx);
        uint32_t c = (uint32_t)lin(N, cx, cy, cz);
        cell[p] = c;
        count[c]++;
    }
    start[0] = 0;
    for (size_t c = 0; c < n3; c++) start[c + 1] = start[c] + count[c];
    uint64_t *fill = (uint64_t *)malloc(sizeof(uint64_t) * n3);
    memcpy(fill, start, sizeof(uint64_t) * n3);
    for (int64_t p = 0; p < n; p++) perm[fill[cell[p]]++] = p;
    free(fill);
    /* ascending ID inside each cell (insertion sort; cells hold few particles) */
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)n3; c++) {
        for (uint64_t i = start[c] + 1; i < start[c + 1]; i++) {
            int64_t v = perm[i];
            uint64_t j = i;
            while (j > start[c] && id[perm[j - 1]] > id[v]) { perm[j] = perm[j - 1]; j--; }
            perm[j] = v;
        }
    }
}

/* ======================================================================== */
/* projections                                                               */
/* ======================================================================== */

/* Ordered scatter.  The reference projections loop over source sites in
   site order (x fastest), build a local cube of contributions per site and
   add it onto the target sites (gevolution.hpp:958-1021 and alike).  Each
   source at plane z writes planes z and z+1 only.  To run in parallel while
   adding onto every target in exactly the sequential site order, the planes
   are split into contiguous chunks, one per thread; a chunk [z0,z1) first
   recomputes the plane-(z0) part of slab z0-1, then sweeps its own slabs in
   order; slab z1-1 writes only its own plane (the next chunk recomputes the
   rest), and the plane-0 part of slab N-1 is applied right after slab 0,
   where it falls in site order.  The sums are therefore identical to those
   of the single-threaded reference loop for any thread count. */
typedef struct { int64_t idx; int upper; double v; } contrib;
typedef int (*source_fn)(const void *ctx, int x, int y, int z, int64_t s, uint64_t b0,
                         uint64_t b1, contrib *out);

enum { PASS_LOWER = 1, PASS_UPPER = 2, PASS_BOTH = 3 };

static void slab_pass(int N, const uint64_t *start, source_fn fn, const void *ctx, double *base,
                      int z, int which)
{
    contrib c[40];
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
            int64_t s = lin(N, x, y, z);
            uint64_t b0 = start[s], b1 = start[s + 1];
            if (b0 == b1) continue;
            int nc = fn(ctx, x, y, z, s, b0, b1, c);
            for (int i = 0; i < nc; i++)
                if (which & (c[i].upper ? PASS_UPPER : PASS_LOWER)) base[c[i].idx] += c[i].v;
        }
}

static void ordered_scatter(int N, const uint64_t *start, source_fn fn, const void *ctx,
                            double *base)
{
    int nchunks = 1;
#ifdef _OPENMP
    nchunks = omp_get_max_threads();
#endif
    if (nchunks > N) nchunks = N;
    #pragma omp parallel for schedule(static, 1)
    for (int k = 0; k < nchunks; k++) {
        int z0 = (int)((int64_t)k * N / nchunks), z1 = (int)((int64_t)(k + 1) * N / nchunks);
        if (z0 > 0) slab_pass(N, start, fn, ctx, base, z0 - 1, PASS_UPPER);
        for (int z = z0; z < z1; z++) {
            slab_pass(N, start, fn, ctx, base, z, z == z1 - 1 ? PASS_LOWER : PASS_BOTH);
            if (z == 0 && N > 1) slab_pass(N, start, fn, ctx, base, N - 1, PASS_UPPER);
        }
    }
}

typedef struct {
    int N;
    double dx, a, mass;
    const double *pos, *q, *phi;
    const int64_t *perm;
    size_t n3;
} proj_ctx;

static void corner_phi(const proj_ctx *P, int x, int y, int z, double lp[8])
{
    const int N = P->N;
    if (!P->phi) {
        for (int i = 0; i < 8; i++) lp[i] = 0.;
        return;
    }
    const double *phi = P->phi;
    lp[0] = at(phi, N, x, y, z, 0, 0, 0);
    lp[1] = at(phi, N, x, y, z, 0, 0, 1);
    lp[2] = at(phi, N, x, y, z, 0, 1, 0);
    lp[3] = at(phi, N, x, y, z, 0, 1, 1);
    lp[4] = at(phi, N, x, y, z, 1, 0, 0);
    lp[5] = at(phi, N, x, y, z, 1, 0, 1);
    lp[6] = at(phi, N, x, y, z, 1, 1, 0);
    lp[7] = at(phi, N, x, y, z, 1, 1, 1);
}

static inline contrib mk(int N, int x, int y, int z, int dx_, int dy_, int dz_, size_t plane,
                         double v)
{
    contrib c;
    c.idx = (int64_t)plane + lin(N, wrapi(x + dx_, N), wrapi(y + dy_, N), wrapi(z + dz_, N));
    c.upper = dz_;
    c.v = v;
    return c;
}

/* projection_T00_project per-site body, gevolution.hpp:960-1020 */
static int t00_source(const void *ctx_, int x, int y, int z, int64_t s, uint64_t b0, uint64_t b1,
                      contrib *out)
{
    (void)s;
    const proj_ctx *P = (const proj_ctx *)ctx_;
    const int N = P->N;
    const double dx = P->dx, a = P->a, mass = P->mass;
    double referPos[3] = {x * dx, y * dx, z * dx};
    double lc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double lp[8];
    corner_phi(P, x, y, z, lp);
    double e = a, f = 0.;
    for (uint64_t k = b0; k < b1; k++) {
        int64_t p = P->perm[k];
        double up[3], dn[3];
        for (int i = 0; i < 3; i++) {
            up[i] = (P->pos[3 * p + i] - referPos[i]) / dx;
            dn[i] = 1.0 - up[i];
        }
        if (P->phi) {
            const double *qq = P->q + 3 * p;
            f = qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2];
            e = sqrt(f + a * a);
            f = 3. * e + f / e;
        }
        lc[0] += dn[0] * dn[1] * dn[2] * (e + f * lp[0]);
        lc[1] += dn[0] * dn[1] * up[2] * (e + f * lp[1]);
        lc[2] += dn[0] * up[1] * dn[2] * (e + f * lp[2]);
        lc[3] += dn[0] * up[1] * up[2] * (e + f * lp[3]);
        lc[4] += up[0] * dn[1] * dn[2] * (e + f * lp[4]);
        lc[5] += up[0] * dn[1] * up[2] * (e + f * lp[5]);
        lc[6] += up[0] * up[1] * dn[2] * (e + f * lp[6]);
        lc[7] += up[0] * up[1] * up[2] * (e + f * lp[7]);
    }
    /* target order of gevolution.hpp:1012-1019 */
    out[0] = mk(N, x, y, z, 0, 0, 0, 0, lc[0] * mass);
    out[1] = mk(N, x, y, z, 0, 0, 1, 0, lc[1] * mass);
    out[2] = mk(N, x, y, z, 0, 1, 0, 0, lc[2] * mass);
    out[3] = mk(N, x, y, z, 0, 1, 1, 0, lc[3] * mass);
    out[4] = mk(N, x, y, z, 1, 0, 0, 0, lc[4] * mass);
    out[5] = mk(N, x, y, z, 1, 0, 1, 0, lc[5] * mass);
    out[6] = mk(N, x, y, z, 1, 1, 0, 0, lc[6] * mass);
    out[7] = mk(N, x, y, z, 1, 1, 1, 0, lc[7] * mass);
    return 8;
}

/* projection_T00_project, gevolution.hpp:926-1022 */
void orc_projection_T00_project(int N, int64_t n, const double *pos, const double *q,
                                const uint64_t *start, const int64_t *perm, double pmass,
                                double *T00, double a, const double *phi, double coeff)
{
    (void)n;
    proj_ctx P;
    P.N = N;
    P.dx = 1.0 / (double)N;
    P.mass = coeff / (P.dx * P.dx * P.dx);
    P.mass *= pmass;
    P.mass /= a;
    P.a = a;
    P.pos = pos; P.q = q; P.phi = phi; P.perm = perm;
    P.n3 = (size_t)N * N * N;
    ordered_scatter(N, start, t00_source, &P, T00);
}

/* LATfield2 scalarProjectionCIC_project (main.cpp:402), restated as
   projection_T00_project with a = 1 and phi = NULL: plain CIC of
   coeff*m/dx^3.  Parity unpinned against LATfield2 (absent). */
void orc_scalar_projection_cic_project(int N, int64_t n, const double *pos,
                                       const uint64_t *start, const int64_t *perm, double pmass,
                                       double *rho, double coeff)
{
    orc_projection_T00_project(N, n, pos, NULL, start, perm, pmass, rho, 1.0, NULL, coeff);
}

/* projection_Tij_project per-site body, gevolution.hpp:1207-1294 */
static int tij_source(const void *ctx_, int x, int y, int z, int64_t s, uint64_t b0, uint64_t b1,
                      contrib *out)
{
    (void)s;
    const proj_ctx *P = (const proj_ctx *)ctx_;
    const int N = P->N;
    const size_t n3 = P->n3;
    const double dx = P->dx, a = P->a, mass = P->mass;
    double referPos[3] = {(double)x * dx, (double)y * dx, (double)z * dx};
    double tij[6] = {0, 0, 0, 0, 0, 0};
    double tii[24];
    for (int i = 0; i < 24; i++) tii[i] = 0.;
    double lp[8];
    corner_phi(P, x, y, z, lp);
    for (uint64_t k = b0; k < b1; k++) {
        int64_t p = P->perm[k];
        double up[3], dn[3];
        for (int i = 0; i < 3; i++) {
            up[i] = (P->pos[3 * p + i] - referPos[i]) / dx;
            dn[i] = 1.0 - up[i];
        }
        const double *qq = P->q + 3 * p;
        double f = qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2];
        double e = sqrt(f + a * a);
        f = 4. + a * a / (f + a * a);
        double w;
        for (int i = 0; i < 3; i++) {
            w = mass * qq[i] * qq[i] / e;
            tii[0 + i * 8] += w * dn[0] * dn[1] * dn[2] * (1. + f * lp[0]);
            tii[1 + i * 8] += w * dn[0] * dn[1] * up[2] * (1. + f * lp[1]);
            tii[2 + i * 8] += w * dn[0] * up[1] * dn[2] * (1. + f * lp[2]);
            tii[3 + i * 8] += w * dn[0] * up[1] * up[2] * (1. + f * lp[3]);
            tii[4 + i * 8] += w * up[0] * dn[1] * dn[2] * (1. + f * lp[4]);
            tii[5 + i * 8] += w * up[0] * dn[1] * up[2] * (1. + f * lp[5]);
            tii[6 + i * 8] += w * up[0] * up[1] * dn[2] * (1. + f * lp[6]);
            tii[7 + i * 8] += w * up[0] * up[1] * up[2] * (1. + f * lp[7]);
        }
        w = mass * qq[0] * qq[1] / e;
        tij[0] += w * dn[2] * (1. + f * 0.25 * (lp[0] + lp[2] + lp[4] + lp[6]));
        tij[1] += w * up[2] * (1. + f * 0.25 * (lp[1] + lp[3] + lp[5] + lp[7]));
        w = mass * qq[0] * qq[2] / e;
        tij[2] += w * dn[1] * (1. + f * 0.25 * (lp[0] + lp[1] + lp[4] + lp[5]));
        tij[3] += w * up[1] * (1. + f * 0.25 * (lp[2] + lp[3] + lp[6] + lp[7]));
        w = mass * qq[1] * qq[2] / e;
        tij[4] += w * dn[0] * (1. + f * 0.25 * (lp[0] + lp[1] + lp[2] + lp[3]));
        tij[5] += w * up[0] * (1. + f * 0.25 * (lp[4] + lp[5] + lp[6] + lp[7]));
    }
    /* component planes: 00=0, 01=1, 02=2, 11=3, 12=4, 22=5; diag(i) = 0,3,5.
       Target order of gevolution.hpp:1277-1294. */
    const size_t dg[3] = {0, 3 * n3, 5 * n3};
    int c = 0;
    for (int i = 0; i < 3; i++) out[c++] = mk(N, x, y, z, 0, 0, 0, dg[i], tii[8 * i]);
    out[c++] = mk(N, x, y, z, 0, 0, 0, 1 * n3, tij[0]);
    out[c++] = mk(N, x, y, z, 0, 0, 0, 2 * n3, tij[2]);
    out[c++] = mk(N, x, y, z, 0, 0, 0, 4 * n3, tij[4]);
    for (int i = 0; i < 3; i++) out[c++] = mk(N, x, y, z, 1, 0, 0, dg[i], tii[4 + 8 * i]);
    out[c++] = mk(N, x, y, z, 1, 0, 0, 4 * n3, tij[5]);
    for (int i = 0; i < 3; i++) out[c++] = mk(N, x, y, z, 0, 1, 0, dg[i], tii[2 + 8 * i]);
    out[c++] = mk(N, x, y, z, 0, 1, 0, 2 * n3, tij[3]);
    for (int i = 0; i < 3; i++) out[c++] = mk(N, x, y, z, 0, 0, 1, dg[i], tii[1 + 8 * i]);
    out[c++] = mk(N, x, y, z, 0, 0, 1, 1 * n3, tij[1]);
    for (int i = 0; i < 3; i++) out[c++] = mk(N, x, y, z, 1, 1, 0, dg[i], tii[6 + 8 * i]);
    for (int i = 0; i < 3; i++) out[c++] = mk(N, x, y, z, 1, 0, 1, dg[i], tii[5 + 8 * i]);
    for (int i = 0; i < 3; i++) out[c++] = mk(N, x, y, z, 0, 1, 1, dg[i], tii[3 + 8 * i]);
    for (int i = 0; i < 3; i++) out[c++] = mk(N, x, y, z, 1, 1, 1, dg[i], tii[7 + 8 * i]);
    return c;
}

/* projection_Tij_project, gevolution.hpp:1172-1297 */
void orc_projection_Tij_project(int N, int64_t n, const double *pos, const double *q,
                                const uint64_t *start, const int64_t *perm, double pmass,
                                double *Tij, double a, const double *phi, double coeff)
{
    (void)n;
    proj_ctx P;
    P.N = N;
    P.dx = 1.0 / (double)N;
    P.mass = coeff / (P.dx * P.dx * P.dx);
    P.mass *= pmass;
    P.mass /= a;
    P.a = a;
    P.pos = pos; P.q = q; P.phi = phi; P.perm = perm;
    P.n3 = (size_t)N * N * N;
    ordered_scatter(N, start, tij_source, &P, Tij);
}

/* projection_T0i_project per-site body, gevolution.hpp:1078-1145 */
static int t0i_source(const void *ctx_, int x, int y, int z, int64_t s, uint64_t b0, uint64_t b1,
                      contrib *out)
{
    (void)s;
    const proj_ctx *P = (const proj_ctx *)ctx_;
    const int N = P->N;
    const size_t n3 = P->n3;
    const double dx = P->dx, mass = P->mass;
    double referPos[3] = {x * dx, y * dx, z * dx};
    double qi[12];
    for (int i = 0; i < 12; i++) qi[i] = 0.;
    double lp[8];
    corner_phi(P, x, y, z, lp);
    for (uint64_t k = b0; k < b1; k++) {
        int64_t p = P->perm[k];
        double up[3], dn[3];
        for (int i = 0; i < 3; i++) {
            up[i] = (P->pos[3 * p + i] - referPos[i]) / dx;
            dn[i] = 1.0 - up[i];
        }
        const double *qq = P->q + 3 * p;
        double w = mass * qq[0];
        qi[0] += w * dn[1] * dn[2];
        qi[1] += w * up[1] * dn[2];
        qi[2] += w * dn[1] * up[2];
        qi[3] += w * up[1] * up[2];
        w = mass * qq[1];
        qi[4] += w * dn[0] * dn[2];
        qi[5] += w * up[0] * dn[2];
        qi[6] += w * dn[0] * up[2];
        qi[7] += w * up[0] * up[2];
        w = mass * qq[2];
        qi[8] += w * dn[0] * dn[1];
        qi[9] += w * up[0] * dn[1];
        qi[10] += w * dn[0] * up[1];
        qi[11] += w * up[0] * up[1];
    }
    /* target order of gevolution.hpp:1129-1144 */
    out[0] = mk(N, x, y, z, 0, 0, 0, 0 * n3, qi[0] * (1. + lp[0] + lp[4]));
    out[1] = mk(N, x, y, z, 0, 0, 0, 1 * n3, qi[4] * (1. + lp[0] + lp[2]));
    out[2] = mk(N, x, y, z, 0, 0, 0, 2 * n3, qi[8] * (1. + lp[0] + lp[1]));
    out[3] = mk(N, x, y, z, 1, 0, 0, 1 * n3, qi[5] * (1. + lp[4] + lp[6]));
    out[4] = mk(N, x, y, z, 1, 0, 0, 2 * n3, qi[9] * (1. + lp[4] + lp[5]));
    out[5] = mk(N, x, y, z, 0, 1, 0, 0 * n3, qi[1] * (1. + lp[2] + lp[6]));
    out[6] = mk(N, x, y, z, 0, 1, 0, 2 * n3, qi[10] * (1. + lp[2] + lp[3]));
    out[7] = mk(N, x, y, z, 0, 0, 1, 0 * n3, qi[2] * (1. + lp[1] + lp[5]));
    out[8] = mk(N, x, y, z, 0, 0, 1, 1 * n3, qi[6] * (1. + lp[1] + lp[3]));
    out[9] = mk(N, x, y, z, 0, 1, 1, 0 * n3, qi[3] * (1. + lp[3] + lp[7]));
    out[10] = mk(N, x, y, z, 1, 0, 1, 1 * n3, qi[7] * (1. + lp[5] + lp[7]));
    out[11] = mk(N, x, y, z, 1, 1, 0, 2 * n3, qi[11] * (1. + lp[6] + lp[7]));
    return 12;
}

/* projection_T0i_project, gevolution.hpp:1045-1147 */
void orc_projection_T0i_project(int N, int64_t n, const double *pos, const double *q,
                                const uint64_t *start, const int64_t *perm, double pmass,
                                double *T0i, const double *phi, double coeff)
{
    (void)n;
    proj_ctx P;
    P.N = N;
    P.dx = 1.0 / (double)N;
    P.mass = coeff / (P.dx * P.dx * P.dx);
    P.mass *= pmass;
    P.a = 1.;
    P.pos = pos; P.q = q; P.phi = phi; P.perm = perm;
    P.n3 = (size_t)N * N * N;
    ordered_scatter(N, start, t0i_source, &P, T0i);
}

/* ======================================================================== */
/* kicks and drifts                                                          */
/* ======================================================================== */

/* cell and fractional offset of one particle: the filed site plus
   ref_dist = modf(pos/dx) (Particles_gevolution.hpp:170 idiom) */
static inline void locate(const double *p, double dx, int c[3], double rd[3])
{
    for (int i = 0; i < 3; i++) {
        double t = p[i] / dx;
        double ip = floor(t);
        c[i] = (int)ip;
        rd[i] = t - ip;
    }
}

#define F(f, dx_, dy_, dz_) at(f, N, c[0], c[1], c[2], dx_, dy_, dz_)
#define BF(comp, dx_, dy_, dz_) at(Bi + (size_t)(comp) * n3, N, c[0], c[1], c[2], dx_, dy_, dz_)

/* update_q, gevolution.hpp:570-678 (GRADIENT_ORDER 1) */
static double update_q_one(int N, double dx, const double *pos, double *q, double dtau,
                           const double *phi, const double *chi, const double *Bi, int nfield,
                           const double *params)
{
    const size_t n3 = (size_t)N * N * N;
    int c[3];
    double rd[3];
    locate(pos, dx, c, rd);
    double gradphi[3] = {0, 0, 0};
    double pgradB[3] = {0, 0, 0};
    double v2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    double e2 = v2 + params[0] * params[0];

    gradphi[0] = (1. - rd[1]) * (1. - rd[2]) * (F(phi, 1, 0, 0) - F(phi, 0, 0, 0));
    gradphi[1] = (1. - rd[0]) * (1. - rd[2]) * (F(phi, 0, 1, 0) - F(phi, 0, 0, 0));
    gradphi[2] = (1. - rd[0]) * (1. - rd[1]) * (F(phi, 0, 0, 1) - F(phi, 0, 0, 0));
    gradphi[0] += rd[1] * (1. - rd[2]) * (F(phi, 1, 1, 0) - F(phi, 0, 1, 0));
    gradphi[1] += rd[0] * (1. - rd[2]) * (F(phi, 1, 1, 0) - F(phi, 1, 0, 0));
    gradphi[2] += rd[0] * (1. - rd[1]) * (F(phi, 1, 0, 1) - F(phi, 1, 0, 0));
    gradphi[0] += (1. - rd[1]) * rd[2] * (F(phi, 1, 0, 1) - F(phi, 0, 0, 1));
    gradphi[1] += (1. - rd[0]) * rd[2] * (F(phi, 0, 1, 1) - F(phi, 0, 0, 1));
    gradphi[2] += (1. - rd[0]) * rd[1] * (F(phi, 0, 1, 1) - F(phi, 0, 1, 0));
    gradphi[0] += rd[1] * rd[2] * (F(phi, 1, 1, 1) - F(phi, 0, 1, 1));
    gradphi[1] += rd[0] * rd[2] * (F(phi, 1, 1, 1) - F(phi, 1, 0, 1));
    gradphi[2] += rd[0] * rd[1] * (F(phi, 1, 1, 1) - F(phi, 1, 1, 0));

    gradphi[0] *= (v2 + e2) / e2;
    gradphi[1] *= (v2 + e2) / e2;
    gradphi[2] *= (v2 + e2) / e2;

    if (nfield >= 2 && chi != NULL) {
        gradphi[0] -= (1. - rd[1]) * (1. - rd[2]) * (F(chi, 1, 0, 0) - F(chi, 0, 0, 0));
        gradphi[1] -= (1. - rd[0]) * (1. - rd[2]) * (F(chi, 0, 1, 0) - F(chi, 0, 0, 0));
        gradphi[2] -= (1. - rd[0]) * (1. - rd[1]) * (F(chi, 0, 0, 1) - F(chi, 0, 0, 0));
        gradphi[0] -= rd[1] * (1. - rd[2]) * (F(chi, 1, 1, 0) - F(chi, 0, 1, 0));
        gradphi[1] -= rd[0] * (1. - rd[2]) * (F(chi, 1, 1, 0) - F(chi, 1, 0, 0));
        gradphi[2] -= rd[0] * (1. - rd[1]) * (F(chi, 1, 0, 1) - F(chi, 1, 0, 0));
        gradphi[0] -= (1. - rd[1]) * rd[2] * (F(chi, 1, 0, 1) - F(chi, 0, 0, 1));
        gradphi[1] -= (1. - rd[0]) * rd[2] * (F(chi, 0, 1, 1) - F(chi, 0, 0, 1));
        gradphi[2] -= (1. - rd[0]) * rd[1] * (F(chi, 0, 1, 1) - F(chi, 0, 1, 0));
        gradphi[0] -= rd[1] * rd[2] * (F(chi, 1, 1, 1) - F(chi, 0, 1, 1));
        gradphi[1] -= rd[0] * rd[2] * (F(chi, 1, 1, 1) - F(chi, 1, 0, 1));
        gradphi[2] -= rd[0] * rd[1] * (F(chi, 1, 1, 1) - F(chi, 1, 1, 0));
    }

    e2 = sqrt(e2);

    if (nfield >= 3 && Bi != NULL) {
        pgradB[0] = ((1. - rd[2]) * (BF(1, 1, 0, 0) - BF(1, 0, 0, 0)) + rd[2] * (BF(1, 1, 0, 1) - BF(1, 0, 0, 1))) * q[1];
        pgradB[0] += ((1. - rd[1]) * (BF(2, 1, 0, 0) - BF(2, 0, 0, 0)) + rd[1] * (BF(2, 1, 1, 0) - BF(2, 0, 1, 0))) * q[2];
        pgradB[0] += (1. - rd[1]) * (1. - rd[2]) * ((rd[0] - 1.) * BF(0, -1, 0, 0) + (1. - 2. * rd[0]) * BF(0, 0, 0, 0) + rd[0] * BF(0, 1, 0, 0)) * q[0];
        pgradB[0] += rd[1] * (1. - rd[2]) * ((rd[0] - 1.) * BF(0, -1, 1, 0) + (1. - 2. * rd[0]) * BF(0, 0, 1, 0) + rd[0] * BF(0, 1, 1, 0)) * q[0];
        pgradB[0] += (1. - rd[1]) * rd[2] * ((rd[0] - 1.) * BF(0, -1, 0, 1) + (1. - 2. * rd[0]) * BF(0, 0, 0, 1) + rd[0] * BF(0, 1, 0, 1)) * q[0];
        pgradB[0] += rd[1] * rd[2] * ((rd[0] - 1.) * BF(0, -1, 1, 1) + (1. - 2. * rd[0]) * BF(0, 0, 1, 1) + rd[0] * BF(0, 1, 1, 1)) * q[0];

        pgradB[1] = ((1. - rd[0]) * (BF(2, 0, 1, 0) - BF(2, 0, 0, 0)) + rd[0] * (BF(2, 1, 1, 0) - BF(2, 1, 0, 0))) * q[2];
        pgradB[1] += ((1. - rd[2]) * (BF(0, 0, 1, 0) - BF(0, 0, 0, 0)) + rd[2] * (BF(0, 0, 1, 1) - BF(0, 0, 0, 1))) * q[0];
        pgradB[1] += (1. - rd[0]) * (1. - rd[2]) * ((rd[1] - 1.) * BF(1, 0, -1, 0) + (1. - 2. * rd[1]) * BF(1, 0, 0, 0) + rd[1] * BF(1, 0, 1, 0)) * q[1];
        pgradB[1] += rd[0] * (1. - rd[2]) * ((rd[1] - 1.) * BF(1, 1, -1, 0) + (1. - 2. * rd[1]) * BF(1, 1, 0, 0) + rd[1] * BF(1, 1, 1, 0)) * q[1];
        pgradB[1] += (1. - rd[0]) * rd[2] * ((rd[1] - 1.) * BF(1, 0, -1, 1) + (1. - 2. * rd[1]) * BF(1, 0, 0, 1) + rd[1] * BF(1, 0, 1, 1)) * q[1];
        pgradB[1] += rd[0] * rd[2] * ((rd[1] - 1.) * BF(1, 1, -1, 1) + (1. - 2. * rd[1]) * BF(1, 1, 0, 1) + rd[1] * BF(1, 1, 1, 1)) * q[1];

        pgradB[2] = ((1. - rd[1]) * (BF(0, 0, 0, 1) - BF(0, 0, 0, 0)) + rd[1] * (BF(0, 0, 1, 1) - BF(0, 0, 1, 0))) * q[0];
        pgradB[2] += ((1. - rd[0]) * (BF(1, 0, 0, 1) - BF(1, 0, 0, 0)) + rd[0] * (BF(1, 1, 0, 1) - BF(1, 1, 0, 0))) * q[1];
        pgradB[2] += (1. - rd[0]) * (1. - rd[1]) * ((rd[2] - 1.) * BF(2, 0, 0, -1) + (1. - 2. * rd[2]) * BF(2, 0, 0, 0) + rd[2] * BF(2, 0, 0, 1)) * q[2];
        pgradB[2] += rd[0] * (1. - rd[1]) * ((rd[2] - 1.) * BF(2, 1, 0, -1) + (1. - 2. * rd[2]) * BF(2, 1, 0, 0) + rd[2] * BF(2, 1, 0, 1)) * q[2];
        pgradB[2] += (1. - rd[0]) * rd[1] * ((rd[2] - 1.) * BF(2, 0, 1, -1) + (1. - 2. * rd[2]) * BF(2, 0, 1, 0) + rd[2] * BF(2, 0, 1, 1)) * q[2];
        pgradB[2] += rd[0] * rd[1] * ((rd[2] - 1.) * BF(2, 1, 1, -1) + (1. - 2. * rd[2]) * BF(2, 1, 1, 0) + rd[2] * BF(2, 1, 1, 1)) * q[2];

        gradphi[0] += pgradB[0] / params[1] / e2;
        gradphi[1] += pgradB[1] / params[1] / e2;
        gradphi[2] += pgradB[2] / params[1] / e2;
    }

    v2 = 0.;
    for (int i = 0; i < 3; i++) {
        q[i] -= dtau * e2 * gradphi[i] / dx;
        v2 += q[i] * q[i];
    }
    return v2 / params[0] / params[0];
}

/* update_q_Newton, gevolution.hpp:709-776 */
static double update_q_newton_one(int N, double dx, const double *pos, double *q, double dtau,
                                  const double *psi, const double *chi, int nfield,
                                  const double *params)
{
    int c[3];
    double rd[3];
    locate(pos, dx, c, rd);
    double g[3];
    g[0] = (1. - rd[1]) * (1. - rd[2]) * (F(psi, 1, 0, 0) - F(psi, 0, 0, 0));
    g[1] = (1. - rd[0]) * (1. - rd[2]) * (F(psi, 0, 1, 0) - F(psi, 0, 0, 0));
    g[2] = (1. - rd[0]) * (1. - rd[1]) * (F(psi, 0, 0, 1) - F(psi, 0, 0, 0));
    g[0] += rd[1] * (1. - rd[2]) * (F(psi, 1, 1, 0) - F(psi, 0, 1, 0));
    g[1] += rd[0] * (1. - rd[2]) * (F(psi, 1, 1, 0) - F(psi, 1, 0, 0));
    g[2] += rd[0] * (1. - rd[1]) * (F(psi, 1, 0, 1) - F(psi, 1, 0, 0));
    g[0] += (1. - rd[1]) * rd[2] * (F(psi, 1, 0, 1) - F(psi, 0, 0, 1));
    g[1] += (1. - rd[0]) * rd[2] * (F(psi, 0, 1, 1) - F(psi, 0, 0, 1));
    g[2] += (1. - rd[0]) * rd[1] * (F(psi, 0, 1, 1) - F(psi, 0, 1, 0));
    g[0] += rd[1] * rd[2] * (F(psi, 1, 1, 1) - F(psi, 0, 1, 1));
    g[1] += rd[0] * rd[2] * (F(psi, 1, 1, 1) - F(psi, 1, 0, 1));
    g[2] += rd[0] * rd[1] * (F(psi, 1, 1, 1) - F(psi, 1, 1, 0));
    if (nfield >= 2 && chi != NULL) {
        g[0] -= (1. - rd[1]) * (1. - rd[2]) * (F(chi, 1, 0, 0) - F(chi, 0, 0, 0));
        g[1] -= (1. - rd[0]) * (1. - rd[2]) * (F(chi, 0, 1, 0) - F(chi, 0, 0, 0));
        g[2] -= (1. - rd[0]) * (1. - rd[1]) * (F(chi, 0, 0, 1) - F(chi, 0, 0, 0));
        g[0] -= rd[1] * (1. - rd[2]) * (F(chi, 1, 1, 0) - F(chi, 0, 1, 0));
        g[1] -= rd[0] * (1. - rd[2]) * (F(chi, 1, 1, 0) - F(chi, 1, 0, 0));
        g[2] -= rd[0] * (1. - rd[1]) * (F(chi, 1, 0, 1) - F(chi, 1, 0, 0));
        g[0] -= (1. - rd[1]) * rd[2] * (F(chi, 1, 0, 1) - F(chi, 0, 0, 1));
        g[1] -= (1. - rd[0]) * rd[2] * (F(chi, 0, 1, 1) - F(chi, 0, 0, 1));
        g[2] -= (1. - rd[0]) * rd[1] * (F(chi, 0, 1, 1) - F(chi, 0, 1, 0));
        g[0] -= rd[1] * rd[2] * (F(chi, 1, 1, 1) - F(chi, 0, 1, 1));
        g[1] -= rd[0] * rd[2] * (F(chi, 1, 1, 1) - F(chi, 1, 0, 1));
        g[2] -= rd[0] * rd[1] * (F(chi, 1, 1, 1) - F(chi, 1, 1, 0));
    }
    double v2 = 0.;
    for (int i = 0; i < 3; i++) {
        q[i] -= dtau * params[0] * g[i] / dx;
        v2 += q[i] * q[i];
    }
    return v2 / params[0] / params[0];
}

/* update_pos, gevolution.hpp:810-871 */
static void update_pos_one(int N, double dx, double *pos, const double *q, double dtau,
                           const double *phi, const double *chi, const double *Bi, int nfield,
                           const double *params)
{
    const size_t n3 = (size_t)N * N * N;
    int c[3];
    double rd[3];
    locate(pos, dx, c, rd);
    double v[3];
    double v2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    double e2 = v2 + params[0] * params[0];
    double ph = 0, ch = 0;
    if (nfield >= 1) {
        ph = F(phi, 0, 0, 0) * (1. - rd[0]) * (1. - rd[1]) * (1. - rd[2]);
        ph += F(phi, 1, 0, 0) * rd[0] * (1. - rd[1]) * (1. - rd[2]);
        ph += F(phi, 0, 1, 0) * (1. - rd[0]) * rd[1] * (1. - rd[2]);
        ph += F(phi, 1, 1, 0) * rd[0] * rd[1] * (1. - rd[2]);
        ph += F(phi, 0, 0, 1) * (1. - rd[0]) * (1. - rd[1]) * rd[2];
        ph += F(phi, 1, 0, 1) * rd[0] * (1. - rd[1]) * rd[2];
        ph += F(phi, 0, 1, 1) * (1. - rd[0]) * rd[1] * rd[2];
        ph += F(phi, 1, 1, 1) * rd[0] * rd[1] * rd[2];
    }
    if (nfield >= 2) {
        ch = F(chi, 0, 0, 0) * (1. - rd[0]) * (1. - rd[1]) * (1. - rd[2]);
        ch += F(chi, 1, 0, 0) * rd[0] * (1. - rd[1]) * (1. - rd[2]);
        ch += F(chi, 0, 1, 0) * (1. - rd[0]) * rd[1] * (1. - rd[2]);
        ch += F(chi, 1, 1, 0) * rd[0] * rd[1] * (1. - rd[2]);
        ch += F(chi, 0, 0, 1) * (1. - rd[0]) * (1. - rd[1]) * rd[2];
        ch += F(chi, 1, 0, 1) * rd[0] * (1. - rd[1]) * rd[2];
        ch += F(chi, 0, 1, 1) * (1. - rd[0]) * rd[1] * rd[2];
        ch += F(chi, 1, 1, 1) * rd[0] * rd[1] * rd[2];
    }
    v2 = (1. + (3. - v2 / e2) * ph - ch) / sqrt(e2);
    v[0] = q[0] * v2;
    v[1] = q[1] * v2;
    v[2] = q[2] * v2;
    if (nfield >= 3) {
        double b[3];
        b[0] = BF(0, 0, 0, 0) * (1. - rd[1]) * (1. - rd[2]);
        b[1] = BF(1, 0, 0, 0) * (1. - rd[0]) * (1. - rd[2]);
        b[2] = BF(2, 0, 0, 0) * (1. - rd[0]) * (1. - rd[1]);
        b[1] += BF(1, 1, 0, 0) * rd[0] * (1. - rd[2]);
        b[2] += BF(2, 1, 0, 0) * rd[0] * (1. - rd[1]);
        b[0] += BF(0, 0, 1, 0) * rd[1] * (1. - rd[2]);
        b[2] += BF(2, 0, 1, 0) * (1. - rd[0]) * rd[1];
        b[0] += BF(0, 0, 0, 1) * (1. - rd[1]) * rd[2];
        b[1] += BF(1, 0, 0, 1) * (1. - rd[0]) * rd[2];
        b[1] += BF(1, 1, 0, 1) * rd[0] * rd[2];
        b[0] += BF(0, 0, 1, 1) * rd[1] * rd[2];
        b[2] += BF(2, 1, 1, 0) * rd[0] * rd[1];
        for (int l = 0; l < 3; l++) pos[l] += dtau * (v[l] + b[l] / params[1]);
    } else {
        for (int l = 0; l < 3; l++) pos[l] += dtau * v[l];
    }
}

#undef F
#undef BF

/* Periodic wrap applied by the re-filing in LATfield2's moveParticles.  A
   coordinate whose cell index floor(p/dx) would round to N is set to 0 so that
   every filed particle satisfies 0 <= floor(p/dx) < N (a measure-zero edge the
   reference leaves to LATfield2; parity unpinned there). */
static inline double wrap_pos(double p, double dx, int N)
{
    if (p >= 1.0) p -= 1.0;
    if (p < 0.0) p += 1.0;
    /* displacements beyond one box length (not produced by sane time steps)
       are reduced modulo 1; a non-finite coordinate is filed at 0 */
    if (!(p >= 0.0 && p < 1.0)) p = p - floor(p);
    if (!(p >= 0.0) || (int)floor(p / dx) >= N) p = 0.0;
    return p;
}

double orc_update_vel(int N, int64_t n, const double *pos, double *q, int method, double dtau,
                      const double *phi, const double *chi, const double *Bi, int nfield,
                      const double params[2])
{
    const double dx = 1.0 / (double)N;
    double vmax = 0.;
    #pragma omp parallel for schedule(static) reduction(max : vmax)
    for (int64_t p = 0; p < n; p++) {
        double r;
        if (method == ORC_Q_NEWTON)
            r = update_q_newton_one(N, dx, pos + 3 * p, q + 3 * p, dtau, phi, chi, nfield, params);
        else
            r = update_q_one(N, dx, pos + 3 * p, q + 3 * p, dtau, phi, chi, Bi, nfield, params);
        if (r > vmax) vmax = r;
    }
    return sqrt(vmax);
}

void orc_move_particles(int N, int64_t n, double *pos, const double *q, int method, double dtau,
                        const double *phi, const double *chi, const double *Bi, int nfield,
                        const double params[2])
{
    const double dx = 1.0 / (double)N;
    #pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        double *pp = pos + 3 * p;
        const double *qq = q + 3 * p;
        if (method == ORC_POS_NEWTON) {
            /* update_pos_Newton, gevolution.hpp:900-903 */
            for (int l = 0; l < 3; l++) pp[l] += dtau * qq[l] / params[0];
        } else {
            update_pos_one(N, dx, pp, qq, dtau, phi, chi, Bi, nfield, params);
        }
        for (int l = 0; l < 3; l++) pp[l] = wrap_pos(pp[l], dx, N);
    }
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

double orc_field_sum(int64_t count, const double *f)
{
    double s = 0.;
    for (int64_t i = 0; i < count; i++) s += f[i];
    return s;
}

/* ======================================================================== */
/* box and cycle                                                             */
/* ======================================================================== */

int orc_box_alloc(orc_box *b, int N, int64_t n)
{
    const size_t n3 = (size_t)N * N * N, nk = (size_t)N * N * (N / 2 + 1) * 2;
    memset(b, 0, sizeof(*b));
    b->n = n;
    b->pos = (double *)calloc((size_t)n * 3 + 1, sizeof(double));
    b->q = (double *)calloc((size_t)n * 3 + 1, sizeof(double));
    b->id = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    b->phi = (double *)calloc(n3, sizeof(double));
    b->chi = (double *)calloc(n3, sizeof(double));
    b->source = (double *)calloc(n3, sizeof(double));
    b->Sij = (double *)calloc(6 * n3, sizeof(double));
    b->Bi = (double *)calloc(3 * n3, sizeof(double));
    b->scalarFT = (double *)calloc(nk, sizeof(double));
    b->SijFT = (double *)calloc(6 * nk, sizeof(double));
    b->BiFT = (double *)calloc(3 * nk, sizeof(double));
    b->cell = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    b->count = (uint32_t *)calloc(n3, sizeof(uint32_t));
    b->start = (uint64_t *)calloc(n3 + 1, sizeof(uint64_t));
    b->perm = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    if (!b->pos || !b->q || !b->id || !b->phi || !b->chi || !b->source || !b->Sij || !b->Bi ||
        !b->scalarFT || !b->SijFT || !b->BiFT || !b->cell || !b->count || !b->start || !b->perm)
        return -1;
    return 0;
}

void orc_box_free(orc_box *b)
{
    free(b->pos); free(b->q); free(b->id); free(b->phi); free(b->chi); free(b->source);
    free(b->Sij); free(b->Bi); free(b->scalarFT); free(b->SijFT); free(b->BiFT);
    free(b->cell); free(b->count); free(b->start); free(b->perm);
    memset(b, 0, sizeof(*b));
}

/* The field stage of one cycle of main.cpp:372-879 (projections, both
   prepareFTsource stencils, the FFTs and k-space solves: main.cpp:372-598)
   for one CDM species, no ncdm, no radiation.  Leaves phi, chi, Bi, BiFT and
   st->T00hom, st->phi_k0 as the reference does at the point where it writes
   its outputs (main.cpp:605-680).  The particle filing must be current. */
void orc_cycle_fields(const orc_sim *sim, orc_box *b, orc_state *st)
{
    const int N = sim->N;
    const size_t n3 = (size_t)N * N * N;
    const double dx = 1.0 / (double)N;
    const double fourpiG = sim->fourpiG;
    const orc_cosmo *cs = &sim->cosmo;
    const double a = st->a;
    double t0 = now_s(), tf;

    /* projections, main.cpp:378-450 */
    memset(b->source, 0, sizeof(double) * n3);
    if (sim->gr_flag > 0)
        orc_projection_T00_project(N, b->n, b->pos, b->q, b->start, b->perm, b->pmass, b->source, a, b->phi, 1.);
    else
        orc_scalar_projection_cic_project(N, b->n, b->pos, b->start, b->perm, b->pmass, b->source, 1.);
    if (sim->vector_flag == 1) {
        memset(b->Bi, 0, sizeof(double) * 3 * n3);
        orc_projection_T0i_project(N, b->n, b->pos, b->q, b->start, b->perm, b->pmass, b->Bi, b->phi, 1.);
    }
    memset(b->Sij, 0, sizeof(double) * 6 * n3);
    orc_projection_Tij_project(N, b->n, b->pos, b->q, b->start, b->perm, b->pmass, b->Sij, a, b->phi, 1.);
    orc_timer_add(ORC_T_PROJECTION, now_s() - t0);
    t0 = now_s();
    double t_fft = 0.;
    int n_fft = 0;
#define ORC_FFT(call)      \
    do {                   \
        tf = now_s();      \
        call;              \
        t_fft += now_s() - tf; \
        n_fft++;           \
    } while (0)

    if (sim->gr_flag > 0) {
        /* main.cpp:457-491 */
        st->T00hom = orc_field_sum((int64_t)n3, b->source) / (double)n3;
        if (st->dtau_old > 0.) {
            double H = orc_hconf(a, fourpiG, cs);
            orc_prepare_ft_source_scalar(N, b->phi, b->chi, b->source, cs->Omega_cdm + cs->Omega_b + 0.0,
                                         b->source, 3. * H * dx * dx / st->dtau_old,
                                         fourpiG * dx * dx / a, 3. * H * H * dx * dx);
            ORC_FFT(orc_fft_r2c(N, b->source, b->scalarFT));
            orc_solve_modified_poisson_ft(N, b->scalarFT, b->scalarFT, 1. / (dx * dx),
                                          3. * orc_hconf(a, fourpiG, cs) / st->dtau_old);
            ORC_FFT(orc_fft_c2r(N, b->scalarFT, b->phi));
        }
    } else {
        /* main.cpp:493-516 */
        ORC_FFT(orc_fft_r2c(N, b->source, b->scalarFT));
        orc_solve_modified_poisson_ft(N, b->scalarFT, b->scalarFT, fourpiG / a, 0.);
        ORC_FFT(orc_fft_c2r(N, b->scalarFT, b->phi));
    }
    st->phi_k0 = b->scalarFT[0]; /* main.cpp:533 */

    /* main.cpp:539-598 */
    orc_prepare_ft_source_tensor(N, b->phi, b->Sij, b->Sij, 2. * fourpiG * dx * dx / a);
    for (int c = 0; c < 6; c++)
        ORC_FFT(orc_fft_r2c(N, b->Sij + c * n3, b->SijFT + (size_t)c * N * N * (N / 2 + 1) * 2));
    orc_project_ft_scalar(N, b->SijFT, b->scalarFT, 0);
    ORC_FFT(orc_fft_c2r(N, b->scalarFT, b->chi));
    const size_t nk2 = (size_t)N * N * (N / 2 + 1) * 2;
    if (sim->vector_flag == 1) {
        for (int c = 0; c < 3; c++) ORC_FFT(orc_fft_r2c(N, b->Bi + c * n3, b->BiFT + c * nk2));
        orc_project_ft_vector(N, b->BiFT, b->BiFT, fourpiG * dx * dx, 0.);
    } else {
        orc_evolve_ft_vector(N, b->SijFT, b->BiFT, a * a * st->dtau_old);
    }
    if (sim->gr_flag > 0)
        for (int c = 0; c < 3; c++) ORC_FFT(orc_fft_c2r(N, b->BiFT + c * nk2, b->Bi + c * n3));
#undef ORC_FFT
    orc_timer_add(ORC_T_GRAVITY, now_s() - t0);
    g_timer[ORC_T_FFT] += t_fft;
    g_count[ORC_T_FFT] += n_fft;
}

/* The particle stage of the cycle (main.cpp:766-822, 867-874): kick, half
   RK4 step, drift, half RK4 step, re-filing, and the time-step update. */
void orc_cycle_particles(const orc_sim *sim, orc_box *b, orc_state *st)
{
    const int N = sim->N;
    const double fourpiG = sim->fourpiG;
    const orc_cosmo *cs = &sim->cosmo;
    double a = st->a;
    double params[2];
    double t0 = now_s();

    params[0] = a;
    params[1] = a * a * N;
    if (sim->gr_flag > 0)
        st->maxvel = orc_update_vel(N, b->n, b->pos, b->q, ORC_Q_GR, (st->dtau + st->dtau_old) / 2.,
                                    b->phi, b->chi, b->Bi, (1. / a < sim->z_relax + 1. ? 3 : 2), params);
    else
        st->maxvel = orc_update_vel(N, b->n, b->pos, b->q, ORC_Q_NEWTON, (st->dtau + st->dtau_old) / 2.,
                                    b->phi, b->chi, b->Bi, 1, params);
    orc_timer_add(ORC_T_UPDATE_Q, now_s() - t0);
    t0 = now_s();
    orc_rungekutta4bg(&a, fourpiG, cs, 0.5 * st->dtau);
    params[0] = a;
    params[1] = a * a * N;
    if (sim->gr_flag > 0)
        orc_move_particles(N, b->n, b->pos, b->q, ORC_POS_GR, st->dtau, b->phi, b->chi, b->Bi,
                           (1. / a < sim->z_relax + 1. ? 3 : 0), params);
    else
        orc_move_particles(N, b->n, b->pos, b->q, ORC_POS_NEWTON, st->dtau, NULL, NULL, NULL, 0, params);
    orc_rungekutta4bg(&a, fourpiG, cs, 0.5 * st->dtau);
    if (sim->gr_flag > 0) st->maxvel /= sqrt(st->maxvel * st->maxvel + 1.0);
    orc_file_particles(N, b->n, b->pos, b->id, b->cell, b->count, b->start, b->perm);
    orc_timer_add(ORC_T_MOVE, now_s() - t0);

    st->tau += st->dtau;
    st->dtau_old = st->dtau;
    st->a = a;
    st->dtau = orc_next_dtau(sim, a);
    st->cycle++;
}

/* One cycle of main.cpp:372-879 for one CDM species, no outputs, no ncdm,
   no radiation.  st->a, tau, dtau, dtau_old, cycle are advanced in place.
   The particle filing must be current on entry (orc_file_particles) and is
   current again on exit. */
void orc_cycle(const orc_sim *sim, orc_box *b, orc_state *st)
{
    orc_cycle_fields(sim, b, st);
    orc_cycle_particles(sim, b, st);
}
