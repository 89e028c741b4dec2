/*
 * gev_oracle_ic.c -- CPU restatement of gevolution 1.2's basic initial
 * condition generator (ic_basic.hpp) and of the counter-based PRNG it draws
 * from (prng_engine.hpp).  TEST INFRASTRUCTURE ONLY (see gev_oracle.h).
 *
 * Scope: the transfer-function path of generateIC_basic (ic_basic.hpp:1626-2259)
 * for one CDM species with baryon treatment "ignore" (baryons folded into the
 * CDM mass), no non-CDM species, no CLASS; GR or Newtonian gauge.
 *
 * Parity notes (each also stated at the function):
 *   - PRNG: bit-exact; pinned by golden vectors written by the reference's own
 *     prng_engine.hpp (tests/golden/sitmo_prng.json).
 *   - Spline: natural cubic spline (GSL's gsl_interp_cspline, GSL absent here);
 *     pinned against scipy's CubicSpline(bc_type="natural") and exact
 *     reproduction of linear data.  Bitwise agreement with GSL unpinned.
 *   - CIC kernel: pinned by closed forms (sum zero; the sc0 template reproduces
 *     the standard kernel exactly).
 *   - Realisation: stream positions pinned by an independent closed form over
 *     the golden-pinned PRNG; amplitudes pinned statistically.  The complex
 *     division by the kernel uses the textbook formula (LATfield2's Imag is
 *     absent: its rounding unpinned).  log(r1) on a float resolves to the C
 *     library's logf in the reference; `cr_log` = 1 uses the correctly rounded
 *     (float)log((double)r1) instead (glibc's logf differs from it by one ulp
 *     on about 0.7% of all 2^32 PRNG outputs).
 */
#include "gev_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ======================================================================== */
/* PRNG                                                                      */
/* ======================================================================== */

/* Threefish-256 with 20 rounds and zero tweak (the Skein paper's block
   cipher), keyed by (seed, 0, 0, 0) and applied to the block counter
   (counter, 0, 0, 0): the block function of sitmo::prng_engine
   (prng_engine.hpp:164-176 seeding, :305-346 encrypt_counter). */
void orc_prng_block(uint32_t seed, uint64_t counter, uint64_t out[4])
{
    static const int R[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
                                {25, 33}, {46, 12}, {58, 22}, {32, 32}};
    uint64_t k[5] = {seed, 0, 0, 0, 0};
    k[4] = 0x1BD11BDAA9FC1A22ull ^ k[0] ^ k[1] ^ k[2] ^ k[3];
    uint64_t x[4] = {counter, 0, 0, 0};
    for (int d = 0; d < 20; d++) {
        if (d % 4 == 0) {
            const int s = d / 4;
            for (int i = 0; i < 4; i++) x[i] += k[(s + i) % 5];
            x[3] += (uint64_t)s;
        }
        /* word pairs alternate (0,1)(2,3) and (0,3)(2,1): the 4-word permutation */
        const int a1 = (d & 1) ? 3 : 1, b1 = (d & 1) ? 1 : 3;
        const int r0 = R[d % 8][0], r1 = R[d % 8][1];
        x[0] += x[a1];
        x[a1] = ((x[a1] << r0) | (x[a1] >> (64 - r0))) ^ x[0];
        x[2] += x[b1];
        x[b1] = ((x[b1] << r1) | (x[b1] >> (64 - r1))) ^ x[2];
    }
    for (int i = 0; i < 4; i++) out[i] = x[i] + k[(5 + i) % 5];
    out[3] += 5;
}

/* Output number `pos` (0-based) of an engine seeded with `seed`.  The engine
   hands out the 8 32-bit halves of each block low half first, and its
   discard(z) is equivalent to z draws (prng_engine.hpp:194-238), so the stream
   is this pure function of the position. */
uint32_t orc_prng_at(uint32_t seed, uint64_t pos)
{
    uint64_t o[4];
    orc_prng_block(seed, pos >> 3, o);
    const uint64_t w = o[(pos & 7) >> 1];
    return (pos & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
}

/* ======================================================================== */
/* natural cubic spline (GSL gsl_interp_cspline semantics)                   */
/* ======================================================================== */

/* c[i] = half the second derivative at knot i; c[0] = c[n-1] = 0.  The
   interior system h_{i-1} c_{i-1} + 2 (h_{i-1} + h_i) c_i + h_i c_{i+1} =
   3 (dy_i/h_i - dy_{i-1}/h_{i-1}) is solved by LDL^T elimination. */
int orc_spline_init(int n, const double *x, const double *y, double *c)
{
    if (n < 2) return -1;
    for (int i = 0; i < n; i++) c[i] = 0.;
    const int m = n - 2;
    if (m <= 0) return 0;
    double *dg = (double *)malloc(sizeof(double) * (size_t)m * 3);
    if (!dg) return -1;
    double *od = dg + m, *g = dg + 2 * m;
    for (int i = 0; i < m; i++) {
        const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
        od[i] = h1;
        dg[i] = 2.0 * (h1 + h0);
        g[i] = 3.0 * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
    }
    /* forward elimination: dg[i] becomes the pivot, g[i] the reduced rhs */
    for (int i = 1; i < m; i++) {
        const double l = od[i - 1] / dg[i - 1];
        dg[i] -= l * od[i - 1];
        g[i] -= l * g[i - 1];
    }
    c[m] = g[m - 1] / dg[m - 1];
    for (int i = m - 2; i >= 0; i--) c[i + 1] = (g[i] - od[i] * c[i + 2]) / dg[i];
    free(dg);
    return 0;
}

/* interval i with x[i] <= xq < x[i+1] by bisection (xq = x[n-1] -> n-2) */
static int spline_index(int n, const double *x, double xq)
{
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        const int mid = (lo + hi) / 2;
        if (x[mid] > xq) hi = mid; else lo = mid;
    }
    return lo;
}

/* y(xq) = y_i + t (b + t (c_i + t d)), t = xq - x_i.  Outside [x0, x_{n-1}]
   returns NaN (GSL raises a domain error there). */
double orc_spline_eval(int n, const double *x, const double *y, const double *c, double xq)
{
    if (!(xq >= x[0] && xq <= x[n - 1])) return NAN;
    const int i = spline_index(n, x, xq);
    const double h = x[i + 1] - x[i];
    const double dy = y[i + 1] - y[i];
    const double t = xq - x[i];
    const double b = dy / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
    const double d = (c[i + 1] - c[i]) / (3.0 * h);
    return y[i] + t * (b + t * (c[i] + t * d));
}

/* ======================================================================== */
/* CIC kernel (ic_basic.hpp:737-1052)                                        */
/* ======================================================================== */

static inline int64_t site(int N, int x, int y, int z)
{
    return (int64_t)x + (int64_t)N * ((int64_t)y + (int64_t)N * z);
}

/* generateCICKernel.  numpcl = 0 or pcldata = NULL: the standard kernel
   (ic_basic.hpp:747-766).  Otherwise the kernel of the tiled template
   (:768-1051): for every template particle and every octant whose CIC cell it
   lies in, the three directional stencils, accumulated in the reference's
   order with its float/double operand types. */
void orc_generate_cic_kernel(int N, double *ker, int64_t numpcl, const float *pcldata, int numtile)
{
    const long ls = N;
    double renorm = (double)(ls * ls);
    memset(ker, 0, sizeof(double) * (size_t)N * N * N);
    if (numpcl == 0 || pcldata == NULL) {
        ker[site(N, 0, 0, 0)] = 6. * renorm;
        ker[site(N, 1 % N, 0, 0)] = -renorm;
        ker[site(N, N - 1, 0, 0)] = -renorm;
        ker[site(N, 0, 1 % N, 0)] = -renorm;
        ker[site(N, 0, 0, 1 % N)] = -renorm;
        ker[site(N, 0, N - 1, 0)] = -renorm;
        ker[site(N, 0, 0, N - 1)] = -renorm;
        return;
    }
    renorm /= (double)(numpcl * (long)numtile * (long)numtile * (long)numtile) / (double)(ls * ls * ls);
    /* octant order of the reference: signs (sx, sy, sz) */
    static const int OCT[8][3] = {{1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1},
                                  {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}, {1, -1, -1}};
    for (int64_t i = 0; i < numpcl; i++) {
        const float *p = pcldata + 3 * i;
        for (int oct = 0; oct < 8; oct++) {
            const int sx = OCT[oct][0], sy = OCT[oct][1], sz = OCT[oct][2];
            /* distance to the octant's cell corner: a float expression for the
               lower corner, a double one (1. - p) for the upper */
            double dxv, dyv, dzv;
            if (sx > 0) dxv = ls * p[0] / numtile; else dxv = ls * (1. - p[0]) / numtile;
            if (sy > 0) dyv = ls * p[1] / numtile; else dyv = ls * (1. - p[1]) / numtile;
            if (sz > 0) dzv = ls * p[2] / numtile; else dzv = ls * (1. - p[2]) / numtile;
            if (dxv >= 1. || dyv >= 1. || dzv >= 1.) continue;
            const float wx = (float)(1. - dxv), wy = (float)(1. - dyv), wz = (float)(1. - dzv);
            const int xp = (N + sx) % N, xm = (N - sx) % N;
            const int yp = (N + sy) % N, ym = (N - sy) % N;
            const int zp = (N + sz) % N, zm = (N - sz) % N;
            float ww, q1;
            /* 0-direction (:915-958) */
            ww = (float)(wy * wz * renorm);
            q1 = (wx > 0.9) ? 2.f : 1.f;
            ker[site(N, 0, 0, 0)] += ww * wy * wz * q1;
            ker[site(N, xp, 0, 0)] -= ww * wy * wz;
            if (wx > 0.9) ker[site(N, xm, 0, 0)] -= ww * wy * wz;
            ker[site(N, 0, 0, zp)] += ww * wy * (1. - wz) * q1;
            ker[site(N, xp, 0, zp)] -= ww * wy * (1. - wz);
            if (wx > 0.9) ker[site(N, xm, 0, zp)] -= ww * wy * (1. - wz);
            ker[site(N, 0, yp, 0)] += ww * (1. - wy) * wz * q1;
            ker[site(N, xp, yp, 0)] -= ww * (1. - wy) * wz;
            if (wx > 0.9) ker[site(N, xm, yp, 0)] -= ww * (1. - wy) * wz;
            ker[site(N, 0, yp, zp)] += ww * (1. - wy) * (1. - wz) * q1;
            ker[site(N, xp, yp, zp)] -= ww * (1. - wy) * (1. - wz);
            if (wx > 0.9) ker[site(N, xm, yp, zp)] -= ww * (1. - wy) * (1. - wz);
            /* 1-direction (:960-1003) */
            ww = (float)(wx * wz * renorm);
            q1 = (wy > 0.9) ? 2.f : 1.f;
            ker[site(N, 0, 0, 0)] += ww * wx * wz * q1;
            ker[site(N, xp, 0, 0)] += ww * (1. - wx) * wz * q1;
            ker[site(N, 0, 0, zp)] += ww * wx * (1. - wz) * q1;
            ker[site(N, xp, 0, zp)] += ww * (1. - wx) * (1. - wz) * q1;
            ker[site(N, 0, yp, 0)] -= ww * wx * wz;
            ker[site(N, xp, yp, 0)] -= ww * (1. - wx) * wz;
            ker[site(N, 0, yp, zp)] -= ww * wx * (1. - wz);
            ker[site(N, xp, yp, zp)] -= ww * (1. - wx) * (1. - wz);
            if (wy > 0.9) {
                ker[site(N, 0, ym, 0)] -= ww * wx * wz;
                ker[site(N, xp, ym, 0)] -= ww * (1. - wx) * wz;
                ker[site(N, 0, ym, zp)] -= ww * wx * (1. - wz);
                ker[site(N, xp, ym, zp)] -= ww * (1. - wx) * (1. - wz);
            }
            /* 2-direction (:1005-1048) */
            ww = (float)(wx * wy * renorm);
            q1 = (wz > 0.9) ? 2.f : 1.f;
            ker[site(N, 0, 0, 0)] += ww * wx * wy * q1;
            ker[site(N, xp, 0, 0)] += ww * (1. - wx) * wy * q1;
            ker[site(N, 0, yp, 0)] += ww * wx * (1. - wy) * q1;
            ker[site(N, xp, yp, 0)] += ww * (1. - wx) * (1. - wy) * q1;
            ker[site(N, 0, 0, zp)] -= ww * wx * wy;
            ker[site(N, xp, 0, zp)] -= ww * (1. - wx) * wy;
            ker[site(N, 0, yp, zp)] -= ww * wx * (1. - wy);
            ker[site(N, xp, yp, zp)] -= ww * (1. - wx) * (1. - wy);
            if (wz > 0.9) {
                ker[site(N, 0, 0, zm)] -= ww * wx * wy;
                ker[site(N, xp, 0, zm)] -= ww * (1. - wx) * wy;
                ker[site(N, 0, yp, zm)] -= ww * wx * (1. - wy);
                ker[site(N, xp, yp, zm)] -= ww * (1. - wx) * (1. - wy);
            }
        }
    }
}

/* ======================================================================== */
/* Gaussian realisation (ic_basic.hpp:1089-1379)                             */
/* ======================================================================== */

typedef struct {
    int n;
    const double *x, *y, *c;
} spl;

typedef struct {
    uint32_t seed;
    uint64_t pos;
} stream;

static inline uint32_t draw(stream *s) { return orc_prng_at(s->seed, s->pos++); }

/* one mode of the realisation (ic_basic.hpp:1149-1166): draws r1 (redrawn
   while zero) and r2, then
   F = e^{+-2 pi i r2} (1 + 7.5 coeff / k2) / K(k) sqrt(-2 log r1) P(sqrt k2) s,
   with K the transformed CIC kernel held in `cell` (ignorekernel: K = 1 and
   no (1 + 7.5 coeff / k2) factor). */
static void mode(double *cell, stream *st, int *used, float k2, float s, double coeff,
                 const spl *sp, int conj, int ignorekernel, int cr_log)
{
    float r1, r2;
    k2 *= 4. * M_PI * M_PI;
    do {
        r1 = (float)draw(st) / (float)0xFFFFFFFFu;
        (*used)++;
    } while (r1 == 0.);
    r2 = (float)draw(st) / (float)0xFFFFFFFFu;
    (*used)++;
    double re = cos(2. * M_PI * r2), im = sin(2. * M_PI * r2);
    if (conj) im = -im;
    if (!ignorekernel) {
        const double f = 1. + 7.5 * coeff / k2;
        re *= f;
        im *= f;
        const double kr = cell[0], ki = cell[1];
        const double den = kr * kr + ki * ki;
        const double nr = (re * kr + im * ki) / den, ni = (im * kr - re * ki) / den;
        re = nr;
        im = ni;
    }
    const float lg = cr_log ? (float)log((double)r1) : logf(r1);
    const double amp = sqrt(-2. * lg);
    const double pk = orc_spline_eval(sp->n, sp->x, sp->y, sp->c, (double)sqrtf(k2));
    re *= amp; im *= amp;
    re *= pk;  im *= pk;
    re *= s;   im *= s;
    cell[0] = re;
    cell[1] = im;
}

/* generateDisplacementField on one process (kymin = kzmin = 0,
   kymax = kzmax = N-1).  potFT: the transformed kernel on entry (unless
   ignorekernel), the realisation on exit; layout kx + (N/2+1)(ky + N kz). */
void orc_generate_displacement_field(int N, double *potFT, double coeff, int nsp,
                                     const double *spx, const double *spy, const double *spc,
                                     uint32_t seed, int ksphere, int deconvolve_f,
                                     int ignorekernel, int cr_log)
{
    const int H = N / 2, kmax = H - 1;
    const uint64_t S = 65536;  /* HUGE_SKIP, ic_basic.hpp:31-33 */
    const spl sp = {nsp, spx, spy, spc};
    float *sinc = (float *)malloc(sizeof(float) * (size_t)N);
    sinc[0] = 1.;
    for (int i = 1; i < N; i++)
        sinc[i] = deconvolve_f == 1 ? (float)(sin(M_PI * (float)i / (float)N) * (float)N / (M_PI * (float)i)) : 1.f;
#define CELL(kx, ky, kz) (potFT + 2 * ((int64_t)(kx) + (int64_t)(H + 1) * ((int64_t)(ky) + (int64_t)N * (kz))))
#define ZERO(p) ((p)[0] = 0., (p)[1] = 0.)
    stream st;
    int i, j, kx, ky, kz;
    /* quadrant ky <= N/2, kz <= N/2 (:1127-1174) */
    st.seed = seed; st.pos = 0;
    ZERO(CELL(0, 0, 0));
    kx = 1;
    for (kz = 0; kz < H + 1; kz++) {
        for (ky = 0, j = 0; ky < H + 1; ky++, j++) {
            for (i = 0; kx < H + 1; kx++) {
                float k2 = (float)(kx * kx) + (float)(ky * ky) + (float)(kz * kz);
                if (kx >= kmax || ky >= kmax || kz >= kmax || (k2 >= kmax * kmax && ksphere > 0))
                    ZERO(CELL(kx, ky, kz));
                else
                    mode(CELL(kx, ky, kz), &st, &i, k2, sinc[kx] * sinc[ky] * sinc[kz], coeff, &sp, 0,
                         ignorekernel, cr_log);
            }
            st.pos += S - (uint64_t)i;
            kx = 0;
        }
        st.pos += S * (S - (uint64_t)j);
    }
    /* quadrant ky > N/2, kz <= N/2 (:1176-1216) */
    st.seed = seed; st.pos = ((S + 0) * S + (uint64_t)(N - (N - 1))) * S;
    for (kz = 0; kz < H + 1; kz++) {
        for (ky = N - 1, j = 0; ky >= H + 1; ky--, j++) {
            for (kx = 0, i = 0; kx < H + 1; kx++) {
                float k2 = (float)(kx * kx) + (float)((N - ky) * (N - ky)) + (float)(kz * kz);
                if (kx >= kmax || (N - ky) >= kmax || kz >= kmax || (k2 >= kmax * kmax && ksphere > 0))
                    ZERO(CELL(kx, ky, kz));
                else
                    mode(CELL(kx, ky, kz), &st, &i, k2, sinc[kx] * sinc[N - ky] * sinc[kz], coeff, &sp, 0,
                         ignorekernel, cr_log);
            }
            st.pos += S - (uint64_t)i;
        }
        st.pos += S * (S - (uint64_t)j);
    }
    /* quadrant ky <= N/2, kz > N/2: kx >= 1 (:1218-1258), then kx = 0 from the
       first quadrant's streams with the conjugate phase (:1259-1291) */
    st.seed = seed; st.pos = ((S + S + (uint64_t)(N - (N - 1))) * S + 0) * S;
    for (kz = N - 1; kz >= H + 1; kz--) {
        for (ky = 0, j = 0; ky < H + 1; ky++, j++) {
            for (kx = 1, i = 0; kx < H + 1; kx++) {
                float k2 = (float)(kx * kx) + (float)(ky * ky) + (float)((N - kz) * (N - kz));
                if (kx >= kmax || ky >= kmax || (N - kz) >= kmax || (k2 >= kmax * kmax && ksphere > 0))
                    ZERO(CELL(kx, ky, kz));
                else
                    mode(CELL(kx, ky, kz), &st, &i, k2, sinc[kx] * sinc[ky] * sinc[N - kz], coeff, &sp, 0,
                         ignorekernel, cr_log);
            }
            st.pos += S - (uint64_t)i;
        }
        st.pos += S * (S - (uint64_t)j);
    }
    st.seed = seed; st.pos = ((uint64_t)(N - (N - 1)) * S + 0) * S;
    for (kz = N - 1; kz >= H + 1; kz--) {
        for (ky = 0, j = 0; ky < H + 1; ky++, j++) {
            float k2 = (float)(ky * ky) + (float)((N - kz) * (N - kz));
            i = 0;
            if (ky >= kmax || (N - kz) >= kmax || (k2 >= kmax * kmax && ksphere > 0))
                ZERO(CELL(0, ky, kz));
            else
                mode(CELL(0, ky, kz), &st, &i, k2, sinc[ky] * sinc[N - kz], coeff, &sp, 1, ignorekernel, cr_log);
            st.pos += S - (uint64_t)i;
        }
        st.pos += S * (S - (uint64_t)j);
    }
    /* quadrant ky > N/2, kz > N/2: kx >= 1 (:1293-1333), then kx = 0 from the
       third quadrant's streams with the conjugate phase (:1334-1366) */
    st.seed = seed; st.pos = ((S + S + S + (uint64_t)(N - (N - 1))) * S + (uint64_t)(N - (N - 1))) * S;
    for (kz = N - 1; kz >= H + 1; kz--) {
        for (ky = N - 1, j = 0; ky >= H + 1; ky--, j++) {
            for (kx = 1, i = 0; kx < H + 1; kx++) {
                float k2 = (float)(kx * kx) + (float)((N - ky) * (N - ky)) + (float)((N - kz) * (N - kz));
                if (kx >= kmax || (N - ky) >= kmax || (N - kz) >= kmax || (k2 >= kmax * kmax && ksphere > 0))
                    ZERO(CELL(kx, ky, kz));
                else
                    mode(CELL(kx, ky, kz), &st, &i, k2, sinc[kx] * sinc[N - ky] * sinc[N - kz], coeff, &sp, 0,
                         ignorekernel, cr_log);
            }
            st.pos += S - (uint64_t)i;
        }
        st.pos += S * (S - (uint64_t)j);
    }
    st.seed = seed; st.pos = ((S + S + (uint64_t)(N - (N - 1))) * S + (uint64_t)(N - (N - 1))) * S;
    for (kz = N - 1; kz >= H + 1; kz--) {
        for (ky = N - 1, j = 0; ky >= H + 1; ky--, j++) {
            float k2 = (float)((N - ky) * (N - ky)) + (float)((N - kz) * (N - kz));
            i = 0;
            if ((N - ky) >= kmax || (N - kz) >= kmax || (k2 >= kmax * kmax && ksphere > 0))
                ZERO(CELL(0, ky, kz));
            else
                mode(CELL(0, ky, kz), &st, &i, k2, sinc[N - ky] * sinc[N - kz], coeff, &sp, 1, ignorekernel,
                     cr_log);
            st.pos += S - (uint64_t)i;
        }
        st.pos += S * (S - (uint64_t)j);
    }
#undef CELL
#undef ZERO
    free(sinc);
}

/* ======================================================================== */
/* particles (ic_basic.hpp:1400-1434, :60-173)                               */
/* ======================================================================== */

/* initializeParticlePositions: particle ID = i + numpart (xt + numtile (yt +
   numtile zt)); written at index ID, pos = (tile + template) / numtile. */
void orc_initialize_particle_positions(int64_t numpart, const float *partdata, int numtile,
                                       double *pos, int64_t *id)
{
    for (long zt = 0; zt < numtile; zt++)
        for (long yt = 0; yt < numtile; yt++)
            for (long xt = 0; xt < numtile; xt++)
                for (int64_t i = 0; i < numpart; i++) {
                    const int64_t ID = i + numpart * (xt + (long)numtile * (yt + (long)numtile * zt));
                    pos[3 * ID + 0] = ((double)xt + partdata[3 * i + 0]) / (double)numtile;
                    pos[3 * ID + 1] = ((double)yt + partdata[3 * i + 1]) / (double)numtile;
                    pos[3 * ID + 2] = ((double)zt + partdata[3 * i + 2]) / (double)numtile;
                    id[ID] = ID;
                }
}

static inline double fat(const double *f, int N, int x, int y, int z)
{
    x %= N; y %= N; z %= N;
    return f[site(N, x, y, z)];
}

/* CIC gradient of a scalar field at the particle (the 12-term sum shared by
   displace_pcls_ic_basic and initialize_q_ic_basic, ic_basic.hpp:70-95,
   :128-153), divided by the lattice resolution. */
static void ic_grad(int N, const double *f, const int c[3], const double rd[3], double g[3])
{
    const int x = c[0], y = c[1], z = c[2];
    const double f000 = fat(f, N, x, y, z), f100 = fat(f, N, x + 1, y, z);
    const double f010 = fat(f, N, x, y + 1, z), f001 = fat(f, N, x, y, z + 1);
    const double f110 = fat(f, N, x + 1, y + 1, z), f101 = fat(f, N, x + 1, y, z + 1);
    const double f011 = fat(f, N, x, y + 1, z + 1), f111 = fat(f, N, x + 1, y + 1, z + 1);
    g[0] = (1. - rd[1]) * (1. - rd[2]) * (f100 - f000);
    g[1] = (1. - rd[0]) * (1. - rd[2]) * (f010 - f000);
    g[2] = (1. - rd[0]) * (1. - rd[1]) * (f001 - f000);
    g[0] += rd[1] * (1. - rd[2]) * (f110 - f010);
    g[1] += rd[0] * (1. - rd[2]) * (f110 - f100);
    g[2] += rd[0] * (1. - rd[1]) * (f101 - f100);
    g[0] += (1. - rd[1]) * rd[2] * (f101 - f001);
    g[1] += (1. - rd[0]) * rd[2] * (f011 - f001);
    g[2] += (1. - rd[0]) * rd[1] * (f011 - f010);
    g[0] += rd[1] * rd[2] * (f111 - f011);
    g[1] += rd[0] * rd[2] * (f111 - f101);
    g[2] += rd[0] * rd[1] * (f111 - f110);
    const double lr = 1.0 / (double)N;
    g[0] /= lr;
    g[1] /= lr;
    g[2] /= lr;
}

static inline void ic_locate(const double *p, double dx, int c[3], double rd[3])
{
    for (int i = 0; i < 3; i++) {
        const double t = p[i] / dx;
        const double ip = floor(t);
        c[i] = (int)ip;
        rd[i] = t - ip;
    }
}

static inline double ic_wrap(double p, double dx, int N)
{
    if (p >= 1.0) p -= 1.0;
    if (p < 0.0) p += 1.0;
    if (!(p >= 0.0 && p < 1.0)) p = p - floor(p);
    if (!(p >= 0.0) || (int)floor(p / dx) >= N) p = 0.0;
    return p;
}

/* moveParticles(displace_pcls_ic_basic, coeff, &field, 1, NULL, &max, MAX, 1)
   (ic_basic.hpp:1995): pos += coeff grad(field); returns the maximum of
   coeff |grad(field)|; positions wrapped into [0,1). */
double orc_displace_pcls_ic(int N, int64_t n, double *pos, double coeff, const double *field)
{
    const double dx = 1.0 / (double)N;
    double mx = 0.;
    #pragma omp parallel for schedule(static) reduction(max : mx)
    for (int64_t p = 0; p < n; p++) {
        double *pp = pos + 3 * p;
        int c[3];
        double rd[3], g[3];
        ic_locate(pp, dx, c, rd);
        ic_grad(N, field, c, rd, g);
        const double out = coeff * sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        if (out > mx) mx = out;
        for (int i = 0; i < 3; i++) pp[i] += coeff * g[i];
        for (int i = 0; i < 3; i++) pp[i] = ic_wrap(pp[i], dx, N);
    }
    return mx;
}

/* updateVel(initialize_q_ic_basic, coeff, &field, 1) (ic_basic.hpp:2065):
   q = -coeff grad(field); returns sqrt(max |q|^2). */
double orc_initialize_q_ic(int N, int64_t n, const double *pos, double *q, double coeff,
                           const double *field)
{
    const double dx = 1.0 / (double)N;
    double vmax = 0.;
    #pragma omp parallel for schedule(static) reduction(max : vmax)
    for (int64_t p = 0; p < n; p++) {
        int c[3];
        double rd[3], g[3], v2 = 0.;
        ic_locate(pos + 3 * p, dx, c, rd);
        ic_grad(N, field, c, rd, g);
        for (int i = 0; i < 3; i++) {
            q[3 * p + i] = -g[i] * coeff;
            v2 += q[3 * p + i] * q[3 * p + i];
        }
        if (v2 > vmax) vmax = v2;
    }
    return sqrt(vmax);
}

/* ======================================================================== */
/* generateIC_basic (transfer-function path, CDM only)                       */
/* ======================================================================== */

double orc_pk_primordial(double k, const orc_ic *ic)
{
    /* ic_basic.hpp:186-189; k_pivot in 1/Mpc */
    return ic->A_s * pow(k / ic->k_pivot, ic->n_s - 1.);
}

/* spline tables for the transfer-function path: phi (pkspline), CDM
   displacement (tk_d1) and CDM velocity potential (tk_t1) as in
   ic_basic.hpp:1721-1760 (gr_flag 1) / 1730-1735 (gr_flag 0) and
   :1952-1983 (baryon_flag 0).  k: k * boxsize as loaded. */
void orc_ic_tables(const orc_sim *sim, const orc_ic *ic, int n, const double *k, const double *d_tot,
                   const double *t_tot, const double *d_cdm, const double *t_cdm, double *y_phi,
                   double *y_disp, double *y_vel)
{
    const orc_cosmo *cs = &sim->cosmo;
    const double fpG = sim->fourpiG;
    const double a = 1. / (1. + ic->z_in);
    const double Ha = orc_hconf(a, fpG, cs);
    const double rescale =
        3. * Ha * Ha * Ha *
        (1. + 0.5 * Ha * Ha *
                  ((1. / orc_hconf(0.98 * a, fpG, cs) / orc_hconf(0.98 * a, fpG, cs)) -
                   (8. / orc_hconf(0.99 * a, fpG, cs) / orc_hconf(0.99 * a, fpG, cs)) +
                   (8. / orc_hconf(1.01 * a, fpG, cs) / orc_hconf(1.01 * a, fpG, cs)) -
                   (1. / orc_hconf(1.02 * a, fpG, cs) / orc_hconf(1.02 * a, fpG, cs))) /
                  0.12);
    for (int i = 0; i < n; i++)
        y_phi[i] = (1.5 * (Ha * Ha - orc_hconf(1., fpG, cs) * orc_hconf(1., fpG, cs) * a * a * cs->Omega_Lambda) *
                        d_tot[i] +
                    rescale * t_tot[i] / k[i] / k[i]) *
                   M_PI * sqrt(orc_pk_primordial(k[i] * cs->h / ic->boxsize, ic) / k[i]) / k[i];
    for (int i = 0; i < n; i++) {
        const double sq = sqrt(orc_pk_primordial(k[i] * cs->h / ic->boxsize, ic) / k[i]);
        if (sim->gr_flag > 0) {
            y_disp[i] = -3. * y_phi[i] / k[i] / k[i] - d_cdm[i] * M_PI * sq / k[i];
        } else {
            /* N-body gauge correction 3 Hconf theta_tot / k^2 (:1731-1733) */
            const double nb = -3. * Ha * M_PI * t_tot[i] * sq / k[i] / k[i] / k[i];
            y_disp[i] = nb - d_cdm[i] * M_PI * sq / k[i];
        }
        y_vel[i] = -a * t_cdm[i] * M_PI * sq / k[i];
    }
}

int orc_generate_ic_basic(const orc_sim *sim, const orc_ic *ic, int64_t numpcl, const float *pcldata,
                          int ntk, const double *k, const double *d_tot, const double *t_tot,
                          const double *d_cdm, const double *t_cdm, orc_box *b, double *maxvel,
                          double *max_disp)
{
    const int N = sim->N;
    const int64_t n3 = (int64_t)N * N * N;
    const double a = 1. / (1. + ic->z_in);
    const double fpG = sim->fourpiG;
    const int64_t np = numpcl * (int64_t)ic->numtile * ic->numtile * ic->numtile;
    if (b->n != np) return -1;
    double *tab = (double *)malloc(sizeof(double) * (size_t)ntk * 6);
    if (!tab) return -1;
    double *yp = tab, *yd = tab + ntk, *yv = tab + 2 * ntk;
    double *cp = tab + 3 * ntk, *cd = tab + 4 * ntk, *cv = tab + 5 * ntk;
    orc_ic_tables(sim, ic, ntk, k, d_tot, t_tot, d_cdm, t_cdm, yp, yd, yv);
    orc_spline_init(ntk, k, yp, cp);
    orc_spline_init(ntk, k, yd, cd);
    orc_spline_init(ntk, k, yv, cv);

    /* displacement field (:1672-1676, :1961, :1965-1966) */
    if (ic->correct_displacement)
        orc_generate_cic_kernel(N, b->source, numpcl, pcldata, ic->numtile);
    else
        orc_generate_cic_kernel(N, b->source, 0, NULL, 1);
    orc_fft_r2c(N, b->source, b->scalarFT);
    orc_generate_displacement_field(N, b->scalarFT, 0., ntk, k, yd, cd, ic->seed, ic->ksphere, 1, 0,
                                    ic->cr_log);
    orc_fft_c2r(N, b->scalarFT, b->chi);

    /* particles (:1967-1983) */
    b->pmass = (sim->cosmo.Omega_cdm + sim->cosmo.Omega_b) / (double)np;
    orc_initialize_particle_positions(numpcl, pcldata, ic->numtile, b->pos, b->id);
    const double md = orc_displace_pcls_ic(N, b->n, b->pos, 1., b->chi);
    if (max_disp) *max_disp = md;

    /* velocities (:2030-2055) */
    if (ic->correct_displacement) orc_generate_cic_kernel(N, b->source, 0, NULL, 1);
    orc_fft_r2c(N, b->source, b->scalarFT);
    orc_generate_displacement_field(N, b->scalarFT, 0., ntk, k, yv, cv, ic->seed, ic->ksphere, 0, 0,
                                    ic->cr_log);
    orc_fft_c2r(N, b->scalarFT, b->chi);
    double mv = orc_initialize_q_ic(N, b->n, b->pos, b->q, 1., b->chi) / a;

    /* phi (:2145-2152, :2175-2176) */
    orc_fft_r2c(N, b->source, b->scalarFT);
    orc_generate_displacement_field(N, b->scalarFT, 0., ntk, k, yp, cp, ic->seed, ic->ksphere, 0, 0,
                                    ic->cr_log);
    orc_fft_c2r(N, b->scalarFT, b->phi);

    /* filing for the projections */
    orc_file_particles(N, b->n, b->pos, b->id, b->cell, b->count, b->start, b->perm);

    /* B (:2231-2240) */
    memset(b->Bi, 0, sizeof(double) * 3 * (size_t)n3);
    orc_projection_T0i_project(N, b->n, b->pos, b->q, b->start, b->perm, b->pmass, b->Bi, b->phi, 1.);
    {
        const int64_t nk = (int64_t)N * N * (N / 2 + 1) * 2;
        for (int c = 0; c < 3; c++) orc_fft_r2c(N, b->Bi + c * n3, b->BiFT + c * nk);
        orc_project_ft_vector(N, b->BiFT, b->BiFT, fpG / (double)N / (double)N, 0.);
        for (int c = 0; c < 3; c++) orc_fft_c2r(N, b->BiFT + c * nk, b->Bi + c * n3);
        /* chi (:2241-2248) */
        memset(b->Sij, 0, sizeof(double) * 6 * (size_t)n3);
        orc_projection_Tij_project(N, b->n, b->pos, b->q, b->start, b->perm, b->pmass, b->Sij, a, b->phi, 1.);
        orc_prepare_ft_source_tensor(N, b->phi, b->Sij, b->Sij, 2. * fpG / a / (double)N / (double)N);
        for (int c = 0; c < 6; c++) orc_fft_r2c(N, b->Sij + c * n3, b->SijFT + c * nk);
        orc_project_ft_scalar(N, b->SijFT, b->scalarFT, 0);
        orc_fft_c2r(N, b->scalarFT, b->chi);
    }
    /* main.cpp:326-330 */
    if (sim->gr_flag > 0) mv /= sqrt(mv * mv + 1.0);
    if (maxvel) *maxvel = mv;
    free(tab);
    return 0;
}

