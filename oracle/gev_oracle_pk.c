/*
 * gev_oracle_pk.c -- CPU restatement of gevolution 1.2's binned power-spectrum
 * estimator (tools.hpp:53-212).  TEST INFRASTRUCTURE ONLY (see gev_oracle.h).
 *
 * Parity: the binning, weights, deconvolution and per-bin moments follow the
 * reference literally; the summation runs over the k grid in the [kz][ky][kx]
 * order of this layout (LATfield2's rKSite order on its transposed k lattice
 * is unpinned, which affects only the last bits of the sums).  Pinned by a
 * single-mode closed form and an independent numpy binning (tests).
 */
#include "gev_oracle.h"

#include <math.h>
#include <stdlib.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ======================================================================== */
/* power spectra (tools.hpp:53-346)                                          */
/* ======================================================================== */

/* extractCrossSpectrum (tools.hpp:53-212) on one process.  f1, f2: complex
   fields of ncomp components in the k layout; symmetric != 0 uses the
   6-component symmetric-tensor sum (off-diagonals doubled, tools.hpp:137-145);
   comp1, comp2 >= 0 select one component of each.  Outputs as the reference's
   root process leaves them (means and scatters per bin). */
void orc_extract_cross_spectrum(int N, const double *f1, const double *f2, int ncomp, int symmetric,
                                int numbins, int deconvolve, int ktype, int comp1, int comp2, double *kbin,
                                double *power, double *kscatter, double *pscatter, int *occupation)
{
    const int H = N / 2;
    const int64_t nk = (int64_t)N * N * (H + 1);
    double *typek2 = (double *)malloc(sizeof(double) * N);
    double *sinc = (double *)malloc(sizeof(double) * N);
    int i;
    if (ktype == 0) { /* KTYPE_GRID */
        for (i = 0; i < N; i++) {
            typek2[i] = 2. * (double)N * sin(M_PI * (double)i / (double)N);
            typek2[i] *= typek2[i];
        }
    } else {
        for (i = 0; i <= N / 2; i++) {
            typek2[i] = 2. * M_PI * (double)i;
            typek2[i] *= typek2[i];
        }
        for (; i < N; i++) {
            typek2[i] = 2. * M_PI * (double)(N - i);
            typek2[i] *= typek2[i];
        }
    }
    sinc[0] = 1.;
    for (i = 1; i <= N / 2; i++)
        sinc[i] = deconvolve ? sin(M_PI * (float)i / (float)N) * (float)N / (M_PI * (float)i) : 1.;
    for (; i < N; i++) sinc[i] = sinc[N - i];
    const double k2max = 3. * typek2[N / 2];
    for (i = 0; i < numbins; i++) {
        kbin[i] = power[i] = kscatter[i] = pscatter[i] = 0.;
        occupation[i] = 0;
    }
#define C1(c, idx) (f1 + 2 * ((int64_t)(c) * nk + (idx)))
#define C2(c, idx) (f2 + 2 * ((int64_t)(c) * nk + (idx)))
    for (int kz = 0; kz < N; kz++)
        for (int ky = 0; ky < N; ky++)
            for (int kx = 0; kx <= H; kx++) {
                if (kx == 0 && ky == 0 && kz == 0) continue;
                const int weight = (kx == 0 || (kx == N / 2 && N % 2 == 0)) ? 1 : 2;
                const int64_t idx = kx + (int64_t)(H + 1) * (ky + (int64_t)N * kz);
                const double k2 = typek2[kx] + typek2[ky] + typek2[kz];
                double s = sinc[kx] * sinc[ky] * sinc[kz];
                s *= s;
                double p;
                /* real part of a * conj(b): a.re b.re - a.im (-b.im) */
#define PRE(a, b) ((a)[0] * (b)[0] - (a)[1] * -(b)[1])
                if (comp1 >= 0 && comp2 >= 0 && comp1 < ncomp && comp2 < ncomp) {
                    p = PRE(C1(comp1, idx), C2(comp2, idx));
                } else if (symmetric) {
                    p = PRE(C1(1, idx), C2(1, idx));
                    p += PRE(C1(2, idx), C2(2, idx));
                    p += PRE(C1(4, idx), C2(4, idx));
                    p *= 2.;
                    p += PRE(C1(0, idx), C2(0, idx));
                    p += PRE(C1(3, idx), C2(3, idx));
                    p += PRE(C1(5, idx), C2(5, idx));
                } else {
                    p = 0.;
                    for (int c = 0; c < ncomp; c++) p += PRE(C1(c, idx), C2(c, idx));
                }
#undef PRE
                i = (int)floor((double)((double)numbins * sqrt(k2 / k2max)));
                if (i < numbins) {
                    kbin[i] += weight * sqrt(k2);
                    kscatter[i] += weight * k2;
                    power[i] += weight * p * k2 * sqrt(k2) / s;
                    pscatter[i] += weight * p * p * k2 * k2 * k2 / s / s;
                    occupation[i] += weight;
                }
            }
#undef C1
#undef C2
    for (i = 0; i < numbins; i++) {
        if (occupation[i] > 0) {
            kscatter[i] = sqrt(kscatter[i] * occupation[i] - kbin[i] * kbin[i]) / occupation[i];
            if (!isfinite(kscatter[i])) kscatter[i] = 0.;
            kbin[i] = kbin[i] / occupation[i];
            power[i] /= occupation[i];
            pscatter[i] = sqrt(pscatter[i] / occupation[i] - power[i] * power[i]);
            if (!isfinite(pscatter[i])) pscatter[i] = 0.;
        }
    }
    free(typek2);
    free(sinc);
}
