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
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

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

/* ======================================================================== */
/* the main loop with power-spectrum outputs                                 */
/* ======================================================================== */

static double omp_wtime_or_clock(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return (double)clock() / CLOCKS_PER_SEC;
#endif
}

/* A number as writePowerSpectrum prints it ("%e" for the table, "%f" for the
   header redshift) and the interpolating pass reads it back (fscanf %le /
   %lf), tools.hpp:283-301,318-332. */
static double text_round_trip(const char *fmt, double v)
{
    char buf[64];
    snprintf(buf, sizeof buf, fmt, v);
    return strtod(buf, NULL);
}

/* writeSpectra's delta spectrum: projection_init(source); T00 with phi (GR,
   output.hpp:2001-2010) or the CIC mass (Newton, output.hpp:1820-1827);
   plan_source forward; extractPowerSpectrum(deconvolve, KTYPE_LINEAR). */
static void delta_spectrum(const orc_sim *sim, orc_box *b, double a, int numbins, double *kbin,
                           double *power, double *kscatter, double *pscatter, int *occupation)
{
    const int N = sim->N;
    memset(b->source, 0, sizeof(double) * (size_t)N * N * N);
    if (sim->gr_flag > 0)
        orc_projection_T00_project(N, b->n, b->pos, b->q, b->start, b->perm, b->pmass, b->source, a, b->phi, 1.);
    else
        orc_scalar_projection_cic_project(N, b->n, b->pos, b->start, b->perm, b->pmass, b->source, 1.);
    orc_fft_r2c(N, b->source, b->scalarFT);
    orc_extract_cross_spectrum(N, b->scalarFT, b->scalarFT, 1, 0, numbins, 1, 1, -1, -1, kbin, power, kscatter,
                               pscatter, occupation);
}

/* Scratch of the output loop: the spectrum of the current write and what
   each output's look-ahead file holds. */
typedef struct {
    int numbins;
    double rescalek, rescalep;
    double *kb, *pw, *ks, *ps, *p2, *pre, *pre_z;
    int *oc, *have;
} pk_writer;

/* writeSpectra + writePowerSpectrum (tools.hpp:268-346 with
   EXACT_OUTPUT_REDSHIFTS) of output j at the current a */
static void write_spectrum(const orc_sim *sim, orc_box *b, const orc_state *st, pk_writer *w, int j, double zt,
                           double *k, double *pk, double *sigma_k, double *sigma_pk, int *occupation,
                           double *z_out, int *cycle_out)
{
    const int nb = w->numbins;
    double aw = st->a;
    const double t0 = omp_wtime_or_clock();
    delta_spectrum(sim, b, st->a, nb, w->kb, w->pw, w->ks, w->ps, w->oc);
    orc_timer_add(ORC_T_SPECTRA, omp_wtime_or_clock() - t0);
    for (int i = 0; i < nb; i++) w->p2[i] = w->pw[i] / w->rescalep;
    if (1. / aw < zt + 1.) {
        if (w->have[j]) { /* the file of the look-ahead write exists: interpolate */
            double weight = w->pre_z[j];
            weight = (weight - zt) / (1. + weight - 1. / aw);
            for (int i = 0; i < nb; i++)
                if (w->oc[i] > 0) w->p2[i] = w->pre[(size_t)j * nb + i];
            for (int i = 0; i < nb; i++) w->p2[i] = (1. - weight) * w->p2[i] + weight * w->pw[i] / w->rescalep;
            aw = 1. / (zt + 1.);
        }
    } else { /* the look-ahead write: what the file holds afterwards */
        w->have[j] = 1;
        w->pre_z[j] = text_round_trip("%f", (1. / aw) - 1.);
        for (int i = 0; i < nb; i++) w->pre[(size_t)j * nb + i] = w->oc[i] > 0 ? text_round_trip("%e", w->p2[i]) : 0.;
    }
    for (int i = 0; i < nb; i++) {
        const size_t o = (size_t)j * nb + i;
        k[o] = w->kb[i] / w->rescalek;
        pk[o] = w->p2[i];
        sigma_k[o] = w->ks[i] / w->rescalek;
        sigma_pk[o] = w->oc[i] > 0 ? w->ps[i] / w->rescalep / sqrt((double)w->oc[i]) : 0.;
        occupation[o] = w->oc[i];
    }
    z_out[j] = (1. / aw) - 1.;
    cycle_out[j] = st->cycle;
}

int orc_run_to_outputs(const orc_sim *sim, orc_box *b, orc_state *st, int num_pk, const double *z_pk,
                       int numbins, double boxsize, int max_cycles, double *k, double *pk,
                       double *sigma_k, double *sigma_pk, int *occupation, double *z_out, int *cycle_out)
{
    const int N = sim->N;
    const orc_cosmo *cs = &sim->cosmo;
    const double n3 = (double)N * (double)N * (double)N;
    /* rescalep of the delta spectrum: output.hpp:2026 (GR, no ncdm) and 1853 (Newton) */
    const double om = sim->gr_flag > 0 ? cs->Omega_cdm + cs->Omega_b + 0.0 : cs->Omega_m;
    pk_writer w;
    w.numbins = numbins;
    w.rescalek = boxsize;
    w.rescalep = n3 * n3 * 2. * M_PI * M_PI * om * om;
    w.kb = malloc(sizeof(double) * numbins);
    w.pw = malloc(sizeof(double) * numbins);
    w.ks = malloc(sizeof(double) * numbins);
    w.ps = malloc(sizeof(double) * numbins);
    w.p2 = malloc(sizeof(double) * numbins);
    w.pre = calloc((size_t)num_pk * numbins, sizeof(double));
    w.pre_z = calloc((size_t)num_pk, sizeof(double));
    w.oc = malloc(sizeof(int) * numbins);
    w.have = calloc((size_t)num_pk, sizeof(int));
    int pkcount = 0, ran = 0, ret = -1;
    while (ran < max_cycles) {
        orc_cycle_fields(sim, b, st);
        /* main.cpp:639-657 */
        if (pkcount < num_pk && 1. / st->a < z_pk[pkcount] + 1.) {
            write_spectrum(sim, b, st, &w, pkcount, z_pk[pkcount], k, pk, sigma_k, sigma_pk, occupation, z_out,
                           cycle_out);
            pkcount++;
        }
        /* main.cpp:659-677 */
        double tmp = st->a;
        orc_rungekutta4bg(&tmp, sim->fourpiG, cs, 0.5 * st->dtau);
        orc_rungekutta4bg(&tmp, sim->fourpiG, cs, 0.5 * st->dtau);
        if (pkcount < num_pk && 1. / tmp < z_pk[pkcount] + 1.)
            write_spectrum(sim, b, st, &w, pkcount, z_pk[pkcount], k, pk, sigma_k, sigma_pk, occupation, z_out,
                           cycle_out);
        /* main.cpp:685-693: the loop ends once every output is written */
        if (pkcount >= num_pk) {
            ret = pkcount;
            break;
        }
        orc_cycle_particles(sim, b, st);
        ran++;
    }
    free(w.kb); free(w.pw); free(w.ks); free(w.ps); free(w.p2); free(w.pre); free(w.pre_z); free(w.oc);
    free(w.have);
    return ret;
}
