/*
 * gev_oracle.h -- FP64 CPU restatement of gevolution 1.2's per-cycle
 * particle-mesh loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation and the timed CPU baseline ("restated reference algorithm on
 * CPU").  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product path never calls it.
 *
 * Conventions (shared with the HIP library, include/gev_pm.h):
 *   - real fields: N^3 doubles, index x + N*(y + N*z)  (dim 0 = x fastest,
 *     as LATfield2's Site iteration);
 *   - tensor fields: 6 planes of N^3 in the order 00,01,02,11,12,22
 *     (LATfield2 "symmetric" component order);  vector fields: 3 planes;
 *   - complex (k-space) fields: N*N*(N/2+1) interleaved (re,im) doubles,
 *     index kx + H*(ky + N*kz), H = N/2+1 -- the numpy rfftn layout of an
 *     array shaped (z,y,x).  kx is the frequency of real dim 0;
 *   - FFTs are unnormalised in both directions (forward sign -1), so
 *     c2r(r2c(f)) = N^3 f  (manual.pdf sec. 4, p.14);
 *   - particles: pos[n][3] in box units [0,1), q[n][3] = canonical momentum
 *     per unit mass (gevolution.hpp:543-546), id[n]; one species, mass m.
 *
 * Parity status of each function is stated in its comment in gev_oracle.c.
 */
#ifndef GEV_ORACLE_H
#define GEV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- cosmology / background (background.hpp:137-213, metadata.hpp:250-266) */
typedef struct {
    double Omega_cdm, Omega_b, Omega_m, Omega_Lambda, Omega_rad;
    double Omega_fld, w0_fld, wa_fld;
    double h;
} orc_cosmo;

double orc_hconf(double a, double fourpiG, const orc_cosmo *c);
void   orc_rungekutta4bg(double *a, double fourpiG, const orc_cosmo *c, double dtau);
double orc_particle_horizon(double a, double fourpiG, const orc_cosmo *c);

/* ---- FFT (LATfield2 PlanFFT semantics over FFTW3 r2c / c2r) -------------- */
void orc_fft_r2c(int N, const double *in, double *out);
void orc_fft_c2r(int N, const double *in, double *out);
/* 1D complex DFT of length n (n = 2^a 3^b), sign -1 (fwd) or +1 (bwd),
   unnormalised; used by tests to pin the 3D transforms. */
void orc_fft1d(int n, const double *in, double *out, int sign);

/* ---- k-space solves (gevolution.hpp:211-535) ---------------------------- */
void orc_solve_modified_poisson_ft(int N, const double *srcFT, double *potFT,
                                   double coeff, double modif);
void orc_project_ft_scalar(int N, const double *SijFT, double *chiFT, int add);
void orc_evolve_ft_vector(int N, const double *SijFT, double *BiFT, double a2dtau);
void orc_project_ft_vector(int N, const double *SiFT, double *BiFT,
                           double coeff, double modif);

/* ---- stencils (gevolution.hpp:56-192) ------------------------------------ */
void orc_prepare_ft_source_tensor(int N, const double *phi, const double *Tij,
                                  double *Sij, double coeff);
void orc_prepare_ft_source_scalar(int N, const double *phi, const double *chi,
                                  const double *source, double bgmodel,
                                  double *result, double coeff, double coeff2,
                                  double coeff3);

/* ---- particles ------------------------------------------------------------ */
/* Cell filing (LATfield2 Particles::moveParticles re-filing): cell index of
   every particle, per-cell counts, CSR start (N^3+1) and perm = particle
   indices grouped by cell, ascending ID inside a cell. */
void orc_file_particles(int N, int64_t n, const double *pos, const int64_t *id,
                        uint32_t *cell, uint32_t *count, uint64_t *start,
                        int64_t *perm);

/* Projections.  The target is accumulated (+=), as in the reference; the
   caller zeroes it first (projection_init).  phi may be NULL. */
void orc_projection_T00_project(int N, int64_t n, const double *pos, const double *q,
                                const uint64_t *start, const int64_t *perm,
                                double pmass, double *T00, double a,
                                const double *phi, double coeff);
void orc_projection_Tij_project(int N, int64_t n, const double *pos, const double *q,
                                const uint64_t *start, const int64_t *perm,
                                double pmass, double *Tij, double a,
                                const double *phi, double coeff);
void orc_projection_T0i_project(int N, int64_t n, const double *pos, const double *q,
                                const uint64_t *start, const int64_t *perm,
                                double pmass, double *T0i, const double *phi,
                                double coeff);
void orc_scalar_projection_cic_project(int N, int64_t n, const double *pos,
                                       const uint64_t *start, const int64_t *perm,
                                       double pmass, double *rho, double coeff);

enum { ORC_Q_GR = 0, ORC_Q_NEWTON = 1 };
enum { ORC_POS_GR = 0, ORC_POS_NEWTON = 1 };

/* Particles::updateVel with update_q / update_q_Newton.  Returns
   sqrt(max over particles of the callback's return value). */
double orc_update_vel(int N, int64_t n, const double *pos, double *q, int method,
                      double dtau, const double *phi, const double *chi,
                      const double *Bi, int nfield, const double params[2]);
/* Particles::moveParticles with update_pos / update_pos_Newton, followed by
   the periodic wrap of the position into [0,1). */
void orc_move_particles(int N, int64_t n, double *pos, const double *q, int method,
                        double dtau, const double *phi, const double *chi,
                        const double *Bi, int nfield, const double params[2]);

double orc_field_sum(int64_t count, const double *f);

/* OpenMP thread count of subsequent calls (tests: results must not depend on it). */
void orc_set_threads(int n);

/* Wall-clock time of the cycle's parts, in main.cpp's BENCHMARK categories
   (main.cpp:907-922): projections, gravity solver (its FFTs also counted
   apart), update momenta, move particles, spectra output; counts of calls. */
enum { ORC_T_PROJECTION = 0, ORC_T_GRAVITY, ORC_T_FFT, ORC_T_UPDATE_Q, ORC_T_MOVE, ORC_T_SPECTRA, ORC_NTIMERS };
void orc_timers_reset(void);
void orc_timers_get(double *seconds, int64_t *counts);
void orc_timer_add(int which, double seconds);
/* STREAM-style triad bandwidth of the host (GB/s, best of reps). */
double orc_stream_triad(int64_t n, int reps);

/* ---- one cycle of main.cpp:372-879 (CDM only, no outputs) ---------------- */
typedef struct {
    int N;
    int gr_flag;       /* 1 = GR, 0 = Newton */
    int vector_flag;   /* 0 = parabolic, 1 = elliptic */
    double Cf, steplimit, z_relax, fourpiG;
    orc_cosmo cosmo;
} orc_sim;

typedef struct {
    double a, tau, dtau, dtau_old, maxvel, T00hom, phi_k0;
    int cycle;
} orc_state;

typedef struct {
    int64_t n;
    double pmass;
    double *pos, *q;
    int64_t *id;
    /* fields */
    double *phi, *chi, *source, *Sij, *Bi;
    double *scalarFT, *SijFT, *BiFT;
    /* filing */
    uint32_t *cell, *count;
    uint64_t *start;
    int64_t *perm;
} orc_box;

int orc_box_alloc(orc_box *b, int N, int64_t n);
void orc_box_free(orc_box *b);
void orc_cycle(const orc_sim *sim, orc_box *b, orc_state *st);
/* the two halves of orc_cycle: field stage (main.cpp:372-598) and particle
   stage (main.cpp:766-874) */
void orc_cycle_fields(const orc_sim *sim, orc_box *b, orc_state *st);
void orc_cycle_particles(const orc_sim *sim, orc_box *b, orc_state *st);
double orc_next_dtau(const orc_sim *sim, double a);

/* The main loop (main.cpp:372-879) with delta power-spectrum outputs at
   z_pk[0..num_pk) (descending; output.hpp:1796-1853 Newton, 1983-2028 GR)
   and EXACT_OUTPUT_REDSHIFTS (main.cpp:638-677; tools.hpp:268-346): stops
   after the field stage of the cycle that writes the last output, as the
   reference's loop breaks (main.cpp:685-693).  Tables [num_pk][numbins]:
   k / boxsize, P (interpolated in redshift through the reference's text
   round trip), sigma(k), sigma(P), count, as writePowerSpectrum prints them;
   z_out[j] the header redshift, cycle_out[j] the cycle.  Returns the number
   of outputs written, -1 if max_cycles cycles ran first. */
int orc_run_to_outputs(const orc_sim *sim, orc_box *b, orc_state *st, int num_pk, const double *z_pk,
                       int numbins, double boxsize, int max_cycles, double *k, double *pk,
                       double *sigma_k, double *sigma_pk, int *occupation, double *z_out, int *cycle_out);

/* ---- initial conditions (ic_basic.hpp, prng_engine.hpp): gev_oracle_ic.c -- */
/* counter-based PRNG: block function and the pure stream position -> output */
void orc_prng_block(uint32_t seed, uint64_t counter, uint64_t out[4]);
uint32_t orc_prng_at(uint32_t seed, uint64_t pos);
/* natural cubic spline: c[n] = half second derivatives */
int orc_spline_init(int n, const double *x, const double *y, double *c);
double orc_spline_eval(int n, const double *x, const double *y, const double *c, double xq);
/* generateCICKernel into a real field (numpcl = 0: the standard kernel) */
void orc_generate_cic_kernel(int N, double *ker, int64_t numpcl, const float *pcldata, int numtile);
/* generateDisplacementField (potFT: kernel in, realisation out) */
void orc_generate_displacement_field(int N, double *potFT, double coeff, int nsp,
                                     const double *spx, const double *spy, const double *spc,
                                     uint32_t seed, int ksphere, int deconvolve_f,
                                     int ignorekernel, int cr_log);
void orc_initialize_particle_positions(int64_t numpart, const float *partdata, int numtile,
                                       double *pos, int64_t *id);
double orc_displace_pcls_ic(int N, int64_t n, double *pos, double coeff, const double *field);
double orc_initialize_q_ic(int N, int64_t n, const double *pos, double *q, double coeff,
                           const double *field);

typedef struct {
    uint32_t seed;
    int ksphere;              /* k-domain = sphere */
    int correct_displacement; /* template-aware CIC kernel */
    int numtile;              /* tiling factor */
    int cr_log;               /* 1: correctly rounded log(r1) instead of logf */
    double z_in, A_s, n_s, k_pivot, boxsize;
} orc_ic;

double orc_pk_primordial(double k, const orc_ic *ic);
/* spline ordinates for phi, CDM displacement and CDM velocity potential */
void orc_ic_tables(const orc_sim *sim, const orc_ic *ic, int n, const double *k, const double *d_tot,
                   const double *t_tot, const double *d_cdm, const double *t_cdm, double *y_phi,
                   double *y_disp, double *y_vel);
/* generateIC_basic, transfer-function path, one CDM species; b allocated for
   numpcl * numtile^3 particles.  maxvel as main.cpp:326-330 leaves it. */
/* extractCrossSpectrum (tools.hpp:53-212); symmetric: 6-component tensor sum */
void orc_extract_cross_spectrum(int N, const double *f1, const double *f2, int ncomp, int symmetric,
                                int numbins, int deconvolve, int ktype, int comp1, int comp2, double *kbin,
                                double *power, double *kscatter, double *pscatter, int *occupation);

int orc_generate_ic_basic(const orc_sim *sim, const orc_ic *ic, int64_t numpcl, const float *pcldata,
                          int ntk, const double *k, const double *d_tot, const double *t_tot,
                          const double *d_cdm, const double *t_cdm, orc_box *b, double *maxvel,
                          double *max_disp);

#ifdef __cplusplus
}
#endif
#endif
