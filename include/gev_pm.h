/*
 * gev_pm.h -- C ABI of the MI355X particle-mesh library (libgevpm.so).
 *
 * One simulation box (fields + one particle species) lives whole in one GPU's
 * HBM inside a gev_ctx.  Each hot-path entry point replaces one call of
 * gevolution 1.2's main loop (main.cpp:372-879) and keeps that call's
 * argument meaning; the reference function it replaces is cited beside it
 * (file:line in the gevolution-1.2 source tree).  LATfield2's halo folds,
 * halo updates and particle migration (projection_*_comm, updateHalo, the MPI
 * part of moveParticles) have no entry point: fields are periodic arrays and
 * indices wrap.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative GEV_E* code; the
 *     message is available from gev_last_error(ctx).  Nothing calls exit().
 *   - Calls are asynchronous on the context's HIP stream unless stated;
 *     functions returning host values synchronise that stream.
 *   - Host buffers are borrowed for the duration of the call only.
 *   - Real fields: N^3 doubles, index x + N*(y + N*z).  Tensor fields: 6
 *     components in the order 00,01,02,11,12,22.  Vector fields: 3.
 *     k-space fields on the host: N*N*(N/2+1) complex (re,im) doubles, index
 *     kx + (N/2+1)*(ky + N*kz) (numpy rfftn layout of a (z,y,x) array).
 *   - Particle positions are in box units [0,1); q is the canonical momentum
 *     per unit mass (gevolution.hpp:543-546).
 *   - A context is not thread-safe; distinct contexts may be used from
 *     distinct threads.
 */
#ifndef GEV_PM_H
#define GEV_PM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gev_ctx gev_ctx;

enum {
    GEV_OK = 0,
    GEV_EINVAL = -1,   /* bad argument */
    GEV_ENOMEM = -2,   /* device allocation failed */
    GEV_EHIP = -3,     /* HIP runtime error */
    GEV_ESTATE = -4    /* call not valid in the current context state */
};

/* fields (LATfield2 Field<> objects of main.cpp:226-246) */
typedef enum {
    GEV_PHI = 0,      /* phi       Field<Real>, 1 comp   */
    GEV_CHI = 1,      /* chi       Field<Real>, 1 comp   */
    GEV_SOURCE = 2,   /* source    Field<Real>, 1 comp   */
    GEV_SIJ = 3,      /* Sij       Field<Real>, 6 comps  */
    GEV_BI = 4,       /* Bi        Field<Real>, 3 comps  */
    GEV_SCALARFT = 5, /* scalarFT  Field<Cplx>, 1 comp   */
    GEV_SIJFT = 6,    /* SijFT     Field<Cplx>, 6 comps  */
    GEV_BIFT = 7      /* BiFT      Field<Cplx>, 3 comps  */
} gev_field;

/* FFT plans, main.cpp:238-246 (PlanFFT<Cplx> plan_x(&real, &kfield)) */
typedef enum {
    GEV_PLAN_SOURCE = 0, /* source <-> scalarFT */
    GEV_PLAN_PHI = 1,    /* phi    <-> scalarFT */
    GEV_PLAN_CHI = 2,    /* chi    <-> scalarFT */
    GEV_PLAN_SIJ = 3,    /* Sij    <-> SijFT    */
    GEV_PLAN_BI = 4      /* Bi     <-> BiFT     */
} gev_plan;

enum { GEV_FFT_FORWARD = 0, GEV_FFT_BACKWARD = 1 };

/* per-particle callbacks of updateVel / moveParticles (gevolution.hpp:570-903)
   are selected by enum; no host function pointers cross the ABI */
enum { GEV_Q_GR = 0, GEV_Q_NEWTON = 1 };     /* update_q, update_q_Newton     */
enum { GEV_POS_GR = 0, GEV_POS_NEWTON = 1 }; /* update_pos, update_pos_Newton */

/* forms of a box (gev_box_desc.form) */
enum {
    GEV_FORM_FP64 = 0, /* plain FP64: the drop-in                                  */
    GEV_FORM_FP32 = 1  /* precision form: FP32 field storage, FFT passes, deposit
                          sums and field stencils; FP64 particles, background,
                          momentum factors and the k-space solves' coefficients,
                          denominators and zero mode */
};

/* flags of a box (gev_box_desc.flags) */
enum {
    GEV_BOX_COMPACT = 1 /* one particle set: a drift writes each particle back
                           in place instead of storing the particles in cell
                           order into a second set (56 bytes less per particle:
                           a 1024^3 FP64 box fits one 288 GB MI355X; particle
                           passes then read through the filing's permutation) */
};

typedef struct {
    int32_t ngrid;      /* N: lattice points per dimension (2^a 3^b, 8..1024)  */
    int32_t form;       /* GEV_FORM_FP64 (0) or GEV_FORM_FP32                  */
    int64_t npart;      /* particle capacity                                   */
    void *stream;       /* hipStream_t to run on; NULL = a new stream          */
    int32_t flags;      /* GEV_BOX_* flags (0: none)                           */
    int32_t reserved0;
} gev_box_desc;

/* ---- lifecycle ---------------------------------------------------------- */

/* Device bytes a context for `desc` allocates (used for admission). */
int64_t gev_ctx_footprint_bytes(const gev_box_desc *desc);
/* Replaces the Field::initialize / PlanFFT construction block of
   main.cpp:209-246 for one box. */
int gev_ctx_create(const gev_box_desc *desc, gev_ctx **out);
int gev_ctx_destroy(gev_ctx *ctx);
const char *gev_last_error(const gev_ctx *ctx);
/* Blocks until all work queued on the context stream has finished. */
int gev_ctx_synchronize(gev_ctx *ctx);

/* ---- data movement (IC upload, parity checks) --------------------------- */

/* n particles, AoS as part_simple: pos[n][3], q[n][3], id[n]; mass is the
   species mass of part_simple_info (ic_basic.hpp:1987).  Files the particles
   into cells (as Particles::addParticle_global does). */
int gev_upload_particles(gev_ctx *ctx, int64_t n, const double *pos, const double *q,
                         const int64_t *id, double mass);
/* Copies particles back in storage order (not sorted by ID).  Every pass
   that drifts (moveParticles, kick_drift, gev_run_cycles) stores the
   particles in the cell order of the filing it read, so the order changes. */
int gev_download_particles(gev_ctx *ctx, double *pos, double *q, int64_t *id);
int64_t gev_particle_count(const gev_ctx *ctx);
/* Per-cell particle counts (N^3 uint32) and the cell index of every particle
   in storage order (n uint32) as filed by the last moveParticles/upload. */
int gev_download_filing(gev_ctx *ctx, uint32_t *cell_count, uint32_t *particle_cell);

/* Host arrays are FP64 in both forms; an FP32 context rounds uploads to float
   and widens downloads. */
int gev_upload_field(gev_ctx *ctx, gev_field f, const double *host);
int gev_download_field(gev_ctx *ctx, gev_field f, double *host);

/* Carries the whole state of src into dst (same ngrid, same device, capacity
   for src's particles), converting between the forms: particles, every real
   and k-space field (BiFT, the carried vector state, included) and the
   species mass; dst refiles its particles.  The initial-condition generator
   runs on an FP64 context; this brings its result into an FP32 one.  No
   counterpart in the reference, which has one precision per build. */
int gev_copy_state(gev_ctx *src, gev_ctx *dst);

/* ---- hot path: one entry per reference call ----------------------------- */

/* projection_init (LATfield2; main.cpp:378,438): zero a real field. */
int gev_projection_init(gev_ctx *ctx, gev_field f);

/* projection_T00_project, gevolution.hpp:926-1022 (main.cpp:385):
   source += T00 deposit.  use_phi = 0 passes phi = NULL. */
int gev_projection_T00_project(gev_ctx *ctx, double a, int use_phi, double coeff);
/* projection_Tij_project, gevolution.hpp:1172-1297 (main.cpp:439): Sij += Tij. */
int gev_projection_Tij_project(gev_ctx *ctx, double a, int use_phi, double coeff);
/* projection_T00_project + projection_Tij_project fused into one particle
   pass (main.cpp:385 and 439 with the same a, phi and coeff). */
int gev_projection_T00_Tij_project(gev_ctx *ctx, double a, double coeff);
/* projection_T0i_project, gevolution.hpp:1045-1147 (main.cpp:427): Bi += T0i. */
int gev_projection_T0i_project(gev_ctx *ctx, int use_phi, double coeff);
/* scalarProjectionCIC_project (LATfield2; main.cpp:402): source += CIC of
   coeff*mass/dx^3. */
int gev_scalar_projection_cic_project(gev_ctx *ctx, double coeff);

/* Sum of a real field's first component (main.cpp:459-462, T00hom*N^3);
   synchronous. */
int gev_field_sum(gev_ctx *ctx, gev_field f, double *sum);

/* prepareFTsource (2), gevolution.hpp:169-192 (main.cpp:472):
   source <- f(phi, chi, source). */
int gev_prepare_ft_source_scalar(gev_ctx *ctx, double bgmodel, double coeff, double coeff2,
                                 double coeff3);
/* prepareFTsource (1), gevolution.hpp:56-147 (main.cpp:539): Sij <- f(phi, Sij). */
int gev_prepare_ft_source_tensor(gev_ctx *ctx, double coeff);

/* PlanFFT<Cplx>::execute (LATfield2 over FFTW3; main.cpp:477,488,544,563,575,593).
   Unnormalised r2c forward / c2r backward; backward preserves the k field. */
int gev_fft_execute(gev_ctx *ctx, gev_plan plan, int direction);

/* solveModifiedPoissonFT(scalarFT, scalarFT, coeff, modif),
   gevolution.hpp:501-535 (main.cpp:483, 506). */
int gev_solve_modified_poisson_ft(gev_ctx *ctx, double coeff, double modif);
/* projectFTscalar(SijFT, scalarFT, add), gevolution.hpp:211-265 (main.cpp:558). */
int gev_project_ft_scalar(gev_ctx *ctx, int add);
/* evolveFTvector(SijFT, BiFT, a2dtau), gevolution.hpp:284-330 (main.cpp:586). */
int gev_evolve_ft_vector(gev_ctx *ctx, double a2dtau);
/* projectFTscalar(SijFT, scalarFT, 0) + evolveFTvector(SijFT, BiFT, a2dtau)
   fused into one pass over SijFT (main.cpp:558 and 586). */
int gev_project_ft_scalar_evolve_vector(gev_ctx *ctx, double a2dtau);
/* projectFTvector(BiFT, BiFT, coeff, modif), gevolution.hpp:350-392
   (main.cpp:580; ic_basic.hpp:2238). */
int gev_project_ft_vector(gev_ctx *ctx, double coeff, double modif);

/* Particles::updateVel(update_q | update_q_Newton, dtau, {phi,chi,Bi}, nfield,
   params) (LATfield2 driver; gevolution.hpp:570-776; main.cpp:771-781).
   params = {a, a^2 N}.  *maxvel (may be NULL; synchronous if not) receives
   sqrt(max over particles of the callback's return). */
int gev_update_vel(gev_ctx *ctx, int method, double dtau, int nfield, const double params[2],
                   double *maxvel);
/* Particles::moveParticles(update_pos | update_pos_Newton, dtau, ..., nfield,
   params) (gevolution.hpp:810-903; main.cpp:794-804) including the periodic
   wrap and re-filing into cells. */
int gev_move_particles(gev_ctx *ctx, int method, double dtau, int nfield,
                       const double params[2]);
/* updateVel followed by moveParticles in one particle pass (main.cpp:771-800).
   The host advances a by rungekutta4bg between the two calls in the reference;
   here the caller passes both parameter sets.  *maxvel as gev_update_vel. */
int gev_kick_drift(gev_ctx *ctx, int q_method, double dtau_kick, int nfield_kick,
                   const double params_kick[2], int pos_method, double dtau_drift,
                   int nfield_drift, const double params_drift[2], double *maxvel);

/* ---- cycle driver (main.cpp:372-879 for one CDM species) ----------------- */

typedef struct {
    double Omega_cdm, Omega_b, Omega_m, Omega_Lambda, Omega_rad;
    double Omega_fld, w0_fld, wa_fld;
    double h;
} gev_cosmology;

typedef struct {
    int32_t gr_flag;     /* 1 = GR, 0 = Newton                       */
    int32_t vector_flag; /* 0 = parabolic, 1 = elliptic              */
    double Cf;           /* Courant factor                           */
    double steplimit;    /* time step limit (fraction of Hubble time)*/
    double z_relax;      /* relaxation redshift (nfield switch)      */
    double fourpiG;      /* 1.5 boxsize^2 / c^2                      */
    gev_cosmology cosmo;
} gev_sim_params;

typedef struct {
    double a, tau, dtau, dtau_old, maxvel, T00hom, phi_k0;
    int32_t cycle;
    int32_t stage;  /* 0: at the start of `cycle`; 1: the field stage of `cycle`
                       has run (gev_run_to_outputs stopped at an output) and its
                       particle stage comes next */
} gev_cycle_state;

/* Conformal Hubble rate, background.hpp:137-140 (no ncdm). */
double gev_hconf(double a, double fourpiG, const gev_cosmology *c);
/* rungekutta4bg, background.hpp:167-177. */
void gev_rungekutta4bg(double *a, double fourpiG, const gev_cosmology *c, double dtau);
/* Time step rule of main.cpp:292-295 / 869-872. */
double gev_next_dtau(const gev_sim_params *p, int ngrid, double a);

/* Runs `ncycles` cycles of the main loop (no outputs) and advances *st as
   main.cpp:825-874 does; a state left at stage 1 by gev_run_to_outputs
   completes its cycle's particle stage as the first of them.  Synchronous at
   the end (reads T00hom, maxvel).
   The cycle applies prepareFTsource inside the first FFT pass and the k-space
   solves inside the z passes, so afterwards source and Sij hold the raw
   deposits, SijFT is scratch and scalarFT is only kept at k = 0; phi, chi,
   Bi, BiFT and the particles are the carried state, as in the reference. */
int gev_run_cycles(gev_ctx *ctx, const gev_sim_params *p, gev_cycle_state *st, int ncycles);

/* ---- basic initial conditions (ic_basic.hpp, prng_engine.hpp) ----------- */

/* Outputs pos0 .. pos0+n-1 of sitmo::prng_engine after seed(seed)
   (prng_engine.hpp:88-380; discard(z) advances the position by z), computed
   on the device; host_out receives n values.  Synchronous. */
int gev_prng_fill(gev_ctx *ctx, uint32_t seed, uint64_t pos0, int64_t n, uint32_t *host_out);

/* loadHomogeneousTemplate (ic_basic.hpp:191-324): Gadget-2 file, type-1
   particle positions divided by the header's BoxSize, each in [0,1].  Sets
   *numpcl; writes 3 floats per particle to out when capacity (particles)
   suffices, else returns GEV_EINVAL.  No context, no device work. */
int gev_load_homogeneous_template(const char *path, float *out, int64_t capacity, int64_t *numpcl);

/* loadTransferFunctions, file form (ic_basic.hpp:494-717): the columns
   labelled k, d_<qname> and t_<qname> of a CLASS-format table; returns
   k * boxsize, d, t * boxsize / h.  Errors as the reference's aborts (no
   file, fewer than 2 rows, columns not found, negative or unordered k). */
int gev_load_transfer_functions(const char *path, const char *qname, double boxsize, double h,
                                double *k, double *tk_d, double *tk_t, int32_t capacity,
                                int32_t *npoints);

/* generateCICKernel (ic_basic.hpp:737-1052) into the real scalar field f
   (GEV_SOURCE in the reference).  numpcl = 0 or pcldata = NULL: the standard
   kernel; otherwise the kernel of the template tiled numtile times. */
int gev_generate_cic_kernel(gev_ctx *ctx, gev_field f, int64_t numpcl, const float *pcldata,
                            int32_t numtile);

/* generateDisplacementField (ic_basic.hpp:1089-1379) on scalarFT in place: on
   entry the transformed CIC kernel (unless ignorekernel), on exit the Gaussian
   realisation whose amplitude is the natural cubic spline through (x[i], y[i])
   (gsl_interp_cspline) at 2 pi |k|.  The Gaussian factor uses the correctly
   rounded single-precision log of r1; the reference's logf may differ by one
   float ulp. */
int gev_generate_displacement_field(gev_ctx *ctx, double coeff, int32_t nknots, const double *x,
                                    const double *y, uint32_t seed, int32_t ksphere,
                                    int32_t deconvolve_f, int32_t ignorekernel);

/* initializeParticlePositions (ic_basic.hpp:1400-1434): numpart * numtile^3
   particles at (tile + template) / numtile with ID = i + numpart (xt +
   numtile (yt + numtile zt)), q = 0, species mass `mass`; filed into cells. */
int gev_initialize_particle_positions(gev_ctx *ctx, int64_t numpart, const float *partdata,
                                      int32_t numtile, double mass);

/* moveParticles(displace_pcls_ic_basic, coeff, &f, 1, NULL, &max, MAX, 1)
   (ic_basic.hpp:60-116, 1995): pos += coeff grad(f), wrapped and re-filed;
   *max_displacement (may be NULL; synchronous if not) = max coeff |grad f|. */
int gev_displace_particles_ic(gev_ctx *ctx, double coeff, gev_field f, double *max_displacement);

/* updateVel(initialize_q_ic_basic, coeff, &f, 1) (ic_basic.hpp:118-173,
   2065): q = -coeff grad(f); *maxvel (may be NULL) = sqrt(max |q|^2). */
int gev_initialize_q_ic(gev_ctx *ctx, double coeff, gev_field f, double *maxvel);

typedef struct {
    uint32_t seed;
    int32_t ksphere;              /* k-domain = sphere                  */
    int32_t correct_displacement; /* template-aware CIC kernel          */
    int32_t numtile;              /* tiling factor of the template      */
    double z_in;                  /* initial redshift                   */
    double A_s, n_s, k_pivot;     /* primordial spectrum; k_pivot 1/Mpc */
    double boxsize;               /* Mpc/h                              */
} gev_ic_params;

/* generateIC_basic (ic_basic.hpp:1626-2259), transfer-function path for one
   CDM species with the baryons folded in (baryon treatment "ignore") and no
   non-CDM species.  k, d_tot, t_tot, d_cdm, t_cdm: ntk rows as returned by
   gev_load_transfer_functions for "tot" and "cdm".  Leaves the particles
   (filed), phi, chi, Bi and BiFT as the reference leaves them; *maxvel as
   main.cpp:326-330 turns it (GR: v / sqrt(v^2 + 1)); *max_displacement in
   box units.  Either pointer may be NULL. */
int gev_generate_ic_basic(gev_ctx *ctx, const gev_sim_params *p, const gev_ic_params *ic,
                          int64_t numpcl, const float *pcldata, int32_t ntk, const double *k,
                          const double *d_tot, const double *t_tot, const double *d_cdm,
                          const double *t_cdm, double *maxvel, double *max_displacement);

/* ---- hibernation and restart (hibernation.hpp:512-612; ic_read.hpp) ----- */

/* hibernate() after a gev_run_cycles call (main.cpp:827-865): writes the
   particles, phi and chi (GR) and B / (a^2 N) (parabolic; a Newtonian run
   first transforms BiFT back), with a, tau, the completed cycle's dtau and
   its index, to one binary file.  Needs a state at stage 0.  Synchronous. */
int gev_hibernate(gev_ctx *ctx, const gev_sim_params *p, const gev_cycle_state *st, const char *path);
/* readIC's restart branch (ic_read.hpp:58-330, 671-716): loads the file into
   a context of the same grid, rescales B by a^2 / N^2 and rebuilds BiFT by a
   forward FFT, and sets *st so that gev_run_cycles continues the run (cycle
   = saved + 1, dtau_old = saved dtau, dtau by ic_read.hpp:298-301). */
int gev_restart(gev_ctx *ctx, const gev_sim_params *p, const char *path, gev_cycle_state *st);

/* ---- power spectra (tools.hpp:53-346; output.hpp:1990-2028) -------------- */

/* extractCrossSpectrum (tools.hpp:53-212) of two k-space fields
   (GEV_SCALARFT, GEV_SIJFT, GEV_BIFT).  comp1, comp2 >= 0 pick one component
   of each; otherwise a 6-component field uses the symmetric-tensor sum
   (off-diagonals doubled) and others the sum over components.  numbins in
   [1, 1024]; deconvolve: divide by the CIC window; ktype 0 = grid momentum
   2N sin(pi n/N), 1 = linear 2 pi n.  Outputs per bin: mean k, mean power
   (weighted by k^3), the k and power scatters, and the mode count, as the
   reference's root process leaves them.  Per-bin sums are deterministic.
   Synchronous. */
int gev_extract_cross_spectrum(gev_ctx *ctx, gev_field f1, gev_field f2, int32_t numbins,
                               int32_t deconvolve, int32_t ktype, int32_t comp1, int32_t comp2,
                               double *kbin, double *power, double *kscatter, double *pscatter,
                               int32_t *occupation);
/* extractPowerSpectrum (tools.hpp:236-239): the cross spectrum of f with itself. */
int gev_extract_power_spectrum(gev_ctx *ctx, gev_field f, int32_t numbins, int32_t deconvolve,
                               int32_t ktype, double *kbin, double *power, double *kscatter,
                               double *pscatter, int32_t *occupation);
/* gev_extract_power_spectrum that allocates nothing: no counterpart in the
   reference; for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid it).  Its device
   scratch is a static array of the library (about 10.5 MB per device, held
   until the process ends) shared by every call on that device, so calls that
   overlap on one device race; every call overwrites what it uses.  Limits:
   numbins <= 256 and N <= 1024 (GEV_EINVAL beyond).  Results equal
   gev_extract_power_spectrum's bit for bit.  Synchronous. */
int gev_extract_power_spectrum_noalloc(gev_ctx *ctx, gev_field f, int32_t numbins, int32_t deconvolve,
                                       int32_t ktype, double *kbin, double *power, double *kscatter,
                                       double *pscatter, int32_t *occupation);

/* writePowerSpectrum (tools.hpp:268-346) with EXACT_OUTPUT_REDSHIFTS: when
   1/a < z_target + 1 and `filename` already holds the spectrum written before
   z_target was crossed, the power is interpolated linearly in redshift to
   z_target.  ASCII table as the reference writes it.  No context. */
int gev_write_power_spectrum(const double *kbin, const double *power, const double *kscatter,
                             const double *pscatter, const int32_t *occupation, int32_t numbins,
                             double rescalek, double rescalep, const char *filename,
                             const char *description, double a, double z_target);

/* Where each output of gev_run_to_outputs was written. */
typedef struct {
    double z;             /* the redshift of the table's header line          */
    int32_t cycle;        /* the cycle whose field stage it followed          */
    int32_t interpolated; /* 1: interpolated in redshift with the look-ahead
                             write of the cycle before                         */
} gev_pk_output_info;

/* The main loop, main.cpp:372-879, with the delta power-spectrum output
   (writeSpectra: output.hpp:1796-1853 Newton, CIC mass; 1983-2028 GR, T00
   with phi) at each of z_pk[0..num_pk) (strictly descending) and
   EXACT_OUTPUT_REDSHIFTS on (main.cpp:638-677; writePowerSpectrum,
   tools.hpp:268-346): an output is written after the field stage of the
   cycle whose a has crossed z_pk[j], interpolated linearly in redshift with
   the spectrum written one cycle before by the look-ahead, whose values pass
   through the reference's text table ("%e" powers, "%f" redshift) exactly as
   the reference reads its own file back.  The loop ends after the field
   stage of the cycle that writes the last output (main.cpp:685-693), with no
   particle update, so st->cycle is the number of completed particle updates.
   Tables [num_pk][numbins] hold what writePowerSpectrum prints: k / boxsize,
   P, sigma(k), sigma(P), count (rows with count 0 are not printed by the
   reference).  numbins in [1, 1024]; boxsize in Mpc/h.  Returns GEV_ESTATE
   if max_cycles particle updates ran before the last output.  Synchronous.
   On return st->stage is 1; a later call (or gev_run_cycles) continues the
   loop exactly as if it had not stopped: it first makes the stopped cycle's
   look-ahead check for its own z_pk[0], then runs that cycle's particle
   stage. */
int gev_run_to_outputs(gev_ctx *ctx, const gev_sim_params *p, gev_cycle_state *st, int32_t num_pk,
                       const double *z_pk, int32_t numbins, double boxsize, int32_t max_cycles,
                       double *k, double *pk, double *sigma_k, double *sigma_pk, int32_t *occupation,
                       gev_pk_output_info *info);

/* One line of the background log, main.cpp:520-537: the cycle, tau and a at
   the cycle's start, Hconf(a)/Hconf(1), phi(k=0) after the cycle's Poisson
   solve (scalarFT(k=0).real()) and T00hom (0 for Newtonian runs). */
typedef struct {
    int32_t cycle;
    int32_t reserved0;
    double tau, a, hconf_ratio, phi_k0, T00hom;
} gev_background_record;

/* gev_run_cycles that also returns one background record per cycle in
   rec[ncycles] (device values copied asynchronously; one synchronisation per
   1024 cycles).  Needs a state at stage 0. */
int gev_run_cycles_recorded(gev_ctx *ctx, const gev_sim_params *p, gev_cycle_state *st, int ncycles,
                            gev_background_record *rec);
/* Appends records to a background file in the reference's format
   (main.cpp:531-533: header before the cycle-0 line).  No context. */
int gev_write_background(const char *path, const gev_background_record *rec, int n);

/* ---- batches: many boxes stepped by the same launches -------------------- */

/* No counterpart in the reference, which runs one simulation per process
   (main.cpp:115-149): a batch holds nbox boxes of the same grid, particle
   count, form and gravity (GR or Newtonian; parabolic B) and runs the main
   loop of main.cpp:372-879 for all of them with one launch per kernel, each
   box with its own cosmology, seed, time step and outputs.  The per-cycle
   host arithmetic of gev_run_cycles (the background step and every call's
   coefficients) runs on the device with the same expressions, so every
   box's results equal those of the same box run alone with gev_run_cycles /
   gev_run_to_outputs bit for bit, and each stage of the cycle can be
   captured once in a HIP graph and replayed (use_graph). */
typedef struct gev_batch gev_batch;

typedef struct {
    int32_t ngrid;     /* N of every box                                */
    int32_t form;      /* GEV_FORM_FP64 or GEV_FORM_FP32                 */
    int32_t nbox;      /* boxes                                          */
    int32_t gr_flag;   /* 1 = GR, 0 = Newton, for every box              */
    int64_t npart;     /* particles of every box                         */
    void *stream;      /* hipStream_t; NULL = a new stream               */
} gev_batch_desc;

int64_t gev_batch_footprint_bytes(const gev_batch_desc *desc);
int gev_batch_create(const gev_batch_desc *desc, gev_batch **out);
int gev_batch_destroy(gev_batch *batch);
const char *gev_batch_last_error(const gev_batch *batch);
/* The single-box context of box b: every entry point above works on it (IC
   generation, uploads, downloads, spectra), on the batch's stream, between
   batch calls; the batch's next cycle starts from whatever it then holds.
   Owned by the batch (gev_ctx_destroy refuses it). */
int gev_batch_box(gev_batch *batch, int32_t b, gev_ctx **view);
/* Box b takes part in the batch's cycles from state *st (stage 0) with its
   own parameters (p->gr_flag equal to the batch's, p->vector_flag 0); its
   particles (npart of them, filed) and fields are what its view holds;
   boxsize in Mpc/h scales its spectrum outputs. */
int gev_batch_start_box(gev_batch *batch, int32_t b, const gev_sim_params *p, const gev_cycle_state *st,
                        double boxsize);
int gev_batch_stop_box(gev_batch *batch, int32_t b);
/* Box b's state as the batch advanced it (maxvel read at the end of runs). */
int gev_batch_get_state(gev_batch *batch, int32_t b, gev_cycle_state *st);
/* ncycles cycles of every running box (no outputs), as gev_run_cycles. */
int gev_batch_run_cycles(gev_batch *batch, int32_t ncycles, int32_t use_graph);
/* The delta power-spectrum outputs of every box, as gev_run_to_outputs
   writes them: at z_pk[0..num_pk) (strictly descending), numbins bins. */
int gev_batch_set_outputs(gev_batch *batch, int32_t num_pk, const double *z_pk, int32_t numbins);
/* One cycle of every running box with its outputs; a box that writes its
   last output stops after that cycle's field stage (finished[b] = 1 for the
   boxes that finished in this call; finished may be NULL). */
int gev_batch_step_outputs(gev_batch *batch, int32_t use_graph, int32_t *finished);
/* Waits for the batch's work; checks every running box's device time
   stepping (a, tau, dtau, cycle) against the host's, and reads maxvel. */
int gev_batch_synchronize(gev_batch *batch);
/* Steps until every running box has written its last output. */
int gev_batch_run_to_outputs(gev_batch *batch, int32_t use_graph, int32_t max_cycles);
/* The batch's tile autotuner (runs once, before the first cycle, if not
   called): times the strided FFT passes' tiles with 64-, 128- and 256-byte
   rows on the batch's FFT scratch over every box and keeps the fastest per
   axis (N <= 256; results do not depend on the pick).  Returns the picks
   (0: the default tile) and pass_ms[2][3], the candidates' times in ms per
   launch (y then z; all zero when the pick came from an earlier batch of the
   same grid, form and box count).  Any pointer may be NULL. */
int gev_batch_autotune(gev_batch *batch, int32_t *row_bytes_y, int32_t *row_bytes_z, float *pass_ms);
/* The [num_pk][numbins] tables of a finished box (see gev_run_to_outputs). */
int gev_batch_get_outputs(gev_batch *batch, int32_t b, double *k, double *pk, double *sigma_k, double *sigma_pk,
                          int32_t *occupation, gev_pk_output_info *info);

/* ---- timing (the -DBENCHMARK categories of main.cpp:71-88) --------------- */

/* When enabled, every kernel launch is bracketed by HIP events on the context
   stream and its duration is accumulated per kernel name. */
int gev_timing_enable(gev_ctx *ctx, int on);
int gev_timing_reset(gev_ctx *ctx);
/* Writes "name calls total_ms\n" lines into buf (NUL-terminated; truncated to
   len). Synchronous. */
int gev_timing_report(gev_ctx *ctx, char *buf, size_t len);

/* ---- bandwidth probe ----------------------------------------------------- */

/* No counterpart in the reference: the device's achievable HBM bandwidth,
   measured with hand-written streaming kernels over two buffers of `bytes`
   each (allocated and freed here, on a stream of its own), `reps` timed
   launches after one untimed.  *gbs = bytes moved per launch (read + written
   for the copies) / mean launch time; *ms = mean launch time.  It is the
   practical ceiling the kernels' HBM rooflines are read against. */
enum { GEV_PROBE_COPY = 0, GEV_PROBE_COPY_NT = 1, GEV_PROBE_READ = 2, GEV_PROBE_WRITE = 3 };
int gev_bandwidth_probe(int64_t bytes, int kind, int reps, double *gbs, double *ms);

#ifdef __cplusplus
}
#endif
#endif
