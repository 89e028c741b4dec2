// gev_internal.h -- context layout and helpers shared by the HIP sources of
// libgevpm.so.  Not part of the public ABI (include/gev_pm.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/gev_pm.h"

#define GEV_EXPORT extern "C" __attribute__((visibility("default")))

// per-cycle background records buffered in pinned memory before one synchronisation
constexpr int REC_SLOTS = 1024;

namespace gev {

// A store that streams past the caches (non-temporal) when nt is set: for
// the large outputs no kernel reads back before they have left the caches
// anyway; small boxes keep plain stores, whose data the next kernel may
// still find in the caches.
template <typename T> __device__ __forceinline__ void store_stream(T *p, T v, bool nt) {
    if (nt) {
        if constexpr (sizeof(T) == 16) {
            using V = double __attribute__((ext_vector_type(2)));
            __builtin_nontemporal_store(*reinterpret_cast<const V *>(&v), reinterpret_cast<V *>(p));
        } else if constexpr (sizeof(T) == 8 && !__is_same(T, double) && !__is_same(T, int64_t)) {
            using V = float __attribute__((ext_vector_type(2)));
            __builtin_nontemporal_store(*reinterpret_cast<const V *>(&v), reinterpret_cast<V *>(p));
        } else {
            __builtin_nontemporal_store(v, p);
        }
    } else {
        *p = v;
    }
}
// the grid size from which the large outputs are streamed
constexpr int STREAM_MIN_N = 256;


// One complex k-space row holds kx in [0, N/2]; rows are padded to a multiple
// of 8 complexes so that every row starts on a 128-byte boundary (FP64).
inline int kpitch(int N) { return ((N / 2 + 1) + 7) / 8 * 8; }

struct TimerRec {
    const char *name;
    hipEvent_t beg, end;
};

struct TimerAcc {
    std::string name;
    long calls = 0;
    double ms = 0.0;
};

// ---------------------------------------------------------------- background

// Conformal Hubble rate, background.hpp:137-140 (no ncdm), the same code on
// the host and the device (-ffp-contract=off on both, IEEE sqrt and division),
// so a, tau and dtau advanced on the device equal the host's bit for bit.  The
// dark-energy fluid term is skipped when Omega_fld is zero: it is then exactly
// zero and adding it changes nothing, and the device's exp / pow need not
// equal the host library's.
__host__ __device__ inline double hconf_of(double a, double fourpiG, double Om_cdm, double Om_b, double Om_L,
                                           double Om_rad, double Om_fld, double w0, double wa) {
    double s = ((Om_cdm + Om_b + 0.0) / a) + (Om_L * a * a) + (Om_rad / a / a);
    if (Om_fld != 0.0) s = s + (Om_fld * exp(3. * wa * (a - 1.)) / pow(a, 1. + 3. * (w0 + wa)));
    return sqrt((2. * fourpiG / 3.) * s);
}

// rungekutta4bg, background.hpp:167-177, over a Hubble-rate function of a
template <typename H> __host__ __device__ inline void rk4bg_with(double &a, double dtau, const H &hconf) {
    double k1a, k2a, k3a, k4a;
    k1a = a * hconf(a);
    k2a = (a + k1a * dtau / 2.) * hconf(a + k1a * dtau / 2.);
    k3a = (a + k2a * dtau / 2.) * hconf(a + k2a * dtau / 2.);
    k4a = (a + k3a * dtau) * hconf(a + k3a * dtau);
    a += dtau * (k1a + 2. * k2a + 2. * k3a + k4a) / 6.;
}

// ---------------------------------------------------------------- batches

// Scalars of one box of a batch for one cycle (device array, one per box).
// The state and the static parameters are set by the host; k_cycle_begin
// derives the cycle's coefficients with the host driver's expressions and
// k_cycle_end advances the state (main.cpp:766-874).
struct BoxCycle {
    // state at the cycle's start (as gev_cycle_state)
    double a, tau, dtau, dtau_old;
    // static per box
    double fourpiG, Cf, steplimit, z_relax, pmass;
    double Om_cdm, Om_b, Om_L, Om_rad, Om_fld, w0, wa;
    // the cycle's coefficients
    double a_half;                      // a after the first half RK4 step (drift)
    double dep_m, dep_m_a;              // deposit mass factors, coeff m / dx^3 (and / a)
    double prep_bg, prep_c, prep_c2, prep_c3;  // prepareFTsource (scalar), main.cpp:472
    double pois_c0, pois_c1;            // solveModifiedPoissonFT: coeff / (-N^3), modif
    double tens_c;                      // prepareFTsource (tensor): 2 fourpiG dx^2 / a
    double a2dtau;                      // evolveFTvector: a^2 dtau_old
    double dtk, pk0, pk1, dtd, pd0, pd1;  // kick and drift
    int32_t nfk, nfd;
    int32_t on;       // the box takes part in this cycle
    int32_t on_phi;   // ... and its phi update runs (GR with dtau_old > 0; Newton always)
    int32_t parity;   // the live particle set: 0 = pos/q/id, 1 = pos2/q2/id2
    int32_t cycle;
    int32_t gr;
    int32_t stop;     // the host ends the box after this cycle's field stage
};

// A launch over the boxes of a batch: box b's arrays are those of box 0
// shifted by b * stride bytes.  cyc == nullptr: one box, the kernel's
// by-value coefficients; otherwise the box index comes from the grid and the
// coefficients from cyc[box].  flag selects the box's enable flag (0: on,
// 1: on_phi).
struct Batch {
    int64_t stride = 0;
    const BoxCycle *cyc = nullptr;
    int flag = 0;
};

__device__ __forceinline__ bool box_on(const Batch &bt, int b) {
    return !bt.cyc || (bt.flag ? bt.cyc[b].on_phi : bt.cyc[b].on) != 0;
}
template <typename T> __host__ __device__ __forceinline__ T *box_ptr(T *p, int64_t stride, int b) {
    return p ? reinterpret_cast<T *>(reinterpret_cast<char *>(const_cast<typename std::remove_const<T>::type *>(p)) +
                                     (int64_t)b * stride)
             : p;
}

}  // namespace gev

struct gev_ctx {
    // device memory: one slab per box (nbox boxes, stride bytes apart); the
    // pointers below are box 0's.  A view (a box of a batch handed out for the
    // single-box entry points) has its pointers shifted and owns nothing.
    void *slab = nullptr;
    int nbox = 1;
    int64_t stride = 0;
    bool view = false;

    int N = 0;
    int P = 0;            // k-space row pitch (complex elements)
    bool compact = false; // one particle set, drifts in place (GEV_BOX_COMPACT)
    int64_t cap = 0;      // particle capacity
    int64_t n = 0;        // particles resident
    double pmass = 0.0;   // species mass (part_simple_info)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;

    // FP32 form: the fields below hold float / float2 elements (the pointers
    // keep their FP64 type as handles; launchers cast with fld<F>()); the
    // particles, the k-space tables and the background stay FP64.
    bool fp32 = false;
    // real fields, N^3 elements per component
    double *phi = nullptr, *chi = nullptr, *source = nullptr, *Sij = nullptr, *Bi = nullptr;
    // complex fields, P*N*N complexes per component
    double2 *scalarFT = nullptr, *SijFT = nullptr, *BiFT = nullptr, *fft_tmp = nullptr;
    // FFT twiddles exp(-2 pi i t/N), t < N (float copy for the FP32 form); k-space tables
    double2 *twiddle = nullptr;
    float2 *twiddle_f = nullptr;
    double *gridk2 = nullptr;   // (2N sin(pi n/N))^2
    double2 *kshift = nullptr;  // 2N sin(pi n/N) exp(-i pi n/N)

    // particles, structure of arrays.  A drifting pass writes every particle
    // at its cell-sorted index into the second set and the sets swap, so the
    // storage order follows the last filing.
    double *pos = nullptr;  // [3][cap]
    double *q = nullptr;    // [3][cap]
    int64_t *id = nullptr;
    double *pos2 = nullptr, *q2 = nullptr;
    int64_t *id2 = nullptr;
    uint32_t *cell = nullptr;   // filed cell of every particle
    uint32_t *perm = nullptr;   // particle indices grouped by cell, ascending ID
    uint32_t *count = nullptr;  // N^3 per-cell counts; the scatter counts them back down to zero
    bool count_clean = false;   // count is all zero (every completed filing leaves it so)
    uint32_t *start = nullptr;  // N^3 + 1 CSR offsets
    uint32_t *scan_part = nullptr;  // scan block partials
    // the filing's ID sort: a flag byte per segment (256 consecutive cells)
    // with a cell of more than 4 particles, the listed cells of dense
    // segments, and their count (sort_n[1])
    uint32_t *sort_segs = nullptr, *sort_big = nullptr, *sort_n = nullptr;
    int64_t scan_nblk = 0;

    // row bytes of the y (0) and z (1) FFT pass tiles; 0: the form's default
    // (a batch's autotuner sets them)
    int col_rb[2] = {0, 0};

    // power-spectrum scratch (tables, workgroup histograms, bin totals) for
    // up to pk_bins bins
    double *pk_scratch = nullptr;
    int pk_bins = 0;
    // the IC generator's host-built tables, copied in stream order
    // (ic_stage_bytes(N) per box)
    char *ic_stage = nullptr;

    // reductions
    double *red_d = nullptr;              // device double scratch (sums)
    unsigned long long *red_u = nullptr;  // device max-as-bits scratch
    double *host_pinned = nullptr;        // pinned readback slot
    double *rec_pinned = nullptr;         // pinned per-cycle (T00hom sum, phi(k=0)) pairs, REC_SLOTS of them

    // state: field known to be all zero (projection_init), so deposits store
    bool zeroed[5] = {false, false, false, false, false};

    std::string err;

    // timing
    bool timing = false;
    std::vector<gev::TimerRec> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<gev::TimerAcc> acc;
};

namespace gev {

// sets ctx->err, returns code
int fail(gev_ctx *ctx, int code, const std::string &msg);
int check_hip(gev_ctx *ctx, hipError_t e, const char *what);

// launch bracketing for the -DBENCHMARK style timers
void timer_begin(gev_ctx *ctx, const char *name);
void timer_end(gev_ctx *ctx);

struct ScopedTimer {
    gev_ctx *c;
    ScopedTimer(gev_ctx *ctx, const char *name) : c(ctx) { timer_begin(ctx, name); }
    ~ScopedTimer() { timer_end(c); }
};

inline int64_t n3(const gev_ctx *c) { return (int64_t)c->N * c->N * c->N; }
// element bytes of a real field and of a complex field in the context's form
inline int64_t rbytes(const gev_ctx *c) { return c->fp32 ? 4 : 8; }
inline int64_t cbytes(const gev_ctx *c) { return c->fp32 ? 8 : 16; }
// a field handle as the element type of the form in use
template <typename F> inline F *fld(double *p) { return reinterpret_cast<F *>(p); }
template <typename F> inline const F *fld(const double *p) { return reinterpret_cast<const F *>(p); }
template <typename C> inline C *cfld(double2 *p) { return reinterpret_cast<C *>(p); }
template <typename C> inline const C *cfld(const double2 *p) { return reinterpret_cast<const C *>(p); }
inline int64_t nk(const gev_ctx *c) { return (int64_t)c->P * c->N * c->N; }
// cells the filing's ID sort lists, at most: cells of more than 4 particles
// in segments (256 consecutive cells) holding more than 2048 particles
inline int64_t sort_big_cap(int64_t cap) { return (cap / 2049 + 1) * 256; }
// component k of a complex field handle (components are nk elements apart in
// the form's element size)
inline double2 *kcomp(const gev_ctx *c, double2 *p, int k) {
    return reinterpret_cast<double2 *>(reinterpret_cast<char *>(p) + (size_t)k * nk(c) * cbytes(c));
}

inline unsigned grid_for(int64_t work, int block) {
    return (unsigned)((work + block - 1) / block);
}

// kernels implemented in the .hip files, called by api.hip
// grid sizes with a compiled FFT plan (radix 2, 3, 4, 8 stages)
#define GEV_FFT_SIZES(X) \
    X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

int fft_forward(gev_ctx *ctx, const double *real, double2 *kfield, int ncomp);
int fft_backward(gev_ctx *ctx, const double2 *kfield, double *real, int ncomp);
// the x and y passes only (the z pass fused with a solve follows), and the
// backward transform of a field whose z pass is already done (in place on
// kfield, which is scratch afterwards)
int fft_forward_xy(gev_ctx *ctx, const double *real, double2 *kfield, int ncomp, const Batch &bt = Batch{});
int fft_backward_yx(gev_ctx *ctx, double2 *kfield, double *real, int ncomp, const Batch &bt = Batch{});
// prepareFTsource fused into the first FFT pass (fftsolve.hip): Sij -> SijFT
// and source (+ phi, chi) -> scalarFT, after the x pass only
// A Batch with a BoxCycle array runs every box of the context's batch (box
// coefficients from the array, the by-value coefficients ignored).
int launch_x_prep_tensor(gev_ctx *c, double coeff, const Batch &bt = Batch{});
int launch_x_prep_scalar(gev_ctx *c, double bg, double coeff, double coeff2, double coeff3,
                         const Batch &bt = Batch{});
int fft_forward_y(gev_ctx *ctx, double2 *kfield, int ncomp, const Batch &bt = Batch{});  // the y pass alone, in place
int fft_forward_z(gev_ctx *ctx, double2 *kfield, int ncomp, const Batch &bt = Batch{});  // the z pass alone, in place
// z pass + solve + inverse z pass (fftsolve.hip); kstore / k0 optional
int launch_zsolve_poisson(gev_ctx *c, const double2 *in, double2 *out, double coeff, double modif,
                          double2 *kstore, double2 *k0, const Batch &bt = Batch{});
int launch_zsolve_tensor(gev_ctx *c, const double2 *in, double2 *out, double a2dtau, bool evolve,
                         double2 *kstore, double2 *k0, const Batch &bt = Batch{});
int launch_zsolve_vector(gev_ctx *c, const double2 *in, double2 *out, double coeff, double modif,
                         double2 *bift);

// initial conditions (ic.hip)
struct RealizeSetup {
    uint8_t *redo;               // N*N row marks (device)
    const float *sinc;           // N deconvolution factors (device)
    const double *sx, *sy, *sc;  // spline knots, ordinates, half second derivatives (device)
    int nsp;
    uint32_t seed;
    int ksphere, ignorekernel;
    double coeff;
};
int launch_prng_fill(gev_ctx *c, uint32_t seed, uint64_t pos0, int64_t n, uint32_t *dev_out);
int launch_realize(gev_ctx *c, const RealizeSetup &s);
int launch_init_particles(gev_ctx *c, const float *dev_tpl, int64_t numpart, int numtile);
int launch_ic_displace(gev_ctx *c, const double *field, double coeff);  // max -> red_u; refiles
int launch_ic_q(gev_ctx *c, const double *field, double coeff);         // max |q|^2 -> red_u
// binned (cross) spectrum of k fields (spectrum.hip); mode 0: components
// (comp1, comp2), 1: symmetric tensor, 2: sum over components
// doubles of power-spectrum scratch for numbins bins at grid N; the bins a
// context's own scratch holds
int64_t spectrum_scratch_doubles(int N, int numbins);
inline int spectrum_bins_cap(int N) { return N / 2 > 64 ? N / 2 : 64; }
// the IC staging area of a box of grid N: up to IC_STAGE_KNOTS spline knots
// (x, y and the coefficients), the N-entry sinc row, N^2 row marks and a
// crystal template of up to IC_STAGE_TEMPLATE particles, each 256-byte
// aligned; larger tables are allocated per call
constexpr int64_t IC_STAGE_KNOTS = 4096, IC_STAGE_TEMPLATE = 4096;
inline int64_t ic_stage_bytes(int N) {
    return 3 * (IC_STAGE_KNOTS * 8 + 256) + ((int64_t)N * 4 + 256) + ((int64_t)N * N + 256) +
           (IC_STAGE_TEMPLATE * 12 + 256);
}
int spectrum(gev_ctx *c, const double2 *f1, const double2 *f2, int ncomp, int mode, int comp1, int comp2,
             int numbins, int deconvolve, int ktype, double *kbin, double *power, double *kscatter,
             double *pscatter, int32_t *occupation);
int launch_put_stencil(gev_ctx *c, double *field, const double v[27]);  // v[(dz+1)*9+(dy+1)*3+dx+1]

int launch_poisson(gev_ctx *ctx, double coeff, double modif);
int launch_project_scalar(gev_ctx *ctx, int add);
int launch_evolve_vector(gev_ctx *ctx, double a2dtau);
int launch_project_scalar_evolve_vector(gev_ctx *ctx, double a2dtau);
int launch_project_vector(gev_ctx *ctx, double coeff, double modif);

int launch_prepare_scalar(gev_ctx *ctx, double bg, double c1, double c2, double c3);
int launch_prepare_tensor(gev_ctx *ctx, double coeff);
int launch_field_sum(gev_ctx *ctx, const double *f, int64_t count, double *out_host);
int launch_fill_zero(gev_ctx *ctx, void *ptr, size_t bytes);
int launch_convert(gev_ctx *ctx, const void *src, bool src32, void *dst, bool dst32, int64_t n);

enum DepositKind { DEP_T00 = 1, DEP_TIJ = 2, DEP_T00_TIJ = 3, DEP_T0I = 4, DEP_CIC = 5 };
int launch_deposit(gev_ctx *ctx, int kind, double a, int use_phi, double coeff);
int launch_deposit_batch(gev_ctx *ctx, bool gr, const Batch &bt);

// the main loop's spectrum outputs (api.hip): writeSpectra's delta spectrum
// and writePowerSpectrum with EXACT_OUTPUT_REDSHIFTS, per box and output
struct PkWriter {
    int num_pk = 0;
    size_t numbins = 0;
    double rescalek = 0., rescalep = 0.;
    std::vector<double> kb, pw, ks, ps, p2, pre, pre_z;
    std::vector<int32_t> oc, have;
    void init(const gev_ctx *c, const gev_sim_params *p, int num_pk, int numbins, double boxsize);
};
// output j at z_pk = zt from the box's current state into table j of the
// [num_pk][numbins] outputs
int write_spectrum(gev_ctx *c, const gev_sim_params *p, const gev_cycle_state *st, PkWriter &w, int j,
                   double zt, double *k, double *pk, double *sk, double *spk, int32_t *occ,
                   gev_pk_output_info *info);
// batches (api.hip): a box's array pointers at its slab; a batch's contexts
void assign_box_pointers(gev_ctx *c, char *base);
int create_ctx(const gev_box_desc *d, int nbox, gev_ctx **out);
int64_t slab_bytes(int N, int64_t cap, bool fp32, bool compact = false);

int launch_refile(gev_ctx *ctx);  // cells from positions, then CSR
int launch_particles_batch(gev_ctx *ctx, bool gr, const Batch &bt);
int launch_particle_update(gev_ctx *ctx, bool kick, int qm, double dtk, int nfk, const double pk[2],
                           bool drift, int pm, double dtd, int nfd, const double pd[2],
                           double *maxvel);

}  // namespace gev
