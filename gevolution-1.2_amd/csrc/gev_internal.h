// gev_internal.h -- context layout and helpers shared by the HIP sources of
// libgevpm.so.  Not part of the public ABI (include/gev_pm.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gev_pm.h"

#define GEV_EXPORT extern "C" __attribute__((visibility("default")))

// per-cycle background records buffered in pinned memory before one synchronisation
constexpr int REC_SLOTS = 1024;

namespace gev {

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

}  // namespace gev

struct gev_ctx {
    int N = 0;
    int P = 0;            // k-space row pitch (complex elements)
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

    // power-spectrum scratch (tables, workgroup histograms, bin totals) for
    // up to pk_bins bins
    double *pk_scratch = nullptr;
    int pk_bins = 0;

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
int fft_forward_xy(gev_ctx *ctx, const double *real, double2 *kfield, int ncomp);
int fft_backward_yx(gev_ctx *ctx, double2 *kfield, double *real, int ncomp);
// prepareFTsource fused into the first FFT pass (fftsolve.hip): Sij -> SijFT
// and source (+ phi, chi) -> scalarFT, after the x pass only
int launch_x_prep_tensor(gev_ctx *c, double coeff);
int launch_x_prep_scalar(gev_ctx *c, double bg, double coeff, double coeff2, double coeff3);
int fft_forward_y(gev_ctx *ctx, double2 *kfield, int ncomp);  // the y pass alone, in place
// z pass + solve + inverse z pass (fftsolve.hip); kstore / k0 optional
int launch_zsolve_poisson(gev_ctx *c, const double2 *in, double2 *out, double coeff, double modif,
                          double2 *kstore, double2 *k0);
int launch_zsolve_tensor(gev_ctx *c, const double2 *in, double2 *out, double a2dtau, bool evolve,
                         double2 *kstore, double2 *k0);
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

int launch_refile(gev_ctx *ctx);  // cells from positions, then CSR
int launch_particle_update(gev_ctx *ctx, bool kick, int qm, double dtk, int nfk, const double pk[2],
                           bool drift, int pm, double dtd, int nfd, const double pd[2],
                           double *maxvel);

}  // namespace gev
