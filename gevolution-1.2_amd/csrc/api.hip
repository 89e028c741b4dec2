// api.hip -- context lifecycle, data movement, the C ABI of include/gev_pm.h,
// the host-side cycle driver (main.cpp:372-879) and the -DBENCHMARK-style
// kernel timers.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "gev_internal.h"

namespace gev {
bool fft_size_supported(int N);

int fail(gev_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int check_hip(gev_ctx *ctx, hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    return fail(ctx, GEV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

static hipEvent_t take_event(gev_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void timer_begin(gev_ctx *ctx, const char *name) {
    if (!ctx->timing) return;
    TimerRec r;
    r.name = name;
    r.beg = take_event(ctx);
    r.end = take_event(ctx);
    if (!r.beg || !r.end) return;
    (void)hipEventRecord(r.beg, ctx->stream);
    ctx->pending.push_back(r);
}

void timer_end(gev_ctx *ctx) {
    if (!ctx->timing || ctx->pending.empty()) return;
    (void)hipEventRecord(ctx->pending.back().end, ctx->stream);
}

static void timers_collect(gev_ctx *ctx) {
    if (ctx->pending.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &r : ctx->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.beg, r.end);
        TimerAcc *a = nullptr;
        for (auto &x : ctx->acc)
            if (x.name == r.name) a = &x;
        if (!a) {
            ctx->acc.push_back(TimerAcc{r.name, 0, 0.0});
            a = &ctx->acc.back();
        }
        a->calls++;
        a->ms += ms;
        ctx->event_pool.push_back(r.beg);
        ctx->event_pool.push_back(r.end);
    }
    ctx->pending.clear();
}

// gridk2 / kshift tables exactly as gevolution.hpp:222-227 computes them
static void host_ktables(int N, std::vector<double> &g, std::vector<double2> &ks) {
    g.resize(N);
    ks.resize(N);
    for (int i = 0; i < N; i++) {
        g[i] = 2. * (double)N * sin(M_PI * (double)i / (double)N);
        double c = cos(M_PI * (double)i / (double)N), s = -sin(M_PI * (double)i / (double)N);
        ks[i] = make_double2(g[i] * c, g[i] * s);
        g[i] *= g[i];
    }
}

static void host_twiddles(int N, std::vector<double2> &tw) {
    tw.resize(N);
    for (int t = 0; t < N; t++) {
        long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)N;
        tw[t] = make_double2((double)cosl(ang), -(double)sinl(ang));
    }
}

struct Sizes {
    int64_t n3, nk, cap, nblk_scan, nred;
};

static Sizes sizes_of(int N, int64_t cap) {
    Sizes s;
    s.n3 = (int64_t)N * N * N;
    s.nk = (int64_t)kpitch(N) * N * N;
    s.cap = cap < 1 ? 1 : cap;
    s.nblk_scan = (s.n3 + 4095) / 4096;
    s.nred = 8 + (s.n3 + 4095) / 4096 + 64;
    return s;
}

int64_t slab_bytes(int N, int64_t cap, bool fp32, bool compact);

static void free_all(gev_ctx *c) {
    if (c->view) return;  // a view owns nothing
    if (c->slab) (void)hipFree(c->slab);
    c->slab = nullptr;
    if (c->host_pinned) (void)hipHostFree(c->host_pinned);
    if (c->rec_pinned) (void)hipHostFree(c->rec_pinned);
    c->host_pinned = c->rec_pinned = nullptr;
}

// ---------------------------------------------------------------- cycle

static double hconf(double a, double fourpiG, const gev_cosmology *c) {
    // background.hpp:137-140 with bg_ncdm = 0
    return hconf_of(a, fourpiG, c->Omega_cdm, c->Omega_b, c->Omega_Lambda, c->Omega_rad, c->Omega_fld, c->w0_fld,
                    c->wa_fld);
}

static void rk4bg(double &a, double fourpiG, const gev_cosmology *c, double dtau) {
    // background.hpp:167-177
    rk4bg_with(a, dtau, [&](double x) { return hconf(x, fourpiG, c); });
}

}  // namespace gev

using namespace gev;


GEV_EXPORT int64_t gev_ctx_footprint_bytes(const gev_box_desc *d) {
    if (!d || d->ngrid < 2 || d->npart < 0) return -1;
    if (d->form != GEV_FORM_FP64 && d->form != GEV_FORM_FP32) return -1;
    if ((d->flags & ~GEV_BOX_COMPACT) != 0) return -1;
    return slab_bytes(d->ngrid, d->npart, d->form == GEV_FORM_FP32, (d->flags & GEV_BOX_COMPACT) != 0);
}

namespace gev {

// The device arrays of one box, each at an offset of one slab: a box's whole
// state is one contiguous range, and the boxes of a batch lie one slab
// stride apart (the offsets are those of box 0).  Offsets are 4 KiB aligned
// (k rows are then 128-byte aligned too).
struct SlabLayout {
    std::vector<std::pair<void **, int64_t>> arrays;
    int64_t bytes = 0;
    void add(void **p, int64_t n) { arrays.push_back({p, n}); }
    int64_t size() const {
        int64_t off = 0;
        for (auto &a : arrays) off = (off + a.second + 4095) / 4096 * 4096;
        return (off + (2 << 20) - 1) / (2 << 20) * (2 << 20);  // slabs start on 2 MiB
    }
    void assign(char *base) const {
        int64_t off = 0;
        for (auto &a : arrays) {
            *a.first = base + off;
            off = (off + a.second + 4095) / 4096 * 4096;
        }
    }
};

// a compact box's second particle set is its only one
static void alias_compact(gev_ctx *c) {
    if (!c->compact) return;
    c->pos2 = c->pos;
    c->q2 = c->q;
    c->id2 = c->id;
}

static void slab_layout(gev_ctx *c, const Sizes &s, SlabLayout &L) {
    const int64_t rb = rbytes(c), cb = cbytes(c);
    L.add((void **)&c->phi, s.n3 * rb);
    L.add((void **)&c->chi, s.n3 * rb);
    L.add((void **)&c->source, s.n3 * rb);
    L.add((void **)&c->Sij, 6 * s.n3 * rb);
    L.add((void **)&c->Bi, 3 * s.n3 * rb);
    L.add((void **)&c->scalarFT, s.nk * cb);
    L.add((void **)&c->SijFT, 6 * s.nk * cb);
    L.add((void **)&c->BiFT, 3 * s.nk * cb);
    L.add((void **)&c->fft_tmp, s.nk * cb);
    L.add((void **)&c->twiddle, c->N * 16);
    if (c->fp32) L.add((void **)&c->twiddle_f, c->N * 8);
    L.add((void **)&c->gridk2, c->N * 8);
    L.add((void **)&c->kshift, c->N * 16);
    L.add((void **)&c->pos, 3 * s.cap * 8);
    L.add((void **)&c->q, 3 * s.cap * 8);
    L.add((void **)&c->id, s.cap * 8);
    if (!c->compact) {
        L.add((void **)&c->pos2, 3 * s.cap * 8);
        L.add((void **)&c->q2, 3 * s.cap * 8);
        L.add((void **)&c->id2, s.cap * 8);
    }
    L.add((void **)&c->cell, s.cap * 4);
    L.add((void **)&c->perm, s.cap * 4);
    L.add((void **)&c->count, s.n3 * 4);
    L.add((void **)&c->start, (s.n3 + 1) * 4);
    L.add((void **)&c->scan_part, s.nblk_scan * 4);
    L.add((void **)&c->sort_segs, (s.n3 / 256 + 1) * 4);
    L.add((void **)&c->sort_big, sort_big_cap(s.cap) * 4);
    L.add((void **)&c->sort_n, 8);
    L.add((void **)&c->red_d, s.nred * 8);
    L.add((void **)&c->red_u, 64);
    c->pk_bins = spectrum_bins_cap(c->N);
    L.add((void **)&c->pk_scratch, spectrum_scratch_doubles(c->N, c->pk_bins) * 8);
    L.add((void **)&c->ic_stage, ic_stage_bytes(c->N));
}

// points c's device arrays at the slab starting at base (a box of a batch)
void assign_box_pointers(gev_ctx *c, char *base) {
    SlabLayout L;
    slab_layout(c, sizes_of(c->N, c->cap), L);
    L.assign(base);
    alias_compact(c);
}

// bytes of one box's slab
int64_t slab_bytes(int N, int64_t cap, bool fp32, bool compact) {
    gev_ctx c;
    c.N = N;
    c.fp32 = fp32;
    c.compact = compact;
    SlabLayout L;
    slab_layout(&c, sizes_of(N, cap), L);
    return L.size();
}

// Creates a context holding nbox boxes of the same shape, one slab each; the
// context's pointers are box 0's.
int create_ctx(const gev_box_desc *d, int nbox, gev_ctx **out) {
    if (!out) return GEV_EINVAL;
    *out = nullptr;
    if (!d || d->npart < 0 || d->npart > 0xFFFFFFFFll || nbox < 1) return GEV_EINVAL;
    if ((d->flags & ~GEV_BOX_COMPACT) != 0 || ((d->flags & GEV_BOX_COMPACT) && nbox > 1)) return GEV_EINVAL;
    if (d->form != GEV_FORM_FP64 && d->form != GEV_FORM_FP32) return GEV_EINVAL;
    if (!fft_size_supported(d->ngrid)) return GEV_EINVAL;
    gev_ctx *c = new (std::nothrow) gev_ctx();
    if (!c) return GEV_ENOMEM;
    c->N = d->ngrid;
    c->fp32 = d->form == GEV_FORM_FP32;
    c->compact = (d->flags & GEV_BOX_COMPACT) != 0;
    c->P = kpitch(c->N);
    c->cap = d->npart;
    (void)hipGetDevice(&c->device);
    if (d->stream) {
        c->stream = (hipStream_t)d->stream;
    } else {
        if (hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return GEV_EHIP;
        }
        c->own_stream = true;
    }
    Sizes s = sizes_of(c->N, c->cap);
    SlabLayout L;
    slab_layout(c, s, L);
    c->nbox = nbox;
    c->stride = L.size();
    bool ok = hipMalloc(&c->slab, (size_t)c->stride * nbox) == hipSuccess;
    if (!ok) c->slab = nullptr;
    if (ok) {
        L.assign((char *)c->slab);
        alias_compact(c);
    }
    if (ok && hipHostMalloc((void **)&c->host_pinned, 64) != hipSuccess) ok = false;
    if (ok && hipHostMalloc((void **)&c->rec_pinned, sizeof(double) * 2 * REC_SLOTS) != hipSuccess) ok = false;
    if (!ok) {
        free_all(c);
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return GEV_ENOMEM;
    }
    c->scan_nblk = s.nblk_scan;
    std::vector<double> g;
    std::vector<double2> ks, tw;
    host_ktables(c->N, g, ks);
    host_twiddles(c->N, tw);
    std::vector<float2> twf;
    if (c->fp32) {
        twf.resize(c->N);
        for (int t = 0; t < c->N; t++) twf[t] = make_float2((float)tw[t].x, (float)tw[t].y);
    }
    // every field starts at zero, like a fresh LATfield2 Field, and the filing
    // is that of no particles: the whole slab of every box is cleared
    int rc = launch_fill_zero(c, c->slab, (size_t)c->stride * nbox);
    for (int b = 0; b < nbox && !rc; b++) {
        rc = rc ? rc : check_hip(c, hipMemcpy(box_ptr(c->gridk2, c->stride, b), g.data(), c->N * 8,
                                              hipMemcpyHostToDevice), "k tables");
        rc = rc ? rc : check_hip(c, hipMemcpy(box_ptr(c->kshift, c->stride, b), ks.data(), c->N * 16,
                                              hipMemcpyHostToDevice), "k tables");
        rc = rc ? rc : check_hip(c, hipMemcpy(box_ptr(c->twiddle, c->stride, b), tw.data(), c->N * 16,
                                              hipMemcpyHostToDevice), "twiddles");
        if (c->fp32)
            rc = rc ? rc : check_hip(c, hipMemcpy(box_ptr(c->twiddle_f, c->stride, b), twf.data(), c->N * 8,
                                                  hipMemcpyHostToDevice), "twiddles");
    }
    if (!rc) c->count_clean = true;
    if (!rc) rc = check_hip(c, hipStreamSynchronize(c->stream), "create sync");
    if (rc) {
        free_all(c);
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

}  // namespace gev

GEV_EXPORT int gev_ctx_create(const gev_box_desc *d, gev_ctx **out) { return create_ctx(d, 1, out); }

GEV_EXPORT int gev_ctx_destroy(gev_ctx *c) {
    if (!c) return GEV_EINVAL;
    if (c->view) return fail(c, GEV_ESTATE, "a box of a batch is freed with its batch (gev_batch_destroy)");
    (void)hipStreamSynchronize(c->stream);
    for (auto &r : c->pending) {
        (void)hipEventDestroy(r.beg);
        (void)hipEventDestroy(r.end);
    }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    free_all(c);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

GEV_EXPORT const char *gev_last_error(const gev_ctx *c) { return c ? c->err.c_str() : "null context"; }

GEV_EXPORT int gev_ctx_synchronize(gev_ctx *c) {
    if (!c) return GEV_EINVAL;
    return check_hip(c, hipStreamSynchronize(c->stream), "synchronize");
}

// ---------------------------------------------------------------- data

GEV_EXPORT int gev_upload_particles(gev_ctx *c, int64_t n, const double *pos, const double *q,
                                    const int64_t *id, double mass) {
    if (!c) return GEV_EINVAL;
    if (n < 0 || n > c->cap) return fail(c, GEV_EINVAL, "particle count exceeds capacity");
    if (n > 0 && (!pos || !q || !id)) return fail(c, GEV_EINVAL, "null particle buffer");
    const double dx = 1.0 / (double)c->N;
    std::vector<double> soa((size_t)6 * (size_t)(n > 0 ? n : 1));
    for (int64_t p = 0; p < n; p++)
        for (int i = 0; i < 3; i++) {
            double x = pos[3 * p + i];
            if (!(x >= 0.0 && x < 1.0) || (int)floor(x / dx) >= c->N)
                return fail(c, GEV_EINVAL, "particle position outside [0,1)");
            soa[(size_t)i * n + p] = x;
            soa[(size_t)(3 + i) * n + p] = q[3 * p + i];
        }
    int rc = check_hip(c, hipStreamSynchronize(c->stream), "upload sync");
    for (int i = 0; i < 3 && !rc && n > 0; i++) {
        rc = check_hip(c, hipMemcpy(c->pos + (size_t)i * c->cap, soa.data() + (size_t)i * n, n * 8,
                                    hipMemcpyHostToDevice), "upload pos");
        if (!rc)
            rc = check_hip(c, hipMemcpy(c->q + (size_t)i * c->cap, soa.data() + (size_t)(3 + i) * n,
                                        n * 8, hipMemcpyHostToDevice), "upload q");
    }
    if (!rc && n > 0) rc = check_hip(c, hipMemcpy(c->id, id, n * 8, hipMemcpyHostToDevice), "upload id");
    if (rc) return rc;
    c->n = n;
    c->pmass = mass;
    return launch_refile(c);
}

GEV_EXPORT int gev_download_particles(gev_ctx *c, double *pos, double *q, int64_t *id) {
    if (!c) return GEV_EINVAL;
    const int64_t n = c->n;
    if (n == 0) return 0;
    std::vector<double> soa((size_t)6 * n);
    int rc = check_hip(c, hipStreamSynchronize(c->stream), "download sync");
    for (int i = 0; i < 3 && !rc; i++) {
        rc = check_hip(c, hipMemcpy(soa.data() + (size_t)i * n, c->pos + (size_t)i * c->cap, n * 8,
                                    hipMemcpyDeviceToHost), "download pos");
        if (!rc)
            rc = check_hip(c, hipMemcpy(soa.data() + (size_t)(3 + i) * n, c->q + (size_t)i * c->cap,
                                        n * 8, hipMemcpyDeviceToHost), "download q");
    }
    if (!rc && id) rc = check_hip(c, hipMemcpy(id, c->id, n * 8, hipMemcpyDeviceToHost), "download id");
    if (rc) return rc;
    for (int64_t p = 0; p < n; p++)
        for (int i = 0; i < 3; i++) {
            if (pos) pos[3 * p + i] = soa[(size_t)i * n + p];
            if (q) q[3 * p + i] = soa[(size_t)(3 + i) * n + p];
        }
    return 0;
}

GEV_EXPORT int64_t gev_particle_count(const gev_ctx *c) { return c ? c->n : -1; }

GEV_EXPORT int gev_download_filing(gev_ctx *c, uint32_t *cell_count, uint32_t *particle_cell) {
    if (!c) return GEV_EINVAL;
    int rc = check_hip(c, hipStreamSynchronize(c->stream), "filing sync");
    if (!rc && cell_count) {
        // per-cell counts are the differences of the CSR offsets
        std::vector<uint32_t> st((size_t)n3(c) + 1);
        rc = check_hip(c, hipMemcpy(st.data(), c->start, (n3(c) + 1) * 4, hipMemcpyDeviceToHost), "counts");
        for (int64_t i = 0; !rc && i < n3(c); i++) cell_count[i] = st[i + 1] - st[i];
    }
    if (!rc && particle_cell && c->n > 0)
        rc = check_hip(c, hipMemcpy(particle_cell, c->cell, c->n * 4, hipMemcpyDeviceToHost), "cells");
    return rc;
}

static int field_desc(gev_ctx *c, gev_field f, void **ptr, int *ncomp, bool *complex_) {
    *complex_ = false;
    switch (f) {
    case GEV_PHI: *ptr = c->phi; *ncomp = 1; return 0;
    case GEV_CHI: *ptr = c->chi; *ncomp = 1; return 0;
    case GEV_SOURCE: *ptr = c->source; *ncomp = 1; return 0;
    case GEV_SIJ: *ptr = c->Sij; *ncomp = 6; return 0;
    case GEV_BI: *ptr = c->Bi; *ncomp = 3; return 0;
    case GEV_SCALARFT: *ptr = c->scalarFT; *ncomp = 1; *complex_ = true; return 0;
    case GEV_SIJFT: *ptr = c->SijFT; *ncomp = 6; *complex_ = true; return 0;
    case GEV_BIFT: *ptr = c->BiFT; *ncomp = 3; *complex_ = true; return 0;
    }
    return fail(c, GEV_EINVAL, "unknown field");
}

GEV_EXPORT int gev_upload_field(gev_ctx *c, gev_field f, const double *host) {
    if (!c || !host) return GEV_EINVAL;
    void *p;
    int nc;
    bool cx;
    int rc = field_desc(c, f, &p, &nc, &cx);
    if (rc) return rc;
    rc = check_hip(c, hipStreamSynchronize(c->stream), "upload sync");
    if (rc) return rc;
    const int H = c->N / 2 + 1;
    if (c->fp32) {
        // the host array is FP64 in both forms: rounded to float here
        if (!cx) {
            if ((int)f <= GEV_BI) c->zeroed[f] = false;
            std::vector<float> t((size_t)nc * n3(c));
            for (size_t i = 0; i < t.size(); i++) t[i] = (float)host[i];
            return check_hip(c, hipMemcpy(p, t.data(), t.size() * 4, hipMemcpyHostToDevice), "upload field");
        }
        std::vector<float2> t((size_t)nc * c->N * c->N * H);
        for (size_t i = 0; i < t.size(); i++) t[i] = make_float2((float)host[2 * i], (float)host[2 * i + 1]);
        return check_hip(c, hipMemcpy2D(p, (size_t)c->P * 8, t.data(), (size_t)H * 8, (size_t)H * 8,
                                        (size_t)nc * c->N * c->N, hipMemcpyHostToDevice),
                         "upload k field");
    }
    if (!cx) {
        if ((int)f <= GEV_BI) c->zeroed[f] = false;
        return check_hip(c, hipMemcpy(p, host, (size_t)nc * n3(c) * 8, hipMemcpyHostToDevice), "upload field");
    }
    return check_hip(c, hipMemcpy2D(p, (size_t)c->P * 16, host, (size_t)H * 16, (size_t)H * 16,
                                    (size_t)nc * c->N * c->N, hipMemcpyHostToDevice),
                     "upload k field");
}

GEV_EXPORT int gev_download_field(gev_ctx *c, gev_field f, double *host) {
    if (!c || !host) return GEV_EINVAL;
    void *p;
    int nc;
    bool cx;
    int rc = field_desc(c, f, &p, &nc, &cx);
    if (rc) return rc;
    rc = check_hip(c, hipStreamSynchronize(c->stream), "download sync");
    if (rc) return rc;
    const int H = c->N / 2 + 1;
    if (c->fp32) {
        if (!cx) {
            std::vector<float> t((size_t)nc * n3(c));
            rc = check_hip(c, hipMemcpy(t.data(), p, t.size() * 4, hipMemcpyDeviceToHost), "download field");
            for (size_t i = 0; !rc && i < t.size(); i++) host[i] = t[i];
            return rc;
        }
        std::vector<float2> t((size_t)nc * c->N * c->N * H);
        rc = check_hip(c, hipMemcpy2D(t.data(), (size_t)H * 8, p, (size_t)c->P * 8, (size_t)H * 8,
                                      (size_t)nc * c->N * c->N, hipMemcpyDeviceToHost),
                       "download k field");
        for (size_t i = 0; !rc && i < t.size(); i++) {
            host[2 * i] = t[i].x;
            host[2 * i + 1] = t[i].y;
        }
        return rc;
    }
    if (!cx)
        return check_hip(c, hipMemcpy(host, p, (size_t)nc * n3(c) * 8, hipMemcpyDeviceToHost), "download field");
    return check_hip(c, hipMemcpy2D(host, (size_t)H * 16, p, (size_t)c->P * 16, (size_t)H * 16,
                                    (size_t)nc * c->N * c->N, hipMemcpyDeviceToHost),
                     "download k field");
}

// The state of one context carried into another of the same grid size and
// device, converting between the forms: particles (FP64 in both), the real
// and k-space fields including the carried BiFT, the species mass, and the
// filing (rebuilt from the positions).  The initial-condition generator needs
// an FP64 context; this brings its result into an FP32 one.
GEV_EXPORT int gev_copy_state(gev_ctx *src, gev_ctx *dst) {
    if (!src || !dst) return GEV_EINVAL;
    if (src == dst) return fail(dst, GEV_EINVAL, "source and destination are the same context");
    if (src->N != dst->N) return fail(dst, GEV_EINVAL, "contexts differ in grid size");
    if (src->device != dst->device) return fail(dst, GEV_EINVAL, "contexts live on different devices");
    if (src->n > dst->cap) return fail(dst, GEV_EINVAL, "particle count exceeds capacity");
    int rc = check_hip(dst, hipStreamSynchronize(src->stream), "copy state: source sync");
    if (!rc) rc = check_hip(dst, hipStreamSynchronize(dst->stream), "copy state: destination sync");
    if (rc) return rc;
    const int64_t n = src->n;
    for (int i = 0; i < 3 && !rc && n > 0; i++) {
        rc = check_hip(dst, hipMemcpyAsync(dst->pos + (size_t)i * dst->cap, src->pos + (size_t)i * src->cap, n * 8,
                                           hipMemcpyDeviceToDevice, dst->stream), "copy pos");
        if (!rc)
            rc = check_hip(dst, hipMemcpyAsync(dst->q + (size_t)i * dst->cap, src->q + (size_t)i * src->cap,
                                               n * 8, hipMemcpyDeviceToDevice, dst->stream), "copy q");
    }
    if (!rc && n > 0)
        rc = check_hip(dst, hipMemcpyAsync(dst->id, src->id, n * 8, hipMemcpyDeviceToDevice, dst->stream), "copy id");
    const struct {
        void *s, *d;
        int64_t reals;
    } f[] = {{src->phi, dst->phi, n3(src)},          {src->chi, dst->chi, n3(src)},
             {src->source, dst->source, n3(src)},    {src->Sij, dst->Sij, 6 * n3(src)},
             {src->Bi, dst->Bi, 3 * n3(src)},        {src->scalarFT, dst->scalarFT, 2 * nk(src)},
             {src->SijFT, dst->SijFT, 12 * nk(src)}, {src->BiFT, dst->BiFT, 6 * nk(src)}};
    for (const auto &e : f)
        if (!rc) rc = launch_convert(dst, e.s, src->fp32, e.d, dst->fp32, e.reals);
    if (rc) return rc;
    for (int i = 0; i < 5; i++) dst->zeroed[i] = src->zeroed[i];
    dst->n = n;
    dst->pmass = src->pmass;
    rc = launch_refile(dst);
    if (!rc) rc = check_hip(dst, hipStreamSynchronize(dst->stream), "copy state sync");
    return rc;
}

// ---------------------------------------------------------------- hot path

GEV_EXPORT int gev_projection_init(gev_ctx *c, gev_field f) {
    if (!c) return GEV_EINVAL;
    void *p;
    int nc;
    bool cx;
    int rc = field_desc(c, f, &p, &nc, &cx);
    if (rc) return rc;
    if (cx) return fail(c, GEV_EINVAL, "projection_init needs a real field");
    rc = launch_fill_zero(c, p, (size_t)nc * n3(c) * rbytes(c));
    if (!rc) c->zeroed[f] = true;
    return rc;
}

GEV_EXPORT int gev_projection_T00_project(gev_ctx *c, double a, int use_phi, double coeff) {
    if (!c) return GEV_EINVAL;
    return launch_deposit(c, DEP_T00, a, use_phi, coeff);
}

GEV_EXPORT int gev_projection_Tij_project(gev_ctx *c, double a, int use_phi, double coeff) {
    if (!c) return GEV_EINVAL;
    return launch_deposit(c, DEP_TIJ, a, use_phi, coeff);
}

GEV_EXPORT int gev_projection_T00_Tij_project(gev_ctx *c, double a, double coeff) {
    if (!c) return GEV_EINVAL;
    return launch_deposit(c, DEP_T00_TIJ, a, 1, coeff);
}

GEV_EXPORT int gev_projection_T0i_project(gev_ctx *c, int use_phi, double coeff) {
    if (!c) return GEV_EINVAL;
    return launch_deposit(c, DEP_T0I, 1.0, use_phi, coeff);
}

GEV_EXPORT int gev_scalar_projection_cic_project(gev_ctx *c, double coeff) {
    if (!c) return GEV_EINVAL;
    return launch_deposit(c, DEP_CIC, 1.0, 0, coeff);
}

GEV_EXPORT int gev_field_sum(gev_ctx *c, gev_field f, double *sum) {
    if (!c || !sum) return GEV_EINVAL;
    void *p;
    int nc;
    bool cx;
    int rc = field_desc(c, f, &p, &nc, &cx);
    if (rc) return rc;
    if (cx) return fail(c, GEV_EINVAL, "field_sum needs a real field");
    return launch_field_sum(c, (const double *)p, n3(c), sum);
}

GEV_EXPORT int gev_prepare_ft_source_scalar(gev_ctx *c, double bgmodel, double coeff,
                                            double coeff2, double coeff3) {
    if (!c) return GEV_EINVAL;
    c->zeroed[GEV_SOURCE] = false;
    return launch_prepare_scalar(c, bgmodel, coeff, coeff2, coeff3);
}

GEV_EXPORT int gev_prepare_ft_source_tensor(gev_ctx *c, double coeff) {
    if (!c) return GEV_EINVAL;
    c->zeroed[GEV_SIJ] = false;
    return launch_prepare_tensor(c, coeff);
}

GEV_EXPORT int gev_fft_execute(gev_ctx *c, gev_plan plan, int direction) {
    if (!c) return GEV_EINVAL;
    double *real;
    double2 *k;
    int nc;
    gev_field rf;
    switch (plan) {
    case GEV_PLAN_SOURCE: real = c->source; k = c->scalarFT; nc = 1; rf = GEV_SOURCE; break;
    case GEV_PLAN_PHI: real = c->phi; k = c->scalarFT; nc = 1; rf = GEV_PHI; break;
    case GEV_PLAN_CHI: real = c->chi; k = c->scalarFT; nc = 1; rf = GEV_CHI; break;
    case GEV_PLAN_SIJ: real = c->Sij; k = c->SijFT; nc = 6; rf = GEV_SIJ; break;
    case GEV_PLAN_BI: real = c->Bi; k = c->BiFT; nc = 3; rf = GEV_BI; break;
    default: return fail(c, GEV_EINVAL, "unknown FFT plan");
    }
    if (direction == GEV_FFT_FORWARD) return fft_forward(c, real, k, nc);
    if (direction == GEV_FFT_BACKWARD) {
        c->zeroed[rf] = false;
        return fft_backward(c, k, real, nc);
    }
    return fail(c, GEV_EINVAL, "unknown FFT direction");
}

GEV_EXPORT int gev_solve_modified_poisson_ft(gev_ctx *c, double coeff, double modif) {
    if (!c) return GEV_EINVAL;
    return launch_poisson(c, coeff, modif);
}

GEV_EXPORT int gev_project_ft_scalar(gev_ctx *c, int add) {
    if (!c) return GEV_EINVAL;
    return launch_project_scalar(c, add);
}

GEV_EXPORT int gev_evolve_ft_vector(gev_ctx *c, double a2dtau) {
    if (!c) return GEV_EINVAL;
    return launch_evolve_vector(c, a2dtau);
}

GEV_EXPORT int gev_project_ft_scalar_evolve_vector(gev_ctx *c, double a2dtau) {
    if (!c) return GEV_EINVAL;
    return launch_project_scalar_evolve_vector(c, a2dtau);
}

GEV_EXPORT int gev_project_ft_vector(gev_ctx *c, double coeff, double modif) {
    if (!c) return GEV_EINVAL;
    return launch_project_vector(c, coeff, modif);
}

static bool valid_params(const double *p) { return p && p[0] != 0.0 && p[1] != 0.0; }

GEV_EXPORT int gev_update_vel(gev_ctx *c, int method, double dtau, int nfield,
                              const double params[2], double *maxvel) {
    if (!c) return GEV_EINVAL;
    if (method != GEV_Q_GR && method != GEV_Q_NEWTON) return fail(c, GEV_EINVAL, "unknown kick");
    if (!valid_params(params)) return fail(c, GEV_EINVAL, "params must be {a, a^2 N}");
    return launch_particle_update(c, true, method, dtau, nfield, params, false, 0, 0.0, 0, nullptr,
                                  maxvel);
}

GEV_EXPORT int gev_move_particles(gev_ctx *c, int method, double dtau, int nfield,
                                  const double params[2]) {
    if (!c) return GEV_EINVAL;
    if (method != GEV_POS_GR && method != GEV_POS_NEWTON) return fail(c, GEV_EINVAL, "unknown drift");
    if (!valid_params(params)) return fail(c, GEV_EINVAL, "params must be {a, a^2 N}");
    return launch_particle_update(c, false, 0, 0.0, 0, nullptr, true, method, dtau, nfield, params,
                                  nullptr);
}

GEV_EXPORT int gev_kick_drift(gev_ctx *c, int qm, double dtk, int nfk, const double pk[2], int pm,
                              double dtd, int nfd, const double pd[2], double *maxvel) {
    if (!c) return GEV_EINVAL;
    if ((qm != GEV_Q_GR && qm != GEV_Q_NEWTON) || (pm != GEV_POS_GR && pm != GEV_POS_NEWTON))
        return fail(c, GEV_EINVAL, "unknown kick/drift method");
    if (!valid_params(pk) || !valid_params(pd)) return fail(c, GEV_EINVAL, "params must be {a, a^2 N}");
    return launch_particle_update(c, true, qm, dtk, nfk, pk, true, pm, dtd, nfd, pd, maxvel);
}

// ---------------------------------------------------------------- driver

GEV_EXPORT double gev_hconf(double a, double fourpiG, const gev_cosmology *c) {
    return hconf(a, fourpiG, c);
}

GEV_EXPORT void gev_rungekutta4bg(double *a, double fourpiG, const gev_cosmology *c, double dtau) {
    rk4bg(*a, fourpiG, c, dtau);
}

GEV_EXPORT double gev_next_dtau(const gev_sim_params *p, int ngrid, double a) {
    // main.cpp:292-295 / 869-872
    const double dx = 1.0 / (double)ngrid;
    if (p->Cf * dx < p->steplimit / hconf(a, p->fourpiG, &p->cosmo)) return p->Cf * dx;
    return p->steplimit / hconf(a, p->fourpiG, &p->cosmo);
}

// The field stage of one cycle of main.cpp:372-879 (one CDM species):
// projections, both prepareFTsource stencils, the FFTs and k-space solves
// (main.cpp:372-598).  Host scalars are read, not advanced; fields and
// particles never leave the device.  T00hom and phi(k=0) are left on the
// device.
// rec: when not null, a pinned slot that receives (T00hom * N^3, phi(k=0)) of
// this cycle, copied asynchronously at the point where main.cpp:520-537
// records the background line
static int cycle_fields(gev_ctx *c, const gev_sim_params *p, const gev_cycle_state *st, double *rec) {
    const int N = c->N;
    const double dx = 1.0 / (double)N;
    const double fourpiG = p->fourpiG;
    const gev_cosmology *cs = &p->cosmo;
    double a = st->a;
    int rc;
    // phi(k=0) and chi(k=0) of the z solves (FP64): element 0 of scalarFT in
    // the FP64 form, FP64 slots of the reduction scratch in the FP32 form
    double2 *const phik0 = c->fp32 ? reinterpret_cast<double2 *>(c->red_d + 2) : c->scalarFT;
    double2 *const chik0 = c->fp32 ? reinterpret_cast<double2 *>(c->red_d + 4) : c->scalarFT;

    // projection_init(source), projection_init(Sij) + T00 and Tij deposits
    // (main.cpp:378-450): the fused deposit stores every cell, so the
    // zero fill is implied rather than executed.
    c->zeroed[GEV_SOURCE] = true;
    c->zeroed[GEV_SIJ] = true;
    if (p->gr_flag > 0) {
        rc = launch_deposit(c, DEP_T00_TIJ, a, 1, 1.0);
    } else {
        rc = launch_deposit(c, DEP_CIC, 1.0, 0, 1.0);
        if (!rc) rc = launch_deposit(c, DEP_TIJ, a, 1, 1.0);
    }
    if (rc) return rc;
    if (p->vector_flag == 1) {
        c->zeroed[GEV_BI] = true;
        rc = launch_deposit(c, DEP_T0I, 1.0, 1, 1.0);
        if (rc) return rc;
    }

    if (p->gr_flag > 0) {
        rc = launch_field_sum(c, c->source, n3(c), nullptr);  // T00hom * N^3 -> red_d[0]
        if (rc) return rc;
        if (st->dtau_old > 0.) {
            // prepareFTsource inside the first FFT pass, solveModifiedPoissonFT
            // inside the z passes (main.cpp:472-488); only phi(k=0) of scalarFT is kept
            const double H = hconf(a, fourpiG, cs);
            rc = launch_x_prep_scalar(c, cs->Omega_cdm + cs->Omega_b + 0.0, 3. * H * dx * dx / st->dtau_old,
                                      fourpiG * dx * dx / a, 3. * H * H * dx * dx);
            if (!rc) rc = fft_forward_y(c, c->scalarFT, 1);
            if (!rc)
                rc = launch_zsolve_poisson(c, c->scalarFT, c->fft_tmp, 1. / (dx * dx),
                                           3. * hconf(a, fourpiG, cs) / st->dtau_old, nullptr, phik0);
            if (!rc) rc = fft_backward_yx(c, c->fft_tmp, c->phi, 1);
            if (rc) return rc;
        }
    } else {
        rc = fft_forward_xy(c, c->source, c->scalarFT, 1);
        if (!rc) rc = launch_zsolve_poisson(c, c->scalarFT, c->fft_tmp, fourpiG / a, 0., nullptr, phik0);
        if (!rc) rc = fft_backward_yx(c, c->fft_tmp, c->phi, 1);
        if (rc) return rc;
    }
    // phi(k=0) for the background record (main.cpp:533).  Without a solve in
    // this cycle it is whatever scalarFT holds at k = 0 (widened in the FP32 form).
    if (c->fp32 && p->gr_flag > 0 && !(st->dtau_old > 0.)) {
        rc = launch_convert(c, c->scalarFT, true, phik0, false, 2);
        if (rc) return rc;
    }
    rc = check_hip(c, hipMemcpyAsync(c->red_d + 1, phik0, sizeof(double),
                                     hipMemcpyDeviceToDevice, c->stream), "phi_k0");
    if (rc) return rc;
    if (rec) {
        rc = check_hip(c, hipMemcpyAsync(rec, c->red_d, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                       "background record");
        if (rc) return rc;
    }

    // prepareFTsource(Sij) inside the first FFT pass, projectFTscalar -> chi
    // and evolveFTvector -> BiFT inside the z passes (main.cpp:539-593).  SijFT is scratch
    // after the solve: the z-inverted chi goes to its component 0 and B to
    // components 1-3; scalarFT keeps only chi(k=0); BiFT, the carried state,
    // is written in full.
    double2 *const zout_chi = c->SijFT, *const zout_b = kcomp(c, c->SijFT, 1);
    rc = launch_x_prep_tensor(c, 2. * fourpiG * dx * dx / a);
    if (!rc) rc = fft_forward_y(c, c->SijFT, 6);
    if (rc) return rc;
    if (p->vector_flag == 1) {
        rc = launch_zsolve_tensor(c, c->SijFT, zout_chi, 0., false, nullptr, chik0);
        if (!rc) rc = fft_backward_yx(c, zout_chi, c->chi, 1);
        if (!rc) rc = fft_forward_xy(c, c->Bi, c->BiFT, 3);
        if (!rc) rc = launch_zsolve_vector(c, c->BiFT, zout_b, fourpiG * dx * dx, 0., c->BiFT);
    } else {
        rc = launch_zsolve_tensor(c, c->SijFT, zout_chi, a * a * st->dtau_old, true, nullptr, chik0);
        if (!rc) rc = fft_backward_yx(c, zout_chi, c->chi, 1);
    }
    if (rc) return rc;
    // scalarFT ends the cycle holding chi(k=0) in both forms
    if (c->fp32 && (rc = launch_convert(c, chik0, false, c->scalarFT, true, 2))) return rc;
    if (p->gr_flag > 0) {
        rc = fft_backward_yx(c, zout_b, c->Bi, 3);
        if (rc) return rc;
    }
    c->zeroed[GEV_PHI] = c->zeroed[GEV_CHI] = c->zeroed[GEV_BI] = false;
    return 0;
}

// The particle stage (main.cpp:766-874): kick with {a, a^2 N}, half RK4 step,
// drift with the advanced a, half RK4 step (one fused particle pass), then
// the time-step update.  maxvel stays on the device.
static int cycle_particles(gev_ctx *c, const gev_sim_params *p, gev_cycle_state *st) {
    const int N = c->N;
    const double fourpiG = p->fourpiG;
    const gev_cosmology *cs = &p->cosmo;
    double a = st->a;
    int rc;
    double pk[2] = {a, a * a * N};
    const double dtk = (st->dtau + st->dtau_old) / 2.;
    const int nfk = p->gr_flag > 0 ? (1. / a < p->z_relax + 1. ? 3 : 2) : 1;
    rk4bg(a, fourpiG, cs, 0.5 * st->dtau);
    double pd[2] = {a, a * a * N};
    const int nfd = p->gr_flag > 0 ? (1. / a < p->z_relax + 1. ? 3 : 0) : 0;
    rc = launch_particle_update(c, true, p->gr_flag > 0 ? 0 : 1, dtk, nfk, pk, true,
                                p->gr_flag > 0 ? 0 : 1, st->dtau, nfd, pd, nullptr);
    if (rc) return rc;
    rk4bg(a, fourpiG, cs, 0.5 * st->dtau);

    st->tau += st->dtau;
    st->dtau_old = st->dtau;
    st->a = a;
    st->dtau = gev_next_dtau(p, N, a);
    st->cycle++;
    return 0;
}

// One cycle of main.cpp:372-879 (no outputs).  Host scalars are advanced
// exactly as the reference does.
static int run_cycle(gev_ctx *c, const gev_sim_params *p, gev_cycle_state *st, double *rec) {
    int rc = cycle_fields(c, p, st, rec);
    return rc ? rc : cycle_particles(c, p, st);
}

// T00hom, phi(k=0) and maxvel of the last cycle into *st (synchronous)
static int read_cycle_scalars(gev_ctx *c, const gev_sim_params *p, gev_cycle_state *st) {
    double h[2];
    unsigned long long vb;
    int rc = check_hip(c, hipStreamSynchronize(c->stream), "cycle sync");
    if (!rc) rc = check_hip(c, hipMemcpy(h, c->red_d, 2 * sizeof(double), hipMemcpyDeviceToHost), "cycle scalars");
    if (!rc) rc = check_hip(c, hipMemcpy(&vb, c->red_u, sizeof(vb), hipMemcpyDeviceToHost), "cycle vmax");
    if (rc) return rc;
    double mv;
    memcpy(&mv, &vb, sizeof(mv));
    mv = sqrt(mv);
    if (p->gr_flag > 0) mv /= sqrt(mv * mv + 1.0);
    st->maxvel = mv;
    st->T00hom = p->gr_flag > 0 ? h[0] / (double)n3(c) : 0.0;
    st->phi_k0 = h[1];
    return 0;
}

GEV_EXPORT int gev_run_cycles(gev_ctx *c, const gev_sim_params *p, gev_cycle_state *st,
                              int ncycles) {
    if (!c || !p || !st || ncycles < 0) return GEV_EINVAL;
    if (p->gr_flag < 0 || p->gr_flag > 1 || p->vector_flag < 0 || p->vector_flag > 1)
        return fail(c, GEV_EINVAL, "bad simulation flags");
    if (!(st->a > 0.) || !(st->dtau > 0.)) return fail(c, GEV_EINVAL, "state needs a > 0, dtau > 0");
    for (int i = 0; i < ncycles; i++) {
        int rc;
        if (st->stage == 1) {  // the field stage of this cycle ran at an output stop
            rc = cycle_particles(c, p, st);
            st->stage = 0;
        } else {
            rc = run_cycle(c, p, st, nullptr);
        }
        if (rc) return rc;
    }
    return ncycles > 0 ? read_cycle_scalars(c, p, st) : 0;
}

// ---------------------------------------------------------------- outputs

namespace gev {

// A number as writePowerSpectrum prints it ("%e" in the table, "%f" for the
// header redshift) and its interpolating pass reads it back (fscanf %le / %lf),
// tools.hpp:283-301,318-332.
double text_round_trip(const char *fmt, double v) {
    char buf[64];
    snprintf(buf, sizeof buf, fmt, v);
    return strtod(buf, nullptr);
}

void PkWriter::init(const gev_ctx *c, const gev_sim_params *p, int num_pk_, int numbins_, double boxsize) {
    const gev_cosmology *cs = &p->cosmo;
    const double n3d = (double)c->N * (double)c->N * (double)c->N;
    // rescalep of the delta spectrum: output.hpp:2026 (GR, no ncdm), 1853 (Newton)
    const double om = p->gr_flag > 0 ? cs->Omega_cdm + cs->Omega_b + 0.0 : cs->Omega_m;
    num_pk = num_pk_;
    numbins = numbins_;
    rescalek = boxsize;
    rescalep = n3d * n3d * 2. * M_PI * M_PI * om * om;
    for (auto *v : {&kb, &pw, &ks, &ps, &p2}) v->assign(numbins, 0.);
    pre.assign((size_t)num_pk * numbins, 0.);
    pre_z.assign(num_pk, 0.);
    oc.assign(numbins, 0);
    have.assign(num_pk, 0);
}

// writeSpectra's delta spectrum (output.hpp:1820-1827 Newton, 2001-2010 GR)
// and writePowerSpectrum (tools.hpp:268-346, EXACT_OUTPUT_REDSHIFTS) of
// output j at the current a, into table j of the outputs
int write_spectrum(gev_ctx *c, const gev_sim_params *p, const gev_cycle_state *st, PkWriter &w, int j,
                   double zt, double *k, double *pk, double *sk, double *spk, int32_t *occ,
                   gev_pk_output_info *info) {
    if (w.numbins != w.kb.size()) return fail(c, GEV_EINVAL, "spectrum writer not set up");
    const int nb = w.numbins;
    int rc = gev_projection_init(c, GEV_SOURCE);
    if (!rc) rc = p->gr_flag > 0 ? gev_projection_T00_project(c, st->a, 1, 1.0)
                                 : gev_scalar_projection_cic_project(c, 1.0);
    if (!rc) rc = gev_fft_execute(c, GEV_PLAN_SOURCE, GEV_FFT_FORWARD);
    if (!rc) rc = gev_extract_power_spectrum(c, GEV_SCALARFT, nb, 1, 1, w.kb.data(), w.pw.data(), w.ks.data(),
                                             w.ps.data(), w.oc.data());
    if (rc) return rc;
    double aw = st->a;
    for (int i = 0; i < nb; i++) w.p2[i] = w.pw[i] / w.rescalep;
    if (1. / aw < zt + 1.) {
        if (w.have[j]) {  // the file of the look-ahead write exists: interpolate
            double weight = w.pre_z[j];
            weight = (weight - zt) / (1. + weight - 1. / aw);
            for (int i = 0; i < nb; i++)
                if (w.oc[i] > 0) w.p2[i] = w.pre[(size_t)j * nb + i];
            for (int i = 0; i < nb; i++) w.p2[i] = (1. - weight) * w.p2[i] + weight * w.pw[i] / w.rescalep;
            aw = 1. / (zt + 1.);
        }
    } else {  // the look-ahead write: what the file holds afterwards
        w.have[j] = 1;
        w.pre_z[j] = text_round_trip("%f", (1. / aw) - 1.);
        for (int i = 0; i < nb; i++) w.pre[(size_t)j * nb + i] = w.oc[i] > 0 ? text_round_trip("%e", w.p2[i]) : 0.;
    }
    for (int i = 0; i < nb; i++) {
        const size_t o = (size_t)j * nb + i;
        k[o] = w.kb[i] / w.rescalek;
        pk[o] = w.p2[i];
        sk[o] = w.ks[i] / w.rescalek;
        spk[o] = w.oc[i] > 0 ? w.ps[i] / w.rescalep / sqrt((double)w.oc[i]) : 0.;
        occ[o] = w.oc[i];
    }
    info[j].z = (1. / aw) - 1.;
    info[j].cycle = st->cycle;
    info[j].interpolated = aw != st->a;
    return 0;
}

}  // namespace gev

GEV_EXPORT int gev_run_to_outputs(gev_ctx *c, const gev_sim_params *p, gev_cycle_state *st, int32_t num_pk,
                                  const double *z_pk, int32_t numbins, double boxsize, int32_t max_cycles,
                                  double *k, double *pk, double *sigma_k, double *sigma_pk, int32_t *occupation,
                                  gev_pk_output_info *info) {
    if (!c || !p || !st) return GEV_EINVAL;
    if (p->gr_flag < 0 || p->gr_flag > 1 || p->vector_flag < 0 || p->vector_flag > 1)
        return fail(c, GEV_EINVAL, "bad simulation flags");
    if (!(st->a > 0.) || !(st->dtau > 0.)) return fail(c, GEV_EINVAL, "state needs a > 0, dtau > 0");
    if (num_pk < 1 || !z_pk || numbins < 1 || numbins > 1024 || !(boxsize > 0.) || max_cycles < 0)
        return fail(c, GEV_EINVAL, "outputs need num_pk >= 1, z_pk, numbins in [1, 1024], boxsize > 0");
    if (!k || !pk || !sigma_k || !sigma_pk || !occupation || !info) return fail(c, GEV_EINVAL, "null output");
    for (int j = 1; j < num_pk; j++)
        if (!(z_pk[j] < z_pk[j - 1])) return fail(c, GEV_EINVAL, "z_pk must be strictly descending");
    const gev_cosmology *cs = &p->cosmo;
    PkWriter w;
    w.init(c, p, num_pk, numbins, boxsize);
    int pkcount = 0, ran = 0, rc = 0;
    // a run stopped at an output resumes at that cycle's look-ahead check
    bool resume = st->stage == 1;
    while (ran < max_cycles) {
        if (!resume) {
            rc = cycle_fields(c, p, st, nullptr);
            // main.cpp:639-657
            if (!rc && pkcount < num_pk && 1. / st->a < z_pk[pkcount] + 1.) {
                rc = write_spectrum(c, p, st, w, pkcount, z_pk[pkcount], k, pk, sigma_k, sigma_pk, occupation,
                                    info);
                pkcount++;
            }
        }
        resume = false;
        // main.cpp:659-677: the look-ahead of EXACT_OUTPUT_REDSHIFTS
        double tmp = st->a;
        rk4bg(tmp, p->fourpiG, cs, 0.5 * st->dtau);
        rk4bg(tmp, p->fourpiG, cs, 0.5 * st->dtau);
        if (!rc && pkcount < num_pk && 1. / tmp < z_pk[pkcount] + 1.)
            rc = write_spectrum(c, p, st, w, pkcount, z_pk[pkcount], k, pk, sigma_k, sigma_pk, occupation, info);
        if (rc) return rc;
        st->stage = 1;
        // main.cpp:685-693: the loop ends once every output is written
        if (pkcount >= num_pk) break;
        rc = cycle_particles(c, p, st);
        if (rc) return rc;
        st->stage = 0;
        ran++;
    }
    rc = read_cycle_scalars(c, p, st);
    if (rc) return rc;
    return pkcount >= num_pk ? 0 : fail(c, GEV_ESTATE, "max_cycles ran before the last output");
}

GEV_EXPORT int gev_run_cycles_recorded(gev_ctx *c, const gev_sim_params *p, gev_cycle_state *st, int ncycles,
                                       gev_background_record *rec) {
    if (!c || !p || !st || ncycles < 0 || (ncycles > 0 && !rec)) return GEV_EINVAL;
    if (p->gr_flag < 0 || p->gr_flag > 1 || p->vector_flag < 0 || p->vector_flag > 1)
        return fail(c, GEV_EINVAL, "bad simulation flags");
    if (!(st->a > 0.) || !(st->dtau > 0.)) return fail(c, GEV_EINVAL, "state needs a > 0, dtau > 0");
    if (st->stage != 0) return fail(c, GEV_ESTATE, "a stopped cycle is completed by gev_run_cycles first");
    const double H0 = hconf(1., p->fourpiG, &p->cosmo);
    const double n3d = (double)n3(c);
    int done = 0;  // records whose device values have been read back
    for (int i = 0; i < ncycles; i++) {
        // host values of the line are those at the cycle's start (main.cpp:533)
        rec[i].cycle = st->cycle;
        rec[i].reserved0 = 0;
        rec[i].tau = st->tau;
        rec[i].a = st->a;
        rec[i].hconf_ratio = hconf(st->a, p->fourpiG, &p->cosmo) / H0;
        int rc = run_cycle(c, p, st, c->rec_pinned + 2 * (i % REC_SLOTS));
        if (rc) return rc;
        if ((i + 1) % REC_SLOTS == 0 || i + 1 == ncycles) {
            rc = check_hip(c, hipStreamSynchronize(c->stream), "record sync");
            if (rc) return rc;
            for (int j = done; j <= i; j++) {
                const double *v = c->rec_pinned + 2 * (j % REC_SLOTS);
                rec[j].T00hom = p->gr_flag > 0 ? v[0] / n3d : 0.0;
                rec[j].phi_k0 = v[1];
            }
            done = i + 1;
        }
    }
    if (ncycles > 0) {
        unsigned long long vb;
        int rc = check_hip(c, hipMemcpy(&vb, c->red_u, sizeof(vb), hipMemcpyDeviceToHost), "cycle vmax");
        if (rc) return rc;
        double mv;
        memcpy(&mv, &vb, sizeof(mv));
        mv = sqrt(mv);
        if (p->gr_flag > 0) mv /= sqrt(mv * mv + 1.0);
        st->maxvel = mv;
        st->T00hom = rec[ncycles - 1].T00hom;
        st->phi_k0 = rec[ncycles - 1].phi_k0;
    }
    return 0;
}

GEV_EXPORT int gev_write_background(const char *path, const gev_background_record *rec, int n) {
    if (!path || n < 0 || (n > 0 && !rec)) return GEV_EINVAL;
    FILE *f = fopen(path, "a");
    if (!f) return GEV_EINVAL;
    for (int i = 0; i < n; i++) {
        if (rec[i].cycle == 0)
            fprintf(f, "# background statistics\n# cycle   tau/boxsize    a             conformal H/H0  phi(k=0)       T00(k=0)\n");
        fprintf(f, " %6d   %e   %e   %e   %e   %e\n", rec[i].cycle, rec[i].tau, rec[i].a, rec[i].hconf_ratio,
                rec[i].phi_k0, rec[i].T00hom);
    }
    fclose(f);
    return 0;
}

// ---------------------------------------------------------------- timing

GEV_EXPORT int gev_timing_enable(gev_ctx *c, int on) {
    if (!c) return GEV_EINVAL;
    if (!on) timers_collect(c);
    c->timing = on != 0;
    return 0;
}

GEV_EXPORT int gev_timing_reset(gev_ctx *c) {
    if (!c) return GEV_EINVAL;
    timers_collect(c);
    c->acc.clear();
    return 0;
}

GEV_EXPORT int gev_timing_report(gev_ctx *c, char *buf, size_t len) {
    if (!c || !buf || len == 0) return GEV_EINVAL;
    timers_collect(c);
    std::string s;
    char line[256];
    for (auto &a : c->acc) {
        snprintf(line, sizeof line, "%s %ld %.6f\n", a.name.c_str(), a.calls, a.ms);
        s += line;
    }
    strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
    return 0;
}
