// ic_api.hip -- C ABI of the basic initial-condition generator (ic_basic.hpp):
// the file loaders, the host set-up the reference does on the CPU (natural
// cubic spline coefficients, the CIC kernel stencil, the transfer-function
// tables) and the generateIC_basic sequence, which runs on the device through
// ic.hip and the cycle entry points.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gev_internal.h"

using namespace gev;

namespace {

// A device table of the IC set-up: a slice of the box's IC staging area, or
// (a table larger than the area holds) an allocation released at scope exit
struct DevMem {
    void *p = nullptr;
    bool owned = false;
    ~DevMem() {
        if (owned && p) (void)hipFree(p);
    }
};

// consecutive slices of the staging area, 256-byte aligned
struct Stage {
    char *at;
    int64_t left;
    explicit Stage(gev_ctx *c) : at(c->ic_stage), left(ic_stage_bytes(c->N)) {}
    int take(gev_ctx *c, DevMem &m, size_t bytes, const char *what) {
        const int64_t need = ((int64_t)bytes + 255) / 256 * 256;
        if (need <= left) {
            m.p = at;
            at += need;
            left -= need;
            return 0;
        }
        m.owned = true;
        return check_hip(c, hipMalloc(&m.p, bytes > 0 ? bytes : 1), what);
    }
};

// host table -> device, in the context's stream order (the callers
// synchronise the stream before their host tables go out of scope)
int dev_copy(gev_ctx *c, Stage &st, DevMem &m, const void *host, size_t bytes, const char *what) {
    int rc = st.take(c, m, bytes, what);
    if (!rc && bytes > 0) rc = check_hip(c, hipMemcpyAsync(m.p, host, bytes, hipMemcpyHostToDevice, c->stream), what);
    return rc;
}

// Natural cubic spline through (x, y) (gsl_interp_cspline): c[i] is half the
// second derivative at knot i, zero at both ends; interior equations
// h_{i-1} c_{i-1} + 2 (h_{i-1} + h_i) c_i + h_i c_{i+1} = 3 (s_i - s_{i-1}),
// s_i = (y_{i+1} - y_i) / h_i, by Gaussian elimination on the tridiagonal.
std::vector<double> spline_coeffs(const double *x, const double *y, int n) {
    std::vector<double> c(n, 0.0);
    const int m = n - 2;
    if (m <= 0) return c;
    std::vector<double> diag(m), off(m), rhs(m);
    for (int i = 0; i < m; i++) {
        const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
        off[i] = h1;
        diag[i] = 2.0 * (h1 + h0);
        rhs[i] = 3.0 * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
    }
    for (int i = 1; i < m; i++) {
        const double l = off[i - 1] / diag[i - 1];
        diag[i] -= l * off[i - 1];
        rhs[i] -= l * rhs[i - 1];
    }
    c[m] = rhs[m - 1] / diag[m - 1];
    for (int i = m - 2; i >= 0; i--) c[i + 1] = (rhs[i] - off[i] * c[i + 2]) / diag[i];
    return c;
}

// The initial-condition generator evaluates its fields in FP64: in the FP32
// form it is run on an FP64 context whose state gev_copy_state then carries over.
int fp64_only(gev_ctx *c) {
    return c->fp32 ? fail(c, GEV_EINVAL,
                          "the initial-condition generator needs an FP64 context (generate there, then gev_copy_state)")
                   : 0;
}

int real_field(gev_ctx *c, gev_field f, double **out) {
    if (int rc = fp64_only(c)) return rc;
    switch (f) {
    case GEV_PHI: *out = c->phi; return 0;
    case GEV_CHI: *out = c->chi; return 0;
    case GEV_SOURCE: *out = c->source; return 0;
    default: return fail(c, GEV_EINVAL, "a real scalar field is required");
    }
}

// The generateCICKernel stencil on the 3x3x3 neighbourhood of the origin,
// v[(dz+1)*9 + (dy+1)*3 + (dx+1)], accumulated in the reference's loop order
// with its operand types (float weights; products with 1. - w in double).
void cic_stencil(int N, int64_t numpcl, const float *pcl, int numtile, double v[27]) {
    for (int i = 0; i < 27; i++) v[i] = 0.;
    auto at = [&](int dx, int dy, int dz) -> double & { return v[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)]; };
    const long ls = N;
    double renorm = (double)(ls * ls);
    if (numpcl == 0 || pcl == nullptr) {  // ic_basic.hpp:747-766
        at(0, 0, 0) = 6. * renorm;
        at(1, 0, 0) = at(-1, 0, 0) = -renorm;
        at(0, 1, 0) = at(0, 0, 1) = at(0, -1, 0) = at(0, 0, -1) = -renorm;
        return;
    }
    renorm /= (double)(numpcl * (long)numtile * (long)numtile * (long)numtile) / (double)(ls * ls * ls);
    // octants in the reference's order (ic_basic.hpp:775-905): corner signs
    const int sgn[8][3] = {{1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1},
                           {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}, {1, -1, -1}};
    for (int64_t i = 0; i < numpcl; i++) {
        for (int oc = 0; oc < 8; oc++) {
            const int sx = sgn[oc][0], sy = sgn[oc][1], sz = sgn[oc][2];
            double d[3];
            for (int a = 0; a < 3; a++) {
                const int s = a == 0 ? sx : (a == 1 ? sy : sz);
                if (s > 0) d[a] = ls * pcl[3 * i + a] / numtile;  // float expression
                else d[a] = ls * (1. - pcl[3 * i + a]) / numtile;  // double expression
            }
            if (d[0] >= 1. || d[1] >= 1. || d[2] >= 1.) continue;
            const float wx = (float)(1. - d[0]), wy = (float)(1. - d[1]), wz = (float)(1. - d[2]);
            float ww, q1;
            // 0-direction (:915-958)
            ww = (float)(wy * wz * renorm);
            q1 = (wx > 0.9) ? 2.f : 1.f;
            at(0, 0, 0) += ww * wy * wz * q1;
            at(sx, 0, 0) -= ww * wy * wz;
            if (wx > 0.9) at(-sx, 0, 0) -= ww * wy * wz;
            at(0, 0, sz) += ww * wy * (1. - wz) * q1;
            at(sx, 0, sz) -= ww * wy * (1. - wz);
            if (wx > 0.9) at(-sx, 0, sz) -= ww * wy * (1. - wz);
            at(0, sy, 0) += ww * (1. - wy) * wz * q1;
            at(sx, sy, 0) -= ww * (1. - wy) * wz;
            if (wx > 0.9) at(-sx, sy, 0) -= ww * (1. - wy) * wz;
            at(0, sy, sz) += ww * (1. - wy) * (1. - wz) * q1;
            at(sx, sy, sz) -= ww * (1. - wy) * (1. - wz);
            if (wx > 0.9) at(-sx, sy, sz) -= ww * (1. - wy) * (1. - wz);
            // 1-direction (:960-1003)
            ww = (float)(wx * wz * renorm);
            q1 = (wy > 0.9) ? 2.f : 1.f;
            at(0, 0, 0) += ww * wx * wz * q1;
            at(sx, 0, 0) += ww * (1. - wx) * wz * q1;
            at(0, 0, sz) += ww * wx * (1. - wz) * q1;
            at(sx, 0, sz) += ww * (1. - wx) * (1. - wz) * q1;
            at(0, sy, 0) -= ww * wx * wz;
            at(sx, sy, 0) -= ww * (1. - wx) * wz;
            at(0, sy, sz) -= ww * wx * (1. - wz);
            at(sx, sy, sz) -= ww * (1. - wx) * (1. - wz);
            if (wy > 0.9) {
                at(0, -sy, 0) -= ww * wx * wz;
                at(sx, -sy, 0) -= ww * (1. - wx) * wz;
                at(0, -sy, sz) -= ww * wx * (1. - wz);
                at(sx, -sy, sz) -= ww * (1. - wx) * (1. - wz);
            }
            // 2-direction (:1005-1048)
            ww = (float)(wx * wy * renorm);
            q1 = (wz > 0.9) ? 2.f : 1.f;
            at(0, 0, 0) += ww * wx * wy * q1;
            at(sx, 0, 0) += ww * (1. - wx) * wy * q1;
            at(0, sy, 0) += ww * wx * (1. - wy) * q1;
            at(sx, sy, 0) += ww * (1. - wx) * (1. - wy) * q1;
            at(0, 0, sz) -= ww * wx * wy;
            at(sx, 0, sz) -= ww * (1. - wx) * wy;
            at(0, sy, sz) -= ww * wx * (1. - wy);
            at(sx, sy, sz) -= ww * (1. - wx) * (1. - wy);
            if (wz > 0.9) {
                at(0, 0, -sz) -= ww * wx * wy;
                at(sx, 0, -sz) -= ww * (1. - wx) * wy;
                at(0, sy, -sz) -= ww * wx * (1. - wy);
                at(sx, sy, -sz) -= ww * (1. - wx) * (1. - wy);
            }
        }
    }
}

double pk_primordial(double k, const gev_ic_params *ic) {
    return ic->A_s * pow(k / ic->k_pivot, ic->n_s - 1.);  // ic_basic.hpp:186-189
}

// Spline ordinates of generateIC_basic's transfer-function path, baryon
// treatment "ignore": phi (ic_basic.hpp:1720-1722), the CDM displacement and
// velocity potential (:1952-1983; Newtonian-gauge correction :1730-1735).
void ic_tables(const gev_sim_params *p, const gev_ic_params *ic, int n, const double *k, const double *d_tot,
               const double *t_tot, const double *d_cdm, const double *t_cdm, std::vector<double> &yphi,
               std::vector<double> &ydisp, std::vector<double> &yvel) {
    const gev_cosmology *cs = &p->cosmo;
    const double G = p->fourpiG;
    const double a = 1. / (1. + ic->z_in);
    auto H = [&](double x) { return gev_hconf(x, G, cs); };
    const double Ha = H(a);
    const double rescale =
        3. * Ha * Ha * Ha *
        (1. + 0.5 * Ha * Ha *
                  ((1. / H(0.98 * a) / H(0.98 * a)) - (8. / H(0.99 * a) / H(0.99 * a)) +
                   (8. / H(1.01 * a) / H(1.01 * a)) - (1. / H(1.02 * a) / H(1.02 * a))) /
                  0.12);
    yphi.resize(n);
    ydisp.resize(n);
    yvel.resize(n);
    for (int i = 0; i < n; i++) {
        const double sq = sqrt(pk_primordial(k[i] * cs->h / ic->boxsize, ic) / k[i]);
        yphi[i] = (1.5 * (Ha * Ha - H(1.) * H(1.) * a * a * cs->Omega_Lambda) * d_tot[i] +
                   rescale * t_tot[i] / k[i] / k[i]) *
                  M_PI * sq / k[i];
    }
    for (int i = 0; i < n; i++) {
        const double sq = sqrt(pk_primordial(k[i] * cs->h / ic->boxsize, ic) / k[i]);
        if (p->gr_flag > 0)
            ydisp[i] = -3. * yphi[i] / k[i] / k[i] - d_cdm[i] * M_PI * sq / k[i];
        else
            ydisp[i] = -3. * Ha * M_PI * t_tot[i] * sq / k[i] / k[i] / k[i] - d_cdm[i] * M_PI * sq / k[i];
        yvel[i] = -a * t_cdm[i] * M_PI * sq / k[i];
    }
}

int read_max(gev_ctx *c, double *out, bool root) {
    unsigned long long bits = 0;
    int rc = check_hip(c, hipStreamSynchronize(c->stream), "max sync");
    if (!rc) rc = check_hip(c, hipMemcpy(&bits, c->red_u, sizeof bits, hipMemcpyDeviceToHost), "max readback");
    if (rc) return rc;
    double v;
    memcpy(&v, &bits, sizeof v);
    *out = root ? sqrt(v) : v;
    return 0;
}

}  // namespace

GEV_EXPORT int gev_prng_fill(gev_ctx *c, uint32_t seed, uint64_t pos0, int64_t n, uint32_t *host_out) {
    if (!c || n < 0 || (n > 0 && !host_out)) return c ? fail(c, GEV_EINVAL, "bad prng arguments") : GEV_EINVAL;
    if (n == 0) return 0;
    DevMem d;
    d.owned = true;
    int rc = check_hip(c, hipMalloc(&d.p, (size_t)n * 4), "prng buffer");
    if (!rc) rc = launch_prng_fill(c, seed, pos0, n, (uint32_t *)d.p);
    if (!rc) rc = check_hip(c, hipStreamSynchronize(c->stream), "prng sync");
    if (!rc) rc = check_hip(c, hipMemcpy(host_out, d.p, (size_t)n * 4, hipMemcpyDeviceToHost), "prng readback");
    return rc;
}

GEV_EXPORT int gev_load_homogeneous_template(const char *path, float *out, int64_t capacity, int64_t *numpcl) {
    if (!path || !numpcl) return GEV_EINVAL;
    FILE *f = fopen(path, "rb");
    if (!f) return GEV_EINVAL;
    // Gadget-2 header block: 256 bytes between two record markers
    int32_t b1 = 0, b2 = 0;
    unsigned char hdr[256];
    uint32_t npart[6];
    int32_t num_files;
    double box;
    int rc = GEV_EINVAL;
    std::vector<float> buf;
    do {
        if (fread(&b1, 4, 1, f) != 1 || b1 != 256) break;
        if (fread(hdr, 256, 1, f) != 1 || fread(&b2, 4, 1, f) != 1 || b2 != b1) break;
        memcpy(npart, hdr, sizeof npart);
        memcpy(&num_files, hdr + 124, 4);
        memcpy(&box, hdr + 128, 8);
        if (num_files != 1 || !(box > 0.) || npart[1] == 0) break;
        if (fread(&b1, 4, 1, f) != 1) break;
        if (npart[0] > 0 && fseek(f, 12L * (long)npart[0], SEEK_CUR)) break;
        buf.resize((size_t)3 * npart[1]);
        if (fread(buf.data(), 4, buf.size(), f) != buf.size()) break;
        long skip = 0;
        for (int t = 2; t < 6; t++) skip += 12L * (long)npart[t];
        if (skip && fseek(f, skip, SEEK_CUR)) break;
        if (fread(&b2, 4, 1, f) != 1 || b2 != b1) break;
        bool okv = true;
        for (auto &v : buf) {
            v /= box;
            if (v < 0. || v > 1.) okv = false;
        }
        if (!okv) break;
        *numpcl = (int64_t)npart[1];
        rc = 0;
    } while (false);
    fclose(f);
    if (rc) return rc;
    if (!out || capacity < *numpcl) return GEV_EINVAL;
    memcpy(out, buf.data(), buf.size() * sizeof(float));
    return 0;
}

GEV_EXPORT int gev_load_transfer_functions(const char *path, const char *qname, double boxsize, double h,
                                           double *k, double *tk_d, double *tk_t, int32_t capacity,
                                           int32_t *npoints) {
    if (!path || !qname || !npoints) return GEV_EINVAL;
    FILE *f = fopen(path, "r");
    if (!f) return GEV_EINVAL;
    std::vector<std::string> lines;
    char line[2048];
    while (fgets(line, sizeof line, f)) lines.push_back(line);
    fclose(f);
    // columns: labels "<n>:k...", "<n>:d_<q>...", "<n>:t_<q>..." in the header
    int kcol = -1, dcol = -1, tcol = -1, ncol = 0;
    const size_t ql = strlen(qname);
    for (const auto &l : lines) {
        const char *s = l.c_str();
        int i = 0;
        for (const char *p = strchr(s, ':'); p; p = strchr(p, ':'), i++) {
            p++;
            if (*p == 'k') kcol = i;
            else if (*p == 'd' && strncmp(p + 2, qname, ql) == 0) dcol = i;
            else if (*p == 't' && strncmp(p + 2, qname, ql) == 0) tcol = i;
        }
        if (kcol >= 0 && dcol >= 0 && tcol >= 0) {
            ncol = i;
            break;
        }
    }
    if (kcol < 0 || dcol < 0 || tcol < 0) return GEV_EINVAL;
    std::vector<double> kk, dd, tt;
    for (const auto &l : lines) {
        if (l.empty() || l[0] == '#') continue;
        std::vector<double> vals;
        const char *s = l.c_str();
        char *end;
        for (int c = 0; c < ncol; c++) {
            const double v = strtod(s, &end);
            if (end == s) break;
            vals.push_back(v);
            s = end;
        }
        if ((int)vals.size() < ncol) continue;
        if (vals[kcol] < 0.) return GEV_EINVAL;
        const double kv = vals[kcol] * boxsize;
        if (!kk.empty() && kk.back() >= kv) return GEV_EINVAL;
        kk.push_back(kv);
        dd.push_back(vals[dcol]);
        tt.push_back(vals[tcol] * boxsize / h);
    }
    if (kk.size() < 2) return GEV_EINVAL;
    *npoints = (int32_t)kk.size();
    if (capacity < *npoints) return GEV_EINVAL;
    if (k) memcpy(k, kk.data(), kk.size() * 8);
    if (tk_d) memcpy(tk_d, dd.data(), dd.size() * 8);
    if (tk_t) memcpy(tk_t, tt.data(), tt.size() * 8);
    return 0;
}

GEV_EXPORT int gev_generate_cic_kernel(gev_ctx *c, gev_field f, int64_t numpcl, const float *pcldata,
                                       int32_t numtile) {
    if (!c) return GEV_EINVAL;
    double *dst;
    int rc = real_field(c, f, &dst);
    if (rc) return rc;
    if (numpcl < 0 || (numpcl > 0 && pcldata && numtile < 1)) return fail(c, GEV_EINVAL, "bad template");
    double v[27];
    cic_stencil(c->N, numpcl, pcldata, numtile, v);
    rc = launch_fill_zero(c, dst, sizeof(double) * n3(c));
    if (!rc) rc = launch_put_stencil(c, dst, v);
    if (f >= GEV_PHI && f <= GEV_BI) c->zeroed[f] = false;
    return rc;
}

GEV_EXPORT int gev_generate_displacement_field(gev_ctx *c, double coeff, int32_t nknots, const double *x,
                                               const double *y, uint32_t seed, int32_t ksphere,
                                               int32_t deconvolve_f, int32_t ignorekernel) {
    if (!c) return GEV_EINVAL;
    if (int rc = fp64_only(c)) return rc;
    if (nknots < 2 || !x || !y) return fail(c, GEV_EINVAL, "spline needs two or more knots");
    for (int i = 1; i < nknots; i++)
        if (!(x[i] > x[i - 1])) return fail(c, GEV_EINVAL, "spline knots must increase");
    const int N = c->N, kmax = N / 2 - 1;
    // every drawn mode evaluates the spline at 2 pi |k| with 1 <= |k| < kmax
    // (sqrt(3) kmax for the cube): GSL's domain error in the reference
    const double kreach = 2. * M_PI * (ksphere ? (double)kmax : sqrt(3.) * kmax);
    if (kmax > 1 && (x[0] > 2. * M_PI * 0.999999 || x[nknots - 1] < kreach))
        return fail(c, GEV_EINVAL, "spline range does not cover the grid's wave numbers");
    std::vector<double> cc = spline_coeffs(x, y, nknots);
    std::vector<float> sinc(N);
    sinc[0] = 1.;
    for (int i = 1; i < N; i++)
        sinc[i] = deconvolve_f == 1 ? (float)(sin(M_PI * (float)i / (float)N) * (float)N / (M_PI * (float)i)) : 1.f;
    DevMem dx, dy, dc, ds, dr;
    Stage stage(c);
    int rc = dev_copy(c, stage, dx, x, nknots * 8, "spline x");
    if (!rc) rc = dev_copy(c, stage, dy, y, nknots * 8, "spline y");
    if (!rc) rc = dev_copy(c, stage, dc, cc.data(), nknots * 8, "spline c");
    if (!rc) rc = dev_copy(c, stage, ds, sinc.data(), N * 4, "sinc");
    if (!rc) rc = stage.take(c, dr, (size_t)N * N, "row marks");
    if (rc) {
        (void)hipStreamSynchronize(c->stream);  // the host tables outlive the copies
        return rc;
    }
    RealizeSetup s{(uint8_t *)dr.p, (const float *)ds.p, (const double *)dx.p, (const double *)dy.p,
                   (const double *)dc.p, nknots, seed, ksphere, ignorekernel, coeff};
    rc = launch_realize(c, s);
    if (!rc) rc = check_hip(c, hipStreamSynchronize(c->stream), "realize sync");
    return rc;
}

GEV_EXPORT int gev_initialize_particle_positions(gev_ctx *c, int64_t numpart, const float *partdata,
                                                 int32_t numtile, double mass) {
    if (!c) return GEV_EINVAL;
    if (numpart < 1 || !partdata || numtile < 1) return fail(c, GEV_EINVAL, "bad template");
    const int64_t n = numpart * (int64_t)numtile * numtile * numtile;
    if (n > c->cap) return fail(c, GEV_EINVAL, "particle count exceeds capacity");
    for (int64_t i = 0; i < 3 * numpart; i++)
        if (!(partdata[i] >= 0.f && partdata[i] < 1.f))
            return fail(c, GEV_EINVAL, "template positions must lie in [0,1)");
    DevMem t;
    Stage stage(c);
    int rc = dev_copy(c, stage, t, partdata, (size_t)numpart * 12, "template");
    if (rc) {
        (void)hipStreamSynchronize(c->stream);
        return rc;
    }
    c->n = n;
    c->pmass = mass;
    rc = launch_init_particles(c, (const float *)t.p, numpart, numtile);
    if (!rc) rc = launch_refile(c);
    if (!rc) rc = check_hip(c, hipStreamSynchronize(c->stream), "init sync");
    return rc;
}

GEV_EXPORT int gev_displace_particles_ic(gev_ctx *c, double coeff, gev_field f, double *max_displacement) {
    if (!c) return GEV_EINVAL;
    double *src;
    int rc = real_field(c, f, &src);
    if (!rc) rc = launch_ic_displace(c, src, coeff);
    if (!rc && max_displacement) rc = read_max(c, max_displacement, false);
    return rc;
}

GEV_EXPORT int gev_initialize_q_ic(gev_ctx *c, double coeff, gev_field f, double *maxvel) {
    if (!c) return GEV_EINVAL;
    double *src;
    int rc = real_field(c, f, &src);
    if (!rc) rc = launch_ic_q(c, src, coeff);
    if (!rc && maxvel) rc = read_max(c, maxvel, true);
    return rc;
}

GEV_EXPORT int gev_generate_ic_basic(gev_ctx *c, const gev_sim_params *p, const gev_ic_params *ic,
                                     int64_t numpcl, const float *pcldata, int32_t ntk, const double *k,
                                     const double *d_tot, const double *t_tot, const double *d_cdm,
                                     const double *t_cdm, double *maxvel, double *max_displacement) {
    if (!c) return GEV_EINVAL;
    if (int rc = fp64_only(c)) return rc;
    if (!p || !ic || !pcldata || numpcl < 1 || ic->numtile < 1 || ntk < 2 || !k || !d_tot || !t_tot ||
        !d_cdm || !t_cdm)
        return fail(c, GEV_EINVAL, "bad initial-condition arguments");
    const int N = c->N;
    const double a = 1. / (1. + ic->z_in);
    const double G = p->fourpiG;
    const int64_t np = numpcl * (int64_t)ic->numtile * ic->numtile * ic->numtile;
    std::vector<double> yphi, ydisp, yvel;
    ic_tables(p, ic, ntk, k, d_tot, t_tot, d_cdm, t_cdm, yphi, ydisp, yvel);
    int rc;
#define TRY(x)           \
    do {                 \
        rc = (x);        \
        if (rc) return rc; \
    } while (0)
    // CDM displacement (ic_basic.hpp:1672-1676, 1961-1966)
    TRY(gev_generate_cic_kernel(c, GEV_SOURCE, ic->correct_displacement ? numpcl : 0, pcldata, ic->numtile));
    TRY(gev_fft_execute(c, GEV_PLAN_SOURCE, GEV_FFT_FORWARD));
    TRY(gev_generate_displacement_field(c, 0., ntk, k, ydisp.data(), ic->seed, ic->ksphere, 1, 0));
    TRY(gev_fft_execute(c, GEV_PLAN_CHI, GEV_FFT_BACKWARD));
    // particles (:1967-1983)
    TRY(gev_initialize_particle_positions(c, numpcl, pcldata, ic->numtile,
                                          (p->cosmo.Omega_cdm + p->cosmo.Omega_b) / (double)np));
    TRY(gev_displace_particles_ic(c, 1., GEV_CHI, max_displacement));
    // CDM velocities (:2030-2055)
    if (ic->correct_displacement) TRY(gev_generate_cic_kernel(c, GEV_SOURCE, 0, nullptr, 1));
    TRY(gev_fft_execute(c, GEV_PLAN_SOURCE, GEV_FFT_FORWARD));
    TRY(gev_generate_displacement_field(c, 0., ntk, k, yvel.data(), ic->seed, ic->ksphere, 0, 0));
    TRY(gev_fft_execute(c, GEV_PLAN_CHI, GEV_FFT_BACKWARD));
    double mv = 0.;
    TRY(gev_initialize_q_ic(c, 1., GEV_CHI, &mv));
    mv /= a;
    // phi (:2145-2152, 2175-2176)
    TRY(gev_fft_execute(c, GEV_PLAN_SOURCE, GEV_FFT_FORWARD));
    TRY(gev_generate_displacement_field(c, 0., ntk, k, yphi.data(), ic->seed, ic->ksphere, 0, 0));
    TRY(gev_fft_execute(c, GEV_PLAN_PHI, GEV_FFT_BACKWARD));
    // B (:2231-2240)
    TRY(gev_projection_init(c, GEV_BI));
    TRY(gev_projection_T0i_project(c, 1, 1.));
    TRY(gev_fft_execute(c, GEV_PLAN_BI, GEV_FFT_FORWARD));
    TRY(gev_project_ft_vector(c, G / (double)N / (double)N, 0.));
    TRY(gev_fft_execute(c, GEV_PLAN_BI, GEV_FFT_BACKWARD));
    // chi (:2241-2248)
    TRY(gev_projection_init(c, GEV_SIJ));
    TRY(gev_projection_Tij_project(c, a, 1, 1.));
    TRY(gev_prepare_ft_source_tensor(c, 2. * G / a / (double)N / (double)N));
    TRY(gev_fft_execute(c, GEV_PLAN_SIJ, GEV_FFT_FORWARD));
    TRY(gev_project_ft_scalar(c, 0));
    TRY(gev_fft_execute(c, GEV_PLAN_CHI, GEV_FFT_BACKWARD));
#undef TRY
    if (p->gr_flag > 0) mv /= sqrt(mv * mv + 1.0);  // main.cpp:326-330
    if (maxvel) *maxvel = mv;
    return check_hip(c, hipStreamSynchronize(c->stream), "ic sync");
}
