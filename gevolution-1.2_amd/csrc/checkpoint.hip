// checkpoint.hip -- hibernation point and restart of one box: hibernate()
// (hibernation.hpp:512-612, called at main.cpp:827-865) and the restart branch
// of readIC (ic_read.hpp:58-330, 671-716).
//
// What is saved follows the reference: the particles, phi and chi (GR), and B
// divided by a^2 N (parabolic vector method), with the time-stepping scalars
// of the restart settings (a, tau, the last dtau, the cycle).  On restart B is
// multiplied by a^2 / N^2 and transformed forward, which rebuilds BiFT
// (BiFT = r2c(B) / N^3, the carried k-space state), dtau_old is the saved
// dtau and dtau follows ic_read.hpp:298-301.  The reference writes HDF5 files
// through LATfield2; this is one binary file of the same content:
//
//   header (struct CkptHeader below), then little-endian doubles / int64:
//   pos[3][n] (SoA), q[3][n], id[n], phi[N^3] and chi[N^3] if GR,
//   B[3][N^3] / (a^2 N) if parabolic.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gev_internal.h"

namespace {

struct CkptHeader {
    char magic[8];  // "GEVHIB1"
    int32_t ngrid, gr_flag, vector_flag, reserved0;
    int64_t npart;
    double pmass, a, tau, dtau, maxvel;
    int32_t cycle, reserved1;
};

const char kMagic[8] = {'G', 'E', 'V', 'H', 'I', 'B', '1', 0};

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

bool put(FILE *f, const void *p, size_t bytes) { return fwrite(p, 1, bytes, f) == bytes; }
bool get(FILE *f, void *p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

}  // namespace

GEV_EXPORT int gev_hibernate(gev_ctx *c, const gev_sim_params *p, const gev_cycle_state *st, const char *path) {
    if (!c) return GEV_EINVAL;
    if (!p || !st || !path) return gev::fail(c, GEV_EINVAL, "bad hibernate arguments");
    if (st->stage != 0) return gev::fail(c, GEV_ESTATE, "a stopped cycle is completed by gev_run_cycles first");
    const int64_t n3 = gev::n3(c), n = c->n;
    const bool parabolic = p->vector_flag == 0;
    int rc;
    // Newtonian runs keep B only in k space (main.cpp:835-836, 853-854)
    if (parabolic && p->gr_flag == 0 && (rc = gev_fft_execute(c, GEV_PLAN_BI, GEV_FFT_BACKWARD))) return rc;
    File out;
    out.f = fopen(path, "wb");
    if (!out.f) return gev::fail(c, GEV_EINVAL, std::string("cannot write ") + path);
    CkptHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, kMagic, 8);
    h.ngrid = c->N;
    h.gr_flag = p->gr_flag;
    h.vector_flag = p->vector_flag;
    h.npart = n;
    h.pmass = c->pmass;
    h.a = st->a;
    h.tau = st->tau;
    h.dtau = st->dtau_old;  // the dtau of the cycle just completed (main.cpp:863)
    h.maxvel = st->maxvel;
    h.cycle = st->cycle - 1;  // the cycle just completed
    bool ok = put(out.f, &h, sizeof h);
    {
        std::vector<double> pos((size_t)3 * (n > 0 ? n : 1)), q(pos.size());
        std::vector<int64_t> id(n > 0 ? n : 1);
        if (n > 0 && (rc = gev_download_particles(c, pos.data(), q.data(), id.data()))) return rc;
        // AoS from the download -> SoA in the file
        std::vector<double> soa((size_t)n);
        for (int d = 0; d < 3 && ok; d++) {
            for (int64_t i = 0; i < n; i++) soa[i] = pos[3 * i + d];
            ok = put(out.f, soa.data(), n * 8);
        }
        for (int d = 0; d < 3 && ok; d++) {
            for (int64_t i = 0; i < n; i++) soa[i] = q[3 * i + d];
            ok = put(out.f, soa.data(), n * 8);
        }
        if (ok) ok = put(out.f, id.data(), n * 8);
    }
    std::vector<double> f((size_t)3 * n3);
    if (ok && p->gr_flag > 0) {
        if ((rc = gev_download_field(c, GEV_PHI, f.data()))) return rc;
        ok = put(out.f, f.data(), n3 * 8);
        if (ok && (rc = gev_download_field(c, GEV_CHI, f.data()))) return rc;
        if (ok) ok = put(out.f, f.data(), n3 * 8);
    }
    if (ok && parabolic) {
        if ((rc = gev_download_field(c, GEV_BI, f.data()))) return rc;
        const double s = st->a * st->a * (double)c->N;  // hibernation.hpp:530-535
        for (auto &v : f) v /= s;
        ok = put(out.f, f.data(), 3 * n3 * 8);
    }
    if (!ok) return gev::fail(c, GEV_EINVAL, std::string("short write to ") + path);
    return 0;
}

GEV_EXPORT int gev_restart(gev_ctx *c, const gev_sim_params *p, const char *path, gev_cycle_state *st) {
    if (!c) return GEV_EINVAL;
    if (!p || !st || !path) return gev::fail(c, GEV_EINVAL, "bad restart arguments");
    File in;
    in.f = fopen(path, "rb");
    if (!in.f) return gev::fail(c, GEV_EINVAL, std::string("cannot read ") + path);
    CkptHeader h;
    if (!get(in.f, &h, sizeof h) || memcmp(h.magic, kMagic, 8) != 0)
        return gev::fail(c, GEV_EINVAL, "not a hibernation file");
    if (h.ngrid != c->N || h.npart > c->cap || h.npart < 0 || h.gr_flag != p->gr_flag ||
        h.vector_flag != p->vector_flag)
        return gev::fail(c, GEV_EINVAL, "hibernation file does not match this context or run");
    const int64_t n = h.npart, n3 = gev::n3(c);
    int rc;
    {
        std::vector<double> soa((size_t)6 * (n > 0 ? n : 1)), pos((size_t)3 * (n > 0 ? n : 1)), q(pos.size());
        std::vector<int64_t> id(n > 0 ? n : 1);
        if (!get(in.f, soa.data(), (size_t)6 * n * 8) || !get(in.f, id.data(), n * 8))
            return gev::fail(c, GEV_EINVAL, "truncated hibernation file");
        for (int64_t i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) {
                pos[3 * i + d] = soa[(size_t)d * n + i];
                q[3 * i + d] = soa[(size_t)(3 + d) * n + i];
            }
        if ((rc = gev_upload_particles(c, n, pos.data(), q.data(), id.data(), h.pmass))) return rc;
    }
    std::vector<double> f((size_t)3 * n3);
    if (h.gr_flag > 0) {
        if (!get(in.f, f.data(), n3 * 8)) return gev::fail(c, GEV_EINVAL, "truncated hibernation file");
        if ((rc = gev_upload_field(c, GEV_PHI, f.data()))) return rc;
        if (!get(in.f, f.data(), n3 * 8)) return gev::fail(c, GEV_EINVAL, "truncated hibernation file");
        if ((rc = gev_upload_field(c, GEV_CHI, f.data()))) return rc;
    }
    if (h.vector_flag == 0) {
        // ic_read.hpp:303-316: B * a^2 / N^2, then forward: BiFT = r2c(B) / N^3
        if (!get(in.f, f.data(), 3 * n3 * 8)) return gev::fail(c, GEV_EINVAL, "truncated hibernation file");
        const double s = h.a * h.a / ((double)c->N * (double)c->N);
        for (auto &v : f) v *= s;
        if ((rc = gev_upload_field(c, GEV_BI, f.data()))) return rc;
        if ((rc = gev_fft_execute(c, GEV_PLAN_BI, GEV_FFT_FORWARD))) return rc;
    }
    // time stepping (ic_read.hpp:295-301, 707-708)
    const double N = (double)c->N;
    const double H = gev_hconf(h.a, p->fourpiG, &p->cosmo);
    st->a = h.a;
    st->tau = h.tau;
    st->dtau_old = h.dtau;
    st->dtau = p->Cf / N < p->steplimit / H ? p->Cf / N : p->steplimit / H;
    st->maxvel = h.maxvel;
    st->T00hom = 0.;
    st->phi_k0 = 0.;
    st->cycle = h.cycle + 1;
    return gev::check_hip(c, hipStreamSynchronize(c->stream), "restart sync");
}
