// batch.hip -- many boxes of one shape stepped together: the main loop of
// main.cpp:372-879 for every box of a batch by the same kernel launches.
//
// A batch holds nbox boxes of the same grid, particle count, form and
// gravity (GR or Newtonian, parabolic B), each box one slab of device memory
// at a fixed stride, so every cycle kernel takes box b's arrays as box 0's
// shifted by b * stride and its coefficients from the box's BoxCycle.  Each
// launch covers all boxes: a batch of 64^3 boxes fills the GPU with the
// workgroups of many boxes where one box alone leaves most of it idle.
//
// The per-cycle host arithmetic of the single-box driver (the background
// step, the coefficients of every call) runs on the device: k_cycle_begin
// derives each box's coefficients from its state with the same expressions
// (the same functions, compiled for both sides with -ffp-contract=off) and
// k_cycle_end advances a, tau, dtau and the live particle set.  Nothing in a
// cycle needs the host, so the field stage and the particle stage are each
// captured once in a HIP graph and replayed every cycle.  The host keeps a
// mirror of every box's state, advanced with the same arithmetic, to decide
// where outputs fall without reading the device; the mirror is checked
// against the device at the end of every run.
//
// Outputs (the delta power spectrum at the redshifts z_pk, EXACT_OUTPUT_
// REDSHIFTS) are written between the two stages of the cycle that needs them,
// through the box's view (a single-box context over the box's slab), exactly
// as gev_run_to_outputs writes them; a box that has written its last output
// stops there and the others go on.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "gev_internal.h"

using namespace gev;

namespace {

constexpr int TPB = 64;

__device__ double next_dtau_dev(const BoxCycle &c, int N, double a) {
    // main.cpp:292-295 / 869-872
    const double dx = 1.0 / (double)N;
    const double H = hconf_of(a, c.fourpiG, c.Om_cdm, c.Om_b, c.Om_L, c.Om_rad, c.Om_fld, c.w0, c.wa);
    if (c.Cf * dx < c.steplimit / H) return c.Cf * dx;
    return c.steplimit / H;
}

// The coefficients of one cycle for every box that takes part, with the
// expressions of the single-box driver (api.hip: cycle_fields,
// cycle_particles, launch_deposit, launch_zsolve_poisson).
__global__ __launch_bounds__(TPB) void k_cycle_begin(BoxCycle *__restrict__ cyc, int nbox, int N) {
    const int b = blockIdx.x * TPB + threadIdx.x;
    if (b >= nbox) return;
    BoxCycle &c = cyc[b];
    if (!c.on) return;
    auto hc = [&](double x) { return hconf_of(x, c.fourpiG, c.Om_cdm, c.Om_b, c.Om_L, c.Om_rad, c.Om_fld, c.w0, c.wa); };
    const double dx = 1.0 / (double)N;
    const double a = c.a;
    double m = 1.0 / (dx * dx * dx);
    m *= c.pmass;
    c.dep_m = m;
    c.dep_m_a = m / a;
    if (c.gr) {
        c.on_phi = c.dtau_old > 0.;
        const double H = hc(a);
        c.prep_bg = c.Om_cdm + c.Om_b + 0.0;
        c.prep_c = c.on_phi ? 3. * H * dx * dx / c.dtau_old : 0.;
        c.prep_c2 = c.fourpiG * dx * dx / a;
        c.prep_c3 = 3. * H * H * dx * dx;
        c.pois_c0 = 1. / (dx * dx) / -((double)N * (double)N * (double)N);
        c.pois_c1 = c.on_phi ? 3. * hc(a) / c.dtau_old : 0.;
    } else {
        c.on_phi = 1;
        c.pois_c0 = c.fourpiG / a / -((double)N * (double)N * (double)N);
        c.pois_c1 = 0.;
    }
    c.tens_c = 2. * c.fourpiG * dx * dx / a;
    c.a2dtau = a * a * c.dtau_old;
    c.pk0 = a;
    c.pk1 = a * a * N;
    c.dtk = (c.dtau + c.dtau_old) / 2.;
    c.nfk = c.gr ? (1. / a < c.z_relax + 1. ? 3 : 2) : 1;
    double ah = a;
    rk4bg_with(ah, 0.5 * c.dtau, hc);
    c.a_half = ah;
    c.pd0 = ah;
    c.pd1 = ah * ah * N;
    c.nfd = c.gr ? (1. / ah < c.z_relax + 1. ? 3 : 0) : 0;
    c.dtd = c.dtau;
}

// The start of the particle stage: clears each box's kick maximum and filing
// list counter (a box that stops after its field stage keeps the maximum of
// its last kick, as the single-box driver does).
__global__ __launch_bounds__(TPB) void k_particle_begin(const BoxCycle *__restrict__ cyc, int nbox, int64_t stride,
                                                        unsigned long long *vmax0, uint32_t *sortn0) {
    const int b = blockIdx.x * TPB + threadIdx.x;
    if (b >= nbox || !cyc[b].on) return;
    *box_ptr(vmax0, stride, b) = 0ull;
    uint32_t *sn = box_ptr(sortn0, stride, b);
    sn[0] = sn[1] = 0u;
}

// The end of the cycle (main.cpp:814, 867-874): the second half RK4 step,
// tau, dtau_old, the next dtau, the cycle count, and the live particle set
// (the drift wrote the other one).
__global__ __launch_bounds__(TPB) void k_cycle_end(BoxCycle *__restrict__ cyc, int nbox, int N) {
    const int b = blockIdx.x * TPB + threadIdx.x;
    if (b >= nbox) return;
    BoxCycle &c = cyc[b];
    if (!c.on) return;
    auto hc = [&](double x) { return hconf_of(x, c.fourpiG, c.Om_cdm, c.Om_b, c.Om_L, c.Om_rad, c.Om_fld, c.w0, c.wa); };
    double a = c.a_half;
    rk4bg_with(a, 0.5 * c.dtau, hc);
    c.tau += c.dtau;
    c.dtau_old = c.dtau;
    c.a = a;
    c.dtau = next_dtau_dev(c, N, a);
    c.cycle++;
    c.parity ^= 1;
}

// the host's mirror of the same arithmetic
void host_advance(gev_cycle_state &st, const gev_sim_params &p, int N) {
    const gev_cosmology &cs = p.cosmo;
    auto hc = [&](double x) {
        return hconf_of(x, p.fourpiG, cs.Omega_cdm, cs.Omega_b, cs.Omega_Lambda, cs.Omega_rad, cs.Omega_fld,
                        cs.w0_fld, cs.wa_fld);
    };
    double a = st.a;
    rk4bg_with(a, 0.5 * st.dtau, hc);
    rk4bg_with(a, 0.5 * st.dtau, hc);
    st.tau += st.dtau;
    st.dtau_old = st.dtau;
    st.a = a;
    st.dtau = gev_next_dtau(&p, N, a);
    st.cycle++;
}

// a after one whole cycle from st (the look-ahead of main.cpp:660-662)
double host_look_ahead(const gev_cycle_state &st, const gev_sim_params &p) {
    const gev_cosmology &cs = p.cosmo;
    auto hc = [&](double x) {
        return hconf_of(x, p.fourpiG, cs.Omega_cdm, cs.Omega_b, cs.Omega_Lambda, cs.Omega_rad, cs.Omega_fld,
                        cs.w0_fld, cs.wa_fld);
    };
    double tmp = st.a;
    rk4bg_with(tmp, 0.5 * st.dtau, hc);
    rk4bg_with(tmp, 0.5 * st.dtau, hc);
    return tmp;
}

}  // namespace

// One box of a batch, as the host tracks it.
struct BoxHost {
    bool running = false;    // takes part in the cycles
    bool finished = false;   // wrote its last output and stopped
    gev_sim_params p{};
    gev_cycle_state st{};
    int parity = 0;          // live particle set (0: pos/q/id)
    double boxsize = 0.;
    int pkcount = 0;
    gev::PkWriter w;
    std::vector<double> k, pk, sk, spk;
    std::vector<int32_t> occ;
    std::vector<gev_pk_output_info> info;
};

struct gev_batch {
    gev_ctx *base = nullptr;         // the nbox slabs; its pointers are box 0's
    std::vector<gev_ctx *> views;    // the single-box view of every box
    std::vector<BoxHost> box;
    BoxCycle *cyc_d = nullptr;       // device per-box scalars
    BoxCycle *cyc_h = nullptr;       // pinned staging of the same
    int gr = 1;
    bool dirty = true;               // the device scalars need the host's
    // outputs, the same for every box
    std::vector<double> z_pk;
    int numbins = 0;
    // the two stages captured once each
    hipGraphExec_t g_fields = nullptr, g_particles = nullptr;
    // the autotuner's pick: row bytes of the y and z FFT pass tiles and the
    // pass times (ms per launch over every box) of each candidate
    bool tuned = false;
    float tune_ms[2][3] = {{0.f}};
    std::string err;
    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
};

namespace {

int batch_hip(gev_batch *B, hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    return B->fail(GEV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

// the view's particle arrays at the box's live set
void view_at_parity(gev_batch *B, int b) {
    gev_ctx *v = B->views[b];
    double *p1 = box_ptr(B->base->pos, B->base->stride, b), *p2 = box_ptr(B->base->pos2, B->base->stride, b);
    double *q1 = box_ptr(B->base->q, B->base->stride, b), *q2 = box_ptr(B->base->q2, B->base->stride, b);
    int64_t *i1 = box_ptr(B->base->id, B->base->stride, b), *i2 = box_ptr(B->base->id2, B->base->stride, b);
    const bool alt = B->box[b].parity != 0;
    v->pos = alt ? p2 : p1;
    v->pos2 = alt ? p1 : p2;
    v->q = alt ? q2 : q1;
    v->q2 = alt ? q1 : q2;
    v->id = alt ? i2 : i1;
    v->id2 = alt ? i1 : i2;
}

// A box's view may have been used between batch calls (a download, or an
// entry point that drifts and so switches the live particle set): the live
// set of every running box is taken from its view again.
void sync_views(gev_batch *B) {
    for (int b = 0; b < (int)B->box.size(); b++) {
        BoxHost &h = B->box[b];
        if (!h.running || h.finished) continue;
        const int par = B->views[b]->pos == box_ptr(B->base->pos, B->base->stride, b) ? 0 : 1;
        if (par != h.parity) {
            h.parity = par;
            B->dirty = true;
        }
    }
}

// the device scalars from the host's state of every box
int upload_scalars(gev_batch *B) {
    gev_ctx *c = B->base;
    for (int b = 0; b < c->nbox; b++) {
        BoxHost &h = B->box[b];
        BoxCycle &y = B->cyc_h[b];
        memset(&y, 0, sizeof y);
        y.on = h.running && !h.finished;
        if (!y.on) continue;
        y.a = h.st.a;
        y.tau = h.st.tau;
        y.dtau = h.st.dtau;
        y.dtau_old = h.st.dtau_old;
        y.cycle = h.st.cycle;
        y.parity = h.parity;
        y.gr = B->gr;
        y.fourpiG = h.p.fourpiG;
        y.Cf = h.p.Cf;
        y.steplimit = h.p.steplimit;
        y.z_relax = h.p.z_relax;
        y.pmass = B->views[b]->pmass;
        y.Om_cdm = h.p.cosmo.Omega_cdm;
        y.Om_b = h.p.cosmo.Omega_b;
        y.Om_L = h.p.cosmo.Omega_Lambda;
        y.Om_rad = h.p.cosmo.Omega_rad;
        y.Om_fld = h.p.cosmo.Omega_fld;
        y.w0 = h.p.cosmo.w0_fld;
        y.wa = h.p.cosmo.wa_fld;
    }
    int rc = batch_hip(B, hipMemcpyAsync(B->cyc_d, B->cyc_h, sizeof(BoxCycle) * c->nbox, hipMemcpyHostToDevice,
                                         c->stream), "batch scalars");
    if (!rc) B->dirty = false;
    return rc;
}

// the field stage of one cycle over every box (main.cpp:372-598)
int field_stage(gev_batch *B) {
    gev_ctx *c = B->base;
    const Batch all{c->stride, B->cyc_d, 0}, phi{c->stride, B->cyc_d, 1};
    hipLaunchKernelGGL(k_cycle_begin, dim3((c->nbox + TPB - 1) / TPB), dim3(TPB), 0, c->stream, B->cyc_d, c->nbox,
                       c->N);
    int rc = launch_deposit_batch(c, B->gr != 0, all);
    if (rc) return rc;
    if (B->gr) {
        rc = launch_x_prep_scalar(c, 0., 0., 0., 0., phi);
        if (!rc) rc = fft_forward_y(c, c->scalarFT, 1, phi);
        if (!rc) rc = launch_zsolve_poisson(c, c->scalarFT, c->fft_tmp, 0., 0., nullptr, nullptr, phi);
        if (!rc) rc = fft_backward_yx(c, c->fft_tmp, c->phi, 1, phi);
    } else {
        rc = fft_forward_xy(c, c->source, c->scalarFT, 1, phi);
        if (!rc) rc = launch_zsolve_poisson(c, c->scalarFT, c->fft_tmp, 0., 0., nullptr, nullptr, phi);
        if (!rc) rc = fft_backward_yx(c, c->fft_tmp, c->phi, 1, phi);
    }
    if (rc) return rc;
    double2 *const zout_chi = c->SijFT, *const zout_b = kcomp(c, c->SijFT, 1);
    rc = launch_x_prep_tensor(c, 0., all);
    if (!rc) rc = fft_forward_y(c, c->SijFT, 6, all);
    if (!rc) rc = launch_zsolve_tensor(c, c->SijFT, zout_chi, 0., true, nullptr, nullptr, all);
    if (!rc) rc = fft_backward_yx(c, zout_chi, c->chi, 1, all);
    if (!rc && B->gr) rc = fft_backward_yx(c, zout_b, c->Bi, 3, all);
    return rc ? rc : check_hip(c, hipGetLastError(), "batch field stage");
}

// the particle stage (main.cpp:766-874) over every box
int particle_stage(gev_batch *B) {
    gev_ctx *c = B->base;
    const Batch all{c->stride, B->cyc_d, 0};
    hipLaunchKernelGGL(k_particle_begin, dim3((c->nbox + TPB - 1) / TPB), dim3(TPB), 0, c->stream, B->cyc_d, c->nbox,
                       c->stride, reinterpret_cast<unsigned long long *>(c->red_u), c->sort_n);
    int rc = launch_particles_batch(c, B->gr != 0, all);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cycle_end, dim3((c->nbox + TPB - 1) / TPB), dim3(TPB), 0, c->stream, B->cyc_d, c->nbox, c->N);
    return check_hip(c, hipGetLastError(), "batch cycle end");
}

int capture(gev_batch *B, int (*stage)(gev_batch *), hipGraphExec_t *out) {
    gev_ctx *c = B->base;
    const bool timing = c->timing;
    c->timing = false;  // no timer events inside a graph
    hipGraph_t g = nullptr;
    int rc = batch_hip(B, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (!rc) {
        int rs = stage(B);
        rc = batch_hip(B, hipStreamEndCapture(c->stream, &g), "end capture");
        if (!rc && rs) rc = B->fail(rs, "batch stage capture: " + c->err);
    }
    if (!rc) rc = batch_hip(B, hipGraphInstantiate(out, g, nullptr, nullptr, 0), "graph instantiate");
    if (g) (void)hipGraphDestroy(g);
    c->timing = timing;
    return rc;
}

int run_stage(gev_batch *B, bool use_graph, bool fields) {
    if (!use_graph) {
        int rc = fields ? field_stage(B) : particle_stage(B);
        return rc ? B->fail(rc, B->base->err) : 0;
    }
    hipGraphExec_t &g = fields ? B->g_fields : B->g_particles;
    if (!g) {
        int rc = capture(B, fields ? field_stage : particle_stage, &g);
        if (rc) return rc;
    }
    return batch_hip(B, hipGraphLaunch(g, B->base->stream), "graph launch");
}

// ---------------------------------------------------------------- autotuner
//
// The strided FFT passes (four y and four z passes per cycle and component
// class) transform N-point columns staged W at a time; W (the tile's row
// bytes) trades the occupancy of a workgroup against whole cache lines per
// row, and the best choice depends on N and on how many boxes share the
// launch.  The tuner times each candidate tile (64-, 128- and 256-byte rows)
// on the batch's own FFT scratch field over every box, and keeps the
// fastest per axis.  Every column is transformed the same way whatever the
// tile, so the pick changes no result.  Picks are kept per (N, form, nbox)
// for the process.

constexpr int TUNE_RB[3] = {64, 128, 256};

std::mutex tune_mutex;
std::map<std::tuple<int, int, int>, std::pair<int, int>> tune_cache;

int autotune(gev_batch *B) {
    gev_ctx *c = B->base;
    B->tuned = true;
    if (c->N > 256) return 0;  // the tile variants exist for N <= 256
    const auto key = std::make_tuple(c->N, (int)c->fp32, c->nbox);
    {
        std::lock_guard<std::mutex> lock(tune_mutex);
        auto it = tune_cache.find(key);
        if (it != tune_cache.end()) {
            c->col_rb[0] = it->second.first;
            c->col_rb[1] = it->second.second;
            return 0;
        }
    }
    // every box on for the timing launches
    BoxCycle *on = nullptr;
    int rc = batch_hip(B, hipMalloc(&on, sizeof(BoxCycle) * c->nbox), "tuner flags");
    std::vector<BoxCycle> h(c->nbox);
    for (auto &y : h) {
        memset(&y, 0, sizeof y);
        y.on = y.on_phi = 1;
    }
    if (!rc) rc = batch_hip(B, hipMemcpy(on, h.data(), sizeof(BoxCycle) * c->nbox, hipMemcpyHostToDevice),
                            "tuner flags");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) rc = batch_hip(B, hipEventCreate(&e0), "tuner event");
    if (!rc) rc = batch_hip(B, hipEventCreate(&e1), "tuner event");
    const Batch all{c->stride, on, 0};
    const bool timing = c->timing;
    c->timing = false;
    int best[2] = {0, 0};
    for (int axis = 0; axis < 2 && !rc; axis++) {
        float best_ms = 0.f;
        for (int v = 0; v < 3 && !rc; v++) {
            c->col_rb[axis] = TUNE_RB[v];
            auto pass = [&]() {
                return axis == 0 ? fft_forward_y(c, c->fft_tmp, 1, all) : fft_forward_z(c, c->fft_tmp, 1, all);
            };
            rc = pass();  // first use of the variant, untimed
            if (!rc) rc = batch_hip(B, hipEventRecord(e0, c->stream), "tuner event");
            for (int r = 0; r < 3 && !rc; r++) rc = pass();
            if (!rc) rc = batch_hip(B, hipEventRecord(e1, c->stream), "tuner event");
            if (!rc) rc = batch_hip(B, hipEventSynchronize(e1), "tuner sync");
            float ms = 0.f;
            if (!rc) rc = batch_hip(B, hipEventElapsedTime(&ms, e0, e1), "tuner time");
            ms /= 3.f;
            B->tune_ms[axis][v] = ms;
            if (!rc && (v == 0 || ms < best_ms)) {
                best_ms = ms;
                best[axis] = TUNE_RB[v];
            }
        }
        c->col_rb[axis] = best[axis];
    }
    c->timing = timing;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (on) (void)hipFree(on);
    if (rc) {
        c->col_rb[0] = c->col_rb[1] = 0;
        return rc;
    }
    std::lock_guard<std::mutex> lock(tune_mutex);
    tune_cache[key] = {best[0], best[1]};
    return 0;
}

bool any_running(const gev_batch *B) {
    for (auto &h : B->box)
        if (h.running && !h.finished) return true;
    return false;
}

// one cycle of every running box; with outputs, every box's output checks
// between the two stages
int batch_cycle(gev_batch *B, bool use_graph, bool outputs) {
    gev_ctx *c = B->base;
    int rc = 0;
    if (!B->tuned && (rc = autotune(B))) return rc;
    sync_views(B);
    if (B->dirty && (rc = upload_scalars(B))) return rc;
    if ((rc = run_stage(B, use_graph, true))) return rc;
    bool stop_any = false;
    if (outputs) {
        const int num_pk = (int)B->z_pk.size();
        for (int b = 0; b < c->nbox && !rc; b++) {
            BoxHost &h = B->box[b];
            if (!h.running || h.finished) continue;
            const double *z = B->z_pk.data();
            auto write = [&](int j) {
                view_at_parity(B, b);
                int r = write_spectrum(B->views[b], &h.p, &h.st, h.w, j, z[j], h.k.data(), h.pk.data(), h.sk.data(),
                                       h.spk.data(), h.occ.data(), h.info.data());
                return r ? B->fail(r, B->views[b]->err) : 0;
            };
            // main.cpp:639-657, the look-ahead 659-677, the end 685-693
            if (h.pkcount < num_pk && 1. / h.st.a < z[h.pkcount] + 1.) {
                rc = write(h.pkcount);
                h.pkcount++;
            }
            const double tmp = host_look_ahead(h.st, h.p);
            if (!rc && h.pkcount < num_pk && 1. / tmp < z[h.pkcount] + 1.) rc = write(h.pkcount);
            if (!rc && h.pkcount >= num_pk) {
                h.finished = true;
                h.st.stage = 1;
                B->cyc_h[b].on = 0;
                stop_any = true;
            }
        }
        if (rc) return rc;
        if (stop_any) {
            // the finished boxes sit out the particle stage
            for (int b = 0; b < c->nbox && !rc; b++)
                if (B->box[b].finished && B->cyc_h[b].on == 0)
                    rc = batch_hip(B, hipMemcpyAsync(&B->cyc_d[b].on, &B->cyc_h[b].on, sizeof(int32_t),
                                                     hipMemcpyHostToDevice, c->stream), "stop box");
            if (rc) return rc;
        }
    }
    if ((rc = run_stage(B, use_graph, false))) return rc;
    for (int b = 0; b < c->nbox; b++) {
        BoxHost &h = B->box[b];
        if (!h.running || h.finished) continue;
        host_advance(h.st, h.p, c->N);
        h.parity ^= 1;
        view_at_parity(B, b);
    }
    return 0;
}

// the device's time stepping equals the host's (a, tau, dtau bit for bit)
int verify_scalars(gev_batch *B) {
    gev_ctx *c = B->base;
    std::vector<BoxCycle> d(c->nbox);
    sync_views(B);
    if (B->dirty) {
        // boxes started since the last cycle: the device takes their state first
        int ru = upload_scalars(B);
        if (ru) return ru;
    }
    int rc = batch_hip(B, hipMemcpyAsync(d.data(), B->cyc_d, sizeof(BoxCycle) * c->nbox, hipMemcpyDeviceToHost,
                                         c->stream), "batch scalars readback");
    if (!rc) rc = batch_hip(B, hipStreamSynchronize(c->stream), "batch sync");
    if (rc) return rc;
    for (int b = 0; b < c->nbox; b++) {
        const BoxHost &h = B->box[b];
        if (!h.running) continue;
        if (d[b].a != h.st.a || d[b].tau != h.st.tau || d[b].dtau != h.st.dtau || d[b].dtau_old != h.st.dtau_old ||
            d[b].cycle != h.st.cycle || d[b].parity != h.parity)
            return B->fail(GEV_ESTATE, "box " + std::to_string(b) + ": device time stepping differs from the host's");
    }
    return 0;
}

// maxvel of box b from its last kick's maximum (main.cpp:816-822)
int read_maxvel_box(gev_batch *B, int b) {
    gev_ctx *c = B->base;
    BoxHost &h = B->box[b];
    if (!h.running || h.st.cycle == 0) return 0;
    unsigned long long vb;
    int rc = batch_hip(B, hipMemcpyAsync(&vb, box_ptr(c->red_u, c->stride, b), sizeof vb, hipMemcpyDeviceToHost,
                                         c->stream), "vmax readback");
    if (!rc) rc = batch_hip(B, hipStreamSynchronize(c->stream), "vmax sync");
    if (rc) return rc;
    double mv;
    memcpy(&mv, &vb, sizeof mv);
    mv = sqrt(mv);
    if (h.p.gr_flag > 0) mv /= sqrt(mv * mv + 1.0);
    h.st.maxvel = mv;
    return 0;
}

int read_maxvel(gev_batch *B) {
    for (int b = 0; b < (int)B->box.size(); b++) {
        int rc = read_maxvel_box(B, b);
        if (rc) return rc;
    }
    return 0;
}

}  // namespace

// ---------------------------------------------------------------- C ABI

GEV_EXPORT int64_t gev_batch_footprint_bytes(const gev_batch_desc *d) {
    if (!d || d->nbox < 1 || d->npart < 0 || (d->form != GEV_FORM_FP64 && d->form != GEV_FORM_FP32)) return -1;
    return slab_bytes(d->ngrid, d->npart, d->form == GEV_FORM_FP32) * d->nbox + (int64_t)sizeof(BoxCycle) * d->nbox;
}

GEV_EXPORT int gev_batch_create(const gev_batch_desc *d, gev_batch **out) {
    if (!out) return GEV_EINVAL;
    *out = nullptr;
    if (!d || d->nbox < 1 || d->gr_flag < 0 || d->gr_flag > 1) return GEV_EINVAL;
    gev_batch *B = new (std::nothrow) gev_batch();
    if (!B) return GEV_ENOMEM;
    gev_box_desc bd{d->ngrid, d->form, d->npart, d->stream, 0, 0};
    int rc = create_ctx(&bd, d->nbox, &B->base);
    if (rc) {
        delete B;
        return rc;
    }
    gev_ctx *c = B->base;
    c->n = d->npart;  // every box holds npart particles when it runs
    B->gr = d->gr_flag;
    B->box.resize(d->nbox);
    B->views.assign(d->nbox, nullptr);
    bool ok = hipMalloc(&B->cyc_d, sizeof(BoxCycle) * d->nbox) == hipSuccess;
    if (!ok) B->cyc_d = nullptr;
    if (ok && hipHostMalloc(&B->cyc_h, sizeof(BoxCycle) * d->nbox) != hipSuccess) {
        B->cyc_h = nullptr;
        ok = false;
    }
    for (int b = 0; b < d->nbox && ok; b++) {
        gev_ctx *v = new (std::nothrow) gev_ctx();
        if (!v) {
            ok = false;
            break;
        }
        v->view = true;
        v->N = c->N;
        v->P = c->P;
        v->cap = c->cap;
        v->fp32 = c->fp32;
        v->stream = c->stream;
        v->own_stream = false;
        v->device = c->device;
        v->scan_nblk = c->scan_nblk;
        v->pk_bins = c->pk_bins;
        v->count_clean = true;
        v->host_pinned = c->host_pinned;
        v->rec_pinned = c->rec_pinned;
        assign_box_pointers(v, (char *)c->slab + (int64_t)b * c->stride);
        B->views[b] = v;
    }
    if (!ok) {
        gev_batch_destroy(B);
        return GEV_ENOMEM;
    }
    *out = B;
    return 0;
}

GEV_EXPORT int gev_batch_destroy(gev_batch *B) {
    if (!B) return GEV_EINVAL;
    if (B->base) (void)hipStreamSynchronize(B->base->stream);
    if (B->g_fields) (void)hipGraphExecDestroy(B->g_fields);
    if (B->g_particles) (void)hipGraphExecDestroy(B->g_particles);
    for (gev_ctx *v : B->views) delete v;
    if (B->cyc_d) (void)hipFree(B->cyc_d);
    if (B->cyc_h) (void)hipHostFree(B->cyc_h);
    if (B->base) {
        B->base->view = false;
        gev_ctx_destroy(B->base);
    }
    delete B;
    return 0;
}

GEV_EXPORT const char *gev_batch_last_error(const gev_batch *B) { return B ? B->err.c_str() : "null batch"; }

GEV_EXPORT int gev_batch_box(gev_batch *B, int32_t b, gev_ctx **view) {
    if (!B || !view || b < 0 || b >= (int32_t)B->views.size()) return GEV_EINVAL;
    *view = B->views[b];
    return 0;
}

GEV_EXPORT int gev_batch_start_box(gev_batch *B, int32_t b, const gev_sim_params *p, const gev_cycle_state *st,
                                   double boxsize) {
    if (!B || !p || !st || b < 0 || b >= (int32_t)B->views.size()) return GEV_EINVAL;
    gev_ctx *v = B->views[b];
    if (p->gr_flag != B->gr) return B->fail(GEV_EINVAL, "the box's gravity differs from the batch's");
    if (p->vector_flag != 0) return B->fail(GEV_EINVAL, "batches run the parabolic vector method");
    if (!(st->a > 0.) || !(st->dtau > 0.) || st->stage != 0) return B->fail(GEV_EINVAL, "state needs a, dtau > 0, stage 0");
    if (v->n != B->base->n) return B->fail(GEV_EINVAL, "every box of the batch holds npart particles");
    if (!v->count_clean) return B->fail(GEV_ESTATE, "the box's filing is not complete");
    BoxHost &h = B->box[b];
    h = BoxHost{};
    h.running = true;
    h.p = *p;
    h.st = *st;
    h.boxsize = boxsize;
    h.parity = v->pos == box_ptr(B->base->pos, B->base->stride, b) ? 0 : 1;
    v->zeroed[GEV_SOURCE] = v->zeroed[GEV_SIJ] = false;
    B->dirty = true;
    return 0;
}

GEV_EXPORT int gev_batch_stop_box(gev_batch *B, int32_t b) {
    if (!B || b < 0 || b >= (int32_t)B->views.size()) return GEV_EINVAL;
    B->box[b].running = false;
    B->box[b].finished = false;
    B->dirty = true;
    return 0;
}

GEV_EXPORT int gev_batch_get_state(gev_batch *B, int32_t b, gev_cycle_state *st) {
    if (!B || !st || b < 0 || b >= (int32_t)B->views.size()) return GEV_EINVAL;
    *st = B->box[b].st;
    return 0;
}

GEV_EXPORT int gev_batch_run_cycles(gev_batch *B, int32_t ncycles, int32_t use_graph) {
    if (!B || ncycles < 0) return GEV_EINVAL;
    for (auto &h : B->box)
        if (h.running && h.finished) return B->fail(GEV_ESTATE, "a finished box is stopped or started again first");
    for (int i = 0; i < ncycles && any_running(B); i++) {
        int rc = batch_cycle(B, use_graph != 0, false);
        if (rc) return rc;
    }
    int rc = verify_scalars(B);
    return rc ? rc : read_maxvel(B);
}

GEV_EXPORT int gev_batch_set_outputs(gev_batch *B, int32_t num_pk, const double *z_pk, int32_t numbins) {
    if (!B || num_pk < 1 || !z_pk || numbins < 1 || numbins > 1024) return GEV_EINVAL;
    for (int j = 1; j < num_pk; j++)
        if (!(z_pk[j] < z_pk[j - 1])) return B->fail(GEV_EINVAL, "z_pk must be strictly descending");
    B->z_pk.assign(z_pk, z_pk + num_pk);
    B->numbins = numbins;
    return 0;
}

GEV_EXPORT int gev_batch_step_outputs(gev_batch *B, int32_t use_graph, int32_t *finished) {
    if (!B || B->z_pk.empty()) return GEV_EINVAL;
    const int nbox = (int)B->box.size(), num_pk = (int)B->z_pk.size();
    for (int b = 0; b < nbox; b++) {
        BoxHost &h = B->box[b];
        if (h.running && !h.finished && h.w.num_pk != num_pk) {
            h.w.init(B->views[b], &h.p, num_pk, B->numbins, h.boxsize);
            const size_t n = (size_t)num_pk * B->numbins;
            h.k.assign(n, 0.);
            h.pk.assign(n, 0.);
            h.sk.assign(n, 0.);
            h.spk.assign(n, 0.);
            h.occ.assign(n, 0);
            h.info.assign(num_pk, gev_pk_output_info{});
        }
    }
    std::vector<bool> before(nbox);
    for (int b = 0; b < nbox; b++) before[b] = B->box[b].finished;
    int rc = any_running(B) ? batch_cycle(B, use_graph != 0, true) : 0;
    if (rc) return rc;
    for (int b = 0; b < nbox && !rc; b++)
        if (B->box[b].finished && !before[b]) rc = read_maxvel_box(B, b);
    if (rc) return rc;
    if (finished)
        for (int b = 0; b < nbox; b++) finished[b] = B->box[b].finished && !before[b];
    return 0;
}

GEV_EXPORT int gev_batch_run_to_outputs(gev_batch *B, int32_t use_graph, int32_t max_cycles) {
    if (!B || B->z_pk.empty() || max_cycles < 0) return GEV_EINVAL;
    for (int i = 0; any_running(B); i++) {
        if (i >= max_cycles) return B->fail(GEV_ESTATE, "max_cycles ran before every box wrote its outputs");
        int rc = gev_batch_step_outputs(B, use_graph, nullptr);
        if (rc) return rc;
    }
    int rc = verify_scalars(B);
    return rc ? rc : read_maxvel(B);
}

GEV_EXPORT int gev_batch_synchronize(gev_batch *B) {
    if (!B) return GEV_EINVAL;
    int rc = verify_scalars(B);
    return rc ? rc : read_maxvel(B);
}

GEV_EXPORT int gev_batch_autotune(gev_batch *B, int32_t *row_bytes_y, int32_t *row_bytes_z, float *pass_ms) {
    if (!B) return GEV_EINVAL;
    int rc = B->tuned ? 0 : autotune(B);
    if (rc) return rc;
    if (row_bytes_y) *row_bytes_y = B->base->col_rb[0];
    if (row_bytes_z) *row_bytes_z = B->base->col_rb[1];
    if (pass_ms) memcpy(pass_ms, B->tune_ms, sizeof B->tune_ms);
    return 0;
}

GEV_EXPORT int gev_batch_get_outputs(gev_batch *B, int32_t b, double *k, double *pk, double *sigma_k, double *sigma_pk,
                                     int32_t *occupation, gev_pk_output_info *info) {
    if (!B || b < 0 || b >= (int32_t)B->box.size()) return GEV_EINVAL;
    const BoxHost &h = B->box[b];
    if (!h.finished) return B->fail(GEV_ESTATE, "box " + std::to_string(b) + " has not written its outputs");
    const size_t n = h.k.size();
    if (k) memcpy(k, h.k.data(), n * 8);
    if (pk) memcpy(pk, h.pk.data(), n * 8);
    if (sigma_k) memcpy(sigma_k, h.sk.data(), n * 8);
    if (sigma_pk) memcpy(sigma_pk, h.spk.data(), n * 8);
    if (occupation) memcpy(occupation, h.occ.data(), n * 4);
    if (info) memcpy(info, h.info.data(), h.info.size() * sizeof(gev_pk_output_info));
    return 0;
}
