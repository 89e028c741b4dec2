// spectrum.hip -- binned (cross) power spectrum of k-space fields,
// extractCrossSpectrum of tools.hpp:53-212, on the device.
//
// The per-bin sums are deterministic (no floating-point atomics).  Along one
// k row (fixed ky, kz) |k| grows with kx, so the row's bin indices are
// non-decreasing: each row splits into runs of equal bin.  A workgroup walks
// its rows one after the other; for every row the first thread of each run
// adds the run's terms in kx order and folds the run total into the
// workgroup's LDS histogram (no two runs of a row share a bin).  Workgroup
// histograms go to global memory and a second kernel sums them per bin in
// workgroup order.  The per-bin means and scatters of the reference's root
// process are then formed on the host.
#include <cmath>
#include <mutex>
#include <vector>

#include "gev_internal.h"

namespace {

constexpr int TPB = 256;
constexpr int NQ = 5;  // kbin, kscatter, power, pscatter, occupation

// C: the k-field element (double2, or float2 in the FP32 form); the products
// and sums are FP64 either way
template <typename C> struct PkArgsT {
    const C *f1, *f2;
    int64_t nkv;
    int ncomp, mode;  // mode 0: components (comp1, comp2); 1: symmetric tensor; 2: sum over components
    int comp1, comp2;
    const double *typek2, *sinc;
    double k2max;
    int N, P, numbins;
    int64_t rows_per_block;
    double *partial;  // gridDim.x * numbins * NQ
};
using PkArgs = PkArgsT<double2>;

// real part of a * conj(b), as the reference's complex product evaluates it
__device__ __forceinline__ double pre(double2 a, double2 b) { return a.x * b.x - a.y * -b.y; }
__device__ __forceinline__ double pre(float2 a, float2 b) {
    return pre(make_double2(a.x, a.y), make_double2(b.x, b.y));
}

template <typename C> __device__ __forceinline__ double cross(const PkArgsT<C> &A, int64_t i) {
    if (A.mode == 0) return pre(A.f1[A.comp1 * A.nkv + i], A.f2[A.comp2 * A.nkv + i]);
    if (A.mode == 1) {
        double p = pre(A.f1[A.nkv + i], A.f2[A.nkv + i]);
        p += pre(A.f1[2 * A.nkv + i], A.f2[2 * A.nkv + i]);
        p += pre(A.f1[4 * A.nkv + i], A.f2[4 * A.nkv + i]);
        p *= 2.;
        p += pre(A.f1[i], A.f2[i]);
        p += pre(A.f1[3 * A.nkv + i], A.f2[3 * A.nkv + i]);
        p += pre(A.f1[5 * A.nkv + i], A.f2[5 * A.nkv + i]);
        return p;
    }
    double p = 0.;
    for (int c = 0; c < A.ncomp; c++) p += pre(A.f1[c * A.nkv + i], A.f2[c * A.nkv + i]);
    return p;
}

template <int MAXH, typename C>
__global__ __launch_bounds__(TPB) void k_spectrum_rows(PkArgsT<C> A) {
    extern __shared__ double hist[];  // numbins * NQ
    __shared__ int s_bin[MAXH];
    __shared__ double s_q[NQ - 1][MAXH];
    __shared__ int s_w[MAXH];
    const int N = A.N, H = N / 2;
    for (int t = threadIdx.x; t < A.numbins * NQ; t += TPB) hist[t] = 0.;
    const int64_t r0 = (int64_t)blockIdx.x * A.rows_per_block;
    const int64_t r1 = min(r0 + A.rows_per_block, (int64_t)N * N);
    for (int64_t r = r0; r < r1; r++) {
        const int ky = (int)(r % N), kz = (int)(r / N);
        __syncthreads();  // previous row's runs are folded
        for (int kx = threadIdx.x; kx <= H; kx += TPB) {
            int bin = -1;
            double q0 = 0., q1 = 0., q2 = 0., q3 = 0.;
            int weight = 0;
            if (!(kx == 0 && ky == 0 && kz == 0)) {
                weight = (kx == 0 || (kx == N / 2 && N % 2 == 0)) ? 1 : 2;
                const double k2 = A.typek2[kx] + A.typek2[ky] + A.typek2[kz];
                double s = A.sinc[kx] * A.sinc[ky] * A.sinc[kz];
                s *= s;
                const double p = cross(A, r * A.P + kx);
                const int i = (int)floor((double)((double)A.numbins * sqrt(k2 / A.k2max)));
                if (i < A.numbins) {
                    bin = i;
                    q0 = weight * sqrt(k2);
                    q1 = weight * k2;
                    q2 = weight * p * k2 * sqrt(k2) / s;
                    q3 = weight * p * p * k2 * k2 * k2 / s / s;
                }
            }
            s_bin[kx] = bin;
            s_q[0][kx] = q0;
            s_q[1][kx] = q1;
            s_q[2][kx] = q2;
            s_q[3][kx] = q3;
            s_w[kx] = bin >= 0 ? weight : 0;
        }
        __syncthreads();
        for (int kx = threadIdx.x; kx <= H; kx += TPB) {
            const int bin = s_bin[kx];
            if (bin < 0 || (kx > 0 && s_bin[kx - 1] == bin)) continue;  // not the head of a run
            double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
            int w = 0;
            for (int j = kx; j <= H && s_bin[j] == bin; j++) {
                a0 += s_q[0][j];
                a1 += s_q[1][j];
                a2 += s_q[2][j];
                a3 += s_q[3][j];
                w += s_w[j];
            }
            double *h = hist + (int64_t)bin * NQ;
            h[0] += a0;
            h[1] += a1;
            h[2] += a2;
            h[3] += a3;
            h[4] += (double)w;
        }
    }
    __syncthreads();
    double *out = A.partial + (int64_t)blockIdx.x * A.numbins * NQ;
    for (int t = threadIdx.x; t < A.numbins * NQ; t += TPB) out[t] = hist[t];
}

// sums the workgroup histograms per bin, in workgroup order
__global__ __launch_bounds__(TPB) void k_spectrum_reduce(const double *__restrict__ partial, int nblk, int numbins,
                                                         double *__restrict__ total) {
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= numbins * NQ) return;
    double s = 0.;
    for (int b = 0; b < nblk; b++) s += partial[(int64_t)b * numbins * NQ + t];
    total[t] = s;
}

}  // namespace

namespace gev {

// tools.hpp:66-104: typek2 and sinc tables, on the host (same expressions)
static void spectrum_tables(int N, int deconvolve, int ktype, double *tk, double *sn) {
    int i;
    if (ktype == 0) {
        for (i = 0; i < N; i++) {
            tk[i] = 2. * (double)N * sin(M_PI * (double)i / (double)N);
            tk[i] *= tk[i];
        }
    } else {
        for (i = 0; i <= N / 2; i++) {
            tk[i] = 2. * M_PI * (double)i;
            tk[i] *= tk[i];
        }
        for (; i < N; i++) {
            tk[i] = 2. * M_PI * (double)(N - i);
            tk[i] *= tk[i];
        }
    }
    sn[0] = 1.;
    for (i = 1; i <= N / 2; i++) sn[i] = deconvolve ? sin(M_PI * (float)i / (float)N) * (float)N / (M_PI * (float)i) : 1.;
    for (; i < N; i++) sn[i] = sn[N - i];
}

// workgroups of the row pass: about 1024, each a run of whole rows
static int spectrum_blocks(int N, int64_t *rows_per_block) {
    const int64_t rows = (int64_t)N * N;
    const int64_t rpb = (rows + 1023) / 1024;
    *rows_per_block = rpb;
    return (int)((rows + rpb - 1) / rpb);
}

int64_t spectrum_scratch_doubles(int N, int numbins) {
    int64_t rpb;
    const int nblk = spectrum_blocks(N, &rpb);
    return 2 * (int64_t)N + ((int64_t)nblk + 1) * numbins * NQ;
}

// The spectrum with its device scratch given: the two tables (2N doubles),
// the workgroup histograms and the bin totals, spectrum_scratch_doubles(N,
// numbins) doubles in all.  tk, sn, tot: host buffers of N, N and numbins * NQ
// doubles.  Copies are ordered on the context stream; synchronous.
static int spectrum_core(gev_ctx *c, const double2 *f1, const double2 *f2, int ncomp, int mode, int comp1,
                         int comp2, int numbins, int deconvolve, int ktype, double *scratch, double *tk, double *sn,
                         double *tot) {
    const int N = c->N;
    spectrum_tables(N, deconvolve, ktype, tk, sn);
    int64_t rpb;
    const int nblk = spectrum_blocks(N, &rpb);
    double *dtk = scratch, *dsn = scratch + N, *dpart = scratch + 2 * N;
    double *dtot = dpart + (int64_t)nblk * numbins * NQ;
    int rc = check_hip(c, hipMemcpyAsync(dtk, tk, N * 8, hipMemcpyHostToDevice, c->stream), "spectrum tables");
    if (!rc) rc = check_hip(c, hipMemcpyAsync(dsn, sn, N * 8, hipMemcpyHostToDevice, c->stream), "spectrum tables");
    if (rc) return rc;
    PkArgs A;
    A.f1 = f1;
    A.f2 = f2;
    A.nkv = nk(c);
    A.ncomp = ncomp;
    A.mode = mode;
    A.comp1 = comp1;
    A.comp2 = comp2;
    A.typek2 = dtk;
    A.sinc = dsn;
    A.k2max = 3. * tk[N / 2];
    A.N = N;
    A.P = c->P;
    A.numbins = numbins;
    A.rows_per_block = rpb;
    A.partial = dpart;
    {
        ScopedTimer t(c, "power_spectrum");
        const size_t shm = (size_t)numbins * NQ * 8;
        if (c->fp32) {
            // the same arguments over float2 fields
            PkArgsT<float2> B;
            B.f1 = cfld<float2>(A.f1);
            B.f2 = cfld<float2>(A.f2);
            B.nkv = A.nkv;
            B.ncomp = A.ncomp;
            B.mode = A.mode;
            B.comp1 = A.comp1;
            B.comp2 = A.comp2;
            B.typek2 = A.typek2;
            B.sinc = A.sinc;
            B.k2max = A.k2max;
            B.N = A.N;
            B.P = A.P;
            B.numbins = A.numbins;
            B.rows_per_block = A.rows_per_block;
            B.partial = A.partial;
            if (N / 2 + 1 <= 257)
                hipLaunchKernelGGL((k_spectrum_rows<257, float2>), dim3(nblk), dim3(TPB), shm, c->stream, B);
            else
                hipLaunchKernelGGL((k_spectrum_rows<513, float2>), dim3(nblk), dim3(TPB), shm, c->stream, B);
        } else if (N / 2 + 1 <= 257) {
            hipLaunchKernelGGL((k_spectrum_rows<257, double2>), dim3(nblk), dim3(TPB), shm, c->stream, A);
        } else {
            hipLaunchKernelGGL((k_spectrum_rows<513, double2>), dim3(nblk), dim3(TPB), shm, c->stream, A);
        }
        hipLaunchKernelGGL(k_spectrum_reduce, dim3((numbins * NQ + TPB - 1) / TPB), dim3(TPB), 0, c->stream,
                           (const double *)dpart, nblk, numbins, dtot);
    }
    rc = check_hip(c, hipGetLastError(), "spectrum launch");
    if (!rc) rc = check_hip(c, hipMemcpyAsync(tot, dtot, (size_t)numbins * NQ * 8, hipMemcpyDeviceToHost, c->stream),
                            "spectrum readback");
    if (!rc) rc = check_hip(c, hipStreamSynchronize(c->stream), "spectrum sync");
    return rc;
}

// the root process's moments (tools.hpp:180-194) from the bin totals
static void spectrum_moments(const double *tot, int numbins, double *kbin, double *power, double *kscatter,
                             double *pscatter, int32_t *occupation) {
    for (int b = 0; b < numbins; b++) {
        const double *t = &tot[(size_t)b * NQ];
        kbin[b] = t[0];
        kscatter[b] = t[1];
        power[b] = t[2];
        pscatter[b] = t[3];
        occupation[b] = (int32_t)t[4];
        if (occupation[b] > 0) {
            kscatter[b] = sqrt(kscatter[b] * occupation[b] - kbin[b] * kbin[b]) / occupation[b];
            if (!std::isfinite(kscatter[b])) kscatter[b] = 0.;
            kbin[b] = kbin[b] / occupation[b];
            power[b] /= occupation[b];
            pscatter[b] = sqrt(pscatter[b] / occupation[b] - power[b] * power[b]);
            if (!std::isfinite(pscatter[b])) pscatter[b] = 0.;
        }
    }
}

// The spectrum on the context's own scratch (allocated with the context for
// up to spectrum_bins_cap(N) bins; more bins take a scratch allocated for the
// call).
int spectrum(gev_ctx *c, const double2 *f1, const double2 *f2, int ncomp, int mode, int comp1, int comp2,
             int numbins, int deconvolve, int ktype, double *kbin, double *power, double *kscatter,
             double *pscatter, int32_t *occupation) {
    const int N = c->N;
    std::vector<double> tk(N), sn(N), tot((size_t)numbins * NQ);
    double *scratch = c->pk_scratch;
    void *temp = nullptr;
    if (numbins > c->pk_bins) {
        int rc = check_hip(c, hipMalloc(&temp, spectrum_scratch_doubles(N, numbins) * 8), "spectrum scratch");
        if (rc) return rc;
        scratch = (double *)temp;
    }
    int rc = spectrum_core(c, f1, f2, ncomp, mode, comp1, comp2, numbins, deconvolve, ktype, scratch, tk.data(),
                           sn.data(), tot.data());
    if (temp) (void)hipFree(temp);
    if (rc) return rc;
    spectrum_moments(tot.data(), numbins, kbin, power, kscatter, pscatter, occupation);
    return 0;
}

}  // namespace gev

// ---------------------------------------------------------------- static scratch

// Device scratch of gev_extract_power_spectrum_noalloc: the tables, the
// workgroup histograms and the bin totals for up to GEV_PK_NOALLOC_BINS bins
// at N <= 1024 (about 10.5 MB per device, held until the process ends).
// Every call on a device shares that device's copy.
#define GEV_PK_NOALLOC_BINS 256
#define GEV_PK_NOALLOC_N 1024
#define GEV_PK_NOALLOC_DOUBLES (2 * GEV_PK_NOALLOC_N + (1024 + 1) * GEV_PK_NOALLOC_BINS * 5)
static __device__ __align__(256) double scratch[GEV_PK_NOALLOC_DOUBLES];

namespace {

// the device address of `scratch` per device: looked up once per device, with
// that device current, on the first call there
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
double *scratch_ptr[MAX_DEVICES] = {};

int static_scratch(gev_ctx *c, double **out) {
    int dev = -1;
    int rc = gev::check_hip(c, hipGetDevice(&dev), "current device");
    if (rc) return rc;
    if (dev < 0 || dev >= MAX_DEVICES) return gev::fail(c, GEV_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        rc = gev::check_hip(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "static scratch address");
        if (rc) return rc;
        scratch_ptr[dev] = (double *)p;
    }
    *out = scratch_ptr[dev];
    return 0;
}

}  // namespace

// ---------------------------------------------------------------- C ABI

static int kfield(gev_ctx *c, gev_field f, const double2 **p, int *ncomp) {
    switch (f) {
    case GEV_SCALARFT: *p = c->scalarFT; *ncomp = 1; return 0;
    case GEV_SIJFT: *p = c->SijFT; *ncomp = 6; return 0;
    case GEV_BIFT: *p = c->BiFT; *ncomp = 3; return 0;
    default: return gev::fail(c, GEV_EINVAL, "a k-space field is required");
    }
}

GEV_EXPORT int gev_extract_cross_spectrum(gev_ctx *c, gev_field f1, gev_field f2, int32_t numbins,
                                          int32_t deconvolve, int32_t ktype, int32_t comp1, int32_t comp2,
                                          double *kbin, double *power, double *kscatter, double *pscatter,
                                          int32_t *occupation) {
    if (!c) return GEV_EINVAL;
    const double2 *a, *b;
    int na, nb;
    int rc = kfield(c, f1, &a, &na);
    if (!rc) rc = kfield(c, f2, &b, &nb);
    if (rc) return rc;
    if (numbins < 1 || numbins > 1024) return gev::fail(c, GEV_EINVAL, "numbins must be in [1, 1024]");
    if (!kbin || !power || !kscatter || !pscatter || !occupation) return gev::fail(c, GEV_EINVAL, "null output");
    int mode;
    if (comp1 >= 0 && comp2 >= 0 && comp1 < na && comp2 < nb) mode = 0;
    else if (na != nb) return gev::fail(c, GEV_EINVAL, "fields differ in components");
    else mode = na == 6 ? 1 : 2;
    return gev::spectrum(c, a, b, na, mode, comp1, comp2, numbins, deconvolve, ktype, kbin, power, kscatter,
                         pscatter, occupation);
}

GEV_EXPORT int gev_extract_power_spectrum(gev_ctx *c, gev_field f, int32_t numbins, int32_t deconvolve,
                                          int32_t ktype, double *kbin, double *power, double *kscatter,
                                          double *pscatter, int32_t *occupation) {
    return gev_extract_cross_spectrum(c, f, f, numbins, deconvolve, ktype, -1, -1, kbin, power, kscatter, pscatter,
                                      occupation);
}

GEV_EXPORT int gev_extract_power_spectrum_noalloc(gev_ctx *c, gev_field f, int32_t numbins, int32_t deconvolve,
                                                  int32_t ktype, double *kbin, double *power, double *kscatter,
                                                  double *pscatter, int32_t *occupation) {
    if (!c) return GEV_EINVAL;
    const double2 *a;
    int na;
    int rc = kfield(c, f, &a, &na);
    if (rc) return rc;
    if (numbins < 1 || numbins > GEV_PK_NOALLOC_BINS)
        return gev::fail(c, GEV_EINVAL, "numbins must be in [1, 256] for the allocation-free spectrum");
    if (c->N > GEV_PK_NOALLOC_N) return gev::fail(c, GEV_EINVAL, "the allocation-free spectrum takes N <= 1024");
    if (!kbin || !power || !kscatter || !pscatter || !occupation) return gev::fail(c, GEV_EINVAL, "null output");
    double *dev = nullptr;
    rc = static_scratch(c, &dev);
    if (rc) return rc;
    // host buffers on the stack: tables and bin totals
    double tk[GEV_PK_NOALLOC_N], sn[GEV_PK_NOALLOC_N], tot[GEV_PK_NOALLOC_BINS * 5];
    rc = gev::spectrum_core(c, a, a, na, na == 6 ? 1 : 2, -1, -1, numbins, deconvolve, ktype, dev, tk, sn, tot);
    if (rc) return rc;
    gev::spectrum_moments(tot, numbins, kbin, power, kscatter, pscatter, occupation);
    return 0;
}

GEV_EXPORT int gev_write_power_spectrum(const double *kbin, const double *power, const double *kscatter,
                                        const double *pscatter, const int32_t *occupation, int32_t numbins,
                                        double rescalek, double rescalep, const char *filename,
                                        const char *description, double a, double z_target) {
    if (!kbin || !power || !kscatter || !pscatter || !occupation || !filename || numbins < 1) return GEV_EINVAL;
    std::vector<double> power2(numbins);
    for (int i = 0; i < numbins; i++) power2[i] = power[i] / rescalep;
    // EXACT_OUTPUT_REDSHIFTS: interpolate with the spectrum written before
    // z_target was crossed (tools.hpp:273-311)
    if (1. / a < z_target + 1.) {
        FILE *in = fopen(filename, "r");
        double weight = 1.;
        if (in) {
            if (fscanf(in, "%*[^\n]\n") == 0 && fscanf(in, "# redshift z=%lf\n", &weight) == 1) {
                weight = (weight - z_target) / (1. + weight - 1. / a);
                if (fscanf(in, "%*[^\n]\n") == 0) {
                    for (int i = 0; i < numbins; i++)
                        if (occupation[i] > 0 && fscanf(in, " %*e %le %*e %*e %*d \n", &power2[i]) != 1) break;
                }
            } else {
                weight = 1.;
            }
            fclose(in);
            for (int i = 0; i < numbins; i++) power2[i] = (1. - weight) * power2[i] + weight * power[i] / rescalep;
            a = 1. / (z_target + 1.);
        }
    }
    FILE *out = fopen(filename, "w");
    if (!out) return GEV_EINVAL;
    fprintf(out, "# %s\n", description ? description : "");
    fprintf(out, "# redshift z=%f\n", (1. / a) - 1.);
    fprintf(out, "# k              Pk             sigma(k)       sigma(Pk)      count\n");
    for (int i = 0; i < numbins; i++)
        if (occupation[i] > 0)
            fprintf(out, "  %e   %e   %e   %e   %d\n", kbin[i] / rescalek, power2[i], kscatter[i] / rescalek,
                    pscatter[i] / rescalep / sqrt((double)occupation[i]), occupation[i]);
    fclose(out);
    return 0;
}
