// kspace.hip -- Fourier-space solves of gevolution.hpp on the padded k layout
// [kz][ky][kx] (row pitch P, kx < N/2+1).  One thread per k point; every
// kernel reads each input component once and writes each output once.
// The gridk2 / kshift tables are computed on the host with the reference's
// own expressions (gevolution.hpp:222-227) and uploaded, and the complex
// arithmetic follows the reference's association order, so results equal a
// CPU evaluation of the same formulas bit for bit (built with
// -ffp-contract=off).
#include "gev_internal.h"

namespace {

constexpr int TPB = 256;

__device__ __forceinline__ double2 c2(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return c2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return c2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return c2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 rmul(double s, double2 a) { return c2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cmulr(double2 a, double s) { return c2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cdivr(double2 a, double s) { return c2(a.x / s, a.y / s); }
__device__ __forceinline__ double2 cconj(double2 a) { return c2(a.x, -a.y); }

struct KIdx {
    int kx, ky, kz;
    bool valid;
};

__device__ __forceinline__ KIdx kindex(int64_t i, int P, int N) {
    KIdx k;
    k.kx = (int)(i % P);
    int64_t r = i / P;
    k.ky = (int)(r % N);
    k.kz = (int)(r / N);
    k.valid = k.kx <= N / 2;
    return k;
}

// solveModifiedPoissonFT, gevolution.hpp:501-535 (coeff already divided by -N^3)
__global__ __launch_bounds__(TPB) void k_poisson(double2 *__restrict__ ft, const double *__restrict__ g,
                                                 int P, int N, int64_t nkv, double coeff, double modif) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nkv) return;
    KIdx k = kindex(i, P, N);
    if (!k.valid) return;
    double2 s = ft[i];
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) {
        ft[i] = modif == 0. ? c2(0., 0.) : cdivr(cmulr(s, coeff), modif);
    } else {
        ft[i] = cdivr(cmulr(s, coeff), g[k.kx] + g[k.ky] + g[k.kz] + modif);
    }
}

// projectFTscalar body, gevolution.hpp:251-260
__device__ __forceinline__ double2 project_scalar(const double2 S[6], double g0, double g1, double g2,
                                                  double2 k0, double2 k1, double2 k2, int N) {
    double2 num = rmul(g1 + g2 - 2. * g0, S[0]);
    num = cadd(num, rmul(g0 + g2 - 2. * g1, S[3]));
    num = cadd(num, rmul(g0 + g1 - 2. * g2, S[5]));
    num = csub(num, cmul(cmul(rmul(6., k0), k1), S[1]));
    num = csub(num, cmul(cmul(rmul(6., k0), k2), S[2]));
    num = csub(num, cmul(cmul(rmul(6., k1), k2), S[4]));
    double den = 2. * (g0 + g1 + g2) * (g0 + g1 + g2) * (double)N;
    return cdivr(num, den);
}

// evolveFTvector body, gevolution.hpp:314-325; adds onto B[3]
__device__ __forceinline__ void evolve_vector(const double2 S[6], double2 B[3], double g0, double g1,
                                              double g2, double2 k0, double2 k1, double2 k2,
                                              double a2dtau) {
    double k4 = g0 + g1 + g2;
    k4 *= k4;
    const double2 pre = c2(0., -2. * a2dtau / k4);
    double2 t, u;
    t = rmul(g1 + g2, S[0]);
    t = csub(t, rmul(g1, S[3]));
    t = csub(t, rmul(g2, S[5]));
    t = csub(t, cmul(cmul(rmul(2., k1), k2), S[4]));
    t = cmul(cconj(k0), t);
    u = cadd(cmul(k1, S[1]), cmul(k2, S[2]));
    t = cadd(t, rmul(g1 + g2 - g0, u));
    B[0] = cadd(B[0], cmul(pre, t));

    t = rmul(g0 + g2, S[3]);
    t = csub(t, rmul(g0, S[0]));
    t = csub(t, rmul(g2, S[5]));
    t = csub(t, cmul(cmul(rmul(2., k0), k2), S[2]));
    t = cmul(cconj(k1), t);
    u = cadd(cmul(k0, S[1]), cmul(k2, S[4]));
    t = cadd(t, rmul(g0 + g2 - g1, u));
    B[1] = cadd(B[1], cmul(pre, t));

    t = rmul(g0 + g1, S[5]);
    t = csub(t, rmul(g0, S[0]));
    t = csub(t, rmul(g1, S[3]));
    t = csub(t, cmul(cmul(rmul(2., k0), k1), S[1]));
    t = cmul(cconj(k2), t);
    u = cadd(cmul(k0, S[2]), cmul(k1, S[4]));
    t = cadd(t, rmul(g0 + g1 - g2, u));
    B[2] = cadd(B[2], cmul(pre, t));
}

// MODE bit 1: projectFTscalar -> chi; bit 2: evolveFTvector -> BiFT
template <int MODE>
__global__ __launch_bounds__(TPB) void k_sij_solve(const double2 *__restrict__ SijFT,
                                                   double2 *__restrict__ chiFT,
                                                   double2 *__restrict__ BiFT,
                                                   const double *__restrict__ g,
                                                   const double2 *__restrict__ ks, int P, int N,
                                                   int64_t nkv, int add, double a2dtau) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nkv) return;
    KIdx k = kindex(i, P, N);
    if (!k.valid) return;
    const bool zero = k.kx == 0 && k.ky == 0 && k.kz == 0;
    double2 S[6];
#pragma unroll
    for (int c = 0; c < 6; c++) S[c] = SijFT[c * nkv + i];
    const double g0 = g[k.kx], g1 = g[k.ky], g2 = g[k.kz];
    const double2 k0 = ks[k.kx], k1 = ks[k.ky], k2 = ks[k.kz];
    if (MODE & 1) {
        if (zero) {
            chiFT[i] = c2(0., 0.);
        } else {
            double2 r = project_scalar(S, g0, g1, g2, k0, k1, k2, N);
            chiFT[i] = add ? cadd(chiFT[i], r) : r;
        }
    }
    if (MODE & 2) {
        if (zero) {
            for (int c = 0; c < 3; c++) BiFT[c * nkv + i] = c2(0., 0.);
        } else {
            double2 B[3] = {BiFT[i], BiFT[nkv + i], BiFT[2 * nkv + i]};
            evolve_vector(S, B, g0, g1, g2, k0, k1, k2, a2dtau);
            for (int c = 0; c < 3; c++) BiFT[c * nkv + i] = B[c];
        }
    }
}

// projectFTvector, gevolution.hpp:350-392 (in place on BiFT)
__global__ __launch_bounds__(TPB) void k_project_vector(double2 *__restrict__ B,
                                                        const double *__restrict__ g,
                                                        const double2 *__restrict__ ks, int P,
                                                        int N, int64_t nkv, double coeff,
                                                        double modif) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nkv) return;
    KIdx k = kindex(i, P, N);
    if (!k.valid) return;
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) {
        for (int c = 0; c < 3; c++) B[c * nkv + i] = c2(0., 0.);
        return;
    }
    const double2 S0 = B[i], S1 = B[nkv + i], S2 = B[2 * nkv + i];
    const double2 k0 = ks[k.kx], k1 = ks[k.ky], k2v = ks[k.kz];
    const double k2 = g[k.kx] + g[k.ky] + g[k.kz];
    double2 tmp = cadd(cadd(cmul(k0, S0), cmul(k1, S1)), cmul(k2v, S2));
    tmp = cdivr(tmp, k2);
    B[i] = cdivr(cmulr(cmulr(csub(S0, cmul(cconj(k0), tmp)), 4.), coeff), k2 + modif);
    B[nkv + i] = cdivr(cmulr(cmulr(csub(S1, cmul(cconj(k1), tmp)), 4.), coeff), k2 + modif);
    B[2 * nkv + i] = cdivr(cmulr(cmulr(csub(S2, cmul(cconj(k2v), tmp)), 4.), coeff), k2 + modif);
}

}  // namespace

namespace gev {

int launch_poisson(gev_ctx *ctx, double coeff, double modif) {
    const int N = ctx->N;
    coeff /= -((long)N * (long)N * (long)N);  // gevolution.hpp:511
    ScopedTimer t(ctx, "solve_modified_poisson_ft");
    hipLaunchKernelGGL(k_poisson, dim3(grid_for(nk(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                       ctx->scalarFT, ctx->gridk2, ctx->P, N, nk(ctx), coeff, modif);
    return check_hip(ctx, hipGetLastError(), "poisson launch");
}

int launch_project_scalar(gev_ctx *ctx, int add) {
    ScopedTimer t(ctx, "project_ft_scalar");
    hipLaunchKernelGGL(k_sij_solve<1>, dim3(grid_for(nk(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                       ctx->SijFT, ctx->scalarFT, ctx->BiFT, ctx->gridk2, ctx->kshift, ctx->P,
                       ctx->N, nk(ctx), add, 0.0);
    return check_hip(ctx, hipGetLastError(), "project scalar launch");
}

int launch_evolve_vector(gev_ctx *ctx, double a2dtau) {
    ScopedTimer t(ctx, "evolve_ft_vector");
    hipLaunchKernelGGL(k_sij_solve<2>, dim3(grid_for(nk(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                       ctx->SijFT, ctx->scalarFT, ctx->BiFT, ctx->gridk2, ctx->kshift, ctx->P,
                       ctx->N, nk(ctx), 0, a2dtau);
    return check_hip(ctx, hipGetLastError(), "evolve vector launch");
}

int launch_project_scalar_evolve_vector(gev_ctx *ctx, double a2dtau) {
    ScopedTimer t(ctx, "project_scalar_evolve_vector");
    hipLaunchKernelGGL(k_sij_solve<3>, dim3(grid_for(nk(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                       ctx->SijFT, ctx->scalarFT, ctx->BiFT, ctx->gridk2, ctx->kshift, ctx->P,
                       ctx->N, nk(ctx), 0, a2dtau);
    return check_hip(ctx, hipGetLastError(), "fused k solve launch");
}

int launch_project_vector(gev_ctx *ctx, double coeff, double modif) {
    ScopedTimer t(ctx, "project_ft_vector");
    hipLaunchKernelGGL(k_project_vector, dim3(grid_for(nk(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                       ctx->BiFT, ctx->gridk2, ctx->kshift, ctx->P, ctx->N, nk(ctx), coeff, modif);
    return check_hip(ctx, hipGetLastError(), "project vector launch");
}

}  // namespace gev
