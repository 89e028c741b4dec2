// kspace.hip -- Fourier-space solves of gevolution.hpp on the padded k layout
// [kz][ky][kx] (row pitch P, kx < N/2+1).  One thread per k point; every
// kernel reads each input component once and writes each output once.
// The gridk2 / kshift tables are computed on the host with the reference's
// own expressions (gevolution.hpp:222-227) and uploaded, and the complex
// arithmetic follows the reference's association order, so results equal a
// CPU evaluation of the same formulas bit for bit (built with
// -ffp-contract=off).
#include "gev_internal.h"
#include "ksolve.h"

namespace {

using namespace gev_ks;
constexpr int TPB = 256;

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

// C: the k-field element (double2, or float2 in the FP32 form); the solves
// evaluate in FP64 either way.

// solveModifiedPoissonFT, gevolution.hpp:501-535 (coeff already divided by -N^3)
template <typename C>
__global__ __launch_bounds__(TPB) void k_poisson(C *__restrict__ ft, const double *__restrict__ g, int P, int N,
                                                 int64_t nkv, double coeff, double modif) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nkv) return;
    KIdx k = kindex(i, P, N);
    if (!k.valid) return;
    const bool zero = k.kx == 0 && k.ky == 0 && k.kz == 0;
    ft[i] = from_d2<C>(poisson_point(d2(ft[i]), zero, zero ? 0. : g[k.kx] + g[k.ky] + g[k.kz], coeff, modif));
}

// MODE bit 1: projectFTscalar -> chi; bit 2: evolveFTvector -> BiFT
template <int MODE, typename C>
__global__ __launch_bounds__(TPB) void k_sij_solve(const C *__restrict__ SijFT,
                                                   C *__restrict__ chiFT,
                                                   C *__restrict__ BiFT,
                                                   const double *__restrict__ g,
                                                   const double2 *__restrict__ ks, int P, int N,
                                                   int64_t nkv, int add, double a2dtau) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nkv) return;
    KIdx k = kindex(i, P, N);
    if (!k.valid) return;
    const bool zero = k.kx == 0 && k.ky == 0 && k.kz == 0;
    const double g0 = g[k.kx], g1 = g[k.ky], g2 = g[k.kz];
    const double2 k0 = ks[k.kx], k1 = ks[k.ky], k2 = ks[k.kz];
    if constexpr (sizeof(C) == 8) {
        // FP32 form: the bodies the fused z pass uses (FP64 coefficients and
        // denominators, float field combinations)
        float2 S[6];
#pragma unroll
        for (int c = 0; c < 6; c++) S[c] = SijFT[c * nkv + i];
        if (MODE & 1) {
            if (zero) {
                chiFT[i] = make_float2(0.f, 0.f);
            } else {
                const float2 r = project_scalar_f32(S, g0, g1, g2, k0, k1, k2, N);
                chiFT[i] = add ? make_float2(chiFT[i].x + r.x, chiFT[i].y + r.y) : r;
            }
        }
        if (MODE & 2) {
            if (zero) {
                for (int c = 0; c < 3; c++) BiFT[c * nkv + i] = make_float2(0.f, 0.f);
            } else {
                float2 B[3] = {BiFT[i], BiFT[nkv + i], BiFT[2 * nkv + i]};
                evolve_vector_f32(S, B, g0, g1, g2, k0, k1, k2, a2dtau);
                for (int c = 0; c < 3; c++) BiFT[c * nkv + i] = B[c];
            }
        }
        return;
    }
    double2 S[6];
#pragma unroll
    for (int c = 0; c < 6; c++) S[c] = d2(SijFT[c * nkv + i]);
    if (MODE & 1) {
        if (zero) {
            chiFT[i] = from_d2<C>(c2(0., 0.));
        } else {
            double2 r = project_scalar(S, g0, g1, g2, k0, k1, k2, N);
            chiFT[i] = from_d2<C>(add ? cadd(d2(chiFT[i]), r) : r);
        }
    }
    if (MODE & 2) {
        if (zero) {
            for (int c = 0; c < 3; c++) BiFT[c * nkv + i] = from_d2<C>(c2(0., 0.));
        } else {
            double2 B[3] = {d2(BiFT[i]), d2(BiFT[nkv + i]), d2(BiFT[2 * nkv + i])};
            evolve_vector(S, B, g0, g1, g2, k0, k1, k2, a2dtau);
            for (int c = 0; c < 3; c++) BiFT[c * nkv + i] = from_d2<C>(B[c]);
        }
    }
}

// projectFTvector, gevolution.hpp:350-392 (in place on BiFT)
template <typename C>
__global__ __launch_bounds__(TPB) void k_project_vector(C *__restrict__ B,
                                                        const double *__restrict__ g,
                                                        const double2 *__restrict__ ks, int P,
                                                        int N, int64_t nkv, double coeff,
                                                        double modif) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nkv) return;
    KIdx k = kindex(i, P, N);
    if (!k.valid) return;
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) {
        for (int c = 0; c < 3; c++) B[c * nkv + i] = from_d2<C>(c2(0., 0.));
        return;
    }
    double2 S[3] = {d2(B[i]), d2(B[nkv + i]), d2(B[2 * nkv + i])};
    project_vector_point(S, ks[k.kx], ks[k.ky], ks[k.kz], g[k.kx] + g[k.ky] + g[k.kz], coeff, modif);
    B[i] = from_d2<C>(S[0]);
    B[nkv + i] = from_d2<C>(S[1]);
    B[2 * nkv + i] = from_d2<C>(S[2]);
}

template <int MODE> void sij_solve(gev_ctx *ctx, int add, double a2dtau) {
    if (ctx->fp32)
        hipLaunchKernelGGL((k_sij_solve<MODE, float2>), dim3(gev::grid_for(gev::nk(ctx), TPB)), dim3(TPB), 0,
                           ctx->stream, gev::cfld<float2>(ctx->SijFT), gev::cfld<float2>(ctx->scalarFT),
                           gev::cfld<float2>(ctx->BiFT), ctx->gridk2, ctx->kshift, ctx->P, ctx->N, gev::nk(ctx), add,
                           a2dtau);
    else
        hipLaunchKernelGGL((k_sij_solve<MODE, double2>), dim3(gev::grid_for(gev::nk(ctx), TPB)), dim3(TPB), 0,
                           ctx->stream, ctx->SijFT, ctx->scalarFT, ctx->BiFT, ctx->gridk2, ctx->kshift, ctx->P,
                           ctx->N, gev::nk(ctx), add, a2dtau);
}

}  // namespace

namespace gev {

int launch_poisson(gev_ctx *ctx, double coeff, double modif) {
    const int N = ctx->N;
    coeff /= -((long)N * (long)N * (long)N);  // gevolution.hpp:511
    ScopedTimer t(ctx, "solve_modified_poisson_ft");
    if (ctx->fp32)
        hipLaunchKernelGGL(k_poisson<float2>, dim3(grid_for(nk(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                           cfld<float2>(ctx->scalarFT), ctx->gridk2, ctx->P, N, nk(ctx), coeff, modif);
    else
        hipLaunchKernelGGL(k_poisson<double2>, dim3(grid_for(nk(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                           ctx->scalarFT, ctx->gridk2, ctx->P, N, nk(ctx), coeff, modif);
    return check_hip(ctx, hipGetLastError(), "poisson launch");
}

int launch_project_scalar(gev_ctx *ctx, int add) {
    ScopedTimer t(ctx, "project_ft_scalar");
    sij_solve<1>(ctx, add, 0.0);
    return check_hip(ctx, hipGetLastError(), "project scalar launch");
}

int launch_evolve_vector(gev_ctx *ctx, double a2dtau) {
    ScopedTimer t(ctx, "evolve_ft_vector");
    sij_solve<2>(ctx, 0, a2dtau);
    return check_hip(ctx, hipGetLastError(), "evolve vector launch");
}

int launch_project_scalar_evolve_vector(gev_ctx *ctx, double a2dtau) {
    ScopedTimer t(ctx, "project_scalar_evolve_vector");
    sij_solve<3>(ctx, 0, a2dtau);
    return check_hip(ctx, hipGetLastError(), "fused k solve launch");
}

int launch_project_vector(gev_ctx *ctx, double coeff, double modif) {
    ScopedTimer t(ctx, "project_ft_vector");
    if (ctx->fp32)
        hipLaunchKernelGGL(k_project_vector<float2>, dim3(grid_for(nk(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                           cfld<float2>(ctx->BiFT), ctx->gridk2, ctx->kshift, ctx->P, ctx->N, nk(ctx), coeff,
                           modif);
    else
        hipLaunchKernelGGL(k_project_vector<double2>, dim3(grid_for(nk(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                           ctx->BiFT, ctx->gridk2, ctx->kshift, ctx->P, ctx->N, nk(ctx), coeff, modif);
    return check_hip(ctx, hipGetLastError(), "project vector launch");
}

}  // namespace gev
