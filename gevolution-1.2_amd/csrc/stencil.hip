// stencil.hip -- real-space stencils of prepareFTsource (gevolution.hpp:56-192),
// the T00hom reduction (main.cpp:459-462) and the projection_init fill.
// One thread per cell; neighbours come through L2 (a 512^2 FP64 plane is
// 2 MiB, so the three planes a z-row of blocks touches stay cached).
// Expression order follows the reference literally (built with
// -ffp-contract=off), so the results equal a CPU evaluation bit for bit.
#include "gev_internal.h"
#include "prep.h"

namespace {

using namespace gev_prep;
constexpr int TPB = 256;

// prepareFTsource (2), gevolution.hpp:169-192
__global__ __launch_bounds__(TPB) void k_prepare_scalar(const double *__restrict__ phi,
                                                        const double *__restrict__ chi,
                                                        double *source, int N, int64_t n3v,
                                                        double bg, double coeff, double coeff2,
                                                        double coeff3) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n3v) return;
    const int64_t r = i / N;
    const Nb c = nb_of((int)(i % N), (int)(r % N), (int)(r / N), N);
    source[i] = prep_scalar(phi, chi[i], source[i], c, i, N, bg, coeff, coeff2, coeff3);
}

// prepareFTsource (1), gevolution.hpp:56-147
__global__ __launch_bounds__(TPB) void k_prepare_tensor(const double *__restrict__ phi,
                                                        double *Sij, int N, int64_t n3v,
                                                        double coeff) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n3v) return;
    const int64_t r = i / N;
    const Nb c = nb_of((int)(i % N), (int)(r % N), (int)(r / N), N);
    double T[6], S[6];
#pragma unroll
    for (int k = 0; k < 6; k++) T[k] = Sij[k * n3v + i];
    prep_tensor(phi, T, c, i, N, coeff, S);
#pragma unroll
    for (int k = 0; k < 6; k++) Sij[k * n3v + i] = S[k];
}

// deterministic two-level sum: fixed per-block tree, then one block over
// the partials in index order
constexpr int SUM_ITEMS = 16;

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[TPB];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(TPB) void k_sum_partial(const double *__restrict__ f, int64_t n,
                                                     double *__restrict__ part) {
    const int64_t base = (int64_t)blockIdx.x * TPB * SUM_ITEMS;
    double v = 0.;
    for (int k = 0; k < SUM_ITEMS; k++) {
        int64_t i = base + (int64_t)k * TPB + threadIdx.x;
        if (i < n) v += f[i];
    }
    double r = block_sum(v);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ __launch_bounds__(TPB) void k_sum_final(const double *__restrict__ part, int64_t nblk,
                                                   double *__restrict__ out) {
    double v = 0.;
    for (int64_t i = threadIdx.x; i < nblk; i += TPB) v += part[i];
    double r = block_sum(v);
    if (threadIdx.x == 0) out[0] = r;
}

__global__ __launch_bounds__(TPB) void k_fill_zero(double2 *__restrict__ p, int64_t n2) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n2) p[i] = make_double2(0., 0.);
}

}  // namespace

namespace gev {

int launch_prepare_scalar(gev_ctx *ctx, double bg, double c1, double c2, double c3) {
    ScopedTimer t(ctx, "prepare_ft_source_scalar");
    hipLaunchKernelGGL(k_prepare_scalar, dim3(grid_for(n3(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                       ctx->phi, ctx->chi, ctx->source, ctx->N, n3(ctx), bg, c1, c2, c3);
    return check_hip(ctx, hipGetLastError(), "prepare scalar launch");
}

int launch_prepare_tensor(gev_ctx *ctx, double coeff) {
    ScopedTimer t(ctx, "prepare_ft_source_tensor");
    hipLaunchKernelGGL(k_prepare_tensor, dim3(grid_for(n3(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                       ctx->phi, ctx->Sij, ctx->N, n3(ctx), coeff);
    return check_hip(ctx, hipGetLastError(), "prepare tensor launch");
}

// sums count doubles into ctx->red_d[0]; copies to *out_host if non-null (sync)
int launch_field_sum(gev_ctx *ctx, const double *f, int64_t count, double *out_host) {
    const int64_t per = (int64_t)TPB * SUM_ITEMS;
    const int64_t nblk = (count + per - 1) / per;
    // partials live in the reduction scratch after slot 0
    {
        ScopedTimer t(ctx, "field_sum");
        hipLaunchKernelGGL(k_sum_partial, dim3((unsigned)nblk), dim3(TPB), 0, ctx->stream, f, count,
                           ctx->red_d + 8);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(TPB), 0, ctx->stream, ctx->red_d + 8, nblk,
                           ctx->red_d);
    }
    int rc = check_hip(ctx, hipGetLastError(), "field sum launch");
    if (rc) return rc;
    if (out_host) {
        rc = check_hip(ctx, hipMemcpyAsync(ctx->host_pinned, ctx->red_d, sizeof(double),
                                           hipMemcpyDeviceToHost, ctx->stream),
                       "field sum readback");
        if (rc) return rc;
        rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "field sum sync");
        if (rc) return rc;
        *out_host = ctx->host_pinned[0];
    }
    return 0;
}

int launch_fill_zero(gev_ctx *ctx, void *ptr, size_t bytes) {
    // all buffers are multiples of 16 bytes
    if (bytes % 16 != 0 || bytes == 0) return fail(ctx, GEV_EINVAL, "fill size must be a positive multiple of 16");
    const int64_t n2 = (int64_t)(bytes / 16);
    ScopedTimer t(ctx, "projection_init");
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n2, TPB)), dim3(TPB), 0, ctx->stream,
                       (double2 *)ptr, n2);
    return check_hip(ctx, hipGetLastError(), "fill launch");
}

}  // namespace gev
