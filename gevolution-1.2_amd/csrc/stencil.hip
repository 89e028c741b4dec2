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

// F: the field element (double, or float in the FP32 form); the stencils
// evaluate in double either way.

// prepareFTsource (2), gevolution.hpp:169-192
template <typename F>
__global__ __launch_bounds__(TPB) void k_prepare_scalar(const F *__restrict__ phi, const F *__restrict__ chi,
                                                        F *source, int N, int64_t n3v, double bg, double coeff,
                                                        double coeff2, double coeff3) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n3v) return;
    const int64_t r = i / N;
    const Nb c = nb_of((int)(i % N), (int)(r % N), (int)(r / N), N);
    source[i] = (F)prep_scalar(phi, (double)chi[i], (double)source[i], c, i, N, bg, coeff, coeff2, coeff3);
}

// prepareFTsource (1), gevolution.hpp:56-147
template <typename F>
__global__ __launch_bounds__(TPB) void k_prepare_tensor(const F *__restrict__ phi, F *Sij, int N, int64_t n3v,
                                                        double coeff) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n3v) return;
    const int64_t r = i / N;
    const Nb c = nb_of((int)(i % N), (int)(r % N), (int)(r / N), N);
    if constexpr (sizeof(F) == 4) {
        // FP32 form: the float stencil the fused first FFT pass uses
        float T[6], S[6];
#pragma unroll
        for (int k = 0; k < 6; k++) T[k] = Sij[k * n3v + i];
        prep_tensor_f32(phi, T, c, i, N, (float)coeff, S);
#pragma unroll
        for (int k = 0; k < 6; k++) Sij[k * n3v + i] = S[k];
    } else {
        double T[6], S[6];
#pragma unroll
        for (int k = 0; k < 6; k++) T[k] = Sij[k * n3v + i];
        prep_tensor(phi, T, c, i, N, coeff, S);
#pragma unroll
        for (int k = 0; k < 6; k++) Sij[k * n3v + i] = (F)S[k];
    }
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

template <typename F>
__global__ __launch_bounds__(TPB) void k_sum_partial(const F *__restrict__ f, int64_t n,
                                                     double *__restrict__ part) {
    const int64_t base = (int64_t)blockIdx.x * TPB * SUM_ITEMS;
    double v = 0.;
    for (int k = 0; k < SUM_ITEMS; k++) {
        int64_t i = base + (int64_t)k * TPB + threadIdx.x;
        if (i < n) v += (double)f[i];
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
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n2; i += (int64_t)gridDim.x * TPB)
        p[i] = make_double2(0., 0.);
}

// element-wise copy between the two forms' field elements
template <typename S, typename D>
__global__ __launch_bounds__(TPB) void k_convert(const S *__restrict__ src, D *__restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) dst[i] = (D)src[i];
}

}  // namespace

namespace gev {

int launch_prepare_scalar(gev_ctx *ctx, double bg, double c1, double c2, double c3) {
    ScopedTimer t(ctx, "prepare_ft_source_scalar");
    if (ctx->fp32)
        hipLaunchKernelGGL(k_prepare_scalar<float>, dim3(grid_for(n3(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                           fld<float>(ctx->phi), fld<float>(ctx->chi), fld<float>(ctx->source), ctx->N, n3(ctx), bg,
                           c1, c2, c3);
    else
        hipLaunchKernelGGL(k_prepare_scalar<double>, dim3(grid_for(n3(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                           ctx->phi, ctx->chi, ctx->source, ctx->N, n3(ctx), bg, c1, c2, c3);
    return check_hip(ctx, hipGetLastError(), "prepare scalar launch");
}

int launch_prepare_tensor(gev_ctx *ctx, double coeff) {
    ScopedTimer t(ctx, "prepare_ft_source_tensor");
    if (ctx->fp32)
        hipLaunchKernelGGL(k_prepare_tensor<float>, dim3(grid_for(n3(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                           fld<float>(ctx->phi), fld<float>(ctx->Sij), ctx->N, n3(ctx), coeff);
    else
        hipLaunchKernelGGL(k_prepare_tensor<double>, dim3(grid_for(n3(ctx), TPB)), dim3(TPB), 0, ctx->stream,
                           ctx->phi, ctx->Sij, ctx->N, n3(ctx), coeff);
    return check_hip(ctx, hipGetLastError(), "prepare tensor launch");
}

// sums count field elements (double, or float in the FP32 form) into
// ctx->red_d[0] in FP64; copies to *out_host if non-null (sync)
int launch_field_sum(gev_ctx *ctx, const double *f, int64_t count, double *out_host) {
    const int64_t per = (int64_t)TPB * SUM_ITEMS;
    const int64_t nblk = (count + per - 1) / per;
    // partials live in the reduction scratch after slot 0
    {
        ScopedTimer t(ctx, "field_sum");
        if (ctx->fp32)
            hipLaunchKernelGGL(k_sum_partial<float>, dim3((unsigned)nblk), dim3(TPB), 0, ctx->stream,
                               fld<float>(f), count, ctx->red_d + 8);
        else
            hipLaunchKernelGGL(k_sum_partial<double>, dim3((unsigned)nblk), dim3(TPB), 0, ctx->stream, f, count,
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
    // grid-stride: a batch's slabs can exceed the 2^32 work-items of one launch dimension
    const int64_t blocks = (n2 + TPB - 1) / TPB;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)(blocks < (1 << 20) ? blocks : (1 << 20))), dim3(TPB), 0,
                       ctx->stream, (double2 *)ptr, n2);
    return check_hip(ctx, hipGetLastError(), "fill launch");
}

// n reals from src (float if src32, else double) to dst (float if dst32),
// rounded to nearest when narrowing; on ctx's stream
int launch_convert(gev_ctx *ctx, const void *src, bool src32, void *dst, bool dst32, int64_t n) {
    if (n <= 0) return 0;
    const dim3 g(grid_for(n, TPB)), b(TPB);
    if (src32 && dst32)
        hipLaunchKernelGGL((k_convert<float, float>), g, b, 0, ctx->stream, (const float *)src, (float *)dst, n);
    else if (src32)
        hipLaunchKernelGGL((k_convert<float, double>), g, b, 0, ctx->stream, (const float *)src, (double *)dst, n);
    else if (dst32)
        hipLaunchKernelGGL((k_convert<double, float>), g, b, 0, ctx->stream, (const double *)src, (float *)dst, n);
    else
        hipLaunchKernelGGL((k_convert<double, double>), g, b, 0, ctx->stream, (const double *)src, (double *)dst,
                           n);
    return check_hip(ctx, hipGetLastError(), "convert launch");
}

}  // namespace gev
