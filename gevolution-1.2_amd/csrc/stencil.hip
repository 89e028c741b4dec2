// stencil.hip -- real-space stencils of prepareFTsource (gevolution.hpp:56-192),
// the T00hom reduction (main.cpp:459-462) and the projection_init fill.
// One thread per cell; neighbours come through L2 (a 512^2 FP64 plane is
// 2 MiB, so the three planes a z-row of blocks touches stay cached).
// Expression order follows the reference literally (built with
// -ffp-contract=off), so the results equal a CPU evaluation bit for bit.
#include "gev_internal.h"

namespace {

constexpr int TPB = 256;

struct Cell {
    int x, y, z, xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ Cell cell_of(int64_t i, int N) {
    Cell c;
    c.x = (int)(i % N);
    int64_t r = i / N;
    c.y = (int)(r % N);
    c.z = (int)(r / N);
    c.xm = c.x == 0 ? N - 1 : c.x - 1;
    c.xp = c.x == N - 1 ? 0 : c.x + 1;
    c.ym = c.y == 0 ? N - 1 : c.y - 1;
    c.yp = c.y == N - 1 ? 0 : c.y + 1;
    c.zm = c.z == 0 ? N - 1 : c.z - 1;
    c.zp = c.z == N - 1 ? 0 : c.z + 1;
    return c;
}

__device__ __forceinline__ int64_t L(int N, int x, int y, int z) {
    return (int64_t)x + (int64_t)N * ((int64_t)y + (int64_t)N * z);
}

// prepareFTsource (2), gevolution.hpp:169-192, PHINONLINEAR on, ORIGINALMETRIC off
__global__ __launch_bounds__(TPB) void k_prepare_scalar(const double *__restrict__ phi,
                                                        const double *__restrict__ chi,
                                                        double *source, int N, int64_t n3v,
                                                        double bg, double coeff, double coeff2,
                                                        double coeff3) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n3v) return;
    Cell c = cell_of(i, N);
    const double p = phi[i];
    const double px = phi[L(N, c.xp, c.y, c.z)], mx = phi[L(N, c.xm, c.y, c.z)];
    const double py = phi[L(N, c.x, c.yp, c.z)], my = phi[L(N, c.x, c.ym, c.z)];
    const double pz = phi[L(N, c.x, c.y, c.zp)], mz = phi[L(N, c.x, c.y, c.zm)];
    double r = coeff2 * (source[i] - bg);
    r *= 1. - 2. * p;
    r += 0.125 * (mx - px) * (mx - px);
    r += 0.125 * (my - py) * (my - py);
    r += 0.125 * (mz - pz) * (mz - pz);
    r += (coeff3 - coeff) * p - coeff3 * chi[i];
    source[i] = r;
}

// prepareFTsource (1), gevolution.hpp:56-147, PHINONLINEAR on, ORIGINALMETRIC off
__global__ __launch_bounds__(TPB) void k_prepare_tensor(const double *__restrict__ phi,
                                                        double *Sij, int N, int64_t n3v,
                                                        double coeff) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n3v) return;
    Cell c = cell_of(i, N);
    const double p = phi[i];
    const double px = phi[L(N, c.xp, c.y, c.z)], mx = phi[L(N, c.xm, c.y, c.z)];
    const double py = phi[L(N, c.x, c.yp, c.z)], my = phi[L(N, c.x, c.ym, c.z)];
    const double pz = phi[L(N, c.x, c.y, c.zp)], mz = phi[L(N, c.x, c.y, c.zm)];
    const double pxy = phi[L(N, c.xp, c.yp, c.z)];
    const double pxz = phi[L(N, c.xp, c.y, c.zp)];
    const double pyz = phi[L(N, c.x, c.yp, c.zp)];
    double s;
    s = coeff * Sij[i];
    s += 0.5 * (px - mx) * (px - mx);
    Sij[i] = s;
    s = coeff * Sij[3 * n3v + i];
    s += 0.5 * (py - my) * (py - my);
    Sij[3 * n3v + i] = s;
    s = coeff * Sij[5 * n3v + i];
    s += 0.5 * (pz - mz) * (pz - mz);
    Sij[5 * n3v + i] = s;
    s = coeff * Sij[n3v + i];
    s += px * py - p * pxy;
    s += 0.5 * p * p;
    s -= 0.5 * px * px;
    s -= 0.5 * py * py;
    s += 0.5 * pxy * pxy;
    Sij[n3v + i] = s;
    s = coeff * Sij[2 * n3v + i];
    s += px * pz - p * pxz;
    s += 0.5 * p * p;
    s -= 0.5 * px * px;
    s -= 0.5 * pz * pz;
    s += 0.5 * pxz * pxz;
    Sij[2 * n3v + i] = s;
    s = coeff * Sij[4 * n3v + i];
    s += py * pz - p * pyz;
    s += 0.5 * p * p;
    s -= 0.5 * py * py;
    s -= 0.5 * pz * pz;
    s += 0.5 * pyz * pyz;
    Sij[4 * n3v + i] = s;
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
    const int64_t n2 = (int64_t)(bytes / 16);
    ScopedTimer t(ctx, "projection_init");
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n2, TPB)), dim3(TPB), 0, ctx->stream,
                       (double2 *)ptr, n2);
    return check_hip(ctx, hipGetLastError(), "fill launch");
}

}  // namespace gev
