// probe.hip -- the device's achievable HBM bandwidth, measured with
// hand-written streaming kernels: the practical ceiling that the kernels'
// rooflines are read against (the published peak is 8 TB/s).
//
// Each kernel moves 16-byte vectors, U per thread with the loads of a thread
// issued before its stores, over buffers much larger than the caches.
#include "gev_internal.h"

namespace {

constexpr int PTPB = 256;
using V4 = unsigned int __attribute__((ext_vector_type(4)));  // 16 bytes

// copy: every byte read once and written once
template <int U, bool NT>
__global__ __launch_bounds__(PTPB) void k_probe_copy(const V4 *__restrict__ src, V4 *__restrict__ dst, int64_t n) {
    const int64_t i0 = (int64_t)blockIdx.x * PTPB * U + threadIdx.x;
    V4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t i = i0 + (int64_t)u * PTPB;
        if (i < n) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t i = i0 + (int64_t)u * PTPB;
        if (i < n) {
            if (NT) __builtin_nontemporal_store(v[u], dst + i);
            else dst[i] = v[u];
        }
    }
}

// read: every byte read once (an xor over each wavefront, one 16-byte store
// per wavefront)
template <int U>
__global__ __launch_bounds__(PTPB) void k_probe_read(const V4 *__restrict__ src, V4 *__restrict__ sink, int64_t n) {
    const int64_t i0 = (int64_t)blockIdx.x * PTPB * U + threadIdx.x;
    V4 acc = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t i = i0 + (int64_t)u * PTPB;
        if (i < n) {
            acc ^= src[i];
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        acc.x ^= (uint32_t)__shfl_xor((int)acc.x, off);
        acc.y ^= (uint32_t)__shfl_xor((int)acc.y, off);
        acc.z ^= (uint32_t)__shfl_xor((int)acc.z, off);
        acc.w ^= (uint32_t)__shfl_xor((int)acc.w, off);
    }
    if ((threadIdx.x & 63) == 0) sink[(int64_t)blockIdx.x * (PTPB / 64) + (threadIdx.x >> 6)] = acc;
}

// write: every byte written once
template <int U>
__global__ __launch_bounds__(PTPB) void k_probe_write(V4 *__restrict__ dst, int64_t n) {
    const int64_t i0 = (int64_t)blockIdx.x * PTPB * U + threadIdx.x;
    const V4 v = {1u, 2u, 3u, 4u};
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t i = i0 + (int64_t)u * PTPB;
        if (i < n) dst[i] = v;
    }
}

constexpr int U = 4;

void launch(int kind, const V4 *src, V4 *dst, V4 *sink, int64_t n, hipStream_t s) {
    const unsigned grid = (unsigned)((n + (int64_t)PTPB * U - 1) / ((int64_t)PTPB * U));
    switch (kind) {
    case GEV_PROBE_COPY: hipLaunchKernelGGL((k_probe_copy<U, false>), dim3(grid), dim3(PTPB), 0, s, src, dst, n); break;
    case GEV_PROBE_COPY_NT: hipLaunchKernelGGL((k_probe_copy<U, true>), dim3(grid), dim3(PTPB), 0, s, src, dst, n); break;
    case GEV_PROBE_READ: hipLaunchKernelGGL((k_probe_read<U>), dim3(grid), dim3(PTPB), 0, s, src, sink, n); break;
    default: hipLaunchKernelGGL((k_probe_write<U>), dim3(grid), dim3(PTPB), 0, s, dst, n); break;
    }
}

}  // namespace

GEV_EXPORT int gev_bandwidth_probe(int64_t bytes, int kind, int reps, double *gbs, double *ms) {
    if (bytes < (1 << 20) || reps < 1 || kind < GEV_PROBE_COPY || kind > GEV_PROBE_WRITE || !gbs || !ms)
        return GEV_EINVAL;
    const int64_t n = bytes / 16;
    V4 *src = nullptr, *dst = nullptr, *sink = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const unsigned grid = (unsigned)((n + (int64_t)PTPB * U - 1) / ((int64_t)PTPB * U));
    int rc = GEV_EHIP;
    if (hipMalloc(&src, n * 16) == hipSuccess && hipMalloc(&dst, n * 16) == hipSuccess &&
        hipMalloc(&sink, (size_t)grid * (PTPB / 64) * 16) == hipSuccess && hipStreamCreate(&s) == hipSuccess &&
        hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
        hipMemsetAsync(src, 1, n * 16, s) == hipSuccess) {
        launch(kind, src, dst, sink, n, s);  // untimed first launch
        float t = 0.f;
        if (hipEventRecord(e0, s) == hipSuccess) {
            for (int r = 0; r < reps; r++) launch(kind, src, dst, sink, n, s);
            if (hipGetLastError() == hipSuccess && hipEventRecord(e1, s) == hipSuccess &&
                hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&t, e0, e1) == hipSuccess && t > 0.f) {
                const double moved = (kind == GEV_PROBE_COPY || kind == GEV_PROBE_COPY_NT ? 2.0 : 1.0) * (double)n * 16.0;
                *ms = (double)t / reps;
                *gbs = moved / (*ms * 1e-3) / 1e9;
                rc = 0;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (sink) (void)hipFree(sink);
    return rc;
}
