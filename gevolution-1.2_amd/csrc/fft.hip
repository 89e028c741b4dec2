// fft.hip -- 3D real-to-complex / complex-to-real FFT for one box, replacing
// LATfield2's PlanFFT<Cplx>::execute over FFTW3 (main.cpp:238-246, 477, 488,
// 544, 563, 575, 593).  Unnormalised in both directions, forward sign -1, so
// backward(forward(f)) = N^3 f (manual.pdf sec. 4, p.14).
//
// Layout: real fields are N^3 (x fastest).  k fields are [kz][ky][kx] with kx
// in [0, N/2] stored in rows of pitch P = kpitch(N) complexes.
//
// A 3D transform is three passes, each one HBM read + one HBM write:
//   forward : r2c along x (real rows -> k rows)      k_r2c_x
//             c2c along y, in place                   k_c2c_cols
//             c2c along z, in place                   k_c2c_cols
//   backward: c2c along z (k field -> scratch)        k_c2c_cols
//             c2c along y, in place on scratch        k_c2c_cols
//             c2r along x (scratch -> real rows)      k_c2r_x
// so the backward transform leaves its k input intact (BiFT is state,
// main.cpp:586/593).
//
// Inside a pass a workgroup stages a tile in LDS and runs a Stockham
// auto-sort FFT with compile-time radix plan (8, then 4, 2, 3).  The x passes
// use the half-length complex transform: z[n] = x[2n] + i x[2n+1].  The
// strided passes load W consecutive columns (W complexes = 128 bytes per row
// in FP64) so every global access is a full cache line.
#include "gev_internal.h"

namespace {

constexpr int TPB = 256;

template <typename T> struct cplx_of;
template <> struct cplx_of<double> { using type = double2; };
template <> struct cplx_of<float> { using type = float2; };

template <typename C> __device__ __forceinline__ C mk2(decltype(C::x) a, decltype(C::x) b) {
    C r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mk2<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mk2<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return mk2<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return mk2<C>(a.x, -a.y); }
// multiply by -i
template <typename C> __device__ __forceinline__ C mul_mi(C a) { return mk2<C>(a.y, -a.x); }
// multiply by +i
template <typename C> __device__ __forceinline__ C mul_pi(C a) { return mk2<C>(-a.y, a.x); }

constexpr int next_radix(int n) {
    return n % 8 == 0 ? 8 : (n % 4 == 0 ? 4 : (n % 2 == 0 ? 2 : (n % 3 == 0 ? 3 : n)));
}

// forward DFT of R points in registers
template <typename C, int R> struct Butterfly;

template <typename C> struct Butterfly<C, 2> {
    __device__ __forceinline__ static void run(C *v) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename C> struct Butterfly<C, 4> {
    __device__ __forceinline__ static void run(C *v) {
        C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        C s13 = cadd(v[1], v[3]), d13 = mul_mi(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    }
};

template <typename C> struct Butterfly<C, 8> {
    __device__ __forceinline__ static void run(C *v) {
        using T = decltype(C::x);
        const T h = (T)0.70710678118654752440084436210484903928;
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        Butterfly<C, 4>::run(e);
        Butterfly<C, 4>::run(o);
        // W8^1 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = (-1 - i)/sqrt2
        C o1 = mk2<C>((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
        C o2 = mul_mi(o[2]);
        C o3 = mk2<C>((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <typename C> struct Butterfly<C, 3> {
    __device__ __forceinline__ static void run(C *v) {
        using T = decltype(C::x);
        const T s = (T)0.86602540378443864676372317075293618347;  // sin(2 pi/3)
        C sum = cadd(v[1], v[2]);
        C dif = csub(v[1], v[2]);
        C t = mk2<C>(v[0].x - (T)0.5 * sum.x, v[0].y - (T)0.5 * sum.y);
        C r = mk2<C>(s * dif.y, -s * dif.x);  // -i s (v1 - v2)
        v[0] = cadd(v[0], sum);
        v[1] = cadd(t, r);
        v[2] = csub(t, r);
    }
};

// LDS element address: interleaved tiles (element e of line w at e*LINES+w)
// or contiguous lines (line w at w*L).
template <int L, int LINES, bool IL> __device__ __forceinline__ int eaddr(int line, int e) {
    return IL ? e * LINES + line : line * L + e;
}

// One Stockham stage of radix R on LINES transforms of length L held in LDS,
// in place (all reads to registers, barrier, all writes).  NS is the product
// of the radices of the previous stages.  tw holds exp(-2 pi i t/L), t < L.
template <typename C, int L, int LINES, bool IL, int R, int NS>
__device__ __forceinline__ void stage(C *lds, const C *tw) {
    constexpr int M = L / R;
    constexpr int JOBS = LINES * M;
    constexpr int J = (JOBS + TPB - 1) / TPB;
    constexpr int TWS = L / (NS * R);
    C v[J][R];
#pragma unroll
    for (int t = 0; t < J; t++) {
        const int job = threadIdx.x + t * TPB;
        if (JOBS % TPB == 0 || job < JOBS) {
            const int line = IL ? job % LINES : job / M;
            const int j = IL ? job / LINES : job % M;
            const int k = j % NS;
#pragma unroll
            for (int r = 0; r < R; r++) v[t][r] = lds[eaddr<L, LINES, IL>(line, j + r * M)];
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[t][r] = cmul(v[t][r], tw[k * r * TWS]);
            }
            Butterfly<C, R>::run(v[t]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < J; t++) {
        const int job = threadIdx.x + t * TPB;
        if (JOBS % TPB == 0 || job < JOBS) {
            const int line = IL ? job % LINES : job / M;
            const int j = IL ? job / LINES : job % M;
            const int k = j % NS;
            const int d = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[eaddr<L, LINES, IL>(line, d + r * NS)] = v[t][r];
        }
    }
    __syncthreads();
}

template <typename C, int L, int LINES, bool IL, int NS>
__device__ __forceinline__ void stages(C *lds, const C *tw) {
    if constexpr (NS < L) {
        constexpr int R = next_radix(L / NS);
        static_assert(R == 2 || R == 3 || R == 4 || R == 8, "FFT length must be 2^a 3^b");
        stage<C, L, LINES, IL, R, NS>(lds, tw);
        stages<C, L, LINES, IL, NS * R>(lds, tw);
    }
}

// lines of the x passes per workgroup: about 2048 complexes of LDS
template <int N> constexpr int x_lines() {
    int m = N / 2, l = 1;
    while (l * 2 * m <= 2048) l *= 2;
    return l;
}

// ---------------------------------------------------------------- kernels

// r2c along x: nrows real rows of N -> nrows k rows (pitch P) of N/2+1.
template <typename T, int N>
__global__ __launch_bounds__(TPB) void k_r2c_x(const T *__restrict__ in,
                                               typename cplx_of<T>::type *__restrict__ out, int P,
                                               int64_t nrows,
                                               const typename cplx_of<T>::type *__restrict__ twN) {
    using C = typename cplx_of<T>::type;
    constexpr int M = N / 2;
    constexpr int LINES = x_lines<N>();
    __shared__ C buf[LINES * M];
    __shared__ C tw[M];
    const int64_t row0 = (int64_t)blockIdx.x * LINES;
    for (int t = threadIdx.x; t < M; t += TPB) tw[t] = twN[2 * t];
    {
        // all loads of this thread in flight before the first LDS write
        constexpr int PER = (LINES * M + TPB - 1) / TPB;
        C v[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int idx = threadIdx.x + k * TPB;
            const int line = idx / M, n = idx % M;
            const int64_t row = row0 + line;
            v[k] = mk2<C>(0, 0);
            if ((PER * TPB == LINES * M || idx < LINES * M) && row < nrows)
                v[k] = reinterpret_cast<const C *>(in + row * N)[n];
        }
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int idx = threadIdx.x + k * TPB;
            if (PER * TPB == LINES * M || idx < LINES * M) buf[idx] = v[k];
        }
    }
    __syncthreads();
    stages<C, M, LINES, false, 1>(buf, tw);
    for (int idx = threadIdx.x; idx < LINES * (M + 1); idx += TPB) {
        const int line = idx / (M + 1), k = idx % (M + 1);
        const int64_t row = row0 + line;
        if (row >= nrows) continue;
        C A = buf[line * M + (k == M ? 0 : k)];
        C B = cconj(buf[line * M + (k == 0 ? 0 : M - k)]);
        C E = mk2<C>((A.x + B.x) * (T)0.5, (A.y + B.y) * (T)0.5);
        C D = mk2<C>((A.x - B.x) * (T)0.5, (A.y - B.y) * (T)0.5);
        C O = mul_mi(D);
        out[row * P + k] = cadd(E, cmul(twN[k], O));
    }
}

// c2r along x: nrows k rows (pitch P) -> nrows real rows of N.  The imaginary
// parts of kx = 0 and kx = N/2 are ignored (FFTW / numpy c2r convention).
template <typename T, int N>
__global__ __launch_bounds__(TPB) void k_c2r_x(const typename cplx_of<T>::type *__restrict__ in,
                                               T *__restrict__ out, int P, int64_t nrows,
                                               const typename cplx_of<T>::type *__restrict__ twN) {
    using C = typename cplx_of<T>::type;
    constexpr int M = N / 2;
    constexpr int LINES = x_lines<N>();
    __shared__ C buf[LINES * M];
    __shared__ C tw[M];
    const int64_t row0 = (int64_t)blockIdx.x * LINES;
    for (int t = threadIdx.x; t < M; t += TPB) tw[t] = twN[2 * t];
    {
        // all loads of this thread in flight before the first LDS write
        constexpr int PER = (LINES * M + TPB - 1) / TPB;
        C xa[PER], xb[PER], wk[PER];
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int idx = threadIdx.x + j * TPB;
            const int line = idx / M, k = idx % M;
            const int64_t row = row0 + line;
            xa[j] = xb[j] = wk[j] = mk2<C>(0, 0);
            if ((PER * TPB == LINES * M || idx < LINES * M) && row < nrows) {
                xa[j] = in[row * P + k];
                xb[j] = in[row * P + (M - k)];
                wk[j] = twN[k];
            }
        }
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int idx = threadIdx.x + j * TPB;
            if (!(PER * TPB == LINES * M || idx < LINES * M)) continue;
            const int k = idx % M;
            C Xk = xa[j], Xm = xb[j];
            if (k == 0) {
                Xk.y = 0;
                Xm.y = 0;
            }
            C S = cadd(Xk, cconj(Xm));
            C D = csub(Xk, cconj(Xm));
            C Z = cadd(S, mul_pi(cmul(cconj(wk[j]), D)));
            buf[idx] = cconj(Z);  // inverse transform through the forward one
        }
    }
    __syncthreads();
    stages<C, M, LINES, false, 1>(buf, tw);
    for (int idx = threadIdx.x; idx < LINES * M; idx += TPB) {
        const int line = idx / M, n = idx % M;
        const int64_t row = row0 + line;
        if (row >= nrows) continue;
        C z = cconj(buf[idx]);
        reinterpret_cast<C *>(out + row * N)[n] = z;
    }
}

// columns per workgroup of the strided passes: 128-byte rows, at most 64 KiB
template <typename T, int N> constexpr int col_tile() {
    int w = 128 / (int)(2 * sizeof(T));
    while (w > 1 && w * N * (int)(2 * sizeof(T)) > 65536) w /= 2;
    return w;
}

// c2c along a strided axis.  Column c = (blockIdx.y, kx0 + w): elements at
// in[blockIdx.y*os + kx + e*es], e < N.  INV = backward (sign +1).
template <typename T, int N, bool INV>
__global__ __launch_bounds__(TPB) void k_c2c_cols(const typename cplx_of<T>::type *in,
                                                  typename cplx_of<T>::type *out, int H,
                                                  int64_t es, int64_t os,
                                                  const typename cplx_of<T>::type *__restrict__ twN) {
    using C = typename cplx_of<T>::type;
    constexpr int W = col_tile<T, N>();
    __shared__ C buf[N * W];
    __shared__ C tw[N];
    const int kx0 = blockIdx.x * W;
    const int64_t base = (int64_t)blockIdx.y * os + kx0;
    for (int t = threadIdx.x; t < N; t += TPB) tw[t] = twN[t];
    {
        // all loads of this thread in flight before the first LDS write; W
        // divides TPB, so the column (and its bounds check) is fixed per thread
        constexpr int PER = (N * W + TPB - 1) / TPB;
        const int w = threadIdx.x % W;
        const bool ok = kx0 + w < H;
        C v[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int idx = threadIdx.x + k * TPB;
            v[k] = mk2<C>(0, 0);
            if (ok && (PER * TPB == N * W || idx < N * W)) v[k] = in[base + (int64_t)(idx / W) * es + w];
        }
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int idx = threadIdx.x + k * TPB;
            if (PER * TPB == N * W || idx < N * W) buf[idx] = INV ? cconj(v[k]) : v[k];
        }
    }
    __syncthreads();
    stages<C, N, W, true, 1>(buf, tw);
    for (int idx = threadIdx.x; idx < N * W; idx += TPB) {
        const int w = idx % W, e = idx / W;
        if (kx0 + w < H) {
            C v = buf[idx];
            out[base + e * es + w] = INV ? cconj(v) : v;
        }
    }
}

template <typename T, int N>
int run_forward(gev_ctx *ctx, const T *real, typename cplx_of<T>::type *k, int ncomp,
                const typename cplx_of<T>::type *tw) {
    const int P = ctx->P, H = N / 2 + 1;
    const int64_t nrows = (int64_t)N * N;
    const int64_t n3v = (int64_t)N * N * N, nkv = (int64_t)P * N * N;
    constexpr int LINES = x_lines<N>();
    constexpr int W = col_tile<T, N>();
    for (int c = 0; c < ncomp; c++) {
        {
            gev::ScopedTimer t(ctx, "fft_r2c_x");
            hipLaunchKernelGGL((k_r2c_x<T, N>), dim3(gev::grid_for(nrows, LINES)), dim3(TPB), 0,
                               ctx->stream, real + c * n3v, k + c * nkv, P, nrows, tw);
        }
        {
            gev::ScopedTimer t(ctx, "fft_c2c_y");
            hipLaunchKernelGGL((k_c2c_cols<T, N, false>), dim3((H + W - 1) / W, N), dim3(TPB), 0,
                               ctx->stream, k + c * nkv, k + c * nkv, H, (int64_t)P,
                               (int64_t)P * N, tw);
        }
        {
            gev::ScopedTimer t(ctx, "fft_c2c_z");
            hipLaunchKernelGGL((k_c2c_cols<T, N, false>), dim3((H + W - 1) / W, N), dim3(TPB), 0,
                               ctx->stream, k + c * nkv, k + c * nkv, H, (int64_t)P * N,
                               (int64_t)P, tw);
        }
    }
    return gev::check_hip(ctx, hipGetLastError(), "fft forward launch");
}

template <typename T, int N>
int run_backward(gev_ctx *ctx, const typename cplx_of<T>::type *k, T *real, int ncomp,
                 typename cplx_of<T>::type *tmp, const typename cplx_of<T>::type *tw) {
    const int P = ctx->P, H = N / 2 + 1;
    const int64_t nrows = (int64_t)N * N;
    const int64_t n3v = (int64_t)N * N * N, nkv = (int64_t)P * N * N;
    constexpr int LINES = x_lines<N>();
    constexpr int W = col_tile<T, N>();
    for (int c = 0; c < ncomp; c++) {
        {
            gev::ScopedTimer t(ctx, "fft_c2c_z_inv");
            hipLaunchKernelGGL((k_c2c_cols<T, N, true>), dim3((H + W - 1) / W, N), dim3(TPB), 0,
                               ctx->stream, k + c * nkv, tmp, H, (int64_t)P * N, (int64_t)P, tw);
        }
        {
            gev::ScopedTimer t(ctx, "fft_c2c_y_inv");
            hipLaunchKernelGGL((k_c2c_cols<T, N, true>), dim3((H + W - 1) / W, N), dim3(TPB), 0,
                               ctx->stream, tmp, tmp, H, (int64_t)P, (int64_t)P * N, tw);
        }
        {
            gev::ScopedTimer t(ctx, "fft_c2r_x");
            hipLaunchKernelGGL((k_c2r_x<T, N>), dim3(gev::grid_for(nrows, LINES)), dim3(TPB), 0,
                               ctx->stream, tmp, real + c * n3v, P, nrows, tw);
        }
    }
    return gev::check_hip(ctx, hipGetLastError(), "fft backward launch");
}

#define GEV_FFT_SIZES(X) \
    X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

}  // namespace

namespace gev {

bool fft_size_supported(int N) {
    switch (N) {
#define CASE(n) case n:
        GEV_FFT_SIZES(CASE)
#undef CASE
        return true;
    default:
        return false;
    }
}

int fft_forward(gev_ctx *ctx, const double *real, double2 *kfield, int ncomp) {
    switch (ctx->N) {
#define CASE(n) \
    case n: return run_forward<double, n>(ctx, real, kfield, ncomp, ctx->twiddle);
        GEV_FFT_SIZES(CASE)
#undef CASE
    default:
        return fail(ctx, GEV_EINVAL, "unsupported FFT size");
    }
}

int fft_backward(gev_ctx *ctx, const double2 *kfield, double *real, int ncomp) {
    switch (ctx->N) {
#define CASE(n) \
    case n: return run_backward<double, n>(ctx, kfield, real, ncomp, ctx->fft_tmp, ctx->twiddle);
        GEV_FFT_SIZES(CASE)
#undef CASE
    default:
        return fail(ctx, GEV_EINVAL, "unsupported FFT size");
    }
}

}  // namespace gev
