// fft_core.h -- device building blocks of the LDS FFT passes: complex helpers,
// radix-2/3/4/8 butterflies and the Stockham auto-sort stage sequence, shared by
// the plain transform passes (fft.hip) and the passes fused with k-space solves
// (fftsolve.hip).  Every stage reads its operands to registers, synchronises,
// and writes them back, so a stage sequence leaves the transform in natural
// order in the same LDS buffer.  NT = threads of the workgroup.
#pragma once
#include <hip/hip_runtime.h>

namespace gev_fft {

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
// or contiguous lines (line w from w*xstride<L>()).
// Contiguous lines carry one padding element after every 16, so the
// strided stores of the Stockham stages spread over the LDS banks; a line
// then takes xstride<L>() elements.
template <int L> constexpr int xstride() { return L + L / 16; }
template <int L, int LINES, bool IL> __device__ __forceinline__ int eaddr(int line, int e) {
    return IL ? e * LINES + line : line * xstride<L>() + e + e / 16;
}

// One Stockham stage of radix R on LINES transforms of length L held in LDS,
// in place (all reads to registers, barrier, all writes).  NS is the product
// of the radices of the previous stages.  tw holds exp(-2 pi i t/L), t < L.
template <typename C, int L, int LINES, bool IL, int R, int NS, int NT>
__device__ __forceinline__ void stage(C *lds, const C *tw) {
    constexpr int M = L / R;
    constexpr int JOBS = LINES * M;
    constexpr int J = (JOBS + NT - 1) / NT;
    constexpr int TWS = L / (NS * R);
    C v[J][R];
#pragma unroll
    for (int t = 0; t < J; t++) {
        const int job = threadIdx.x + t * NT;
        if (JOBS % NT == 0 || job < JOBS) {
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
        const int job = threadIdx.x + t * NT;
        if (JOBS % NT == 0 || job < JOBS) {
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

template <typename C, int L, int LINES, bool IL, int NS, int NT = 256>
__device__ __forceinline__ void stages(C *lds, const C *tw) {
    if constexpr (NS < L) {
        constexpr int R = next_radix(L / NS);
        static_assert(R == 2 || R == 3 || R == 4 || R == 8, "FFT length must be 2^a 3^b");
        stage<C, L, LINES, IL, R, NS, NT>(lds, tw);
        stages<C, L, LINES, IL, NS * R, NT>(lds, tw);
    }
}

// lines of the x passes per workgroup: about 2048 complexes of LDS
template <int N> constexpr int x_lines() {
    int m = N / 2, l = 1;
    while (l * 2 * m <= 2048) l *= 2;
    return l;
}

// columns per workgroup of the strided passes: GEV_COL_BYTES-byte rows
// (default 128), at most 64 KiB
#ifndef GEV_COL_BYTES
#define GEV_COL_BYTES 128
#endif
#ifndef GEV_COL_BYTES_F32
#define GEV_COL_BYTES_F32 128
#endif
// threads per workgroup of the strided passes (FP64 / FP32 elements)
#ifndef GEV_COL_TPB
#define GEV_COL_TPB 256
#endif
#ifndef GEV_COL_TPB_F32
#define GEV_COL_TPB_F32 512
#endif
template <typename T> constexpr int col_tpb() { return sizeof(T) == 4 ? GEV_COL_TPB_F32 : GEV_COL_TPB; }
// RB: row bytes of a tile (0: the default of the form); the autotuner of a
// batch may pick 64 or 256 instead
template <typename T, int N, int RB = 0> constexpr int col_tile() {
    int w = (RB ? RB : (sizeof(T) == 4 ? GEV_COL_BYTES_F32 : GEV_COL_BYTES)) / (int)(2 * sizeof(T));
    while (w > 1 && w * N * (int)(2 * sizeof(T)) > 65536) w /= 2;
    return w;
}

}  // namespace gev_fft
