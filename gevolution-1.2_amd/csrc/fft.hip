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
#include "fft_core.h"

namespace {

using namespace gev_fft;
constexpr int TPB = 256;

// ---------------------------------------------------------------- kernels

// r2c along x: nrows real rows of N -> nrows k rows (pitch P) of N/2+1.
template <typename T, int N>
__global__ __launch_bounds__(TPB) void k_r2c_x(const T *__restrict__ in,
                                               typename cplx_of<T>::type *__restrict__ out, int P,
                                               int64_t nrows,
                                               const typename cplx_of<T>::type *__restrict__ twN,
                                               int64_t in_cs, int64_t out_cs, gev::Batch bt) {
    using C = typename cplx_of<T>::type;
    if (!gev::box_on(bt, blockIdx.z)) return;
    in += blockIdx.z * in_cs;  // component (or box of a batch) blockIdx.z
    out += blockIdx.z * out_cs;
    constexpr int M = N / 2;
    constexpr int LINES = x_lines<N>();
    __shared__ C buf[LINES * xstride<M>()];
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
            if (PER * TPB == LINES * M || idx < LINES * M) buf[eaddr<M, LINES, false>(idx / M, idx % M)] = v[k];
        }
    }
    __syncthreads();
    stages<C, M, LINES, false, 1>(buf, tw);
    for (int idx = threadIdx.x; idx < LINES * (M + 1); idx += TPB) {
        const int line = idx / (M + 1), k = idx % (M + 1);
        const int64_t row = row0 + line;
        if (row >= nrows) continue;
        C A = buf[eaddr<M, LINES, false>(line, k == M ? 0 : k)];
        C B = cconj(buf[eaddr<M, LINES, false>(line, k == 0 ? 0 : M - k)]);
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
                                               const typename cplx_of<T>::type *__restrict__ twN,
                                               int64_t in_cs, int64_t out_cs, gev::Batch bt) {
    using C = typename cplx_of<T>::type;
    if (!gev::box_on(bt, blockIdx.z)) return;
    in += blockIdx.z * in_cs;  // component (or box of a batch) blockIdx.z
    out += blockIdx.z * out_cs;
    constexpr int M = N / 2;
    constexpr int LINES = x_lines<N>();
    __shared__ C buf[LINES * xstride<M>()];
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
            buf[eaddr<M, LINES, false>(idx / M, k)] = cconj(Z);  // inverse transform through the forward one
        }
    }
    __syncthreads();
    stages<C, M, LINES, false, 1>(buf, tw);
    for (int idx = threadIdx.x; idx < LINES * M; idx += TPB) {
        const int line = idx / M, n = idx % M;
        const int64_t row = row0 + line;
        if (row >= nrows) continue;
        C z = cconj(buf[eaddr<M, LINES, false>(line, n)]);
        reinterpret_cast<C *>(out + row * N)[n] = z;
    }
}

// c2c along a strided axis.  Column c = (blockIdx.y, kx0 + w): elements at
// in[blockIdx.y*os + kx + e*es], e < N.  INV = backward (sign +1).  RB: the
// tile's row bytes (0: the form's default; every column is transformed the
// same way whatever the tile).
template <typename T, int N, bool INV, int RB = 0>
__global__ __launch_bounds__(col_tpb<T>()) void k_c2c_cols(const typename cplx_of<T>::type *in,
                                                  typename cplx_of<T>::type *out, int H,
                                                  int64_t es, int64_t os,
                                                  const typename cplx_of<T>::type *__restrict__ twN,
                                                  int64_t cs, gev::Batch bt) {
    using C = typename cplx_of<T>::type;
    constexpr int TPB = col_tpb<T>();
    if (!gev::box_on(bt, blockIdx.z)) return;
    in += blockIdx.z * cs;  // component (or box of a batch) blockIdx.z
    out += blockIdx.z * cs;
    constexpr int W = col_tile<T, N, RB>();
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
    stages<C, N, W, true, 1, TPB>(buf, tw);
    for (int idx = threadIdx.x; idx < N * W; idx += TPB) {
        const int w = idx % W, e = idx / W;
        if (kx0 + w < H) {
            C v = buf[idx];
            out[base + e * es + w] = INV ? cconj(v) : v;
        }
    }
}

// Timer names of a pass launched over n components (n <= MAXC): the plain
// name for one, name[n] for several (the per-launch byte count scales by n).
constexpr int MAXC = 6;
enum { P_R2C_X, P_C2C_Y, P_C2C_Z, P_C2C_Z_INV, P_C2C_Y_INV, P_C2R_X };
const char *pass_name(int pass, int n) {
#define GEV_NAMES(s) {s, s "[2]", s "[3]", s "[4]", s "[5]", s "[6]"}
    static const char *const names[6][MAXC] = {GEV_NAMES("fft_r2c_x"),     GEV_NAMES("fft_c2c_y"),
                                               GEV_NAMES("fft_c2c_z"),     GEV_NAMES("fft_c2c_z_inv"),
                                               GEV_NAMES("fft_c2c_y_inv"), GEV_NAMES("fft_c2r_x")};
#undef GEV_NAMES
    return names[pass][n - 1];
}

// One strided pass (axis: 0 = y, 1 = z) with the context's tile for that
// axis: ctx->col_rb[axis] row bytes (0: the default).  The row variants
// exist for N <= 256, the grids batches run.
template <typename T, int N, bool INV>
void launch_cols(gev_ctx *ctx, int axis, unsigned nb, const typename cplx_of<T>::type *in,
                 typename cplx_of<T>::type *out, int H, int64_t es, int64_t os,
                 const typename cplx_of<T>::type *tw, int64_t cs, const gev::Batch &bt) {
    const int rb = N <= 256 ? ctx->col_rb[axis] : 0;
#define GEV_COLS(R)                                                                                             \
    hipLaunchKernelGGL((k_c2c_cols<T, N, INV, R>), dim3((H + col_tile<T, N, R>() - 1) / col_tile<T, N, R>(), N, nb), \
                       dim3(col_tpb<T>()), 0, ctx->stream, in, out, H, es, os, tw, cs, bt)
    if constexpr (N <= 256) {
        if (rb == 64) {
            GEV_COLS(64);
            return;
        }
        if (rb == 256) {
            GEV_COLS(256);
            return;
        }
    }
    GEV_COLS(0);
#undef GEV_COLS
}

// passes: bit 0 = x (r2c), bit 1 = y, bit 2 = z, all components pass by pass.
// The x pass is one launch over up to MAXC components (blockIdx.z).  Over the
// boxes of a batch (bt.cyc set) every pass runs one launch per component with
// the box in blockIdx.z.
template <typename T, int N>
int run_forward(gev_ctx *ctx, const T *real, typename cplx_of<T>::type *k, int ncomp,
                const typename cplx_of<T>::type *tw, int passes, const gev::Batch &bt) {
    using C = typename cplx_of<T>::type;
    const int P = ctx->P, H = N / 2 + 1;
    const int64_t nrows = (int64_t)N * N;
    const int64_t n3v = (int64_t)N * N * N, nkv = (int64_t)P * N * N;
    constexpr int LINES = x_lines<N>();
    const bool batch = bt.cyc != nullptr;
    const unsigned nb = batch ? (unsigned)ctx->nbox : 1u;
    const int64_t rs = batch ? ctx->stride / (int64_t)sizeof(T) : 0, ks = batch ? ctx->stride / (int64_t)sizeof(C) : 0;
    for (int c0 = 0; c0 < ncomp; c0 += MAXC) {
        const int nc = ncomp - c0 < MAXC ? ncomp - c0 : MAXC;
        if (passes & 1) {
            gev::ScopedTimer t(ctx, pass_name(P_R2C_X, batch ? 1 : nc));
            if (batch) {
                for (int c = c0; c < c0 + nc; c++)
                    hipLaunchKernelGGL((k_r2c_x<T, N>), dim3(gev::grid_for(nrows, LINES), 1, nb), dim3(TPB), 0,
                                       ctx->stream, real + c * n3v, k + c * nkv, P, nrows, tw, rs, ks, bt);
            } else {
                hipLaunchKernelGGL((k_r2c_x<T, N>), dim3(gev::grid_for(nrows, LINES), 1, nc), dim3(TPB), 0,
                                   ctx->stream, real + c0 * n3v, k + c0 * nkv, P, nrows, tw, n3v, nkv, bt);
            }
        }
        // the column passes run one launch per component (measured faster
        // than one launch over the components)
        for (int c = c0; c < c0 + nc && (passes & 2); c++) {
            gev::ScopedTimer t(ctx, pass_name(P_C2C_Y, 1));
            launch_cols<T, N, false>(ctx, 0, nb, k + c * nkv, k + c * nkv, H, (int64_t)P, (int64_t)P * N, tw, ks, bt);
        }
        for (int c = c0; c < c0 + nc && (passes & 4); c++) {
            gev::ScopedTimer t(ctx, pass_name(P_C2C_Z, 1));
            launch_cols<T, N, false>(ctx, 1, nb, k + c * nkv, k + c * nkv, H, (int64_t)P * N, (int64_t)P, tw, ks, bt);
        }
    }
    return gev::check_hip(ctx, hipGetLastError(), "fft forward launch");
}

// from_y: k already went through the inverse z pass; the remaining passes run
// in place on k itself (tmp unused), the x pass as one launch over up to MAXC
// components.  Otherwise every component goes z, y, x through tmp.
template <typename T, int N>
int run_backward(gev_ctx *ctx, const typename cplx_of<T>::type *k, T *real, int ncomp,
                 typename cplx_of<T>::type *tmp, const typename cplx_of<T>::type *tw, bool from_y,
                 const gev::Batch &bt) {
    using C = typename cplx_of<T>::type;
    const int P = ctx->P, H = N / 2 + 1;
    const int64_t nrows = (int64_t)N * N;
    const int64_t n3v = (int64_t)N * N * N, nkv = (int64_t)P * N * N;
    constexpr int LINES = x_lines<N>();
    const bool batch = bt.cyc != nullptr;
    const unsigned nb = batch ? (unsigned)ctx->nbox : 1u;
    const int64_t rs = batch ? ctx->stride / (int64_t)sizeof(T) : 0, ks = batch ? ctx->stride / (int64_t)sizeof(C) : 0;
    if (from_y) {
        auto *kk = const_cast<C *>(k);
        for (int c0 = 0; c0 < ncomp; c0 += MAXC) {
            const int nc = ncomp - c0 < MAXC ? ncomp - c0 : MAXC;
            for (int c = c0; c < c0 + nc; c++) {
                gev::ScopedTimer t(ctx, pass_name(P_C2C_Y_INV, 1));
                launch_cols<T, N, true>(ctx, 0, nb, kk + c * nkv, kk + c * nkv, H, (int64_t)P, (int64_t)P * N, tw, ks,
                                        bt);
            }
            {
                gev::ScopedTimer t(ctx, pass_name(P_C2R_X, batch ? 1 : nc));
                if (batch) {
                    for (int c = c0; c < c0 + nc; c++)
                        hipLaunchKernelGGL((k_c2r_x<T, N>), dim3(gev::grid_for(nrows, LINES), 1, nb), dim3(TPB), 0,
                                           ctx->stream, kk + c * nkv, real + c * n3v, P, nrows, tw, ks, rs, bt);
                } else {
                    hipLaunchKernelGGL((k_c2r_x<T, N>), dim3(gev::grid_for(nrows, LINES), 1, nc), dim3(TPB), 0,
                                       ctx->stream, kk + c0 * nkv, real + c0 * n3v, P, nrows, tw, nkv, n3v, bt);
                }
            }
        }
        return gev::check_hip(ctx, hipGetLastError(), "fft backward launch");
    }
    for (int c = 0; c < ncomp; c++) {
        {
            gev::ScopedTimer t(ctx, pass_name(P_C2C_Z_INV, 1));
            launch_cols<T, N, true>(ctx, 1, nb, k + c * nkv, tmp, H, (int64_t)P * N, (int64_t)P, tw, ks, bt);
        }
        {
            gev::ScopedTimer t(ctx, pass_name(P_C2C_Y_INV, 1));
            launch_cols<T, N, true>(ctx, 0, nb, tmp, tmp, H, (int64_t)P, (int64_t)P * N, tw, ks, bt);
        }
        {
            gev::ScopedTimer t(ctx, pass_name(P_C2R_X, 1));
            hipLaunchKernelGGL((k_c2r_x<T, N>), dim3(gev::grid_for(nrows, LINES), 1, nb), dim3(TPB), 0,
                               ctx->stream, tmp, real + c * n3v, P, nrows, tw, ks, rs, bt);
        }
    }
    return gev::check_hip(ctx, hipGetLastError(), "fft backward launch");
}


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

// the passes in the context's form: FP64, or FP32 (float fields, float2 k fields)
static int forward(gev_ctx *ctx, const double *real, double2 *kfield, int ncomp, int passes,
                   const Batch &bt = Batch{}) {
    switch (ctx->N) {
#define CASE(n)                                                                                                  \
    case n:                                                                                                      \
        return ctx->fp32 ? run_forward<float, n>(ctx, fld<float>(real), cfld<float2>(kfield), ncomp, ctx->twiddle_f, \
                                                 passes, bt)                                                     \
                         : run_forward<double, n>(ctx, real, kfield, ncomp, ctx->twiddle, passes, bt);
        GEV_FFT_SIZES(CASE)
#undef CASE
    default:
        return fail(ctx, GEV_EINVAL, "unsupported FFT size");
    }
}

static int backward(gev_ctx *ctx, const double2 *kfield, double *real, int ncomp, bool from_y,
                    const Batch &bt = Batch{}) {
    switch (ctx->N) {
#define CASE(n)                                                                                             \
    case n:                                                                                                 \
        return ctx->fp32 ? run_backward<float, n>(ctx, cfld<float2>(kfield), fld<float>(real), ncomp,       \
                                                  cfld<float2>(ctx->fft_tmp), ctx->twiddle_f, from_y, bt)   \
                         : run_backward<double, n>(ctx, kfield, real, ncomp, ctx->fft_tmp, ctx->twiddle, from_y, bt);
        GEV_FFT_SIZES(CASE)
#undef CASE
    default:
        return fail(ctx, GEV_EINVAL, "unsupported FFT size");
    }
}

int fft_forward(gev_ctx *ctx, const double *real, double2 *kfield, int ncomp) {
    return forward(ctx, real, kfield, ncomp, 7);
}

int fft_backward(gev_ctx *ctx, const double2 *kfield, double *real, int ncomp) {
    return backward(ctx, kfield, real, ncomp, false);
}

int fft_forward_xy(gev_ctx *ctx, const double *real, double2 *kfield, int ncomp, const Batch &bt) {
    return forward(ctx, real, kfield, ncomp, 3, bt);
}

int fft_forward_y(gev_ctx *ctx, double2 *kfield, int ncomp, const Batch &bt) {
    return forward(ctx, nullptr, kfield, ncomp, 2, bt);
}

int fft_forward_z(gev_ctx *ctx, double2 *kfield, int ncomp, const Batch &bt) {
    return forward(ctx, nullptr, kfield, ncomp, 4, bt);
}

int fft_backward_yx(gev_ctx *ctx, double2 *kfield, double *real, int ncomp, const Batch &bt) {
    return backward(ctx, kfield, real, ncomp, true, bt);
}

}  // namespace gev
