// fftsolve.hip -- FFT passes fused with the work around them in a cycle.
//
// First pass: r2c along x applies prepareFTsource (gevolution.hpp:56-192) to
// the rows it stages, so the prepared source is never stored (main.cpp:472-477,
// 539-544).
//
// Last pass: the c2c along z of the forward transform fused with the
// Fourier-space solves of gevolution.hpp and the first pass of the backward
// transform (c2c along z, inverse), for the cycle sequences of main.cpp:
//   source -> solveModifiedPoissonFT -> phi                  (main.cpp:477-488)
//   Sij    -> projectFTscalar + evolveFTvector -> chi, B     (main.cpp:544-593)
//   Sij    -> projectFTscalar -> chi                         (main.cpp:558-563)
//   Bi     -> projectFTvector -> B                           (main.cpp:575-593)
// A workgroup owns W consecutive kx columns at one ky.  It transforms each
// input component's columns along z in LDS and keeps the results in
// registers (PTS points per thread), evaluates the solve per k point with the
// same per-point bodies as the standalone kernels (ksolve.h), and runs the
// inverse z transform of every output component before one store.  Each
// component column is read once and written once; the k-space intermediate
// of the unfused sequence (z pass out, solve in/out, z-inverse in) never
// reaches HBM.  The butterflies are those of fft.hip, so the fused result
// equals the unfused one bit for bit.
#include "fft_core.h"
#include "gev_internal.h"
#include "ksolve.h"
#include "prep.h"

namespace {

using namespace gev_fft;
using namespace gev_ks;
using namespace gev_prep;

enum { ZS_POISSON = 0, ZS_TENSOR = 1, ZS_TENSOR_SCALAR = 2, ZS_VECTOR = 3 };

template <int KIND> struct ZSolve;
template <> struct ZSolve<ZS_POISSON> { static constexpr int NIN = 1, NOUT = 1; };
template <> struct ZSolve<ZS_TENSOR> { static constexpr int NIN = 6, NOUT = 4; };
template <> struct ZSolve<ZS_TENSOR_SCALAR> { static constexpr int NIN = 6, NOUT = 1; };
template <> struct ZSolve<ZS_VECTOR> { static constexpr int NIN = 3, NOUT = 3; };

// Columns per workgroup and points per thread.  Every thread keeps NIN
// components of its points in registers through the solve, so the tensor
// solve (6 inputs + the carried B) takes fewer points per thread, and with
// them more threads, to keep its register count -- and the waves per SIMD
// that hide the column loads -- in check.  The FP32 form (C = float2) has
// its own defaults.  GEV_ZS_COLS / GEV_ZS_PTS_TENSOR (and the _F32 pair)
// override them.
#ifndef GEV_ZS_COLS
#define GEV_ZS_COLS 4
#endif
#ifndef GEV_ZS_PTS_TENSOR
#define GEV_ZS_PTS_TENSOR 4
#endif
#ifndef GEV_ZS_COLS_F32
#define GEV_ZS_COLS_F32 8
#endif
#ifndef GEV_ZS_PTS_TENSOR_F32
#define GEV_ZS_PTS_TENSOR_F32 8
#endif
// the Poisson and vector solves (1 and 3 components) have their own column counts
#ifndef GEV_ZS_COLS_S
#define GEV_ZS_COLS_S 4
#endif
#ifndef GEV_ZS_COLS_S_F32
#define GEV_ZS_COLS_S_F32 4
#endif
template <typename C> constexpr bool is_f32() { return sizeof(C) == 8; }
template <int N, int KIND, typename C> constexpr int zs_cols() {
    constexpr bool T = KIND == ZS_TENSOR || KIND == ZS_TENSOR_SCALAR;
    if (N >= 768) return 2;
    if (is_f32<C>()) return T ? GEV_ZS_COLS_F32 : GEV_ZS_COLS_S_F32;
    return T ? GEV_ZS_COLS : GEV_ZS_COLS_S;
}
// The tensor solve prefetches the carried B into registers during the last
// input transform where the registers allow it; at N = 384 and above 512 the
// prefetch would spill (compiler resource report), so B is read at the solve.
// (FP32 form at 512^3: 3.04 against 3.21 ms with B read at the solve.)
#ifndef GEV_ZS_AHEAD
#define GEV_ZS_AHEAD 1
#endif
constexpr int ZS_AHEAD = GEV_ZS_AHEAD;
#ifndef GEV_ZS_PREFETCH_F32
#define GEV_ZS_PREFETCH_F32 1
#endif
template <int N, typename C = double2> constexpr bool zs_prefetch_b() {
    return ZS_AHEAD == 1 && N <= 512 && N != 384 && (GEV_ZS_PREFETCH_F32 || !is_f32<C>());
}
template <int N, int KIND, typename C> constexpr int zs_pts() {
    if (KIND != ZS_TENSOR && KIND != ZS_TENSOR_SCALAR) return 4;
    if (!is_f32<C>()) return GEV_ZS_PTS_TENSOR;
    return N >= 768 ? 4 : GEV_ZS_PTS_TENSOR_F32;  // two columns above 512: 1024 threads would spill
}
template <int N, int KIND, typename C> constexpr int zs_threads() {
    int t = N * zs_cols<N, KIND, C>() / zs_pts<N, KIND, C>();
    if (t > 1024) t = 1024;
    if (t < 64) t = 64;
    return (t + 63) / 64 * 64;
}

// C: complex element of the k fields (double2, or float2 in the FP32 form).
// The FFT stages run in C; the per-k solves run in FP64 either way (the k
// tables, the denominators and the zero mode stay FP64).
template <typename C> struct ZArgsT {
    const C *in;         // NIN components, after the x and y passes
    C *out;              // NOUT components, after the inverse z pass
    C *bift;             // B state: read (tensor) / written (tensor, vector)
    C *kstore;           // optional: solved k values of output 0 (full field)
    double2 *k0;         // optional: output 0 at k = 0 (one element, FP64)
    const double *g;     // gridk2 per axis index
    const double2 *ks;   // kshift per axis index
    const C *tw;         // exp(-2 pi i t / N), t < N
    int64_t nkv;         // complexes per component
    int P, H;
    double c0, c1;       // Poisson: coeff/(-N^3), modif; tensor: a^2 dtau; vector: coeff, modif
    gev::Batch bt;       // over the boxes of a batch: box blockIdx.z, coefficients from its BoxCycle
};
using ZArgs = ZArgsT<double2>;

template <int N, int KIND, typename C>
__global__ __launch_bounds__((zs_threads<N, KIND, C>())) void k_zsolve(ZArgsT<C> A) {
    constexpr int W = zs_cols<N, KIND, C>();
    constexpr int NT = zs_threads<N, KIND, C>();
    constexpr int NIN = ZSolve<KIND>::NIN, NOUT = ZSolve<KIND>::NOUT;
    constexpr int PTS = (N * W + NT - 1) / NT;
    constexpr bool EXACT = PTS * NT == N * W;
    static_assert(NT % W == 0, "a thread keeps one column");
    __shared__ C buf[N * W];
    __shared__ C tw[N];
    if (A.bt.cyc) {
        const int b = blockIdx.z;
        if (!gev::box_on(A.bt, b)) return;
        const gev::BoxCycle &bc = A.bt.cyc[b];
        A.in = gev::box_ptr(A.in, A.bt.stride, b);
        A.out = gev::box_ptr(A.out, A.bt.stride, b);
        A.bift = gev::box_ptr(A.bift, A.bt.stride, b);
        A.kstore = gev::box_ptr(A.kstore, A.bt.stride, b);
        A.k0 = gev::box_ptr(A.k0, A.bt.stride, b);
        if (KIND == ZS_POISSON) {
            A.c0 = bc.pois_c0;
            A.c1 = bc.pois_c1;
        } else if (KIND == ZS_TENSOR) {
            A.c0 = bc.a2dtau;
        }
    }

    // XCD-aware tile order: hardware workgroup ids go round-robin over the 8
    // XCDs, so tiles l and l+1 (which share 128-byte rows) are given to ids
    // 8 apart, i.e. to consecutive workgroups of one XCD (same L2)
    const int ntx = gridDim.x;
    const int64_t T = (int64_t)ntx * gridDim.y;
    int64_t L = (int64_t)blockIdx.y * ntx + blockIdx.x;
    if (T % 8 == 0) L = (L % 8) * (T / 8) + L / 8;
    const int kx0 = (int)(L % ntx) * W;
    const int ky = (int)(L / ntx);
    const int64_t es = (int64_t)A.P * N;                       // z stride
    const int64_t base = (int64_t)ky * A.P + kx0;
    const int w = threadIdx.x % W;
    const int kx = kx0 + w;
    const bool ok = kx < A.H;

    for (int t = threadIdx.x; t < N; t += NT) tw[t] = A.tw[t];

    // column loads of input component c (element kz = idx / W of column w)
    auto load_comp = [&](int c, C *v) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < PTS; j++) {
            const int idx = threadIdx.x + j * NT;
            v[j] = mk2<C>(0, 0);
            if (ok && (EXACT || idx < N * W)) v[j] = A.in[c * A.nkv + base + (int64_t)(idx / W) * es + w];
        }
    };
    // the carried B of the tensor solve, fetched while the last input transforms
    C Bold[KIND == ZS_TENSOR ? 3 : 1][PTS];
    auto load_bold = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < PTS; j++) {
            const int idx = threadIdx.x + j * NT;
            const int64_t at = base + (int64_t)(idx / W) * es + w;
#pragma unroll
            for (int c = 0; c < (KIND == ZS_TENSOR ? 3 : 1); c++) {
                Bold[c][j] = mk2<C>(0, 0);
                if (KIND == ZS_TENSOR && ok && (EXACT || idx < N * W)) Bold[c][j] = A.bift[c * A.nkv + at];
            }
        }
    };

    C S[NIN][PTS];
    // input components in flight ahead of the one being transformed: the
    // next one, or the next two (ZS_AHEAD = 2, which reads B at the solve
    // instead of prefetching it)
    C v[ZS_AHEAD][PTS];
#pragma unroll
    for (int c = 0; c < ZS_AHEAD && c < NIN; c++) load_comp(c, v[c]);
#pragma unroll
    for (int c = 0; c < NIN; c++) {
#pragma unroll
        for (int j = 0; j < PTS; j++) {
            const int idx = threadIdx.x + j * NT;
            if (EXACT || idx < N * W) buf[idx] = v[c % ZS_AHEAD][j];
        }
        // the following components' loads stay in flight through this
        // component's transform (the stage barriers wait for LDS traffic only)
        if (c + ZS_AHEAD < NIN) load_comp(c + ZS_AHEAD, v[c % ZS_AHEAD]);
        else if (c + 1 == NIN && KIND == ZS_TENSOR && zs_prefetch_b<N, C>()) load_bold();
        __syncthreads();
        stages<C, N, W, true, 1, NT>(buf, tw);
        // each thread reads back only the points it wrote, so the next
        // component may overwrite them without another barrier
#pragma unroll
        for (int j = 0; j < PTS; j++) {
            const int idx = threadIdx.x + j * NT;
            S[c][j] = (EXACT || idx < N * W) ? buf[idx] : mk2<C>(0, 0);
        }
    }

    // the solve, per k point (kx, ky, kz = element index), in FP64; the
    // outputs are kept in the storage type (k = 0 of output 0 also in FP64)
    C O[NOUT][PTS];
    const double gx = A.g[ok ? kx : 0];
    const double2 kxs = A.ks[ok ? kx : 0];
    const double gy = A.g[ky];
    const double2 kys = A.ks[ky];
#pragma unroll
    for (int j = 0; j < PTS; j++) {
        const int idx = threadIdx.x + j * NT;
        const int kz = idx / W;
        const bool live = ok && (EXACT || idx < N * W);
        const double gz = A.g[live ? kz : 0];
        const double2 kzs = A.ks[live ? kz : 0];
        const bool zero = kx == 0 && ky == 0 && kz == 0;
        const int64_t at = base + (int64_t)kz * es + w;
        double2 o0;
        if constexpr (KIND == ZS_POISSON) {
            o0 = poisson_point(d2(S[0][j]), zero, zero ? 0. : gx + gy + gz, A.c0, A.c1);
        } else if constexpr ((KIND == ZS_TENSOR || KIND == ZS_TENSOR_SCALAR) && is_f32<C>()) {
            // FP32 form: FP64 coefficients and denominators, float field combinations
            float2 Sf[6];
#pragma unroll
            for (int c = 0; c < 6; c++) Sf[c] = S[c][j];
            const float2 chi = zero ? make_float2(0.f, 0.f) : project_scalar_f32(Sf, gx, gy, gz, kxs, kys, kzs, N);
            o0 = make_double2(chi.x, chi.y);
            if constexpr (KIND == ZS_TENSOR) {
                float2 B[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
                if (live) {
                    if (!zero) {
                        if constexpr (zs_prefetch_b<N, C>()) {
                            B[0] = Bold[0][j];
                            B[1] = Bold[KIND == ZS_TENSOR ? 1 : 0][j];
                            B[2] = Bold[KIND == ZS_TENSOR ? 2 : 0][j];
                        } else {
                            B[0] = A.bift[at];
                            B[1] = A.bift[A.nkv + at];
                            B[2] = A.bift[2 * A.nkv + at];
                        }
                        evolve_vector_f32(Sf, B, gx, gy, gz, kxs, kys, kzs, A.c0);
                    }
                    A.bift[at] = B[0];
                    A.bift[A.nkv + at] = B[1];
                    A.bift[2 * A.nkv + at] = B[2];
                }
#pragma unroll
                for (int c = 0; c < 3; c++) O[1 + c][j] = B[c];
            }
        } else if constexpr (KIND == ZS_TENSOR || KIND == ZS_TENSOR_SCALAR) {
            double2 Sp[6];
#pragma unroll
            for (int c = 0; c < 6; c++) Sp[c] = d2(S[c][j]);
            o0 = zero ? c2(0., 0.) : project_scalar(Sp, gx, gy, gz, kxs, kys, kzs, N);
            if constexpr (KIND == ZS_TENSOR) {
                double2 B[3] = {c2(0., 0.), c2(0., 0.), c2(0., 0.)};
                if (live) {
                    if (!zero) {
                        if constexpr (zs_prefetch_b<N, C>()) {
                            B[0] = d2(Bold[0][j]);
                            B[1] = d2(Bold[KIND == ZS_TENSOR ? 1 : 0][j]);
                            B[2] = d2(Bold[KIND == ZS_TENSOR ? 2 : 0][j]);
                        } else {
                            B[0] = d2(A.bift[at]);
                            B[1] = d2(A.bift[A.nkv + at]);
                            B[2] = d2(A.bift[2 * A.nkv + at]);
                        }
                        evolve_vector(Sp, B, gx, gy, gz, kxs, kys, kzs, A.c0);
                    }
                    A.bift[at] = from_d2<C>(B[0]);
                    A.bift[A.nkv + at] = from_d2<C>(B[1]);
                    A.bift[2 * A.nkv + at] = from_d2<C>(B[2]);
                }
#pragma unroll
                for (int c = 0; c < 3; c++) O[1 + c][j] = from_d2<C>(B[c]);
            }
        } else {
            double2 Sv[3] = {d2(S[0][j]), d2(S[1][j]), d2(S[2][j])};
            if (zero) {
                Sv[0] = Sv[1] = Sv[2] = c2(0., 0.);
            } else {
                project_vector_point(Sv, kxs, kys, kzs, gx + gy + gz, A.c0, A.c1);
            }
            if (live && A.bift) {
                A.bift[at] = from_d2<C>(Sv[0]);
                A.bift[A.nkv + at] = from_d2<C>(Sv[1]);
                A.bift[2 * A.nkv + at] = from_d2<C>(Sv[2]);
            }
            o0 = Sv[0];
#pragma unroll
            for (int c = 1; c < 3; c++) O[c][j] = from_d2<C>(Sv[c]);
        }
        O[0][j] = from_d2<C>(o0);
        if (live && A.kstore) A.kstore[at] = O[0][j];
        if (zero && A.k0) A.k0[0] = o0;
    }

    // inverse z transform of every output (conjugate, forward stages, conjugate)
#pragma unroll
    for (int o = 0; o < NOUT; o++) {
#pragma unroll
        for (int j = 0; j < PTS; j++) {
            const int idx = threadIdx.x + j * NT;
            if (EXACT || idx < N * W) buf[idx] = cconj(O[o][j]);
        }
        __syncthreads();
        stages<C, N, W, true, 1, NT>(buf, tw);
#pragma unroll
        for (int j = 0; j < PTS; j++) {
            const int idx = threadIdx.x + j * NT;
            if (ok && (EXACT || idx < N * W))
                A.out[o * A.nkv + base + (int64_t)(idx / W) * es + w] = cconj(buf[idx]);
        }
    }
}

// ---------------------------------------------------------------- first pass

// r2c along x with prepareFTsource applied while the rows are staged: the
// deposited source (1 or 6 components) and phi (and chi) are read, the
// prepared values never reach HBM.  Same half-length transform and
// post-processing as k_r2c_x.
template <typename F, typename C> struct XArgsT {
    const F *phi, *chi;       // chi: scalar form only
    const F *src;             // NC deposited components, N^3 each
    C *out;                   // NC k fields after the x pass
    const C *tw;
    int P;
    int64_t nrows, n3v, nkv;
    double c0, c1, c2, c3;  // tensor: coeff; scalar: bg, coeff, coeff2, coeff3
    gev::Batch bt;          // over the boxes of a batch: box blockIdx.y, coefficients from its BoxCycle
};
using XArgs = XArgsT<double, double2>;

// Rows per workgroup: the NC prepared components of a thread's points are
// held in registers until each is transformed, so the 6-component form
// takes at most GEV_XPREP_PTS points per thread and component.
#ifndef GEV_XPREP_PTS
#define GEV_XPREP_PTS 4
#endif
template <int N, int NC> constexpr int xprep_lines() {
    int l = x_lines<N>();
    while (NC > 1 && l > 1 && l * (N / 2) > GEV_XPREP_PTS * 256) l /= 2;
    return l;
}

template <int N, int NC, typename F, typename C>
__global__ __launch_bounds__(256) void k_r2c_x_prep(XArgsT<F, C> A) {
    using R = decltype(C::x);
    constexpr int TPB = 256;
    constexpr int M = N / 2;
    constexpr int LINES = xprep_lines<N, NC>();
    constexpr int PER = (LINES * M + TPB - 1) / TPB;
    constexpr bool EXACT = PER * TPB == LINES * M;
    __shared__ C buf[LINES * xstride<M>()];
    __shared__ C tw[M];
    if (A.bt.cyc) {
        const int b = blockIdx.y;
        if (!gev::box_on(A.bt, b)) return;
        const gev::BoxCycle &bc = A.bt.cyc[b];
        A.phi = gev::box_ptr(A.phi, A.bt.stride, b);
        A.chi = gev::box_ptr(A.chi, A.bt.stride, b);
        A.src = gev::box_ptr(A.src, A.bt.stride, b);
        A.out = gev::box_ptr(A.out, A.bt.stride, b);
        if (NC == 6) {
            A.c0 = bc.tens_c;
        } else {
            A.c0 = bc.prep_bg;
            A.c1 = bc.prep_c;
            A.c2 = bc.prep_c2;
            A.c3 = bc.prep_c3;
        }
    }
    const int64_t row0 = (int64_t)blockIdx.x * LINES;
    for (int t = threadIdx.x; t < M; t += TPB) tw[t] = A.tw[2 * t];
    C v[NC][PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int idx = threadIdx.x + k * TPB;
        const int line = idx / M, n = idx % M;
        const int64_t row = row0 + line;
#pragma unroll
        for (int c = 0; c < NC; c++) v[c][k] = mk2<C>(0, 0);
        if (!((EXACT || idx < LINES * M) && row < A.nrows)) continue;
        const int y = (int)(row % N), z = (int)(row / N);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int x = 2 * n + h;
            const int64_t i = row * N + x;
            const Nb c = nb_of(x, y, z, N);
            if constexpr (NC == 6 && is_f32<C>()) {
                float T[6], S[6];
#pragma unroll
                for (int q = 0; q < 6; q++) T[q] = A.src[q * A.n3v + i];
                prep_tensor_f32(A.phi, T, c, i, N, (float)A.c0, S);
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    if (h == 0) v[q][k].x = S[q];
                    else v[q][k].y = S[q];
                }
            } else if constexpr (NC == 6) {
                double T[6], S[6];
#pragma unroll
                for (int q = 0; q < 6; q++) T[q] = A.src[q * A.n3v + i];
                prep_tensor(A.phi, T, c, i, N, A.c0, S);
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    if (h == 0) v[q][k].x = (R)S[q];
                    else v[q][k].y = (R)S[q];
                }
            } else {
                const double r = prep_scalar(A.phi, (double)A.chi[i], (double)A.src[i], c, i, N, A.c0, A.c1, A.c2, A.c3);
                if (h == 0) v[0][k].x = (R)r;
                else v[0][k].y = (R)r;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NC; q++) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int idx = threadIdx.x + k * TPB;
            if (EXACT || idx < LINES * M) buf[eaddr<M, LINES, false>(idx / M, idx % M)] = v[q][k];
        }
        __syncthreads();
        stages<C, M, LINES, false, 1>(buf, tw);
        for (int idx = threadIdx.x; idx < LINES * (M + 1); idx += TPB) {
            const int line = idx / (M + 1), k = idx % (M + 1);
            const int64_t row = row0 + line;
            if (row >= A.nrows) continue;
            const C Ak = buf[eaddr<M, LINES, false>(line, k == M ? 0 : k)];
            const C Bk = cconj(buf[eaddr<M, LINES, false>(line, k == 0 ? 0 : M - k)]);
            const C E = mk2<C>((Ak.x + Bk.x) * (R)0.5, (Ak.y + Bk.y) * (R)0.5);
            const C D = mk2<C>((Ak.x - Bk.x) * (R)0.5, (Ak.y - Bk.y) * (R)0.5);
            const C O = mul_mi(D);
            A.out[q * A.nkv + row * A.P + k] = cadd(E, cmul(A.tw[k], O));
        }
        __syncthreads();  // the next component overwrites buf
    }
}

template <int N, int NC, typename F, typename C> int launch_x(gev_ctx *c, const XArgsT<F, C> &A) {
    constexpr int LINES = xprep_lines<N, NC>();
    const unsigned nbox = A.bt.cyc ? (unsigned)c->nbox : 1u;
    hipLaunchKernelGGL((k_r2c_x_prep<N, NC, F, C>), dim3(gev::grid_for(A.nrows, LINES), nbox), dim3(256), 0,
                       c->stream, A);
    return gev::check_hip(c, hipGetLastError(), "fused x pass launch");
}

// the context's fields as F / C (double / double2, or float / float2 in the
// FP32 form)
template <typename F, typename C> XArgsT<F, C> xargs_as(const XArgs &A) {
    XArgsT<F, C> B{};
    B.phi = gev::fld<F>(const_cast<double *>(A.phi));
    B.chi = gev::fld<F>(const_cast<double *>(A.chi));
    B.src = gev::fld<F>(const_cast<double *>(A.src));
    B.out = gev::cfld<C>(A.out);
    B.tw = gev::cfld<C>(const_cast<double2 *>(A.tw));
    B.P = A.P;
    B.nrows = A.nrows;
    B.n3v = A.n3v;
    B.nkv = A.nkv;
    B.c0 = A.c0;
    B.c1 = A.c1;
    B.c2 = A.c2;
    B.c3 = A.c3;
    B.bt = A.bt;
    return B;
}

template <int NC> int launch_x_nc(gev_ctx *c, XArgs A) {
    A.tw = c->fp32 ? reinterpret_cast<const double2 *>(c->twiddle_f) : c->twiddle;
    A.P = c->P;
    A.nrows = (int64_t)c->N * c->N;
    A.n3v = gev::n3(c);
    A.nkv = gev::nk(c);
    switch (c->N) {
#define CASE(n) \
    case n: return c->fp32 ? launch_x<n, NC>(c, xargs_as<float, float2>(A)) : launch_x<n, NC>(c, A);
        GEV_FFT_SIZES(CASE)
#undef CASE
    default:
        return gev::fail(c, GEV_EINVAL, "unsupported FFT size");
    }
}

// ---------------------------------------------------------------- last pass

template <int N, int KIND, typename C> int launch_n(gev_ctx *c, const ZArgsT<C> &A) {
    constexpr int W = zs_cols<N, KIND, C>();
    const int H = N / 2 + 1;
    const unsigned nbox = A.bt.cyc ? (unsigned)c->nbox : 1u;
    hipLaunchKernelGGL((k_zsolve<N, KIND, C>), dim3((H + W - 1) / W, N, nbox), dim3(zs_threads<N, KIND, C>()), 0,
                       c->stream, A);
    return gev::check_hip(c, hipGetLastError(), "fused z solve launch");
}

template <typename C> ZArgsT<C> zargs_as(const ZArgs &A) {
    ZArgsT<C> B{};
    B.in = gev::cfld<C>(const_cast<double2 *>(A.in));
    B.out = gev::cfld<C>(A.out);
    B.bift = gev::cfld<C>(A.bift);
    B.kstore = gev::cfld<C>(A.kstore);
    B.k0 = A.k0;
    B.g = A.g;
    B.ks = A.ks;
    B.tw = gev::cfld<C>(const_cast<double2 *>(A.tw));
    B.nkv = A.nkv;
    B.P = A.P;
    B.H = A.H;
    B.c0 = A.c0;
    B.c1 = A.c1;
    B.bt = A.bt;
    return B;
}

template <int KIND> int launch_kind(gev_ctx *c, ZArgs A) {
    A.g = c->gridk2;
    A.ks = c->kshift;
    A.tw = c->fp32 ? reinterpret_cast<const double2 *>(c->twiddle_f) : c->twiddle;
    A.nkv = gev::nk(c);
    A.P = c->P;
    A.H = c->N / 2 + 1;
    switch (c->N) {
#define CASE(n) \
    case n: return c->fp32 ? launch_n<n, KIND>(c, zargs_as<float2>(A)) : launch_n<n, KIND>(c, A);
        GEV_FFT_SIZES(CASE)
#undef CASE
    default:
        return gev::fail(c, GEV_EINVAL, "unsupported FFT size");
    }
}

}  // namespace

namespace gev {

int launch_zsolve_poisson(gev_ctx *c, const double2 *in, double2 *out, double coeff, double modif,
                          double2 *kstore, double2 *k0, const Batch &bt) {
    gev::ScopedTimer t(c, "fft_z_poisson_z");
    ZArgs A{};
    A.bt = bt;
    A.in = in;
    A.out = out;
    A.kstore = kstore;
    A.k0 = k0;
    A.c0 = coeff / -((double)c->N * (double)c->N * (double)c->N);  // gevolution.hpp:511
    A.c1 = modif;
    return launch_kind<ZS_POISSON>(c, A);
}

int launch_zsolve_tensor(gev_ctx *c, const double2 *in, double2 *out, double a2dtau, bool evolve,
                         double2 *kstore, double2 *k0, const Batch &bt) {
    gev::ScopedTimer t(c, evolve ? "fft_z_tensor_z" : "fft_z_tensor_scalar_z");
    ZArgs A{};
    A.bt = bt;
    A.in = in;
    A.out = out;
    A.bift = c->BiFT;
    A.kstore = kstore;
    A.k0 = k0;
    A.c0 = a2dtau;
    return evolve ? launch_kind<ZS_TENSOR>(c, A) : launch_kind<ZS_TENSOR_SCALAR>(c, A);
}

int launch_x_prep_tensor(gev_ctx *c, double coeff, const Batch &bt) {
    ScopedTimer t(c, "prepare_tensor_fft_r2c_x");
    XArgs A{};
    A.bt = bt;
    A.phi = c->phi;
    A.src = c->Sij;
    A.out = c->SijFT;
    A.c0 = coeff;
    return launch_x_nc<6>(c, A);
}

int launch_x_prep_scalar(gev_ctx *c, double bg, double coeff, double coeff2, double coeff3, const Batch &bt) {
    ScopedTimer t(c, "prepare_scalar_fft_r2c_x");
    XArgs A{};
    A.bt = bt;
    A.phi = c->phi;
    A.chi = c->chi;
    A.src = c->source;
    A.out = c->scalarFT;
    A.c0 = bg;
    A.c1 = coeff;
    A.c2 = coeff2;
    A.c3 = coeff3;
    return launch_x_nc<1>(c, A);
}

int launch_zsolve_vector(gev_ctx *c, const double2 *in, double2 *out, double coeff, double modif,
                         double2 *bift) {
    gev::ScopedTimer t(c, "fft_z_vector_z");
    ZArgs A{};
    A.in = in;
    A.out = out;
    A.bift = bift;
    A.c0 = coeff;
    A.c1 = modif;
    return launch_kind<ZS_VECTOR>(c, A);
}

}  // namespace gev
