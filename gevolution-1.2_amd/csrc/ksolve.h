// ksolve.h -- per-k-point bodies of the Fourier-space solves of gevolution.hpp
// (solveModifiedPoissonFT, projectFTscalar, evolveFTvector, projectFTvector),
// shared by the standalone k-space kernels (kspace.hip) and the solves fused
// into the last FFT pass (fftsolve.hip).  The complex arithmetic follows the
// reference's association order; with -ffp-contract=off both users produce the
// same bits for the same inputs.
#pragma once
#include <hip/hip_runtime.h>

namespace gev_ks {

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

// a k value in storage (double2, or float2 in the FP32 form) to FP64 and back:
// the solves below always evaluate in FP64
__device__ __forceinline__ double2 d2(double2 v) { return v; }
__device__ __forceinline__ double2 d2(float2 v) { return make_double2(v.x, v.y); }
template <typename C> __device__ __forceinline__ C from_d2(double2 v);
template <> __device__ __forceinline__ double2 from_d2<double2>(double2 v) { return v; }
template <> __device__ __forceinline__ float2 from_d2<float2>(double2 v) {
    return make_float2((float)v.x, (float)v.y);
}

// solveModifiedPoissonFT body, gevolution.hpp:501-535 (coeff already divided by -N^3)
__device__ __forceinline__ double2 poisson_point(double2 s, bool zero, double k2, double coeff,
                                                 double modif) {
    if (zero) return modif == 0. ? c2(0., 0.) : cdivr(cmulr(s, coeff), modif);
    return cdivr(cmulr(s, coeff), k2 + modif);
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

// projectFTvector body, gevolution.hpp:350-392 (k != 0)
__device__ __forceinline__ void project_vector_point(double2 S[3], double2 k0, double2 k1, double2 k2v,
                                                     double k2, double coeff, double modif) {
    const double2 S0 = S[0], S1 = S[1], S2 = S[2];
    double2 tmp = cadd(cadd(cmul(k0, S0), cmul(k1, S1)), cmul(k2v, S2));
    tmp = cdivr(tmp, k2);
    S[0] = cdivr(cmulr(cmulr(csub(S0, cmul(cconj(k0), tmp)), 4.), coeff), k2 + modif);
    S[1] = cdivr(cmulr(cmulr(csub(S1, cmul(cconj(k1), tmp)), 4.), coeff), k2 + modif);
    S[2] = cdivr(cmulr(cmulr(csub(S2, cmul(cconj(k2v), tmp)), 4.), coeff), k2 + modif);
}

// ---- the FP32 form's tensor solves: the k-dependent coefficients and the
// denominators (k -> 0) are evaluated in FP64 and rounded once; only the
// combination linear in the float field values runs in float.  Same terms
// as project_scalar / evolve_vector above, regrouped by source component.

__device__ __forceinline__ float2 f2(double2 v) { return make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ float2 fmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 faddmul(float2 acc, float2 c, float2 s) {
    const float2 p = fmulc(c, s);
    return make_float2(acc.x + p.x, acc.y + p.y);
}
__device__ __forceinline__ float2 faddmulr(float2 acc, float c, float2 s) {
    return make_float2(acc.x + c * s.x, acc.y + c * s.y);
}

// projectFTscalar body, gevolution.hpp:251-260 (k != 0)
__device__ __forceinline__ float2 project_scalar_f32(const float2 S[6], double g0, double g1, double g2, double2 k0,
                                                     double2 k1, double2 k2, int N) {
    const double inv = 1. / (2. * (g0 + g1 + g2) * (g0 + g1 + g2) * (double)N);
    float2 num = make_float2(0.f, 0.f);
    num = faddmulr(num, (float)((g1 + g2 - 2. * g0) * inv), S[0]);
    num = faddmulr(num, (float)((g0 + g2 - 2. * g1) * inv), S[3]);
    num = faddmulr(num, (float)((g0 + g1 - 2. * g2) * inv), S[5]);
    num = faddmul(num, f2(rmul(-inv, cmul(rmul(6., k0), k1))), S[1]);
    num = faddmul(num, f2(rmul(-inv, cmul(rmul(6., k0), k2))), S[2]);
    num = faddmul(num, f2(rmul(-inv, cmul(rmul(6., k1), k2))), S[4]);
    return num;
}

// evolveFTvector body, gevolution.hpp:314-325 (k != 0); adds onto B[3].  With
// p = -2 a^2 dtau / k^4 and the component axes (i, j, l) = (0, 1, 2), (1, 0, 2)
// and (2, 0, 1):  B_i += i p [ conj(k_i) ((g_j + g_l) S_ii - g_j S_jj - g_l S_ll
//                              - 2 k_j k_l S_jl) + (g_j + g_l - g_i)(k_j S_ij + k_l S_il) ]
__device__ __forceinline__ void evolve_vector_f32(const float2 S[6], float2 B[3], double g0, double g1, double g2,
                                                  double2 k0, double2 k1, double2 k2, double a2dtau) {
    double k4 = g0 + g1 + g2;
    k4 *= k4;
    const double p = -2. * a2dtau / k4;
    const double g[3] = {g0, g1, g2};
    const double2 kk[3] = {k0, k1, k2};
    // symmetric index of (a, b) in S: 00 01 02 11 12 22
    constexpr int SI[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int j = i == 0 ? 1 : 0, l = i == 2 ? 1 : 2;
        const double2 ip_ck = cmul(c2(0., p), cconj(kk[i]));       // i p conj(k_i)
        const double2 ipg = c2(0., p * (g[j] + g[l] - g[i]));       // i p (g_j + g_l - g_i)
        float2 acc = B[i];
        acc = faddmul(acc, f2(rmul(g[j] + g[l], ip_ck)), S[SI[i][i]]);
        acc = faddmul(acc, f2(rmul(-g[j], ip_ck)), S[SI[j][j]]);
        acc = faddmul(acc, f2(rmul(-g[l], ip_ck)), S[SI[l][l]]);
        acc = faddmul(acc, f2(cmul(ip_ck, rmul(-2., cmul(kk[j], kk[l])))), S[SI[j][l]]);
        acc = faddmul(acc, f2(cmul(ipg, kk[j])), S[SI[i][j]]);
        acc = faddmul(acc, f2(cmul(ipg, kk[l])), S[SI[i][l]]);
        B[i] = acc;
    }
}

}  // namespace gev_ks
