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

}  // namespace gev_ks
