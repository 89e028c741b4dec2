// prep.h -- per-site bodies of prepareFTsource (gevolution.hpp:56-192,
// PHINONLINEAR on, ORIGINALMETRIC off), shared by the standalone stencil
// kernels (stencil.hip) and the first FFT pass that applies them on the fly
// (fftsolve.hip).  Expression order follows the reference literally; with
// -ffp-contract=off both users give the same bits.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gev_prep {

struct Nb {
    int x, y, z, xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ Nb nb_of(int x, int y, int z, int N) {
    Nb c;
    c.x = x;
    c.y = y;
    c.z = z;
    c.xm = x == 0 ? N - 1 : x - 1;
    c.xp = x == N - 1 ? 0 : x + 1;
    c.ym = y == 0 ? N - 1 : y - 1;
    c.yp = y == N - 1 ? 0 : y + 1;
    c.zm = z == 0 ? N - 1 : z - 1;
    c.zp = z == N - 1 ? 0 : z + 1;
    return c;
}

__device__ __forceinline__ int64_t L(int N, int x, int y, int z) {
    return (int64_t)x + (int64_t)N * ((int64_t)y + (int64_t)N * z);
}

// prepareFTsource (2), gevolution.hpp:169-192: the new source value at site i
// (phi stored as F: double, or float in the FP32 form; arithmetic in double)
template <typename F>
__device__ __forceinline__ double prep_scalar(const F *__restrict__ phi, double chi_i, double src_i,
                                              const Nb &c, int64_t i, int N, double bg, double coeff,
                                              double coeff2, double coeff3) {
    const double p = phi[i];
    const double px = phi[L(N, c.xp, c.y, c.z)], mx = phi[L(N, c.xm, c.y, c.z)];
    const double py = phi[L(N, c.x, c.yp, c.z)], my = phi[L(N, c.x, c.ym, c.z)];
    const double pz = phi[L(N, c.x, c.y, c.zp)], mz = phi[L(N, c.x, c.y, c.zm)];
    double r = coeff2 * (src_i - bg);
    r *= 1. - 2. * p;
    r += 0.125 * (mx - px) * (mx - px);
    r += 0.125 * (my - py) * (my - py);
    r += 0.125 * (mz - pz) * (mz - pz);
    r += (coeff3 - coeff) * p - coeff3 * chi_i;
    return r;
}

// prepareFTsource (1), gevolution.hpp:56-147: S[6] (order 00,01,02,11,12,22)
// from the deposited T[6] at site i
template <typename F>
__device__ __forceinline__ void prep_tensor(const F *__restrict__ phi, const double T[6], const Nb &c,
                                            int64_t i, int N, double coeff, double S[6]) {
    const double p = phi[i];
    const double px = phi[L(N, c.xp, c.y, c.z)], mx = phi[L(N, c.xm, c.y, c.z)];
    const double py = phi[L(N, c.x, c.yp, c.z)], my = phi[L(N, c.x, c.ym, c.z)];
    const double pz = phi[L(N, c.x, c.y, c.zp)], mz = phi[L(N, c.x, c.y, c.zm)];
    const double pxy = phi[L(N, c.xp, c.yp, c.z)];
    const double pxz = phi[L(N, c.xp, c.y, c.zp)];
    const double pyz = phi[L(N, c.x, c.yp, c.zp)];
    double s;
    s = coeff * T[0];
    s += 0.5 * (px - mx) * (px - mx);
    S[0] = s;
    s = coeff * T[3];
    s += 0.5 * (py - my) * (py - my);
    S[3] = s;
    s = coeff * T[5];
    s += 0.5 * (pz - mz) * (pz - mz);
    S[5] = s;
    s = coeff * T[1];
    s += px * py - p * pxy;
    s += 0.5 * p * p;
    s -= 0.5 * px * px;
    s -= 0.5 * py * py;
    s += 0.5 * pxy * pxy;
    S[1] = s;
    s = coeff * T[2];
    s += px * pz - p * pxz;
    s += 0.5 * p * p;
    s -= 0.5 * px * px;
    s -= 0.5 * pz * pz;
    s += 0.5 * pxz * pxz;
    S[2] = s;
    s = coeff * T[4];
    s += py * pz - p * pyz;
    s += 0.5 * p * p;
    s -= 0.5 * py * py;
    s -= 0.5 * pz * pz;
    s += 0.5 * pyz * pyz;
    S[4] = s;
}

// prepareFTsource (1) in the FP32 form: the same terms as prep_tensor over the
// float fields, evaluated in float
__device__ __forceinline__ void prep_tensor_f32(const float *__restrict__ phi, const float T[6], const Nb &c,
                                                int64_t i, int N, float coeff, float S[6]) {
    const float p = phi[i];
    const float px = phi[L(N, c.xp, c.y, c.z)], mx = phi[L(N, c.xm, c.y, c.z)];
    const float py = phi[L(N, c.x, c.yp, c.z)], my = phi[L(N, c.x, c.ym, c.z)];
    const float pz = phi[L(N, c.x, c.y, c.zp)], mz = phi[L(N, c.x, c.y, c.zm)];
    const float pxy = phi[L(N, c.xp, c.yp, c.z)];
    const float pxz = phi[L(N, c.xp, c.y, c.zp)];
    const float pyz = phi[L(N, c.x, c.yp, c.zp)];
    const float hp = 0.5f * p * p, hx = 0.5f * px * px, hy = 0.5f * py * py, hz = 0.5f * pz * pz;
    S[0] = coeff * T[0] + 0.5f * (px - mx) * (px - mx);
    S[3] = coeff * T[3] + 0.5f * (py - my) * (py - my);
    S[5] = coeff * T[5] + 0.5f * (pz - mz) * (pz - mz);
    S[1] = coeff * T[1] + (px * py - p * pxy) + hp - hx - hy + 0.5f * pxy * pxy;
    S[2] = coeff * T[2] + (px * pz - p * pxz) + hp - hx - hz + 0.5f * pxz * pxz;
    S[4] = coeff * T[4] + (py * pz - p * pyz) + hp - hy - hz + 0.5f * pyz * pyz;
}

}  // namespace gev_prep
