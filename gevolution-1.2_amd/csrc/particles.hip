// particles.hip -- particle-mesh kernels of one box:
//   * CIC projections (gevolution.hpp:926-1297, LATfield2
//     scalarProjectionCIC_project) as an owner-computes gather: one thread
//     per target cell visits the (up to) 8 source cells whose particles
//     deposit into it, in the order a sequential site loop would add them, and
//     stores the sum once.  No atomics; the result is bit-identical to the
//     reference's sequential scatter for the same in-cell particle order.
//   * re-filing (the cell bookkeeping of LATfield2 Particles::moveParticles):
//     per-cell counts, CSR offsets by a three-phase scan, and a permutation
//     listing each cell's particles in ascending ID.
//   * updateVel / moveParticles with update_q(_Newton) / update_pos(_Newton)
//     (gevolution.hpp:570-903), separately or fused into one particle pass.
#include "gev_internal.h"

namespace {

constexpr int TPB = 256;

__device__ __forceinline__ int64_t L(int N, int x, int y, int z) {
    return (int64_t)x + (int64_t)N * ((int64_t)y + (int64_t)N * z);
}
__device__ __forceinline__ int wrapi(int i, int N) { return i < 0 ? i + N : (i >= N ? i - N : i); }

// ------------------------------------------------------------ projections

// KIND: 1 = T00, 2 = Tij, 3 = T00 + Tij, 4 = T0i, 5 = plain CIC (a = 1, phi = NULL)
template <int KIND, bool USE_PHI, bool ACCUM>
__global__ __launch_bounds__(TPB) void k_deposit(
    const double *__restrict__ pos, const double *__restrict__ q, int64_t cap,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ perm,
    const double *__restrict__ phi, double *__restrict__ out_s, double *__restrict__ out_t, int N,
    double a, double mass_s, double mass_t) {
    constexpr bool DO_T00 = KIND == 1 || KIND == 3 || KIND == 5;
    constexpr bool DO_TIJ = KIND == 2 || KIND == 3;
    constexpr bool DO_T0I = KIND == 4;
    const int64_t n3v = (int64_t)N * N * N;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= n3v) return;
    const int x = (int)(t % N);
    const int y = (int)((t / N) % N);
    const int z = (int)(t / ((int64_t)N * N));
    const double dx = 1.0 / (double)N;
    const int xp = x == N - 1 ? 0 : x + 1, yp = y == N - 1 ? 0 : y + 1, zp = z == N - 1 ? 0 : z + 1;

    // phi at the target and its upper neighbours (zero without phi)
    double ph = 0., phx = 0., phy = 0., phz = 0., phxy = 0., phxz = 0., phyz = 0.;
    if (USE_PHI) {
        ph = phi[t];
        phx = phi[L(N, xp, y, z)];
        phy = phi[L(N, x, yp, z)];
        phz = phi[L(N, x, y, zp)];
        phxy = phi[L(N, xp, yp, z)];
        phxz = phi[L(N, xp, y, zp)];
        phyz = phi[L(N, x, yp, zp)];
    }
    // plaquette sums in the reference's operand order (gevolution.hpp:1263-1272)
    const double plq01 = ph + phy + phx + phxy;
    const double plq02 = ph + phz + phx + phxz;
    const double plq12 = ph + phz + phy + phyz;
    // link factors of projection_T0i_project (gevolution.hpp:1129-1144)
    const double lnk0 = 1. + ph + phx, lnk1 = 1. + ph + phy, lnk2 = 1. + ph + phz;

    double acc00 = 0., acc_d[3] = {0., 0., 0.}, acc_o[3] = {0., 0., 0.}, acc_v[3] = {0., 0., 0.};
    if (ACCUM) {
        if (DO_T00) acc00 = out_s[t];
        if (DO_TIJ) {
            acc_d[0] = out_t[t];
            acc_d[1] = out_t[3 * n3v + t];
            acc_d[2] = out_t[5 * n3v + t];
            acc_o[0] = out_t[n3v + t];
            acc_o[1] = out_t[2 * n3v + t];
            acc_o[2] = out_t[4 * n3v + t];
        }
        if (DO_T0I) {
            acc_v[0] = out_t[t];
            acc_v[1] = out_t[n3v + t];
            acc_v[2] = out_t[2 * n3v + t];
        }
    }

    const double *px = pos, *py = pos + cap, *pz = pos + 2 * cap;
    const double *qx = q, *qy = q + cap, *qz = q + 2 * cap;

    // Sources s = t - d, d in {0,1}^3, visited in increasing (wrapped) site
    // index: per axis the lower neighbour first unless it wraps to N-1.
#pragma unroll 1
    for (int iz = 0; iz < 2; iz++) {
        const int dz = z > 0 ? 1 - iz : iz;
        const int sz = wrapi(z - dz, N);
#pragma unroll 1
        for (int iy = 0; iy < 2; iy++) {
            const int dy = y > 0 ? 1 - iy : iy;
            const int sy = wrapi(y - dy, N);
#pragma unroll 1
            for (int ix = 0; ix < 2; ix++) {
                const int dxx = x > 0 ? 1 - ix : ix;
                const int sx = wrapi(x - dxx, N);
                // components this source feeds at t
                const bool f_o01 = DO_TIJ && dxx == 0 && dy == 0;
                const bool f_o02 = DO_TIJ && dxx == 0 && dz == 0;
                const bool f_o12 = DO_TIJ && dy == 0 && dz == 0;
                const bool f_v0 = DO_T0I && dxx == 0;
                const bool f_v1 = DO_T0I && dy == 0;
                const bool f_v2 = DO_T0I && dz == 0;
                const int64_t s = L(N, sx, sy, sz);
                const uint32_t b0 = start[s], b1 = start[s + 1];
                if (b0 == b1) continue;
                const double rx = sx * dx, ry = sy * dx, rz = sz * dx;
                double lc = 0., ld[3] = {0., 0., 0.}, lo01 = 0., lo02 = 0., lo12 = 0.;
                double lv0 = 0., lv1 = 0., lv2 = 0.;
                for (uint32_t k = b0; k < b1; k++) {
                    const uint32_t p = perm[k];
                    const double up0 = (px[p] - rx) / dx, up1 = (py[p] - ry) / dx, up2 = (pz[p] - rz) / dx;
                    const double dn0 = 1.0 - up0, dn1 = 1.0 - up1, dn2 = 1.0 - up2;
                    const double W0 = dxx ? up0 : dn0, W1 = dy ? up1 : dn1, W2 = dz ? up2 : dn2;
                    if (DO_T00) {
                        double e = a, f = 0.;
                        if (KIND != 5 && USE_PHI) {
                            const double q0 = qx[p], q1 = qy[p], q2 = qz[p];
                            f = q0 * q0 + q1 * q1 + q2 * q2;
                            e = sqrt(f + a * a);
                            f = 3. * e + f / e;
                        }
                        lc += W0 * W1 * W2 * (e + f * ph);
                    }
                    if (DO_TIJ) {
                        const double q0 = qx[p], q1 = qy[p], q2 = qz[p];
                        double f = q0 * q0 + q1 * q1 + q2 * q2;
                        const double e = sqrt(f + a * a);
                        f = 4. + a * a / (f + a * a);
                        const double qq[3] = {q0, q1, q2};
#pragma unroll
                        for (int i = 0; i < 3; i++) {
                            const double w = mass_t * qq[i] * qq[i] / e;
                            ld[i] += w * W0 * W1 * W2 * (1. + f * ph);
                        }
                        if (f_o01) lo01 += mass_t * q0 * q1 / e * W2 * (1. + f * 0.25 * plq01);
                        if (f_o02) lo02 += mass_t * q0 * q2 / e * W1 * (1. + f * 0.25 * plq02);
                        if (f_o12) lo12 += mass_t * q1 * q2 / e * W0 * (1. + f * 0.25 * plq12);
                    }
                    if (DO_T0I) {
                        if (f_v0) lv0 += mass_t * qx[p] * W1 * W2;
                        if (f_v1) lv1 += mass_t * qy[p] * W0 * W2;
                        if (f_v2) lv2 += mass_t * qz[p] * W0 * W1;
                    }
                }
                if (DO_T00) acc00 += lc * mass_s;
                if (DO_TIJ) {
                    acc_d[0] += ld[0];
                    acc_d[1] += ld[1];
                    acc_d[2] += ld[2];
                    if (f_o01) acc_o[0] += lo01;
                    if (f_o02) acc_o[1] += lo02;
                    if (f_o12) acc_o[2] += lo12;
                }
                if (DO_T0I) {
                    if (f_v0) acc_v[0] += lv0 * lnk0;
                    if (f_v1) acc_v[1] += lv1 * lnk1;
                    if (f_v2) acc_v[2] += lv2 * lnk2;
                }
            }
        }
    }
    if (DO_T00) out_s[t] = acc00;
    if (DO_TIJ) {
        out_t[t] = acc_d[0];
        out_t[3 * n3v + t] = acc_d[1];
        out_t[5 * n3v + t] = acc_d[2];
        out_t[n3v + t] = acc_o[0];
        out_t[2 * n3v + t] = acc_o[1];
        out_t[4 * n3v + t] = acc_o[2];
    }
    if (DO_T0I) {
        out_t[t] = acc_v[0];
        out_t[n3v + t] = acc_v[1];
        out_t[2 * n3v + t] = acc_v[2];
    }
}

// ------------------------------------------------------------ filing

// Cell of one coordinate: floor(p/dx) with a true division.
__device__ __forceinline__ int cell1(double p, double dx) { return (int)floor(p / dx); }

__global__ __launch_bounds__(TPB) void k_count_cells(const double *__restrict__ pos, int64_t cap,
                                                     int64_t n, int N, uint32_t *__restrict__ count,
                                                     uint32_t *__restrict__ cell,
                                                     uint32_t *__restrict__ slot) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    const double dx = 1.0 / (double)N;
    const int cx = cell1(pos[p], dx), cy = cell1(pos[cap + p], dx), cz = cell1(pos[2 * cap + p], dx);
    const uint32_t c = (uint32_t)L(N, cx, cy, cz);
    cell[p] = c;
    slot[p] = atomicAdd(count + c, 1u);
}

// three-phase exclusive scan of per-cell counts into CSR offsets
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_BLOCK = TPB * SCAN_ITEMS;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t sh[TPB];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < TPB; off <<= 1) {
        uint32_t u = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += u;
        __syncthreads();
    }
    uint32_t incl = sh[threadIdx.x];
    *total = sh[TPB - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(TPB) void k_scan_reduce(const uint32_t *__restrict__ in, int64_t n,
                                                     uint32_t *__restrict__ part) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
        if (base + k < n) s += in[base + k];
    __shared__ uint32_t sh[TPB];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// exclusive scan of the block partials in place (single workgroup)
__global__ __launch_bounds__(TPB) void k_scan_partials(uint32_t *__restrict__ part, int64_t nblk) {
    const int64_t per = (nblk + TPB - 1) / TPB;
    const int64_t b = (int64_t)threadIdx.x * per;
    uint32_t s = 0;
    for (int64_t k = 0; k < per; k++)
        if (b + k < nblk) s += part[b + k];
    uint32_t total;
    uint32_t off = block_exclusive_scan(s, &total);
    for (int64_t k = 0; k < per; k++)
        if (b + k < nblk) {
            uint32_t v = part[b + k];
            part[b + k] = off;
            off += v;
        }
}

__global__ __launch_bounds__(TPB) void k_scan_final(const uint32_t *__restrict__ in, int64_t n,
                                                    const uint32_t *__restrict__ part,
                                                    uint32_t *__restrict__ out) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan(s, &total) + part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = off;
        off += v[k];
    }
    // the last block also writes out[n] = grand total
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == TPB - 1) out[n] = off;
}

__global__ __launch_bounds__(TPB) void k_scatter(const uint32_t *__restrict__ cell,
                                                 const uint32_t *__restrict__ slot,
                                                 const uint32_t *__restrict__ start, int64_t n,
                                                 uint32_t *__restrict__ perm) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    perm[start[cell[p]] + slot[p]] = (uint32_t)p;
}

// ascending ID inside every cell (insertion sort; cells hold few particles)
__global__ __launch_bounds__(TPB) void k_sort_cells(const uint32_t *__restrict__ start,
                                                    const int64_t *__restrict__ id, int64_t n3v,
                                                    uint32_t *__restrict__ perm) {
    const int64_t c = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (c >= n3v) return;
    const uint32_t b0 = start[c], b1 = start[c + 1];
    if (b1 - b0 < 2) return;
    for (uint32_t i = b0 + 1; i < b1; i++) {
        const uint32_t v = perm[i];
        const int64_t key = id[v];
        uint32_t j = i;
        while (j > b0 && id[perm[j - 1]] > key) {
            perm[j] = perm[j - 1];
            j--;
        }
        perm[j] = v;
    }
}

// ------------------------------------------------------------ kick / drift

struct Nbr {
    int x, y, z, xm, xp, ym, yp, zm, zp;
};

struct FieldView {
    const double *f;
    int N;
    Nbr c;
    __device__ __forceinline__ double operator()(int dx, int dy, int dz) const {
        const int X = dx < 0 ? c.xm : (dx > 0 ? c.xp : c.x);
        const int Y = dy < 0 ? c.ym : (dy > 0 ? c.yp : c.y);
        const int Z = dz < 0 ? c.zm : (dz > 0 ? c.zp : c.z);
        return f[L(N, X, Y, Z)];
    }
};

// update_q, gevolution.hpp:570-678 (GRADIENT_ORDER 1); returns v^2/a^2
__device__ __forceinline__ double update_q(double q[3], const double rd[3], FieldView phi,
                                           FieldView chi, FieldView B0, FieldView B1, FieldView B2,
                                           int nfield, const double params[2], double dtau,
                                           double dx) {
    double gradphi[3], pgradB[3];
    double v2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    double e2 = v2 + params[0] * params[0];

    gradphi[0] = (1. - rd[1]) * (1. - rd[2]) * (phi(1, 0, 0) - phi(0, 0, 0));
    gradphi[1] = (1. - rd[0]) * (1. - rd[2]) * (phi(0, 1, 0) - phi(0, 0, 0));
    gradphi[2] = (1. - rd[0]) * (1. - rd[1]) * (phi(0, 0, 1) - phi(0, 0, 0));
    gradphi[0] += rd[1] * (1. - rd[2]) * (phi(1, 1, 0) - phi(0, 1, 0));
    gradphi[1] += rd[0] * (1. - rd[2]) * (phi(1, 1, 0) - phi(1, 0, 0));
    gradphi[2] += rd[0] * (1. - rd[1]) * (phi(1, 0, 1) - phi(1, 0, 0));
    gradphi[0] += (1. - rd[1]) * rd[2] * (phi(1, 0, 1) - phi(0, 0, 1));
    gradphi[1] += (1. - rd[0]) * rd[2] * (phi(0, 1, 1) - phi(0, 0, 1));
    gradphi[2] += (1. - rd[0]) * rd[1] * (phi(0, 1, 1) - phi(0, 1, 0));
    gradphi[0] += rd[1] * rd[2] * (phi(1, 1, 1) - phi(0, 1, 1));
    gradphi[1] += rd[0] * rd[2] * (phi(1, 1, 1) - phi(1, 0, 1));
    gradphi[2] += rd[0] * rd[1] * (phi(1, 1, 1) - phi(1, 1, 0));

    gradphi[0] *= (v2 + e2) / e2;
    gradphi[1] *= (v2 + e2) / e2;
    gradphi[2] *= (v2 + e2) / e2;

    if (nfield >= 2) {
        gradphi[0] -= (1. - rd[1]) * (1. - rd[2]) * (chi(1, 0, 0) - chi(0, 0, 0));
        gradphi[1] -= (1. - rd[0]) * (1. - rd[2]) * (chi(0, 1, 0) - chi(0, 0, 0));
        gradphi[2] -= (1. - rd[0]) * (1. - rd[1]) * (chi(0, 0, 1) - chi(0, 0, 0));
        gradphi[0] -= rd[1] * (1. - rd[2]) * (chi(1, 1, 0) - chi(0, 1, 0));
        gradphi[1] -= rd[0] * (1. - rd[2]) * (chi(1, 1, 0) - chi(1, 0, 0));
        gradphi[2] -= rd[0] * (1. - rd[1]) * (chi(1, 0, 1) - chi(1, 0, 0));
        gradphi[0] -= (1. - rd[1]) * rd[2] * (chi(1, 0, 1) - chi(0, 0, 1));
        gradphi[1] -= (1. - rd[0]) * rd[2] * (chi(0, 1, 1) - chi(0, 0, 1));
        gradphi[2] -= (1. - rd[0]) * rd[1] * (chi(0, 1, 1) - chi(0, 1, 0));
        gradphi[0] -= rd[1] * rd[2] * (chi(1, 1, 1) - chi(0, 1, 1));
        gradphi[1] -= rd[0] * rd[2] * (chi(1, 1, 1) - chi(1, 0, 1));
        gradphi[2] -= rd[0] * rd[1] * (chi(1, 1, 1) - chi(1, 1, 0));
    }

    e2 = sqrt(e2);

    if (nfield >= 3) {
        pgradB[0] = ((1. - rd[2]) * (B1(1, 0, 0) - B1(0, 0, 0)) + rd[2] * (B1(1, 0, 1) - B1(0, 0, 1))) * q[1];
        pgradB[0] += ((1. - rd[1]) * (B2(1, 0, 0) - B2(0, 0, 0)) + rd[1] * (B2(1, 1, 0) - B2(0, 1, 0))) * q[2];
        pgradB[0] += (1. - rd[1]) * (1. - rd[2]) * ((rd[0] - 1.) * B0(-1, 0, 0) + (1. - 2. * rd[0]) * B0(0, 0, 0) + rd[0] * B0(1, 0, 0)) * q[0];
        pgradB[0] += rd[1] * (1. - rd[2]) * ((rd[0] - 1.) * B0(-1, 1, 0) + (1. - 2. * rd[0]) * B0(0, 1, 0) + rd[0] * B0(1, 1, 0)) * q[0];
        pgradB[0] += (1. - rd[1]) * rd[2] * ((rd[0] - 1.) * B0(-1, 0, 1) + (1. - 2. * rd[0]) * B0(0, 0, 1) + rd[0] * B0(1, 0, 1)) * q[0];
        pgradB[0] += rd[1] * rd[2] * ((rd[0] - 1.) * B0(-1, 1, 1) + (1. - 2. * rd[0]) * B0(0, 1, 1) + rd[0] * B0(1, 1, 1)) * q[0];

        pgradB[1] = ((1. - rd[0]) * (B2(0, 1, 0) - B2(0, 0, 0)) + rd[0] * (B2(1, 1, 0) - B2(1, 0, 0))) * q[2];
        pgradB[1] += ((1. - rd[2]) * (B0(0, 1, 0) - B0(0, 0, 0)) + rd[2] * (B0(0, 1, 1) - B0(0, 0, 1))) * q[0];
        pgradB[1] += (1. - rd[0]) * (1. - rd[2]) * ((rd[1] - 1.) * B1(0, -1, 0) + (1. - 2. * rd[1]) * B1(0, 0, 0) + rd[1] * B1(0, 1, 0)) * q[1];
        pgradB[1] += rd[0] * (1. - rd[2]) * ((rd[1] - 1.) * B1(1, -1, 0) + (1. - 2. * rd[1]) * B1(1, 0, 0) + rd[1] * B1(1, 1, 0)) * q[1];
        pgradB[1] += (1. - rd[0]) * rd[2] * ((rd[1] - 1.) * B1(0, -1, 1) + (1. - 2. * rd[1]) * B1(0, 0, 1) + rd[1] * B1(0, 1, 1)) * q[1];
        pgradB[1] += rd[0] * rd[2] * ((rd[1] - 1.) * B1(1, -1, 1) + (1. - 2. * rd[1]) * B1(1, 0, 1) + rd[1] * B1(1, 1, 1)) * q[1];

        pgradB[2] = ((1. - rd[1]) * (B0(0, 0, 1) - B0(0, 0, 0)) + rd[1] * (B0(0, 1, 1) - B0(0, 1, 0))) * q[0];
        pgradB[2] += ((1. - rd[0]) * (B1(0, 0, 1) - B1(0, 0, 0)) + rd[0] * (B1(1, 0, 1) - B1(1, 0, 0))) * q[1];
        pgradB[2] += (1. - rd[0]) * (1. - rd[1]) * ((rd[2] - 1.) * B2(0, 0, -1) + (1. - 2. * rd[2]) * B2(0, 0, 0) + rd[2] * B2(0, 0, 1)) * q[2];
        pgradB[2] += rd[0] * (1. - rd[1]) * ((rd[2] - 1.) * B2(1, 0, -1) + (1. - 2. * rd[2]) * B2(1, 0, 0) + rd[2] * B2(1, 0, 1)) * q[2];
        pgradB[2] += (1. - rd[0]) * rd[1] * ((rd[2] - 1.) * B2(0, 1, -1) + (1. - 2. * rd[2]) * B2(0, 1, 0) + rd[2] * B2(0, 1, 1)) * q[2];
        pgradB[2] += rd[0] * rd[1] * ((rd[2] - 1.) * B2(1, 1, -1) + (1. - 2. * rd[2]) * B2(1, 1, 0) + rd[2] * B2(1, 1, 1)) * q[2];

        gradphi[0] += pgradB[0] / params[1] / e2;
        gradphi[1] += pgradB[1] / params[1] / e2;
        gradphi[2] += pgradB[2] / params[1] / e2;
    }

    v2 = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        q[i] -= dtau * e2 * gradphi[i] / dx;
        v2 += q[i] * q[i];
    }
    return v2 / params[0] / params[0];
}

// update_q_Newton, gevolution.hpp:709-776
__device__ __forceinline__ double update_q_newton(double q[3], const double rd[3], FieldView psi,
                                                  FieldView chi, int nfield, const double params[2],
                                                  double dtau, double dx) {
    double g[3];
    g[0] = (1. - rd[1]) * (1. - rd[2]) * (psi(1, 0, 0) - psi(0, 0, 0));
    g[1] = (1. - rd[0]) * (1. - rd[2]) * (psi(0, 1, 0) - psi(0, 0, 0));
    g[2] = (1. - rd[0]) * (1. - rd[1]) * (psi(0, 0, 1) - psi(0, 0, 0));
    g[0] += rd[1] * (1. - rd[2]) * (psi(1, 1, 0) - psi(0, 1, 0));
    g[1] += rd[0] * (1. - rd[2]) * (psi(1, 1, 0) - psi(1, 0, 0));
    g[2] += rd[0] * (1. - rd[1]) * (psi(1, 0, 1) - psi(1, 0, 0));
    g[0] += (1. - rd[1]) * rd[2] * (psi(1, 0, 1) - psi(0, 0, 1));
    g[1] += (1. - rd[0]) * rd[2] * (psi(0, 1, 1) - psi(0, 0, 1));
    g[2] += (1. - rd[0]) * rd[1] * (psi(0, 1, 1) - psi(0, 1, 0));
    g[0] += rd[1] * rd[2] * (psi(1, 1, 1) - psi(0, 1, 1));
    g[1] += rd[0] * rd[2] * (psi(1, 1, 1) - psi(1, 0, 1));
    g[2] += rd[0] * rd[1] * (psi(1, 1, 1) - psi(1, 1, 0));
    if (nfield >= 2) {
        g[0] -= (1. - rd[1]) * (1. - rd[2]) * (chi(1, 0, 0) - chi(0, 0, 0));
        g[1] -= (1. - rd[0]) * (1. - rd[2]) * (chi(0, 1, 0) - chi(0, 0, 0));
        g[2] -= (1. - rd[0]) * (1. - rd[1]) * (chi(0, 0, 1) - chi(0, 0, 0));
        g[0] -= rd[1] * (1. - rd[2]) * (chi(1, 1, 0) - chi(0, 1, 0));
        g[1] -= rd[0] * (1. - rd[2]) * (chi(1, 1, 0) - chi(1, 0, 0));
        g[2] -= rd[0] * (1. - rd[1]) * (chi(1, 0, 1) - chi(1, 0, 0));
        g[0] -= (1. - rd[1]) * rd[2] * (chi(1, 0, 1) - chi(0, 0, 1));
        g[1] -= (1. - rd[0]) * rd[2] * (chi(0, 1, 1) - chi(0, 0, 1));
        g[2] -= (1. - rd[0]) * rd[1] * (chi(0, 1, 1) - chi(0, 1, 0));
        g[0] -= rd[1] * rd[2] * (chi(1, 1, 1) - chi(0, 1, 1));
        g[1] -= rd[0] * rd[2] * (chi(1, 1, 1) - chi(1, 0, 1));
        g[2] -= rd[0] * rd[1] * (chi(1, 1, 1) - chi(1, 1, 0));
    }
    double v2 = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        q[i] -= dtau * params[0] * g[i] / dx;
        v2 += q[i] * q[i];
    }
    return v2 / params[0] / params[0];
}

// update_pos, gevolution.hpp:810-871
__device__ __forceinline__ void update_pos(double pos[3], const double q[3], const double rd[3],
                                           FieldView phi, FieldView chi, FieldView B0,
                                           FieldView B1, FieldView B2, int nfield,
                                           const double params[2], double dtau) {
    double v[3];
    double v2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    double e2 = v2 + params[0] * params[0];
    double ph = 0., ch = 0.;
    if (nfield >= 1) {
        ph = phi(0, 0, 0) * (1. - rd[0]) * (1. - rd[1]) * (1. - rd[2]);
        ph += phi(1, 0, 0) * rd[0] * (1. - rd[1]) * (1. - rd[2]);
        ph += phi(0, 1, 0) * (1. - rd[0]) * rd[1] * (1. - rd[2]);
        ph += phi(1, 1, 0) * rd[0] * rd[1] * (1. - rd[2]);
        ph += phi(0, 0, 1) * (1. - rd[0]) * (1. - rd[1]) * rd[2];
        ph += phi(1, 0, 1) * rd[0] * (1. - rd[1]) * rd[2];
        ph += phi(0, 1, 1) * (1. - rd[0]) * rd[1] * rd[2];
        ph += phi(1, 1, 1) * rd[0] * rd[1] * rd[2];
    }
    if (nfield >= 2) {
        ch = chi(0, 0, 0) * (1. - rd[0]) * (1. - rd[1]) * (1. - rd[2]);
        ch += chi(1, 0, 0) * rd[0] * (1. - rd[1]) * (1. - rd[2]);
        ch += chi(0, 1, 0) * (1. - rd[0]) * rd[1] * (1. - rd[2]);
        ch += chi(1, 1, 0) * rd[0] * rd[1] * (1. - rd[2]);
        ch += chi(0, 0, 1) * (1. - rd[0]) * (1. - rd[1]) * rd[2];
        ch += chi(1, 0, 1) * rd[0] * (1. - rd[1]) * rd[2];
        ch += chi(0, 1, 1) * (1. - rd[0]) * rd[1] * rd[2];
        ch += chi(1, 1, 1) * rd[0] * rd[1] * rd[2];
    }
    v2 = (1. + (3. - v2 / e2) * ph - ch) / sqrt(e2);
    v[0] = q[0] * v2;
    v[1] = q[1] * v2;
    v[2] = q[2] * v2;
    if (nfield >= 3) {
        double b[3];
        b[0] = B0(0, 0, 0) * (1. - rd[1]) * (1. - rd[2]);
        b[1] = B1(0, 0, 0) * (1. - rd[0]) * (1. - rd[2]);
        b[2] = B2(0, 0, 0) * (1. - rd[0]) * (1. - rd[1]);
        b[1] += B1(1, 0, 0) * rd[0] * (1. - rd[2]);
        b[2] += B2(1, 0, 0) * rd[0] * (1. - rd[1]);
        b[0] += B0(0, 1, 0) * rd[1] * (1. - rd[2]);
        b[2] += B2(0, 1, 0) * (1. - rd[0]) * rd[1];
        b[0] += B0(0, 0, 1) * (1. - rd[1]) * rd[2];
        b[1] += B1(0, 0, 1) * (1. - rd[0]) * rd[2];
        b[1] += B1(1, 0, 1) * rd[0] * rd[2];
        b[0] += B0(0, 1, 1) * rd[1] * rd[2];
        b[2] += B2(1, 1, 0) * rd[0] * rd[1];
#pragma unroll
        for (int l = 0; l < 3; l++) pos[l] += dtau * (v[l] + b[l] / params[1]);
    } else {
#pragma unroll
        for (int l = 0; l < 3; l++) pos[l] += dtau * v[l];
    }
}

// periodic wrap of the re-filing step; a coordinate whose cell would round
// to N is set to 0 so that 0 <= floor(p/dx) < N for every filed particle
__device__ __forceinline__ double wrap_pos(double p, double dx, int N) {
    if (p >= 1.0) p -= 1.0;
    if (p < 0.0) p += 1.0;
    // beyond one box length: reduce modulo 1; non-finite: file at 0
    if (!(p >= 0.0 && p < 1.0)) p = p - floor(p);
    if (!(p >= 0.0) || (int)floor(p / dx) >= N) p = 0.0;
    return p;
}

struct StepArgs {
    int nfk, nfd;
    double dtk, dtd;
    double pk[2], pd[2];
};

// Per particle: locate, optional kick, optional drift (+wrap + count into
// the new cell).  QM / PM select the callbacks (0 = GR, 1 = Newton).
template <bool KICK, int QM, bool DRIFT, int PM>
__global__ __launch_bounds__(TPB) void k_update(double *__restrict__ pos, double *__restrict__ q,
                                                int64_t cap, int64_t n, int N,
                                                const double *__restrict__ phi,
                                                const double *__restrict__ chi,
                                                const double *__restrict__ Bi, StepArgs sa,
                                                unsigned long long *__restrict__ vmax,
                                                uint32_t *__restrict__ count,
                                                uint32_t *__restrict__ cell,
                                                uint32_t *__restrict__ slot) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    double r = 0.;
    if (p < n) {
        const double dx = 1.0 / (double)N;
        const int64_t n3v = (int64_t)N * N * N;
        double x[3] = {pos[p], pos[cap + p], pos[2 * cap + p]};
        double qq[3] = {q[p], q[cap + p], q[2 * cap + p]};
        // filed site and ref_dist = modf(pos/dx)
        double rd[3];
        int ci[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double t = x[i] / dx;
            const double fl = floor(t);
            ci[i] = (int)fl;
            rd[i] = t - fl;
        }
        Nbr nb;
        nb.x = ci[0];
        nb.y = ci[1];
        nb.z = ci[2];
        nb.xm = nb.x == 0 ? N - 1 : nb.x - 1;
        nb.xp = nb.x == N - 1 ? 0 : nb.x + 1;
        nb.ym = nb.y == 0 ? N - 1 : nb.y - 1;
        nb.yp = nb.y == N - 1 ? 0 : nb.y + 1;
        nb.zm = nb.z == 0 ? N - 1 : nb.z - 1;
        nb.zp = nb.z == N - 1 ? 0 : nb.z + 1;
        FieldView fphi{phi, N, nb}, fchi{chi, N, nb};
        FieldView fb0{Bi, N, nb}, fb1{Bi + n3v, N, nb}, fb2{Bi + 2 * n3v, N, nb};
        if (KICK) {
            if (QM == 0)
                r = update_q(qq, rd, fphi, fchi, fb0, fb1, fb2, sa.nfk, sa.pk, sa.dtk, dx);
            else
                r = update_q_newton(qq, rd, fphi, fchi, sa.nfk, sa.pk, sa.dtk, dx);
            q[p] = qq[0];
            q[cap + p] = qq[1];
            q[2 * cap + p] = qq[2];
        }
        if (DRIFT) {
            if (PM == 0) {
                update_pos(x, qq, rd, fphi, fchi, fb0, fb1, fb2, sa.nfd, sa.pd, sa.dtd);
            } else {
#pragma unroll
                for (int l = 0; l < 3; l++) x[l] += sa.dtd * qq[l] / sa.pd[0];  // gevolution.hpp:902
            }
#pragma unroll
            for (int l = 0; l < 3; l++) x[l] = wrap_pos(x[l], dx, N);
            pos[p] = x[0];
            pos[cap + p] = x[1];
            pos[2 * cap + p] = x[2];
            const uint32_t c = (uint32_t)L(N, cell1(x[0], dx), cell1(x[1], dx), cell1(x[2], dx));
            cell[p] = c;
            slot[p] = atomicAdd(count + c, 1u);
        }
    }
    if (KICK) {
        // max of the callback returns (>= 0, so the bit patterns order as integers)
        unsigned long long b = __double_as_longlong(r);
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(b, off);
            b = o > b ? o : b;
        }
        __shared__ unsigned long long sh[TPB / 64];
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = b;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = sh[0];
            for (int w = 1; w < TPB / 64; w++) m = sh[w] > m ? sh[w] : m;
            atomicMax(vmax, m);
        }
    }
}

template <bool KICK, int QM, bool DRIFT, int PM>
void launch_update_t(gev_ctx *ctx, const StepArgs &sa) {
    hipLaunchKernelGGL((k_update<KICK, QM, DRIFT, PM>), dim3(gev::grid_for(ctx->n, TPB)), dim3(TPB),
                       0, ctx->stream, ctx->pos, ctx->q, ctx->cap, ctx->n, ctx->N, ctx->phi,
                       ctx->chi, ctx->Bi, sa, ctx->red_u, ctx->count, ctx->cell, ctx->slot);
}

template <int KIND, bool USE_PHI, bool ACCUM>
void launch_dep_t(gev_ctx *ctx, double *out_s, double *out_t, double a, double ms, double mt) {
    hipLaunchKernelGGL((k_deposit<KIND, USE_PHI, ACCUM>), dim3(gev::grid_for(gev::n3(ctx), TPB)),
                       dim3(TPB), 0, ctx->stream, ctx->pos, ctx->q, ctx->cap, ctx->start, ctx->perm,
                       ctx->phi, out_s, out_t, ctx->N, a, ms, mt);
}

template <int KIND>
void launch_dep_k(gev_ctx *ctx, bool use_phi, bool accum, double *out_s, double *out_t, double a,
                  double ms, double mt) {
    if (use_phi) {
        if (accum) launch_dep_t<KIND, true, true>(ctx, out_s, out_t, a, ms, mt);
        else launch_dep_t<KIND, true, false>(ctx, out_s, out_t, a, ms, mt);
    } else {
        if (accum) launch_dep_t<KIND, false, true>(ctx, out_s, out_t, a, ms, mt);
        else launch_dep_t<KIND, false, false>(ctx, out_s, out_t, a, ms, mt);
    }
}

}  // namespace

namespace gev {

int launch_deposit(gev_ctx *ctx, int kind, double a, int use_phi, double coeff) {
    const double dx = 1.0 / (double)ctx->N;
    // mass factors of gevolution.hpp:945-947 / 1064-1065 / 1191-1193
    double m = coeff / (dx * dx * dx);
    m *= ctx->pmass;
    double m_a = m / a;
    switch (kind) {
    case DEP_T00: {
        ScopedTimer t(ctx, "projection_T00_project");
        launch_dep_k<1>(ctx, use_phi, !ctx->zeroed[GEV_SOURCE], ctx->source, nullptr, a, m_a, m_a);
        ctx->zeroed[GEV_SOURCE] = false;
        break;
    }
    case DEP_TIJ: {
        ScopedTimer t(ctx, "projection_Tij_project");
        launch_dep_k<2>(ctx, use_phi, !ctx->zeroed[GEV_SIJ], nullptr, ctx->Sij, a, m_a, m_a);
        ctx->zeroed[GEV_SIJ] = false;
        break;
    }
    case DEP_T00_TIJ: {
        ScopedTimer t(ctx, "projection_T00_Tij_project");
        const bool acc = !ctx->zeroed[GEV_SOURCE] || !ctx->zeroed[GEV_SIJ];
        if (acc && (ctx->zeroed[GEV_SOURCE] != ctx->zeroed[GEV_SIJ])) {
            // mixed state: accumulate into both (the zeroed one reads zeros)
        }
        launch_dep_k<3>(ctx, true, acc, ctx->source, ctx->Sij, a, m_a, m_a);
        ctx->zeroed[GEV_SOURCE] = false;
        ctx->zeroed[GEV_SIJ] = false;
        break;
    }
    case DEP_T0I: {
        ScopedTimer t(ctx, "projection_T0i_project");
        launch_dep_k<4>(ctx, use_phi, !ctx->zeroed[GEV_BI], nullptr, ctx->Bi, 1.0, m, m);
        ctx->zeroed[GEV_BI] = false;
        break;
    }
    case DEP_CIC: {
        ScopedTimer t(ctx, "scalar_projection_cic_project");
        // projection_T00_project with a = 1 and phi = NULL
        launch_dep_k<5>(ctx, false, !ctx->zeroed[GEV_SOURCE], ctx->source, nullptr, 1.0, m, m);
        ctx->zeroed[GEV_SOURCE] = false;
        break;
    }
    default:
        return fail(ctx, GEV_EINVAL, "unknown deposit kind");
    }
    return check_hip(ctx, hipGetLastError(), "deposit launch");
}

// scan count[] into start[] and build perm[] from cell[] / slot[]
static int finish_filing(gev_ctx *ctx) {
    const int64_t n3v = n3(ctx);
    const int64_t nblk = (n3v + SCAN_BLOCK - 1) / SCAN_BLOCK;
    {
        ScopedTimer t(ctx, "filing_scan");
        hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nblk), dim3(TPB), 0, ctx->stream, ctx->count,
                           n3v, ctx->scan_part);
        hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(TPB), 0, ctx->stream, ctx->scan_part, nblk);
        hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nblk), dim3(TPB), 0, ctx->stream, ctx->count,
                           n3v, ctx->scan_part, ctx->start);
    }
    if (ctx->n > 0) {
        ScopedTimer t(ctx, "filing_scatter");
        hipLaunchKernelGGL(k_scatter, dim3(grid_for(ctx->n, TPB)), dim3(TPB), 0, ctx->stream,
                           ctx->cell, ctx->slot, ctx->start, ctx->n, ctx->perm);
    }
    {
        ScopedTimer t(ctx, "filing_sort");
        hipLaunchKernelGGL(k_sort_cells, dim3(grid_for(n3v, TPB)), dim3(TPB), 0, ctx->stream,
                           ctx->start, ctx->id, n3v, ctx->perm);
    }
    return check_hip(ctx, hipGetLastError(), "filing launch");
}

int launch_refile(gev_ctx *ctx) {
    int rc = check_hip(ctx, hipMemsetAsync(ctx->count, 0, sizeof(uint32_t) * n3(ctx), ctx->stream),
                       "count clear");
    if (rc) return rc;
    if (ctx->n > 0) {
        ScopedTimer t(ctx, "filing_count");
        hipLaunchKernelGGL(k_count_cells, dim3(grid_for(ctx->n, TPB)), dim3(TPB), 0, ctx->stream,
                           ctx->pos, ctx->cap, ctx->n, ctx->N, ctx->count, ctx->cell, ctx->slot);
    }
    rc = check_hip(ctx, hipGetLastError(), "count launch");
    if (rc) return rc;
    return finish_filing(ctx);
}

int launch_particle_update(gev_ctx *ctx, bool kick, int qm, double dtk, int nfk, const double pk[2],
                           bool drift, int pm, double dtd, int nfd, const double pd[2],
                           double *maxvel) {
    StepArgs sa;
    sa.nfk = nfk;
    sa.nfd = nfd;
    sa.dtk = dtk;
    sa.dtd = dtd;
    sa.pk[0] = kick ? pk[0] : 1.0;
    sa.pk[1] = kick ? pk[1] : 1.0;
    sa.pd[0] = drift ? pd[0] : 1.0;
    sa.pd[1] = drift ? pd[1] : 1.0;
    int rc;
    if (kick) {
        rc = check_hip(ctx, hipMemsetAsync(ctx->red_u, 0, sizeof(unsigned long long), ctx->stream),
                       "vmax clear");
        if (rc) return rc;
    }
    if (drift) {
        rc = check_hip(ctx, hipMemsetAsync(ctx->count, 0, sizeof(uint32_t) * n3(ctx), ctx->stream),
                       "count clear");
        if (rc) return rc;
    }
    if (ctx->n > 0) {
        const char *name = kick && drift ? "kick_drift" : (kick ? "update_vel" : "move_particles");
        ScopedTimer t(ctx, name);
        if (kick && drift) {
            if (qm == 0 && pm == 0) launch_update_t<true, 0, true, 0>(ctx, sa);
            else if (qm == 1 && pm == 1) launch_update_t<true, 1, true, 1>(ctx, sa);
            else if (qm == 0 && pm == 1) launch_update_t<true, 0, true, 1>(ctx, sa);
            else launch_update_t<true, 1, true, 0>(ctx, sa);
        } else if (kick) {
            if (qm == 0) launch_update_t<true, 0, false, 0>(ctx, sa);
            else launch_update_t<true, 1, false, 0>(ctx, sa);
        } else if (drift) {
            if (pm == 0) launch_update_t<false, 0, true, 0>(ctx, sa);
            else launch_update_t<false, 0, true, 1>(ctx, sa);
        }
    }
    rc = check_hip(ctx, hipGetLastError(), "particle update launch");
    if (rc) return rc;
    if (drift) {
        rc = finish_filing(ctx);
        if (rc) return rc;
    }
    if (kick && maxvel) {
        rc = check_hip(ctx, hipMemcpyAsync(ctx->host_pinned, ctx->red_u, sizeof(double),
                                           hipMemcpyDeviceToHost, ctx->stream),
                       "vmax readback");
        if (rc) return rc;
        rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "vmax sync");
        if (rc) return rc;
        *maxvel = sqrt(ctx->host_pinned[0]);
    }
    return 0;
}

}  // namespace gev
