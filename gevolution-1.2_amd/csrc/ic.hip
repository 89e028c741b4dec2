// ic.hip -- device side of the basic initial-condition generator
// (ic_basic.hpp): the counter-based PRNG (prng_engine.hpp), the Gaussian
// realisation of generateDisplacementField, the particle lattice of
// initializeParticlePositions and the two particle callbacks
// displace_pcls_ic_basic / initialize_q_ic_basic.
//
// PRNG.  sitmo::prng_engine is Threefish-256 (20 rounds, zero tweak) keyed by
// the seed and run on a 256-bit block counter; output number p is 32-bit half
// (p & 7) of block p >> 3.  discard() only moves p, so every mode of the
// realisation can compute its own draws: a thread per k point, no stream
// state.
//
// Realisation.  The reference walks the four (ky, kz) quadrants of the k grid
// with one stream each, every row starting HUGE_SKIP (65536) draws after the
// previous one, every drawn mode taking two draws (r1, r2), r1 redrawn while
// zero (ic_basic.hpp:1127-1366).  Drawn modes form a prefix of each row, so a
// mode's draws sit at a closed-form offset unless an earlier r1 of the row was
// zero (probability 2^-32 per draw).  The main kernel assumes no redraw and
// marks the row when its own r1 is zero; a second kernel recomputes marked
// rows sequentially, exactly as the reference loop.
#include "gev_internal.h"

namespace {

constexpr int TPB = 256;
constexpr uint64_t HUGE_SKIP = 65536;

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

// Threefish-256, 20 rounds, key (seed, 0, 0, 0), tweak 0, plaintext (counter, 0, 0, 0)
__device__ __forceinline__ void threefish_block(uint32_t seed, uint64_t counter, uint64_t o[4]) {
    const uint64_t k0 = seed, k4 = 0x1BD11BDAA9FC1A22ull ^ k0;
    const uint64_t ks[5] = {k0, 0, 0, 0, k4};
    uint64_t x0 = counter, x1 = 0, x2 = 0, x3 = 0;
    constexpr int R0[8] = {14, 52, 23, 5, 25, 46, 58, 32};
    constexpr int R1[8] = {16, 57, 40, 37, 33, 12, 22, 32};
#pragma unroll
    for (int d = 0; d < 20; d++) {
        if (d % 4 == 0) {
            const int s = d / 4;
            x0 += ks[s % 5];
            x1 += ks[(s + 1) % 5];
            x2 += ks[(s + 2) % 5];
            x3 += ks[(s + 3) % 5] + (uint64_t)s;
        }
        if (d % 2 == 0) {
            x0 += x1; x1 = rotl64(x1, R0[d % 8]) ^ x0;
            x2 += x3; x3 = rotl64(x3, R1[d % 8]) ^ x2;
        } else {
            x0 += x3; x3 = rotl64(x3, R0[d % 8]) ^ x0;
            x2 += x1; x1 = rotl64(x1, R1[d % 8]) ^ x2;
        }
    }
    o[0] = x0 + ks[0];
    o[1] = x1 + ks[1];
    o[2] = x2 + ks[2];
    o[3] = x3 + ks[3] + 5;
}

__device__ __forceinline__ uint32_t draw_at(uint32_t seed, uint64_t pos) {
    uint64_t o[4];
    threefish_block(seed, pos >> 3, o);
    const int w = (int)(pos & 7);
    const uint64_t v = w < 2 ? o[0] : (w < 4 ? o[1] : (w < 6 ? o[2] : o[3]));
    return (w & 1) ? (uint32_t)(v >> 32) : (uint32_t)v;
}

__global__ __launch_bounds__(TPB) void k_prng_fill(uint32_t seed, uint64_t pos0, int64_t n,
                                                   uint32_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    out[i] = draw_at(seed, pos0 + (uint64_t)i);
}

struct RealArgs {
    double2 *ft;          // kernel in (unless ignorekernel), realisation out
    uint8_t *redo;        // N*N row marks (ky + N kz)
    const float *sinc;    // N entries
    const double *sx, *sy, *sc;  // spline knots, ordinates, half second derivatives
    int nsp;
    int N, P;
    int64_t nkv;
    uint32_t seed;
    int ksphere, ignorekernel;
    double coeff;
};

// natural cubic spline, the bisection and cubic of the host setup
__device__ __forceinline__ double spline_eval(const RealArgs &A, double xq) {
    if (!(xq >= A.sx[0] && xq <= A.sx[A.nsp - 1])) return __builtin_nan("");
    int lo = 0, hi = A.nsp - 1;
    while (hi > lo + 1) {
        const int mid = (lo + hi) / 2;
        if (A.sx[mid] > xq) hi = mid; else lo = mid;
    }
    const double h = A.sx[lo + 1] - A.sx[lo];
    const double dy = A.sy[lo + 1] - A.sy[lo];
    const double t = xq - A.sx[lo];
    const double b = dy / h - h * (A.sc[lo + 1] + 2.0 * A.sc[lo]) / 3.0;
    const double d = (A.sc[lo + 1] - A.sc[lo]) / (3.0 * h);
    return A.sy[lo] + t * (b + t * (A.sc[lo] + t * d));
}

// one mode's value from its draws (ic_basic.hpp:1149-1166).  log(r1) is the
// correctly rounded single-precision logarithm.
__device__ __forceinline__ double2 mode_value(const RealArgs &A, double2 ker, uint32_t u1, uint32_t u2,
                                              float k2, float s, bool conj) {
    const float r1 = (float)u1 / 4294967296.0f;
    const float r2 = (float)u2 / 4294967296.0f;
    k2 = (float)((double)k2 * (4. * M_PI * M_PI));
    double re = cos(2. * M_PI * r2), im = sin(2. * M_PI * r2);
    if (conj) im = -im;
    if (!A.ignorekernel) {
        const double f = 1. + 7.5 * A.coeff / (double)k2;
        re *= f;
        im *= f;
        const double den = ker.x * ker.x + ker.y * ker.y;
        const double nr = (re * ker.x + im * ker.y) / den, ni = (im * ker.x - re * ker.y) / den;
        re = nr;
        im = ni;
    }
    const float lg = (float)log((double)r1);
    const double amp = sqrt(-2. * (double)lg);
    const double pk = spline_eval(A, (double)(float)sqrt((double)k2));
    re *= amp; im *= amp;
    re *= pk;  im *= pk;
    re *= (double)s; im *= (double)s;
    return make_double2(re, im);
}

struct Row {
    uint64_t base;   // stream position of the row's first drawn mode
    int kx0;         // first kx of the row's stream
    uint64_t base0;  // kz > N/2: stream position of the kx = 0 mode
    bool has0;
    int yq, zq;      // |ky|, |kz| as lattice frequencies
};

// stream layout of generateDisplacementField on one process
__device__ __forceinline__ Row row_of(int ky, int kz, int N) {
    const int H = N / 2;
    Row r;
    const bool yhi = ky > H, zhi = kz > H;
    r.yq = yhi ? N - ky : ky;
    r.zq = zhi ? N - kz : kz;
    r.has0 = zhi;
    r.base0 = 0;
    const uint64_t S = HUGE_SKIP;
    if (!yhi && !zhi) {
        r.base = ((uint64_t)kz * S + (uint64_t)ky) * S;
        r.kx0 = (ky == 0 && kz == 0) ? 1 : 0;
    } else if (yhi && !zhi) {
        r.base = ((S + (uint64_t)kz) * S + (uint64_t)(N - ky)) * S;
        r.kx0 = 0;
    } else if (!yhi) {
        r.base = ((2 * S + (uint64_t)(N - kz)) * S + (uint64_t)ky) * S;
        r.kx0 = 1;
        r.base0 = ((uint64_t)(N - kz) * S + (uint64_t)ky) * S;
    } else {
        r.base = ((3 * S + (uint64_t)(N - kz)) * S + (uint64_t)(N - ky)) * S;
        r.kx0 = 1;
        r.base0 = ((2 * S + (uint64_t)(N - kz)) * S + (uint64_t)(N - ky)) * S;
    }
    return r;
}

__device__ __forceinline__ bool drawn(int kx, int yq, int zq, int kmax, int ksphere, float *k2) {
    *k2 = (float)(kx * kx) + (float)(yq * yq) + (float)(zq * zq);
    if (kx >= kmax || yq >= kmax || zq >= kmax) return false;
    return !(ksphere > 0 && *k2 >= (float)(kmax * kmax));
}

__global__ __launch_bounds__(TPB) void k_realize(RealArgs A) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= A.nkv) return;
    const int N = A.N, H = N / 2, kmax = H - 1;
    const int kx = (int)(i % A.P);
    if (kx > H) return;
    const int64_t r = i / A.P;
    const int ky = (int)(r % N), kz = (int)(r / N);
    const Row row = row_of(ky, kz, N);
    float k2;
    if ((kx == 0 && ky == 0 && kz == 0) || !drawn(kx, row.yq, row.zq, kmax, A.ksphere, &k2)) {
        A.ft[i] = make_double2(0., 0.);
        return;
    }
    const bool special = kx == 0 && row.has0;
    const uint64_t pos = special ? row.base0 : row.base + 2 * (uint64_t)(kx - row.kx0);
    // pos is even: r1 and r2 are the two halves of one 64-bit block word
    uint64_t o[4];
    threefish_block(A.seed, pos >> 3, o);
    const int w = (int)(pos & 7) >> 1;
    const uint64_t v = w == 0 ? o[0] : (w == 1 ? o[1] : (w == 2 ? o[2] : o[3]));
    const uint32_t u1 = (uint32_t)v, u2 = (uint32_t)(v >> 32);
    if (u1 == 0) A.redo[(int64_t)kz * N + ky] = 1;  // a redraw shifts this row: recomputed below
    const float s = A.sinc[kx] * A.sinc[row.yq] * A.sinc[row.zq];
    A.ft[i] = mode_value(A, A.ft[i], u1, u2, k2, s, special);
}

// sequential walk of one marked row, as the reference loop does it
__global__ __launch_bounds__(TPB) void k_realize_redo(RealArgs A, const double2 *__restrict__ ker) {
    const int64_t rix = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const int N = A.N, H = N / 2, kmax = H - 1;
    if (rix >= (int64_t)N * N || !A.redo[rix]) return;
    const int ky = (int)(rix % N), kz = (int)(rix / N);
    const Row row = row_of(ky, kz, N);
    const int64_t rowoff = ((int64_t)kz * N + ky) * A.P;
    uint64_t p = row.base;
    for (int kx = row.kx0; kx <= H; kx++) {
        float k2;
        if ((kx == 0 && ky == 0 && kz == 0) || !drawn(kx, row.yq, row.zq, kmax, A.ksphere, &k2)) break;
        uint32_t u1;
        do u1 = draw_at(A.seed, p++); while (u1 == 0);
        const uint32_t u2 = draw_at(A.seed, p++);
        const float s = A.sinc[kx] * A.sinc[row.yq] * A.sinc[row.zq];
        A.ft[rowoff + kx] = mode_value(A, ker[rowoff + kx], u1, u2, k2, s, false);
    }
    float k2;
    if (row.has0 && drawn(0, row.yq, row.zq, kmax, A.ksphere, &k2)) {
        uint64_t q = row.base0;
        uint32_t u1;
        do u1 = draw_at(A.seed, q++); while (u1 == 0);
        const uint32_t u2 = draw_at(A.seed, q++);
        const float s = A.sinc[0] * A.sinc[row.yq] * A.sinc[row.zq];
        A.ft[rowoff] = mode_value(A, ker[rowoff], u1, u2, k2, s, true);
    }
}

// initializeParticlePositions: particle ID = i + numpart (xt + numtile (yt + numtile zt)),
// stored at index ID; q = 0
__global__ __launch_bounds__(TPB) void k_init_particles(const float *__restrict__ tpl, int64_t numpart,
                                                        int numtile, int64_t n, int64_t cap,
                                                        double *__restrict__ pos, double *__restrict__ q,
                                                        int64_t *__restrict__ id) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    const int64_t i = p % numpart, t = p / numpart;
    const int64_t xt = t % numtile, yt = (t / numtile) % numtile, zt = t / ((int64_t)numtile * numtile);
    pos[p] = ((double)xt + (double)tpl[3 * i + 0]) / (double)numtile;
    pos[cap + p] = ((double)yt + (double)tpl[3 * i + 1]) / (double)numtile;
    pos[2 * cap + p] = ((double)zt + (double)tpl[3 * i + 2]) / (double)numtile;
    q[p] = q[cap + p] = q[2 * cap + p] = 0.;
    id[p] = p;
}

__device__ __forceinline__ double fat(const double *f, int N, int x, int y, int z) {
    x = x >= N ? x - N : x;
    y = y >= N ? y - N : y;
    z = z >= N ? z - N : z;
    return f[(int64_t)x + (int64_t)N * ((int64_t)y + (int64_t)N * z)];
}

// CIC gradient / lattice resolution at the particle (ic_basic.hpp:70-95, 128-153)
__device__ __forceinline__ void ic_grad(const double *f, int N, const double p[3], double g[3]) {
    const double dx = 1.0 / (double)N;
    int c[3];
    double rd[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double t = p[i] / dx;
        const double ip = floor(t);
        c[i] = (int)ip;
        rd[i] = t - ip;
    }
    const int x = c[0], y = c[1], z = c[2];
    const double f000 = fat(f, N, x, y, z), f100 = fat(f, N, x + 1, y, z);
    const double f010 = fat(f, N, x, y + 1, z), f001 = fat(f, N, x, y, z + 1);
    const double f110 = fat(f, N, x + 1, y + 1, z), f101 = fat(f, N, x + 1, y, z + 1);
    const double f011 = fat(f, N, x, y + 1, z + 1), f111 = fat(f, N, x + 1, y + 1, z + 1);
    g[0] = (1. - rd[1]) * (1. - rd[2]) * (f100 - f000);
    g[1] = (1. - rd[0]) * (1. - rd[2]) * (f010 - f000);
    g[2] = (1. - rd[0]) * (1. - rd[1]) * (f001 - f000);
    g[0] += rd[1] * (1. - rd[2]) * (f110 - f010);
    g[1] += rd[0] * (1. - rd[2]) * (f110 - f100);
    g[2] += rd[0] * (1. - rd[1]) * (f101 - f100);
    g[0] += (1. - rd[1]) * rd[2] * (f101 - f001);
    g[1] += (1. - rd[0]) * rd[2] * (f011 - f001);
    g[2] += (1. - rd[0]) * rd[1] * (f011 - f010);
    g[0] += rd[1] * rd[2] * (f111 - f011);
    g[1] += rd[0] * rd[2] * (f111 - f101);
    g[2] += rd[0] * rd[1] * (f111 - f110);
    g[0] /= dx;
    g[1] /= dx;
    g[2] /= dx;
}

__device__ __forceinline__ double wrap_pos(double p, double dx, int N) {
    if (p >= 1.0) p -= 1.0;
    if (p < 0.0) p += 1.0;
    if (!(p >= 0.0 && p < 1.0)) p = p - floor(p);
    if (!(p >= 0.0) || (int)floor(p / dx) >= N) p = 0.0;
    return p;
}

// block maximum of a non-negative double, folded into *out by its bit pattern
__device__ __forceinline__ void block_max(double v, unsigned long long *out) {
    __shared__ double part[TPB / 64];
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = part[0];
        for (int w = 1; w < TPB / 64; w++) m = fmax(m, part[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// moveParticles(displace_pcls_ic_basic, coeff, &field, ...) with the MAX output
__global__ __launch_bounds__(TPB) void k_ic_displace(double *__restrict__ pos, int64_t cap, int64_t n,
                                                     int N, const double *__restrict__ f, double coeff,
                                                     unsigned long long *__restrict__ vmax) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    double out = 0.;
    if (p < n) {
        double x[3] = {pos[p], pos[cap + p], pos[2 * cap + p]}, g[3];
        ic_grad(f, N, x, g);
        out = coeff * sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const double dx = 1.0 / (double)N;
#pragma unroll
        for (int i = 0; i < 3; i++) pos[i * cap + p] = wrap_pos(x[i] + coeff * g[i], dx, N);
    }
    block_max(out, vmax);
}

// updateVel(initialize_q_ic_basic, coeff, &field, 1): q = -coeff grad, max |q|^2
__global__ __launch_bounds__(TPB) void k_ic_q(const double *__restrict__ pos, double *__restrict__ q,
                                              int64_t cap, int64_t n, int N, const double *__restrict__ f,
                                              double coeff, unsigned long long *__restrict__ vmax) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    double v2 = 0.;
    if (p < n) {
        double x[3] = {pos[p], pos[cap + p], pos[2 * cap + p]}, g[3];
        ic_grad(f, N, x, g);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double v = -g[i] * coeff;
            q[i * cap + p] = v;
            v2 += v * v;
        }
    }
    block_max(v2, vmax);
}

struct Stencil27 {
    double v[27];
};

// writes the 3x3x3 neighbourhood of the origin (periodic) of a real field
__global__ void k_put_stencil(double *__restrict__ f, int N, Stencil27 s) {
    const int t = threadIdx.x;
    if (t >= 27) return;
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const int x = (dx + N) % N, y = (dy + N) % N, z = (dz + N) % N;
    f[(int64_t)x + (int64_t)N * ((int64_t)y + (int64_t)N * z)] = s.v[t];
}

}  // namespace

namespace gev {

int launch_put_stencil(gev_ctx *c, double *field, const double v[27]) {
    Stencil27 s;
    for (int i = 0; i < 27; i++) s.v[i] = v[i];
    hipLaunchKernelGGL(k_put_stencil, dim3(1), dim3(64), 0, c->stream, field, c->N, s);
    return check_hip(c, hipGetLastError(), "stencil launch");
}

int launch_prng_fill(gev_ctx *c, uint32_t seed, uint64_t pos0, int64_t n, uint32_t *dev_out) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_prng_fill, dim3(grid_for(n, TPB)), dim3(TPB), 0, c->stream, seed, pos0, n, dev_out);
    return check_hip(c, hipGetLastError(), "prng launch");
}

int launch_realize(gev_ctx *c, const RealizeSetup &s) {
    const int N = c->N;
    ScopedTimer t(c, "ic_realize");
    RealArgs A;
    A.ft = c->scalarFT;
    A.redo = s.redo;
    A.sinc = s.sinc;
    A.sx = s.sx;
    A.sy = s.sy;
    A.sc = s.sc;
    A.nsp = s.nsp;
    A.N = N;
    A.P = c->P;
    A.nkv = nk(c);
    A.seed = s.seed;
    A.ksphere = s.ksphere;
    A.ignorekernel = s.ignorekernel;
    A.coeff = s.coeff;
    int rc = launch_fill_zero(c, s.redo, (size_t)N * N);
    if (rc) return rc;
    // the redo pass needs the kernel values the first pass overwrites
    if (!s.ignorekernel) {
        rc = check_hip(c, hipMemcpyAsync(c->fft_tmp, c->scalarFT, sizeof(double2) * nk(c),
                                         hipMemcpyDeviceToDevice, c->stream), "kernel copy");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_realize, dim3(grid_for(nk(c), TPB)), dim3(TPB), 0, c->stream, A);
    rc = check_hip(c, hipGetLastError(), "realize launch");
    if (rc) return rc;
    hipLaunchKernelGGL(k_realize_redo, dim3(grid_for((int64_t)N * N, TPB)), dim3(TPB), 0, c->stream, A,
                       (const double2 *)c->fft_tmp);
    return check_hip(c, hipGetLastError(), "realize redo launch");
}

int launch_init_particles(gev_ctx *c, const float *dev_tpl, int64_t numpart, int numtile) {
    if (c->n == 0) return 0;
    hipLaunchKernelGGL(k_init_particles, dim3(grid_for(c->n, TPB)), dim3(TPB), 0, c->stream, dev_tpl, numpart,
                       numtile, c->n, c->cap, c->pos, c->q, c->id);
    return check_hip(c, hipGetLastError(), "init particles launch");
}

int launch_ic_displace(gev_ctx *c, const double *field, double coeff) {
    int rc = launch_fill_zero(c, c->red_u, 16);  // the fill works in 16-byte units; red_u holds 64
    if (rc || c->n == 0) return rc;
    hipLaunchKernelGGL(k_ic_displace, dim3(grid_for(c->n, TPB)), dim3(TPB), 0, c->stream, c->pos, c->cap, c->n,
                       c->N, field, coeff, (unsigned long long *)c->red_u);
    rc = check_hip(c, hipGetLastError(), "ic displace launch");
    return rc ? rc : launch_refile(c);
}

int launch_ic_q(gev_ctx *c, const double *field, double coeff) {
    int rc = launch_fill_zero(c, c->red_u, 16);  // the fill works in 16-byte units; red_u holds 64
    if (rc || c->n == 0) return rc;
    hipLaunchKernelGGL(k_ic_q, dim3(grid_for(c->n, TPB)), dim3(TPB), 0, c->stream, c->pos, c->q, c->cap, c->n,
                       c->N, field, coeff, (unsigned long long *)c->red_u);
    return check_hip(c, hipGetLastError(), "ic q launch");
}

}  // namespace gev
