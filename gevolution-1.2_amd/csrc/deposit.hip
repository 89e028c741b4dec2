// deposit.hip -- CIC projections of one box (gevolution.hpp:926-1297 and
// LATfield2's scalarProjectionCIC_project) as an owner-computes gather.
//
// A workgroup owns a tile of 16x4x4 target cells; its particles come from the
// 17x5x5 source cells [x0-1, x0+16) x [y0-1, y0+4) x [z0-1, z0+4) (periodic).
// The source cells are enumerated in ascending global site index and their
// particles (each cell's run of the filing permutation, ascending ID) are
// staged into LDS as per-particle records -- CIC weights and the energy /
// stress factors, computed once per particle instead of once per target.
// Then every thread sums, for its target cell, the contributions of its 8
// source cells in ascending site index, exactly the order in which the
// reference's sequential site loop adds its local cubes onto that target
// (gevolution.hpp:1012-1019, 1129-1144, 1277-1294).  The target is written
// once; there are no atomics, and the result is bit-identical to the
// sequential scatter for the same in-cell particle order.
//
// Tiles whose source cells hold more particles than the LDS holds are
// processed in chunks of whole source cells (still in site order); a single
// cell larger than a chunk is read directly from global memory.
//
// Clustered chunks (FP64 T00 + Tij): a target thread walks all particles of
// its 8 source cells, so one dense cell keeps its wavefront busy while the
// other lanes idle.  The dense cells of such a chunk are summed first, one
// lane per (cell, target) pair and 8 cells per wavefront, in the same
// particle order; the partial sums replace the cell's records in LDS and the
// target threads add them in site order as before, so the result does not
// change by a bit.
#include "gev_internal.h"

namespace {

// Tile shape (compile-time; GEV_DEP_TX/TY/TZ/CAP override the defaults).
#ifndef GEV_DEP_TX
#define GEV_DEP_TX 16
#endif
#ifndef GEV_DEP_TY
#define GEV_DEP_TY 4
#endif
#ifndef GEV_DEP_TZ
#define GEV_DEP_TZ 4
#endif
#ifndef GEV_DEP_CAP
#define GEV_DEP_CAP 448
#endif
#ifndef GEV_DEP_XCD
#define GEV_DEP_XCD 1
#endif
#ifndef GEV_DEP_ROWS
#define GEV_DEP_ROWS 1
#endif
constexpr int TX = GEV_DEP_TX, TY = GEV_DEP_TY, TZ = GEV_DEP_TZ;
constexpr int TPB = TX * TY * TZ;   // one target cell per thread
constexpr int SX = TX + 1, SY = TY + 1, SZ = TZ + 1;
constexpr int NSRC = SX * SY * SZ;  // source cells of a tile (425 for 16x4x4)
// particle records per chunk: the FP32 form's records are half the size, so
// it stages more of them in the same LDS (clustered tiles take fewer chunks)
#ifndef GEV_DEP_CAP_F32
#define GEV_DEP_CAP_F32 448
#endif
template <typename R> constexpr int dep_cap() { return sizeof(R) == 4 ? GEV_DEP_CAP_F32 : GEV_DEP_CAP; }
// The FP32 form's tiles above this many particles in their source region
// take the fixed-point particle-parallel path, and so do its tiles above CAP
// particles whose region has a cell of at least GEV_DEP_FIXED_CELL (0: none);
// the other tiles above CAP take the chunked gather.  In the fixed-point
// path, a region with a cell of at least GEV_FX_SPREAD particles (0: never)
// spreads each wavefront's lanes over the tile's records (the pass 2 loop).
// At 512^3 these take the FP32 deposit from 8.09 to 7.64 ms at z = 3 and
// from 13.02 to 11.58 ms at z = 0 (448 / 0 / 0: the fixed-point path for
// every tile above CAP, records in order).
#ifndef GEV_DEP_FIXED_MIN
#define GEV_DEP_FIXED_MIN 896
#endif
#ifndef GEV_DEP_FIXED_CELL
#define GEV_DEP_FIXED_CELL 16
#endif
#ifndef GEV_FX_SPREAD
#define GEV_FX_SPREAD 16
#endif
// Source cells with at least this many particles in a chunk are "dense"
// (T00 + Tij deposit): their partial sums for all their in-tile targets are
// computed first by lanes that each take one (cell, target) pair, and the
// target threads read them instead of walking the cell's particles (0: off).
#ifndef GEV_DEP_DENSE
#define GEV_DEP_DENSE 4
#endif
// (the FP32 form: off; its dense tiles take the fixed-point path, and
// without the dense phase its gather keeps 4 wavefronts per SIMD)
#ifndef GEV_DEP_DENSE_F32
#define GEV_DEP_DENSE_F32 0
#endif
template <typename R> constexpr int dense_min() { return sizeof(R) == 4 ? GEV_DEP_DENSE_F32 : GEV_DEP_DENSE; }
// ... and the dense phase runs only in chunks with a cell of at least this
// many particles (below that the target threads' walks are balanced enough)
#ifndef GEV_DEP_DENSE_TRIG
#define GEV_DEP_DENSE_TRIG 8
#endif
// the count of a cell whose partial sums the dense phase has formed carries
// this bit
constexpr uint32_t DENSE_MARK = 0x80000000u;
// A dense cell's partial sums are stored over its own records (record
// arrays 0..11, rows r0 .. r0+n-1): lc, ld0, ld1, ld2 of target offset d in
// arrays 0..3 (d < 4) or 4..7 (d >= 4), row r0 + (d & 3); lo01 of the
// offsets 0 and 4, lo02 of 0 and 2, lo12 of 0 and 1 in arrays 8, 9, 10, row
// r0 + (which of the two).  Four records hold them all.
static_assert(GEV_DEP_DENSE == 0 || GEV_DEP_DENSE >= 4, "a dense cell's records hold its partial sums");
static_assert(GEV_DEP_DENSE_F32 == 0 || GEV_DEP_DENSE_F32 >= 4, "a dense cell's records hold its partial sums");
static_assert(NSRC <= 2 * TPB, "two source cells per thread");
static_assert(TPB % 64 == 0 && TPB <= 1024, "whole wavefronts");

// cell index; N^3 <= 2^30 fits 32 bits
__device__ __forceinline__ uint32_t L(int N, int x, int y, int z) {
    return (uint32_t)x + (uint32_t)N * ((uint32_t)y + (uint32_t)N * (uint32_t)z);
}

// x / dx with dx = 1/N.  For N a power of two dx is exact and the quotient
// equals x * N bit for bit, so the multiply replaces the division.
template <bool POW2> __device__ __forceinline__ double over_dx(double x, double dx, double Nd) {
    return POW2 ? x * Nd : x / dx;
}
__device__ __forceinline__ int wrapi(int i, int N) { return i < 0 ? i + N : (i >= N ? i - N : i); }

// Source slot s along one axis is coordinate lo0 - 1 + s (wrapped).  In
// ascending global order the slots come as 0,1,..,n-1 unless lo0 == 0, where
// slot 0 (coordinate N-1) comes last.
__device__ __forceinline__ int order_pos(int s, int lo0, int n) { return lo0 > 0 ? s : (s == 0 ? n - 1 : s - 1); }
__device__ __forceinline__ int order_slot(int p, int lo0, int n) { return lo0 > 0 ? p : (p == n - 1 ? 0 : p + 1); }

template <int KIND> struct Kind {
    static constexpr bool T00 = KIND == 1 || KIND == 3 || KIND == 5;
    static constexpr bool TIJ = KIND == 2 || KIND == 3;
    static constexpr bool T0I = KIND == 4;
    static constexpr bool QE = (T00 && KIND != 5) || TIJ;  // needs e = sqrt(q^2 + a^2)
};

// per-particle record: everything a target needs from one particle.  R is
// the arithmetic type of the records, the partial sums and the accumulators:
// double in the FP64 form, float in the FP32 form (whose fields are float).
// A record is always evaluated in double from the FP64 particle and then
// rounded to R.
template <typename R> struct Rec {
    R up0, up1, up2;      // (pos - source corner) / dx
    R e, f00;             // T00: e and f (gevolution.hpp:989-991)
    R fij;                // Tij: f (gevolution.hpp:1238)
    R w0, w1, w2;         // Tij: m q_i q_i / e   |  T0i: m q_i
    R w01, w02, w12;      // Tij: m q_i q_j / e
};

// position and momentum of one particle
struct PQ {
    double x, y, z, q0, q1, q2;
};

template <int KIND>
__device__ __forceinline__ PQ load_pq(const double *__restrict__ pos, const double *__restrict__ q, int64_t cap,
                                      uint32_t p) {
    using K = Kind<KIND>;
    PQ v;
    v.x = pos[p];
    v.y = pos[cap + p];
    v.z = pos[2 * cap + p];
    v.q0 = v.q1 = v.q2 = 0.;
    if (K::QE || K::T0I) {
        v.q0 = q[p];
        v.q1 = q[cap + p];
        v.q2 = q[2 * cap + p];
    }
    return v;
}

template <int KIND, bool USE_PHI, bool POW2>
__device__ __forceinline__ Rec<double> make_rec(const PQ &v, double rx, double ry, double rz, double dx, double Nd,
                                                double a, double mass_t) {
    using K = Kind<KIND>;
    Rec<double> r;
    r.up0 = over_dx<POW2>(v.x - rx, dx, Nd);
    r.up1 = over_dx<POW2>(v.y - ry, dx, Nd);
    r.up2 = over_dx<POW2>(v.z - rz, dx, Nd);
    r.e = a;
    r.f00 = 0.;
    r.fij = 0.;
    r.w0 = r.w1 = r.w2 = r.w01 = r.w02 = r.w12 = 0.;
    const double q0 = v.q0, q1 = v.q1, q2 = v.q2;
    if (K::QE) {
        const double fq = q0 * q0 + q1 * q1 + q2 * q2;
        const double e = sqrt(fq + a * a);
        if (K::T00 && KIND != 5 && USE_PHI) {
            r.e = e;
            r.f00 = 3. * e + fq / e;
        }
        if (K::TIJ) {
            r.fij = 4. + a * a / (fq + a * a);
            r.w0 = mass_t * q0 * q0 / e;
            r.w1 = mass_t * q1 * q1 / e;
            r.w2 = mass_t * q2 * q2 / e;
            r.w01 = mass_t * q0 * q1 / e;
            r.w02 = mass_t * q0 * q2 / e;
            r.w12 = mass_t * q1 * q2 / e;
        }
    }
    if (K::T0I) {
        r.w0 = mass_t * q0;
        r.w1 = mass_t * q1;
        r.w2 = mass_t * q2;
    }
    return r;
}

// The FP32 form's record: the CIC weights from the FP64 position (rounded
// once), the momentum-dependent factors evaluated in float.
template <int KIND, bool USE_PHI, bool POW2>
__device__ __forceinline__ Rec<float> make_rec_f32(const PQ &v, double rx, double ry, double rz, double dx, double Nd,
                                                   float a, float mass_t) {
    using K = Kind<KIND>;
    Rec<float> r;
    r.up0 = (float)over_dx<POW2>(v.x - rx, dx, Nd);
    r.up1 = (float)over_dx<POW2>(v.y - ry, dx, Nd);
    r.up2 = (float)over_dx<POW2>(v.z - rz, dx, Nd);
    r.e = a;
    r.f00 = 0.f;
    r.fij = 0.f;
    r.w0 = r.w1 = r.w2 = r.w01 = r.w02 = r.w12 = 0.f;
    const float q0 = (float)v.q0, q1 = (float)v.q1, q2 = (float)v.q2;
    if (K::QE) {
        const float fq = q0 * q0 + q1 * q1 + q2 * q2;
        const float e = sqrtf(fq + a * a), me = mass_t / e;
        if (K::T00 && KIND != 5 && USE_PHI) {
            r.e = e;
            r.f00 = 3.f * e + fq / e;
        }
        if (K::TIJ) {
            r.fij = 4.f + a * a / (fq + a * a);
            r.w0 = me * q0 * q0;
            r.w1 = me * q1 * q1;
            r.w2 = me * q2 * q2;
            r.w01 = me * q0 * q1;
            r.w02 = me * q0 * q2;
            r.w12 = me * q1 * q2;
        }
    }
    if (K::T0I) {
        r.w0 = mass_t * q0;
        r.w1 = mass_t * q1;
        r.w2 = mass_t * q2;
    }
    return r;
}

// the record of the form whose arithmetic type is R
template <typename R, int KIND, bool USE_PHI, bool POW2>
__device__ __forceinline__ Rec<R> record(const PQ &v, double rx, double ry, double rz, double dx, double Nd, double a,
                                         double mass_t) {
    if constexpr (sizeof(R) == 4) return make_rec_f32<KIND, USE_PHI, POW2>(v, rx, ry, rz, dx, Nd, (float)a, (float)mass_t);
    else return make_rec<KIND, USE_PHI, POW2>(v, rx, ry, rz, dx, Nd, a, mass_t);
}

// LDS image of the records (structure of arrays)
template <int KIND, typename R> struct Stage {
    R up0[dep_cap<R>()], up1[dep_cap<R>()], up2[dep_cap<R>()];
    R e[Kind<KIND>::T00 ? dep_cap<R>() : 1], f00[Kind<KIND>::T00 ? dep_cap<R>() : 1];
    R fij[Kind<KIND>::TIJ ? dep_cap<R>() : 1];
    R w0[(Kind<KIND>::TIJ || Kind<KIND>::T0I) ? dep_cap<R>() : 1];
    R w1[(Kind<KIND>::TIJ || Kind<KIND>::T0I) ? dep_cap<R>() : 1];
    R w2[(Kind<KIND>::TIJ || Kind<KIND>::T0I) ? dep_cap<R>() : 1];
    R w01[Kind<KIND>::TIJ ? dep_cap<R>() : 1], w02[Kind<KIND>::TIJ ? dep_cap<R>() : 1], w12[Kind<KIND>::TIJ ? dep_cap<R>() : 1];

    // stores a record (evaluated in double or in R), rounded to R
    template <typename Q> __device__ __forceinline__ void put(int i, const Rec<Q> &r) {
        using K = Kind<KIND>;
        up0[i] = (R)r.up0;
        up1[i] = (R)r.up1;
        up2[i] = (R)r.up2;
        if (K::T00) {
            e[i] = (R)r.e;
            f00[i] = (R)r.f00;
        }
        if (K::TIJ) fij[i] = (R)r.fij;
        if (K::TIJ || K::T0I) {
            w0[i] = (R)r.w0;
            w1[i] = (R)r.w1;
            w2[i] = (R)r.w2;
        }
        if (K::TIJ) {
            w01[i] = (R)r.w01;
            w02[i] = (R)r.w02;
            w12[i] = (R)r.w12;
        }
    }
    // row r of record array k (the T00 + Tij stage: twelve record arrays of
    // dep_cap<R>() values each, back to back)
    __device__ __forceinline__ R &slot(int k, int r) {
        static_assert(KIND == 3, "the T00 + Tij stage");
        return (&up0[0])[k * dep_cap<R>() + r];
    }
    __device__ __forceinline__ R slot(int k, int r) const {
        static_assert(KIND == 3, "the T00 + Tij stage");
        return (&up0[0])[k * dep_cap<R>() + r];
    }
    __device__ __forceinline__ Rec<R> get(int i) const {
        using K = Kind<KIND>;
        Rec<R> r;
        r.up0 = up0[i];
        r.up1 = up1[i];
        r.up2 = up2[i];
        r.e = K::T00 ? e[i] : (R)0;
        r.f00 = K::T00 ? f00[i] : (R)0;
        r.fij = K::TIJ ? fij[i] : (R)0;
        r.w0 = (K::TIJ || K::T0I) ? w0[i] : (R)0;
        r.w1 = (K::TIJ || K::T0I) ? w1[i] : (R)0;
        r.w2 = (K::TIJ || K::T0I) ? w2[i] : (R)0;
        r.w01 = K::TIJ ? w01[i] : (R)0;
        r.w02 = K::TIJ ? w02[i] : (R)0;
        r.w12 = K::TIJ ? w12[i] : (R)0;
        return r;
    }
};

// target-side values and the running per-source partial sums
template <typename R> struct Target {
    R ph, plq01, plq02, plq12, lnk0, lnk1, lnk2;
};

template <typename R> struct Partial {
    R lc, ld0, ld1, ld2, lo01, lo02, lo12, lv0, lv1, lv2;
    __device__ __forceinline__ void clear() {
        lc = ld0 = ld1 = ld2 = lo01 = lo02 = lo12 = lv0 = lv1 = lv2 = (R)0;
    }
};

// one particle's contribution to one target (expression order of
// gevolution.hpp:995-1009, 1241-1272, 1107-1126)
template <int KIND, typename R>
__device__ __forceinline__ void add_particle(Partial<R> &P, const Rec<R> &r, const Target<R> &T, int dxx, int dy,
                                             int dz) {
    using K = Kind<KIND>;
    const R one = (R)1, quarter = (R)0.25;
    const R W0 = dxx ? r.up0 : one - r.up0;
    const R W1 = dy ? r.up1 : one - r.up1;
    const R W2 = dz ? r.up2 : one - r.up2;
    if (K::T00) P.lc += W0 * W1 * W2 * (r.e + r.f00 * T.ph);
    if (K::TIJ) {
        P.ld0 += r.w0 * W0 * W1 * W2 * (one + r.fij * T.ph);
        P.ld1 += r.w1 * W0 * W1 * W2 * (one + r.fij * T.ph);
        P.ld2 += r.w2 * W0 * W1 * W2 * (one + r.fij * T.ph);
        if (dxx == 0 && dy == 0) P.lo01 += r.w01 * W2 * (one + r.fij * quarter * T.plq01);
        if (dxx == 0 && dz == 0) P.lo02 += r.w02 * W1 * (one + r.fij * quarter * T.plq02);
        if (dy == 0 && dz == 0) P.lo12 += r.w12 * W0 * (one + r.fij * quarter * T.plq12);
    }
    if (K::T0I) {
        if (dxx == 0) P.lv0 += r.w0 * W1 * W2;
        if (dy == 0) P.lv1 += r.w1 * W0 * W2;
        if (dz == 0) P.lv2 += r.w2 * W0 * W1;
    }
}

template <typename R> struct Acc {
    R s00, d0, d1, d2, o01, o02, o12, v0, v1, v2;
};

template <int KIND, typename R>
__device__ __forceinline__ void add_source(Acc<R> &A, const Partial<R> &P, const Target<R> &T, int dxx, int dy,
                                           int dz, R mass_s) {
    using K = Kind<KIND>;
    if (K::T00) A.s00 += P.lc * mass_s;
    if (K::TIJ) {
        A.d0 += P.ld0;
        A.d1 += P.ld1;
        A.d2 += P.ld2;
        if (dxx == 0 && dy == 0) A.o01 += P.lo01;
        if (dxx == 0 && dz == 0) A.o02 += P.lo02;
        if (dy == 0 && dz == 0) A.o12 += P.lo12;
    }
    if (K::T0I) {
        if (dxx == 0) A.v0 += P.lv0 * T.lnk0;
        if (dy == 0) A.v1 += P.lv1 * T.lnk1;
        if (dz == 0) A.v2 += P.lv2 * T.lnk2;
    }
}

// does source offset d feed any component of this KIND at all?
template <int KIND> __device__ __forceinline__ bool feeds(int dxx, int dy, int dz) {
    if (KIND == 4) return dxx == 0 || dy == 0 || dz == 0;
    return true;
}

// One source cell's contribution to this thread's target with the offset
// (DX, DY, DZ) fixed at compile time (same operations, same order as the
// reference's local cube; the conditional terms of add_particle / add_source
// fold away).  c: the source cell's index in the tile's source region.
// RANGE = false: the caller knows c is in the chunk [cb, ce) (one chunk).
template <int KIND, int DX, int DY, int DZ, bool RANGE = true, typename ST, typename R>
__device__ __forceinline__ void add_source_cell(Acc<R> &A, const ST &st, const Target<R> &T, int c, int cb, int ce,
                                                uint32_t base, const uint32_t *s_off, const uint32_t *s_cnt,
                                                R mass_s) {
    if (!feeds<KIND>(DX, DY, DZ) || (RANGE && (c < cb || c >= ce))) return;
    const uint32_t n = s_cnt[c];
    if (n == 0) return;
    Partial<R> P;
    P.clear();
    const uint32_t r0 = s_off[c] - base;
    if (KIND == 3 && dense_min<R>() > 0 && (n & DENSE_MARK)) {
        // a dense cell: its partial sums for this offset, already summed
        if constexpr (KIND == 3) {
            constexpr int d = DX | (DY << 1) | (DZ << 2), a0 = 4 * (d >> 2), row = d & 3;
            P.lc = st.slot(a0, (int)r0 + row);
            P.ld0 = st.slot(a0 + 1, (int)r0 + row);
            P.ld1 = st.slot(a0 + 2, (int)r0 + row);
            P.ld2 = st.slot(a0 + 3, (int)r0 + row);
            if (DX == 0 && DY == 0) P.lo01 = st.slot(8, (int)r0 + DZ);
            if (DX == 0 && DZ == 0) P.lo02 = st.slot(9, (int)r0 + DY);
            if (DY == 0 && DZ == 0) P.lo12 = st.slot(10, (int)r0 + DX);
        }
    } else {
#pragma unroll 1
        for (uint32_t i = 0; i < n; i++) add_particle<KIND>(P, st.get((int)(r0 + i)), T, DX, DY, DZ);
    }
    add_source<KIND>(A, P, T, DX, DY, DZ, mass_s);
}

// the target's sums, written once
// (the target sums of a large box are streamed past the caches: 6.86
// against 6.98 ms at 512^3 early in the run)
template <int KIND, typename F>
__device__ __forceinline__ void store_target(const Acc<F> &A, F *__restrict__ out_s, F *__restrict__ out_t,
                                             uint32_t t, int64_t n3v) {
    using K = Kind<KIND>;
    const bool nt = n3v >= (int64_t)gev::STREAM_MIN_N * gev::STREAM_MIN_N * gev::STREAM_MIN_N;
    auto put_out = [nt](F *p, F v) { gev::store_stream(p, v, nt); };
    if (K::T00) put_out(out_s + t, (F)A.s00);
    if (K::TIJ) {
        put_out(out_t + t, (F)A.d0);
        put_out(out_t + 3 * n3v + t, (F)A.d1);
        put_out(out_t + 5 * n3v + t, (F)A.d2);
        put_out(out_t + n3v + t, (F)A.o01);
        put_out(out_t + 2 * n3v + t, (F)A.o02);
        put_out(out_t + 4 * n3v + t, (F)A.o12);
    }
    if (K::T0I) {
        put_out(out_t + t, (F)A.v0);
        put_out(out_t + n3v + t, (F)A.v1);
        put_out(out_t + 2 * n3v + t, (F)A.v2);
    }
}

// Runs CALL(dx, dy, dz) with the offset of a source from its target,
// d = dx | dy << 1 | dz << 2, as compile-time values.  (A switch over
// force-inlined template calls: every accumulator stays in registers.)
#define GEV_BY_OFFSET(d, CALL) \
    switch (d) {               \
    case 0: CALL(0, 0, 0); break; \
    case 1: CALL(1, 0, 0); break; \
    case 2: CALL(0, 1, 0); break; \
    case 3: CALL(1, 1, 0); break; \
    case 4: CALL(0, 0, 1); break; \
    case 5: CALL(1, 0, 1); break; \
    case 6: CALL(0, 1, 1); break; \
    default: CALL(1, 1, 1); break; \
    }

// the target-side values of cell (x, y, z): phi there, its plaquette sums
// (gevolution.hpp:1263-1272) and link factors (1129-1144)
template <bool USE_PHI, typename F>
__device__ __forceinline__ Target<F> load_target(const F *__restrict__ phi, int N, int x, int y, int z) {
    Target<F> T;
    T.ph = (F)0;
    F phx = 0, phy = 0, phz = 0, phxy = 0, phxz = 0, phyz = 0;
    if (USE_PHI) {
        const int xp = x == N - 1 ? 0 : x + 1, yp = y == N - 1 ? 0 : y + 1, zp = z == N - 1 ? 0 : z + 1;
        T.ph = phi[L(N, x, y, z)];
        phx = phi[L(N, xp, y, z)];
        phy = phi[L(N, x, yp, z)];
        phz = phi[L(N, x, y, zp)];
        phxy = phi[L(N, xp, yp, z)];
        phxz = phi[L(N, xp, y, zp)];
        phyz = phi[L(N, x, yp, zp)];
    }
    T.plq01 = T.ph + phy + phx + phxy;
    T.plq02 = T.ph + phz + phx + phxz;
    T.plq12 = T.ph + phz + phy + phyz;
    T.lnk0 = (F)1 + T.ph + phx;
    T.lnk1 = (F)1 + T.ph + phy;
    T.lnk2 = (F)1 + T.ph + phz;
    return T;
}

// The dense cells of a chunk (T00 + Tij): lane (q, d) of a wavefront takes
// cell q of the wavefront's group of 8 and the target at offset d from it,
// and sums that target's partial (add_particle over the cell's records in
// order, exactly the sum the target thread would form), then stores it in
// the cell's record space (the layout above), and marks the
// cell's count.  Each cell's 8 lanes sit in one wavefront, so its records
// and count are read before they are overwritten.  The caller synchronises
// the workgroup before and after.
template <int KIND, bool USE_PHI, typename F>
__device__ __forceinline__ void dense_partials(Stage<KIND, F> &st, const uint16_t *s_dlist, int ndense,
                                               const uint32_t *s_off, uint32_t *s_cnt, uint32_t base,
                                               const F *__restrict__ phi, int N, int x0, int y0, int z0, int nsx,
                                               int nsy, int nsz, int nx, int ny, int nz) {
    if constexpr (KIND == 3) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int q = lane >> 3, d = lane & 7;
        const int dxx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
        for (int g = wave * 8; g < ndense; g += TPB / 8) {
            const int k = g + q;
            const int c = k < ndense ? (int)s_dlist[k] : 0;
            const int px = c % nsx, py = (c / nsx) % nsy, pz = c / (nsx * nsy);
            // source slot s along an axis is coordinate lo0 - 1 + s; the
            // target at offset 1 (0) along that axis is the next (same) one
            const int lx = order_slot(px, x0, nsx) - 1 + dxx, ly = order_slot(py, y0, nsy) - 1 + dy;
            const int lz = order_slot(pz, z0, nsz) - 1 + dz;
            const bool on = k < ndense && lx >= 0 && lx < nx && ly >= 0 && ly < ny && lz >= 0 && lz < nz;
            const uint32_t m = on ? s_cnt[c] : 0u, r0 = on ? s_off[c] - base : 0u;
            Target<F> T;
            T.ph = T.plq01 = T.plq02 = T.plq12 = (F)0;
            T.lnk0 = T.lnk1 = T.lnk2 = (F)1;
            if (on) T = load_target<USE_PHI>(phi, N, x0 + lx, y0 + ly, z0 + lz);
            Partial<F> P;
            P.clear();
#pragma unroll 1
            for (uint32_t i = 0; i < m; i++) add_particle<KIND>(P, st.get((int)(r0 + i)), T, dxx, dy, dz);
            // every lane of this wavefront has read its cell's records
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (on) {
                const int a0 = 4 * (d >> 2), row = (int)r0 + (d & 3);
                st.slot(a0, row) = P.lc;
                st.slot(a0 + 1, row) = P.ld0;
                st.slot(a0 + 2, row) = P.ld1;
                st.slot(a0 + 3, row) = P.ld2;
                if (dxx == 0 && dy == 0) st.slot(8, (int)r0 + dz) = P.lo01;
                if (dxx == 0 && dz == 0) st.slot(9, (int)r0 + dy) = P.lo02;
                if (dy == 0 && dz == 0) st.slot(10, (int)r0 + dxx) = P.lo12;
            }
            // the cell's count tells the target threads to read the slots
            if (k < ndense && d == 0) s_cnt[c] |= DENSE_MARK;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// the particles 0 .. m-1 of the records, in order, onto one partial sum
template <int KIND, int DX, int DY, int DZ, typename ST, typename R>
__device__ __forceinline__ void add_records(Partial<R> &P, const ST &st, const Target<R> &T, uint32_t m) {
#pragma unroll 1
    for (uint32_t i = 0; i < m; i++) add_particle<KIND>(P, st.get((int)i), T, DX, DY, DZ);
}

// FULL: N is a multiple of the tile, so the tile and source-region extents
// are compile-time constants (index arithmetic without integer division).
// BATCH: the launch covers the boxes of a batch (its own instantiation, so
// the one-box kernel is unchanged).
// (FP64, N a multiple of the tile: the register allocator is held to 3
// wavefronts per SIMD, which the kernel's LDS allows anyway; left to itself
// it gives the dense phase 2.)
template <int KIND, bool USE_PHI, bool ACCUM, bool POW2, bool FULL, typename F, bool BATCH = false>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(sizeof(F) == 8 && FULL ? 3 : 1, 8))) void k_deposit_tiled(
    const double *__restrict__ pos, const double *__restrict__ q, int64_t cap,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ perm,
    const F *__restrict__ phi, F *__restrict__ out_s, F *__restrict__ out_t, int N,
    int ntx, int nty, int ntz, double a, double mass_s, double mass_t, const double *__restrict__ pos_b,
    const double *__restrict__ q_b, gev::Batch bt) {
    using K = Kind<KIND>;
    if constexpr (BATCH) {
        // box blockIdx.y of a batch: its arrays, live particle set and coefficients
        const int b = blockIdx.y;
        if (!gev::box_on(bt, b)) return;
        const gev::BoxCycle &bc = bt.cyc[b];
        if (bc.parity) {
            pos = pos_b;
            q = q_b;
        }
        pos = gev::box_ptr(pos, bt.stride, b);
        q = gev::box_ptr(q, bt.stride, b);
        start = gev::box_ptr(start, bt.stride, b);
        perm = gev::box_ptr(perm, bt.stride, b);
        phi = gev::box_ptr(phi, bt.stride, b);
        out_s = gev::box_ptr(out_s, bt.stride, b);
        out_t = gev::box_ptr(out_t, bt.stride, b);
        a = KIND == 5 ? 1.0 : bc.a;
        mass_s = mass_t = KIND == 5 ? bc.dep_m : bc.dep_m_a;
    }
    constexpr int CAP = dep_cap<F>();
    // the records, or (FP32 form, dense tiles) the fixed-point accumulators
    // and the target values
    union StageOrFixed {
        Stage<KIND, F> st;
        struct {
            unsigned long long acc[7 * TPB];
            float T[4 * TPB];
        } fx;
    };
    __shared__ StageOrFixed su;
    Stage<KIND, F> &st = su.st;
    __shared__ uint32_t s_off[NSRC + 1];
    __shared__ uint32_t s_gbeg[NSRC];
    __shared__ uint32_t s_cnt[NSRC];
    __shared__ uint32_t s_scan[TPB / 64];
    __shared__ uint32_t s_rmax;  // FP32 form: the region's largest source cell
    if (sizeof(F) == 4 && threadIdx.x == 0) s_rmax = 0;
    __shared__ uint16_t s_rc[CAP];
    // the chunk's dense source cells (T00 + Tij), in no particular order (not
    // in the batched kernel, whose prologue leaves no registers for the
    // dense phase at 3 wavefronts per SIMD)
    constexpr int DENSE = dense_min<F>();
    constexpr bool DEN = KIND == 3 && DENSE > 0 && !BATCH;
    __shared__ uint16_t s_dlist[DEN ? CAP / (DEN ? DENSE : 1) + 1 : 1];
    __shared__ uint32_t s_nd, s_dmax;
    if (DEN && threadIdx.x == 0) s_nd = s_dmax = 0;

    // XCD-aware tile order: consecutive workgroups of one XCD (dispatched
    // round-robin over the 8 XCDs) take neighbouring tiles and share L2
    const uint32_t nblk = (uint32_t)ntx * nty * ntz;
    uint32_t tile = blockIdx.x;
    if (GEV_DEP_XCD && nblk % 8 == 0) tile = (blockIdx.x % 8) * (nblk / 8) + blockIdx.x / 8;
    const int tx = (int)(tile % ntx), ty = (int)((tile / ntx) % nty), tz = (int)(tile / ((uint32_t)ntx * nty));
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    const int nx = FULL ? TX : min(TX, N - x0), ny = FULL ? TY : min(TY, N - y0);
    const int nz = FULL ? TZ : min(TZ, N - z0);
    const int nsx = nx + 1, nsy = ny + 1, nsz = nz + 1;
    const int nsrc = nsx * nsy * nsz;
    // no wrapped coordinate in this tile or its source region below it: the
    // ascending site order of the 8 sources is the plain one for every target
    const bool interior = x0 > 0 && y0 > 0 && z0 > 0;
    const double dx = 1.0 / (double)N, Nd = (double)N;
    const int64_t n3v = (int64_t)N * N * N;

    // 2. this thread's target (one per thread; the tile has TPB cells)
    const int lx = threadIdx.x % TX, ly = (threadIdx.x / TX) % TY, lz = threadIdx.x / (TX * TY);
    const bool active = lx < nx && ly < ny && lz < nz;
    const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
    Target<F> T;
    T.ph = T.plq01 = T.plq02 = T.plq12 = (F)0;
    T.lnk0 = T.lnk1 = T.lnk2 = (F)1;
    Acc<F> A;
    A.s00 = A.d0 = A.d1 = A.d2 = A.o01 = A.o02 = A.o12 = A.v0 = A.v1 = A.v2 = (F)0;
    const F mass_sf = (F)mass_s;
    const uint32_t t = active ? L(N, x, y, z) : 0u;
    if (active) {
        T = load_target<USE_PHI>(phi, N, x, y, z);
        if (ACCUM) {
            if (K::T00) A.s00 = out_s[t];
            if (K::TIJ) {
                A.d0 = out_t[t];
                A.d1 = out_t[3 * n3v + t];
                A.d2 = out_t[5 * n3v + t];
                A.o01 = out_t[n3v + t];
                A.o02 = out_t[2 * n3v + t];
                A.o12 = out_t[4 * n3v + t];
            }
            if (K::T0I) {
                A.v0 = out_t[t];
                A.v1 = out_t[n3v + t];
                A.v2 = out_t[2 * n3v + t];
            }
        }
    }
    // Fast path: a full interior tile whose source region fits one chunk.
    // Each of the region's SY x SZ rows (SX cells, ascending x at one y and
    // z) is one contiguous run of the filing permutation, so the records
    // follow from the row extents (read once per wavefront) without a
    // scan, and a record's source cell is its particle's filed cell (the
    // same integer as the region slot, so the same corner bits).  One
    // barrier between the staging and the accumulation.
    if (GEV_DEP_ROWS && FULL && interior) {
        constexpr int NROW = SY * SZ;
        static_assert(NROW <= 64, "one lane per row");
        // lane i of every wavefront reads row i's extent (one vector load
        // latency), then the extents are made wave-uniform
        const int lane = threadIdx.x & 63;
        uint32_t rb = 0, re = 0;
        if (lane < NROW) {
            const uint32_t g = L(N, x0 - 1, y0 - 1 + lane % SY, z0 - 1 + lane / SY);
            rb = start[g];
            re = start[g + SX];
        }
        uint32_t roff[NROW + 1], delta[NROW];  // record r of row i: permutation index r + delta[i]
        roff[0] = 0;
#pragma unroll
        for (int i = 0; i < NROW; i++) {
            const uint32_t b = __builtin_amdgcn_readlane(rb, i);
            roff[i + 1] = roff[i] + (__builtin_amdgcn_readlane(re, i) - b);
            delta[i] = b - roff[i];
        }
        const uint32_t total = roff[NROW];
        if (total <= (uint32_t)CAP) {
            constexpr int PR = (CAP + TPB - 1) / TPB;
            uint32_t pidx[PR];
#pragma unroll
            for (int k = 0; k < PR; k++) {
                const uint32_t r = threadIdx.x + k * TPB;
                pidx[k] = 0;
                if (r < total) {
                    uint32_t d = delta[0];
#pragma unroll
                    for (int i = 1; i < NROW; i++) d = r >= roff[i] ? delta[i] : d;
                    pidx[k] = perm[r + d];
                }
            }
            // record offset and count of every source cell: row i's cells by thread i
            if (threadIdx.x < NROW) {
                const int i = threadIdx.x;
                uint32_t d = delta[0];
#pragma unroll
                for (int k = 1; k < NROW; k++) d = i == k ? delta[k] : d;
                const uint32_t g = L(N, x0 - 1, y0 - 1 + i % SY, z0 - 1 + i / SY);
                uint32_t sv[SX + 1];
#pragma unroll
                for (int k = 0; k <= SX; k++) sv[k] = start[g + k];
#pragma unroll
                for (int k = 0; k < SX; k++) {
                    s_off[i * SX + k] = sv[k] - d;
                    s_cnt[i * SX + k] = sv[k + 1] - sv[k];
                    // (s_nd was cleared above by lane 0 of this wavefront)
                    if (DEN && sv[k + 1] - sv[k] >= (uint32_t)DENSE) {
                        s_dlist[atomicAdd(&s_nd, 1u)] = (uint16_t)(i * SX + k);
                        atomicMax(&s_dmax, sv[k + 1] - sv[k]);
                    }
                }
            }
            PQ pq[PR];
#pragma unroll
            for (int k = 0; k < PR; k++)
                if (threadIdx.x + k * TPB < total) pq[k] = load_pq<KIND>(pos, q, cap, pidx[k]);
#pragma unroll
            for (int k = 0; k < PR; k++) {
                const uint32_t r = threadIdx.x + k * TPB;
                if (r < total) {
                    const double rx = (int)floor(over_dx<POW2>(pq[k].x, dx, Nd)) * dx;
                    const double ry = (int)floor(over_dx<POW2>(pq[k].y, dx, Nd)) * dx;
                    const double rz = (int)floor(over_dx<POW2>(pq[k].z, dx, Nd)) * dx;
                    st.put((int)r, record<F, KIND, USE_PHI, POW2>(pq[k], rx, ry, rz, dx, Nd, a, mass_t));
                }
            }
            __syncthreads();
            if (DEN && (uint32_t)__builtin_amdgcn_readfirstlane(s_dmax) >= (uint32_t)GEV_DEP_DENSE_TRIG) {
                dense_partials<KIND, USE_PHI>(st, s_dlist, (int)s_nd, s_off, s_cnt, 0u, phi, N, x0, y0, z0, SX, SY,
                                              SZ, TX, TY, TZ);
                __syncthreads();
            }
            if (active) {
                const int cc = (lz * SY + ly) * SX + lx;  // source (dx, dy, dz) = (1, 1, 1)
                add_source_cell<KIND, 1, 1, 1, false>(A, st, T, cc, 0, NSRC, 0u, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 0, 1, 1, false>(A, st, T, cc + 1, 0, NSRC, 0u, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 1, 0, 1, false>(A, st, T, cc + SX, 0, NSRC, 0u, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 0, 0, 1, false>(A, st, T, cc + SX + 1, 0, NSRC, 0u, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 1, 1, 0, false>(A, st, T, cc + SX * SY, 0, NSRC, 0u, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 0, 1, 0, false>(A, st, T, cc + SX * SY + 1, 0, NSRC, 0u, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 1, 0, 0, false>(A, st, T, cc + SX * SY + SX, 0, NSRC, 0u, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 0, 0, 0, false>(A, st, T, cc + SX * SY + SX + 1, 0, NSRC, 0u, s_off, s_cnt, mass_sf);
                store_target<KIND>(A, out_s, out_t, t, n3v);
            }
            return;
        }
    }

    // 1. source cells in ascending site order: CSR begin and count (two
    //    cells per thread, all four loads in flight together)
    {
        uint32_t g[2], b[2], e[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = threadIdx.x + h * TPB;
            g[h] = 0;
            if (c < nsrc) {
                const int px = c % nsx, py = (c / nsx) % nsy, pz = c / (nsx * nsy);
                const int gx = wrapi(x0 - 1 + order_slot(px, x0, nsx), N);
                const int gy = wrapi(y0 - 1 + order_slot(py, y0, nsy), N);
                const int gz = wrapi(z0 - 1 + order_slot(pz, z0, nsz), N);
                g[h] = L(N, gx, gy, gz);
            }
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            b[h] = start[g[h]];
            e[h] = start[g[h] + 1];
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = threadIdx.x + h * TPB;
            if (c < nsrc) {
                s_gbeg[c] = b[h];
                s_cnt[c] = e[h] - b[h];
            }
        }
    }
    __syncthreads();
    // exclusive scan of the counts, two consecutive cells per thread:
    // shuffle scan inside each wavefront, wavefront totals through LDS
    {
        const int c = 2 * threadIdx.x;
        const uint32_t c0 = c < nsrc ? s_cnt[c] : 0u, c1 = c + 1 < nsrc ? s_cnt[c + 1] : 0u;
        const uint32_t v = c0 + c1;
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        uint32_t incl = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(incl, off);
            if (lane >= off) incl += u;
        }
        if (lane == 63) s_scan[wave] = incl;
        if constexpr (sizeof(F) == 4) {
            // the largest source cell of the region (the FP32 form's path and record order)
            uint32_t mx = c0 > c1 ? c0 : c1;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const uint32_t o = (uint32_t)__shfl_xor((int)mx, off);
                mx = o > mx ? o : mx;
            }
            if (lane == 0) atomicMax(&s_rmax, mx);
        }
        __syncthreads();
        uint32_t before = 0;
#pragma unroll
        for (int w = 0; w < TPB / 64; w++)
            if (w < wave) before += s_scan[w];
        const uint32_t excl = before + incl - v;
        if (c < nsrc) s_off[c] = excl;
        if (c + 1 < nsrc) s_off[c + 1] = excl + c0;
        if (threadIdx.x == TPB - 1) s_off[nsrc] = before + incl;
        __syncthreads();
    }

    // FP32 form, a tile whose source region holds more than CAP particles
    // and is crowded or has a dense cell (the clustered tiles of late
    // cycles; see GEV_DEP_FIXED_MIN): particle-parallel accumulation
    // into 64-bit fixed point in LDS.  In the gather each target thread walks
    // its 8 source cells' particles, so a dense cell keeps its wavefront busy
    // while the other lanes idle; here every thread takes particles in turn
    // and adds each one's contributions to its (up to 8) targets with integer
    // atomics, whose sums do not depend on the order: the result is
    // bit-reproducible.  Every contribution is a float rounded once to the
    // tile's fixed-point scale for its component, a power of two chosen so
    // that the sum of all the tile's contribution bounds stays below 2^62.
    if constexpr (sizeof(F) == 4 && (KIND == 1 || KIND == 2 || KIND == 3)) {
        if (s_off[nsrc] > (uint32_t)max(CAP, GEV_DEP_FIXED_MIN) ||
            (GEV_DEP_FIXED_CELL > 0 && s_off[nsrc] > (uint32_t)CAP && s_rmax >= (uint32_t)GEV_DEP_FIXED_CELL)) {
            constexpr int NCOMP = 7;  // T00, T_00, T_01, T_02, T_11, T_12, T_22
            unsigned long long *s_acc = su.fx.acc;
            float *s_T = su.fx.T;  // ph, plq01, plq02, plq12 of every target
            __shared__ float s_max[TPB / 64][NCOMP];
            __shared__ int s_k[NCOMP];
            const uint32_t total = s_off[nsrc];
            for (int i = threadIdx.x; i < NCOMP * TPB; i += TPB) s_acc[i] = 0ull;
            s_T[threadIdx.x] = T.ph;
            s_T[TPB + threadIdx.x] = T.plq01;
            s_T[2 * TPB + threadIdx.x] = T.plq02;
            s_T[3 * TPB + threadIdx.x] = T.plq12;
            // the record r of the region: source cell, slot, particle
            auto locate = [&](uint32_t r, int &c, int &sx, int &sy, int &sz, uint32_t &pidx) {
                int lo = 0, hi = nsrc - 1;  // the last source cell whose offset is <= r
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= r) lo = mid;
                    else hi = mid - 1;
                }
                c = lo;
                const int px = c % nsx, py = (c / nsx) % nsy, pz = c / (nsx * nsy);
                sx = order_slot(px, x0, nsx);
                sy = order_slot(py, y0, nsy);
                sz = order_slot(pz, z0, nsz);
                pidx = perm[s_gbeg[c] + (r - s_off[c])];
            };
            // pass 1: per-component bounds of one contribution, maximised over the tile
            float phm = active ? fabsf(T.ph) : 0.f;
            float plm = active ? fmaxf(fabsf(T.plq01), fmaxf(fabsf(T.plq02), fabsf(T.plq12))) : 0.f;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                phm = fmaxf(phm, __shfl_xor(phm, off));
                plm = fmaxf(plm, __shfl_xor(plm, off));
            }
            if ((threadIdx.x & 63) == 0) {
                s_max[threadIdx.x >> 6][0] = phm;
                s_max[threadIdx.x >> 6][1] = plm;
            }
            __syncthreads();
            phm = s_max[0][0];
            plm = s_max[0][1];
#pragma unroll
            for (int w = 1; w < TPB / 64; w++) {
                phm = fmaxf(phm, s_max[w][0]);
                plm = fmaxf(plm, s_max[w][1]);
            }
            __syncthreads();
            float bmax[NCOMP] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            const float ms = (float)mass_s;
            for (uint32_t r = threadIdx.x; r < total; r += TPB) {
                int c, sx, sy, sz;
                uint32_t pidx;
                locate(r, c, sx, sy, sz, pidx);
                const PQ v = load_pq<KIND>(pos, q, cap, pidx);
                const double rx = wrapi(x0 - 1 + sx, N) * dx, ry = wrapi(y0 - 1 + sy, N) * dx;
                const double rz = wrapi(z0 - 1 + sz, N) * dx;
                const Rec<float> R = make_rec_f32<KIND, USE_PHI, POW2>(v, rx, ry, rz, dx, Nd, (float)a, (float)mass_t);
                if (K::T00) bmax[0] = fmaxf(bmax[0], fabsf(ms * (R.e + R.f00 * phm)));
                if (K::TIJ) {
                    const float dg = 1.f + R.fij * phm, od = 1.f + R.fij * 0.25f * plm;
                    bmax[1] = fmaxf(bmax[1], fabsf(R.w0 * dg));
                    bmax[2] = fmaxf(bmax[2], fabsf(R.w01 * od));
                    bmax[3] = fmaxf(bmax[3], fabsf(R.w02 * od));
                    bmax[4] = fmaxf(bmax[4], fabsf(R.w1 * dg));
                    bmax[5] = fmaxf(bmax[5], fabsf(R.w12 * od));
                    bmax[6] = fmaxf(bmax[6], fabsf(R.w2 * dg));
                }
            }
#pragma unroll
            for (int cc = 0; cc < NCOMP; cc++) {
                float m = bmax[cc];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
                if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6][cc] = m;
            }
            __syncthreads();
            if (threadIdx.x < NCOMP) {
                float m = s_max[0][threadIdx.x];
#pragma unroll
                for (int w = 1; w < TPB / 64; w++) m = fmaxf(m, s_max[w][threadIdx.x]);
                // scale 2^k with total * bound * 2^k < 2^61
                int e = 0;
                (void)frexp((double)m * (double)total, &e);  // m * total < 2^e
                // (clamped to float's exponent range: a component whose bound is
                // below 2^-59 / total is scaled by 2^120 and keeps float precision)
                s_k[threadIdx.x] = m > 0.f ? max(-100, min(120, 61 - e)) : 0;
            }
            __syncthreads();
            float sc[NCOMP];
#pragma unroll
            for (int cc = 0; cc < NCOMP; cc++) sc[cc] = ldexpf(1.f, s_k[cc]);
            // pass 2: every record's contributions to its in-tile targets.
            // In a tile with a dense cell, lane l of a wavefront takes record
            // l * K + it at its step it (K = ceil(total / 64)), so the lanes
            // of one atomic add sit K records apart rather than in the dense
            // cell, whose particles all add to the same 8 targets
            const bool spread = GEV_FX_SPREAD > 0 && s_rmax >= (uint32_t)GEV_FX_SPREAD;
            const uint32_t KSP = (total + 63) / 64;
            const uint32_t r_first = spread ? (threadIdx.x & 63) * KSP + (threadIdx.x >> 6) : threadIdx.x;
            const uint32_t r_step = spread ? TPB / 64 : TPB;
            const uint32_t r_end = spread ? (threadIdx.x & 63) * KSP + KSP : total;
            for (uint32_t r = r_first; r < r_end; r += r_step) {
                if (r >= total) break;
                int c, sx, sy, sz;
                uint32_t pidx;
                locate(r, c, sx, sy, sz, pidx);
                const PQ v = load_pq<KIND>(pos, q, cap, pidx);
                const double rx = wrapi(x0 - 1 + sx, N) * dx, ry = wrapi(y0 - 1 + sy, N) * dx;
                const double rz = wrapi(z0 - 1 + sz, N) * dx;
                const Rec<float> R = make_rec_f32<KIND, USE_PHI, POW2>(v, rx, ry, rz, dx, Nd, (float)a, (float)mass_t);
#pragma unroll
                for (int d = 0; d < 8; d++) {
                    const int ex = d & 1, ey = (d >> 1) & 1, ez = d >> 2;
                    const int tx_ = sx - 1 + ex, ty_ = sy - 1 + ey, tz_ = sz - 1 + ez;
                    if (tx_ < 0 || tx_ >= nx || ty_ < 0 || ty_ >= ny || tz_ < 0 || tz_ >= nz) continue;
                    const int tl = tx_ + TX * (ty_ + TY * tz_);
                    const float W0 = ex ? R.up0 : 1.f - R.up0;
                    const float W1 = ey ? R.up1 : 1.f - R.up1;
                    const float W2 = ez ? R.up2 : 1.f - R.up2;
                    const float ph = s_T[tl];
                    auto add = [&](int cc, float x) __attribute__((always_inline)) {
                        atomicAdd(s_acc + cc * TPB + tl, (unsigned long long)__float2ll_rn(x * sc[cc]));
                    };
                    if (K::T00) add(0, W0 * W1 * W2 * (R.e + R.f00 * ph) * ms);
                    if (K::TIJ) {
                        const float dg = 1.f + R.fij * ph;
                        add(1, R.w0 * W0 * W1 * W2 * dg);
                        add(4, R.w1 * W0 * W1 * W2 * dg);
                        add(6, R.w2 * W0 * W1 * W2 * dg);
                        if (ex == 0 && ey == 0) add(2, R.w01 * W2 * (1.f + R.fij * 0.25f * s_T[TPB + tl]));
                        if (ex == 0 && ez == 0) add(3, R.w02 * W1 * (1.f + R.fij * 0.25f * s_T[2 * TPB + tl]));
                        if (ey == 0 && ez == 0) add(5, R.w12 * W0 * (1.f + R.fij * 0.25f * s_T[3 * TPB + tl]));
                    }
                }
            }
            __syncthreads();
            if (active) {
                auto val = [&](int cc) {
                    return (float)ldexp((double)(long long)s_acc[cc * TPB + threadIdx.x], -s_k[cc]);
                };
                if (K::T00) A.s00 += val(0);
                if (K::TIJ) {
                    A.d0 += val(1);
                    A.o01 += val(2);
                    A.o02 += val(3);
                    A.d1 += val(4);
                    A.o12 += val(5);
                    A.d2 += val(6);
                }
                store_target<KIND>(A, out_s, out_t, t, n3v);
            }
            return;
        }
    }

    // 3. chunks of whole source cells, in site order (one chunk in the
    //    common case: the 425 source cells hold fewer than CAP particles)
    int cb = 0;
    while (cb < nsrc) {
        const uint32_t base = s_off[cb];
        int ce;
        bool direct = false;
        if (s_off[cb + 1] - base > (uint32_t)CAP) {
            ce = cb + 1;
            direct = true;
        } else if (s_off[nsrc] - base <= (uint32_t)CAP) {
            ce = nsrc;
        } else {
            int lo = cb + 1, hi = nsrc;  // largest ce with s_off[ce] - base <= CAP
            while (lo < hi) {
                const int mid = (lo + hi + 1) / 2;
                if (s_off[mid] - base <= (uint32_t)CAP) lo = mid;
                else hi = mid - 1;
            }
            ce = lo;
        }
        if (!direct) {
            // record -> source cell map, then one thread per record with all
            // permutation loads, then all position/momentum loads, in flight
            const uint32_t nrec = s_off[ce] - base;
            for (int c = cb + threadIdx.x; c < ce; c += TPB) {
                const uint32_t r0 = s_off[c] - base, n = s_cnt[c];
                for (uint32_t i = 0; i < n; i++) s_rc[r0 + i] = (uint16_t)c;
                if (DEN && n >= (uint32_t)DENSE) {
                    s_dlist[atomicAdd(&s_nd, 1u)] = (uint16_t)c;
                    atomicMax(&s_dmax, n);
                }
            }
            __syncthreads();
            constexpr int PR = (CAP + TPB - 1) / TPB;
            uint32_t pidx[PR];
            int pc[PR];
#pragma unroll
            for (int k = 0; k < PR; k++) {
                const uint32_t r = threadIdx.x + k * TPB;
                pc[k] = 0;
                pidx[k] = 0;
                if (r < nrec) {
                    pc[k] = s_rc[r];
                    pidx[k] = perm[s_gbeg[pc[k]] + (r + base - s_off[pc[k]])];
                }
            }
            PQ pq[PR];
#pragma unroll
            for (int k = 0; k < PR; k++)
                if (threadIdx.x + k * TPB < nrec) pq[k] = load_pq<KIND>(pos, q, cap, pidx[k]);
#pragma unroll
            for (int k = 0; k < PR; k++) {
                const uint32_t r = threadIdx.x + k * TPB;
                if (r < nrec) {
                    const int c = pc[k];
                    const int px = c % nsx, py = (c / nsx) % nsy, pz = c / (nsx * nsy);
                    const double rx = wrapi(x0 - 1 + order_slot(px, x0, nsx), N) * dx;
                    const double ry = wrapi(y0 - 1 + order_slot(py, y0, nsy), N) * dx;
                    const double rz = wrapi(z0 - 1 + order_slot(pz, z0, nsz), N) * dx;
                    st.put((int)r, record<F, KIND, USE_PHI, POW2>(pq[k], rx, ry, rz, dx, Nd, a, mass_t));
                }
            }
            __syncthreads();
            if (DEN && (uint32_t)__builtin_amdgcn_readfirstlane(s_dmax) >= (uint32_t)GEV_DEP_DENSE_TRIG) {
                dense_partials<KIND, USE_PHI>(st, s_dlist, (int)s_nd, s_off, s_cnt, base, phi, N, x0, y0, z0, nsx,
                                              nsy, nsz, nx, ny, nz);
                __syncthreads();
            }
            if (active && interior) {
                // tiles away from coordinate 0: the 8 sources come in the
                // fixed order (dz, dy, dx) = 111, 110 .. 000 reversed per
                // axis, i.e. site order, with compile-time offsets
                const int cc = (lz * nsy + ly) * nsx + lx;  // source (dx, dy, dz) = (1, 1, 1)
                add_source_cell<KIND, 1, 1, 1>(A, st, T, cc, cb, ce, base, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 0, 1, 1>(A, st, T, cc + 1, cb, ce, base, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 1, 0, 1>(A, st, T, cc + nsx, cb, ce, base, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 0, 0, 1>(A, st, T, cc + nsx + 1, cb, ce, base, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 1, 1, 0>(A, st, T, cc + nsx * nsy, cb, ce, base, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 0, 1, 0>(A, st, T, cc + nsx * nsy + 1, cb, ce, base, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 1, 0, 0>(A, st, T, cc + nsx * nsy + nsx, cb, ce, base, s_off, s_cnt, mass_sf);
                add_source_cell<KIND, 0, 0, 0>(A, st, T, cc + nsx * nsy + nsx + 1, cb, ce, base, s_off, s_cnt,
                                               mass_sf);
            } else if (active) {
                // the 8 sources in ascending site order: along each axis the
                // lower neighbour (offset 1) comes first, except at
                // coordinate 0 where it wraps to N-1 and comes last
#pragma unroll 1
                for (int k = 0; k < 8; k++) {
                    const int ix = k & 1, iy = (k >> 1) & 1, iz = k >> 2;
                    const int dxx = x > 0 ? 1 - ix : ix, dy = y > 0 ? 1 - iy : iy, dz = z > 0 ? 1 - iz : iz;
                    const int c = (order_pos(lz + 1 - dz, z0, nsz) * nsy + order_pos(ly + 1 - dy, y0, nsy)) * nsx +
                                  order_pos(lx + 1 - dxx, x0, nsx);
#define GEV_SRC(X, Y, Z) add_source_cell<KIND, X, Y, Z>(A, st, T, c, cb, ce, base, s_off, s_cnt, mass_sf)
                    GEV_BY_OFFSET(dxx | (dy << 1) | (dz << 2), GEV_SRC)
#undef GEV_SRC
                }
            }
        } else {
            // one source cell holds more particles than a chunk: they pass
            // through the records in pieces of CAP (in order), and each
            // target it feeds carries its partial sum across the pieces
            const int c = cb;
            const uint32_t n = s_cnt[c];
            int d = -1;  // offset of source c from this thread's target
            if (active) {
#pragma unroll 1
                for (int k = 0; k < 8; k++) {
                    const int ix = k & 1, iy = (k >> 1) & 1, iz = k >> 2;
                    const int dxx = x > 0 ? 1 - ix : ix, dy = y > 0 ? 1 - iy : iy, dz = z > 0 ? 1 - iz : iz;
                    const int ck = (order_pos(lz + 1 - dz, z0, nsz) * nsy + order_pos(ly + 1 - dy, y0, nsy)) * nsx +
                                   order_pos(lx + 1 - dxx, x0, nsx);
                    if (ck == c && feeds<KIND>(dxx, dy, dz)) d = dxx | (dy << 1) | (dz << 2);
                }
            }
            const int px = c % nsx, py = (c / nsx) % nsy, pz = c / (nsx * nsy);
            const double rx = wrapi(x0 - 1 + order_slot(px, x0, nsx), N) * dx;
            const double ry = wrapi(y0 - 1 + order_slot(py, y0, nsy), N) * dx;
            const double rz = wrapi(z0 - 1 + order_slot(pz, z0, nsz), N) * dx;
            Partial<F> P;
            P.clear();
            for (uint32_t s0 = 0; s0 < n; s0 += (uint32_t)CAP) {
                const uint32_t m = min((uint32_t)CAP, n - s0);
                for (uint32_t r = threadIdx.x; r < m; r += TPB)
                    st.put((int)r, record<F, KIND, USE_PHI, POW2>(load_pq<KIND>(pos, q, cap, perm[s_gbeg[c] + s0 + r]),
                                                                  rx, ry, rz, dx, Nd, a, mass_t));
                __syncthreads();
#define GEV_REC(X, Y, Z) add_records<KIND, X, Y, Z>(P, st, T, m)
                if (d >= 0) {
                    GEV_BY_OFFSET(d, GEV_REC)
                }
#undef GEV_REC
                __syncthreads();
            }
#define GEV_ADD(X, Y, Z) add_source<KIND>(A, P, T, X, Y, Z, mass_sf)
            if (d >= 0) {
                GEV_BY_OFFSET(d, GEV_ADD)
            }
#undef GEV_ADD
        }
        if (DEN && threadIdx.x == 0) s_nd = s_dmax = 0;  // every thread has read them
        __syncthreads();
        cb = ce;
    }

    if (active) store_target<KIND>(A, out_s, out_t, t, n3v);
}

template <int KIND, bool USE_PHI, bool ACCUM, typename F>
void launch_t(gev_ctx *ctx, F *out_s, F *out_t, double a, double ms, double mt, const gev::Batch &bt) {
    const int N = ctx->N;
    const int ntx = (N + TX - 1) / TX, nty = (N + TY - 1) / TY, ntz = (N + TZ - 1) / TZ;
    const int64_t nblk = (int64_t)ntx * nty * ntz;
    const unsigned nbox = bt.cyc ? (unsigned)ctx->nbox : 1u;
    const bool pow2 = (N & (N - 1)) == 0, full = N % TX == 0 && N % TY == 0 && N % TZ == 0;
#define GEV_DEP_LAUNCH(P2, FL, B)                                                                             \
    hipLaunchKernelGGL((k_deposit_tiled<KIND, USE_PHI, ACCUM, P2, FL, F, B>), dim3((unsigned)nblk, nbox), dim3(TPB), \
                       0, ctx->stream, ctx->pos, ctx->q, ctx->cap, ctx->start, ctx->perm, gev::fld<F>(ctx->phi),     \
                       out_s, out_t, N, ntx, nty, ntz, a, ms, mt, ctx->pos2, ctx->q2, bt)
    if (bt.cyc) {
        // the batched cycle deposits in store mode only
        if constexpr (!ACCUM) {
            if (pow2 && full) GEV_DEP_LAUNCH(true, true, true);
            else if (full) GEV_DEP_LAUNCH(false, true, true);
            else GEV_DEP_LAUNCH(false, false, true);
        }
    } else if (pow2 && full) {
        GEV_DEP_LAUNCH(true, true, false);
    } else if (full) {
        GEV_DEP_LAUNCH(false, true, false);
    } else {
        GEV_DEP_LAUNCH(false, false, false);
    }
#undef GEV_DEP_LAUNCH
}

template <int KIND, typename F>
void launch_f(gev_ctx *ctx, bool use_phi, bool accum, F *out_s, F *out_t, double a, double ms, double mt,
              const gev::Batch &bt) {
    if (use_phi) {
        if (accum) launch_t<KIND, true, true>(ctx, out_s, out_t, a, ms, mt, bt);
        else launch_t<KIND, true, false>(ctx, out_s, out_t, a, ms, mt, bt);
    } else {
        if (accum) launch_t<KIND, false, true>(ctx, out_s, out_t, a, ms, mt, bt);
        else launch_t<KIND, false, false>(ctx, out_s, out_t, a, ms, mt, bt);
    }
}

template <int KIND>
void launch_k(gev_ctx *ctx, bool use_phi, bool accum, double *out_s, double *out_t, double a, double ms,
              double mt, const gev::Batch &bt = gev::Batch{}) {
    if (ctx->fp32)
        launch_f<KIND>(ctx, use_phi, accum, gev::fld<float>(out_s), gev::fld<float>(out_t), a, ms, mt, bt);
    else
        launch_f<KIND>(ctx, use_phi, accum, out_s, out_t, a, ms, mt, bt);
}

}  // namespace

namespace gev {

int launch_deposit(gev_ctx *ctx, int kind, double a, int use_phi, double coeff) {
    const double dx = 1.0 / (double)ctx->N;
    // mass factors of gevolution.hpp:945-947 / 1064-1065 / 1191-1193
    double m = coeff / (dx * dx * dx);
    m *= ctx->pmass;
    const double m_a = m / a;
    switch (kind) {
    case DEP_T00: {
        ScopedTimer t(ctx, "projection_T00_project");
        launch_k<1>(ctx, use_phi, !ctx->zeroed[GEV_SOURCE], ctx->source, nullptr, a, m_a, m_a);
        ctx->zeroed[GEV_SOURCE] = false;
        break;
    }
    case DEP_TIJ: {
        ScopedTimer t(ctx, "projection_Tij_project");
        launch_k<2>(ctx, use_phi, !ctx->zeroed[GEV_SIJ], nullptr, ctx->Sij, a, m_a, m_a);
        ctx->zeroed[GEV_SIJ] = false;
        break;
    }
    case DEP_T00_TIJ: {
        // a zeroed target reads back zeros when the other one must accumulate
        ScopedTimer t(ctx, "projection_T00_Tij_project");
        const bool acc = !ctx->zeroed[GEV_SOURCE] || !ctx->zeroed[GEV_SIJ];
        launch_k<3>(ctx, true, acc, ctx->source, ctx->Sij, a, m_a, m_a);
        ctx->zeroed[GEV_SOURCE] = false;
        ctx->zeroed[GEV_SIJ] = false;
        break;
    }
    case DEP_T0I: {
        ScopedTimer t(ctx, "projection_T0i_project");
        launch_k<4>(ctx, use_phi, !ctx->zeroed[GEV_BI], nullptr, ctx->Bi, 1.0, m, m);
        ctx->zeroed[GEV_BI] = false;
        break;
    }
    case DEP_CIC: {
        // projection_T00_project with a = 1 and phi = NULL
        ScopedTimer t(ctx, "scalar_projection_cic_project");
        launch_k<5>(ctx, false, !ctx->zeroed[GEV_SOURCE], ctx->source, nullptr, 1.0, m, m);
        ctx->zeroed[GEV_SOURCE] = false;
        break;
    }
    default:
        return fail(ctx, GEV_EINVAL, "unknown deposit kind");
    }
    return check_hip(ctx, hipGetLastError(), "deposit launch");
}

// The field stage's deposits over the boxes of a batch (store mode; every
// box's coefficients from its BoxCycle): T00 + Tij (GR) or the CIC mass and
// Tij (Newton).
int launch_deposit_batch(gev_ctx *ctx, bool gr, const Batch &bt) {
    if (gr) {
        ScopedTimer t(ctx, "projection_T00_Tij_project");
        launch_k<3>(ctx, true, false, ctx->source, ctx->Sij, 0., 0., 0., bt);
    } else {
        {
            ScopedTimer t(ctx, "scalar_projection_cic_project");
            launch_k<5>(ctx, false, false, ctx->source, nullptr, 1., 0., 0., bt);
        }
        ScopedTimer t(ctx, "projection_Tij_project");
        launch_k<2>(ctx, true, false, nullptr, ctx->Sij, 0., 0., 0., bt);
    }
    return check_hip(ctx, hipGetLastError(), "deposit launch");
}

}  // namespace gev
