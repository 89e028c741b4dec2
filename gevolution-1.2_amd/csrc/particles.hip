// particles.hip -- particle kernels of one box (the CIC projections are in
// deposit.hip):
//   * re-filing (the cell bookkeeping of LATfield2 Particles::moveParticles):
//     per-cell counts, CSR offsets by a three-phase scan, and a permutation
//     listing each cell's particles in ascending ID.
//   * updateVel / moveParticles with update_q(_Newton) / update_pos(_Newton)
//     (gevolution.hpp:570-903), separately or fused into one particle pass.
#include <utility>

#include "gev_internal.h"

namespace {

constexpr int TPB = 256;

__device__ __forceinline__ int64_t L(int N, int x, int y, int z) {
    return (int64_t)x + (int64_t)N * ((int64_t)y + (int64_t)N * z);
}
__device__ __forceinline__ int wrapi(int i, int N) { return i < 0 ? i + N : (i >= N ? i - N : i); }

// ------------------------------------------------------------ filing

// x / dx with dx = 1/N.  For N a power of two dx is exact and the quotient
// equals x * N bit for bit, so the multiply replaces the division.
template <bool POW2> __device__ __forceinline__ double over_dx(double x, double dx, double Nd) {
    return POW2 ? x * Nd : x / dx;
}

// Cell of one coordinate: floor(p/dx) with a true division.
template <bool POW2> __device__ __forceinline__ int cell1(double p, double dx, double Nd) {
    return (int)floor(over_dx<POW2>(p, dx, Nd));
}

__global__ __launch_bounds__(TPB) void k_count_cells(const double *__restrict__ pos, int64_t cap,
                                                     int64_t n, int N, uint32_t *__restrict__ count,
                                                     uint32_t *__restrict__ cell) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    const double dx = 1.0 / (double)N;
    const int cx = cell1<false>(pos[p], dx, N), cy = cell1<false>(pos[cap + p], dx, N);
    const int cz = cell1<false>(pos[2 * cap + p], dx, N);
    const uint32_t c = (uint32_t)L(N, cx, cy, cz);
    cell[p] = c;
    atomicAdd(count + c, 1u);
}

// three-phase exclusive scan of per-cell counts into CSR offsets.  A block
// covers SCAN_BLOCK = 4096 counts as SCAN_PASSES passes of TPB x 4 counts;
// every load and store is a coalesced 16-byte access (integer sums are exact,
// so the grouping does not change the result).
constexpr int SCAN_PASSES = 4;
constexpr int SCAN_BLOCK = TPB * 4 * SCAN_PASSES;

// counts i .. i+3 (i a multiple of 4; in is 16-byte aligned), zero past n
__device__ __forceinline__ uint4 load4(const uint32_t *__restrict__ in, int64_t i, int64_t n) {
    if (i + 3 < n) return *reinterpret_cast<const uint4 *>(in + i);
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (i < n) v.x = in[i];
    if (i + 1 < n) v.y = in[i + 1];
    if (i + 2 < n) v.z = in[i + 2];
    return v;
}

// exclusive scan of one value per thread over the workgroup: shuffle scan
// inside each wavefront, wavefront totals through LDS
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t s_w[TPB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < TPB / 64; w++) {
        const uint32_t t = s_w[w];
        if (w < wave) before += t;
        tot += t;
    }
    __syncthreads();  // s_w is reused by the next call
    *total = tot;
    return before + incl - v;
}

__global__ __launch_bounds__(TPB) void k_scan_reduce(const uint32_t *__restrict__ in, int64_t n,
                                                     uint32_t *__restrict__ part, gev::Batch bt) {
    if (!gev::box_on(bt, blockIdx.y)) return;
    in = gev::box_ptr(in, bt.stride, blockIdx.y);
    part = gev::box_ptr(part, bt.stride, blockIdx.y);
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_BLOCK;
    uint4 v[SCAN_PASSES];
#pragma unroll
    for (int k = 0; k < SCAN_PASSES; k++) v[k] = load4(in, b0 + 4 * ((int64_t)k * TPB + threadIdx.x), n);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PASSES; k++) s += v[k].x + v[k].y + v[k].z + v[k].w;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ uint32_t sh[TPB / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < TPB / 64; w++) t += sh[w];
        part[blockIdx.x] = t;
    }
}

// exclusive scan of the block partials in place (single workgroup)
__global__ __launch_bounds__(TPB) void k_scan_partials(uint32_t *__restrict__ part, int64_t nblk, gev::Batch bt) {
    if (!gev::box_on(bt, blockIdx.y)) return;
    part = gev::box_ptr(part, bt.stride, blockIdx.y);
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

__global__ __launch_bounds__(TPB) void k_scan_final(const uint32_t *__restrict__ cnt, int64_t n,
                                                    const uint32_t *__restrict__ part,
                                                    uint32_t *__restrict__ out, gev::Batch bt) {
    if (!gev::box_on(bt, blockIdx.y)) return;
    cnt = gev::box_ptr(cnt, bt.stride, blockIdx.y);
    part = gev::box_ptr(part, bt.stride, blockIdx.y);
    out = gev::box_ptr(out, bt.stride, blockIdx.y);
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_BLOCK;
    uint4 v[SCAN_PASSES];
#pragma unroll
    for (int k = 0; k < SCAN_PASSES; k++) v[k] = load4(cnt, b0 + 4 * ((int64_t)k * TPB + threadIdx.x), n);
    uint32_t carry = part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_PASSES; k++) {
        const int64_t i = b0 + 4 * ((int64_t)k * TPB + threadIdx.x);
        uint32_t total;
        const uint32_t off = carry + block_exclusive_scan(v[k].x + v[k].y + v[k].z + v[k].w, &total);
        const uint4 o = make_uint4(off, off + v[k].x, off + v[k].x + v[k].y, off + v[k].x + v[k].y + v[k].z);
        if (i + 3 < n) {
            *reinterpret_cast<uint4 *>(out + i) = o;
        } else {
            if (i < n) out[i] = o.x;
            if (i + 1 < n) out[i + 1] = o.y;
            if (i + 2 < n) out[i + 2] = o.z;
        }
        carry += total;
    }
    // the last block also writes out[n] = grand total
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = carry;
}

// Runs of equal cells among the active lanes of a wavefront (consecutive
// active lanes with the same c): each lane's offset in its run, and the
// run's length at its first lane.  Particles come in cell order (storage
// order is the last filing's), so a dense cell's particles share a few
// runs, and one atomic per run replaces one per particle.
struct CellRun {
    int head;      // lane of the run's first lane
    uint32_t off;  // this lane's offset in the run
    uint32_t len;  // run length (valid at the head lane)
};
__device__ __forceinline__ CellRun cell_run(uint32_t c) {
    const uint64_t active = __ballot(1);
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t below = active & ((1ull << lane) - 1ull);
    const int prev = below ? 63 - __clzll(below) : -1;  // previous active lane
    const uint32_t cprev = __shfl(c, prev < 0 ? lane : prev);
    const bool head = prev < 0 || cprev != c;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1ull));
    CellRun r;
    r.head = 63 - __clzll(upto);
    r.off = (uint32_t)__popcll(active & ((1ull << lane) - 1ull) & ~((1ull << r.head) - 1ull));
    // the run ends before the next head (or after the last active lane)
    const uint64_t later = heads & ~upto;
    const int end = later ? __ffsll((long long)later) - 1 : 64;
    r.len = (uint32_t)__popcll(active & ~((1ull << lane) - 1ull) & (end == 64 ? ~0ull : ((1ull << end) - 1ull)));
    return r;
}

// every particle takes a place of its cell, counting the cell's count back
// down (so every completed filing leaves count[] all zero for the next one);
// the in-cell order is then fixed by the ID sort.  A place outside the
// cell's CSR range (counts not zero before the counting pass) is dropped.
// One atomic per run of equal cells in a wavefront.
__global__ __launch_bounds__(TPB) void k_scatter(const uint32_t *__restrict__ cell, const uint32_t *__restrict__ start,
                                                 uint32_t *__restrict__ count, int64_t n,
                                                 uint32_t *__restrict__ perm, gev::Batch bt) {
    if (!gev::box_on(bt, blockIdx.y)) return;
    cell = gev::box_ptr(cell, bt.stride, blockIdx.y);
    start = gev::box_ptr(start, bt.stride, blockIdx.y);
    count = gev::box_ptr(count, bt.stride, blockIdx.y);
    perm = gev::box_ptr(perm, bt.stride, blockIdx.y);
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    const uint32_t c = cell[p];
    const CellRun r = cell_run(c);
    uint32_t k = 0;
    if (r.off == 0) k = atomicSub(count + c, r.len);
    k = __shfl(k, r.head) - r.off;  // this particle's count value, as if it had counted down alone
    const uint32_t b = start[c];
    if (k - 1u < start[c + 1] - b) perm[b + k - 1u] = (uint32_t)p;
}

// Ascending ID inside every cell.  A workgroup takes TPB consecutive cells,
// whose particles are one contiguous run of the permutation.  Cells of up to
// SORT_TINY particles are sorted by their own thread in registers (a stable
// compare-exchange network).  If larger cells remain and the segment's run
// fits SORT_SEG entries, their entries are copied to LDS with their IDs,
// every such particle's rank inside its cell (ID, ties by position) is
// counted there by one thread per particle, and the entries are written back
// in rank order.  A denser segment lists its larger cells for k_sort_big (a
// workgroup per cell).  (A thread-per-cell insertion sort over global memory
// made dense cells slow: every comparison was two dependent global loads.)
constexpr uint32_t SORT_TINY = 4;
// Permutation entries of a segment sorted in LDS.  1024 (17 KB of LDS) lets
// 8 workgroups share a CU, twice as many as 2048: the segment sort is
// latency-bound (cell starts, then permutation and IDs, then the ranks), and
// at 512^3, z = 0 it takes 1.64 against 2.85 ms; the few denser segments
// go to k_sort_big.
#ifndef GEV_SORT_SEG
#define GEV_SORT_SEG 1024
#endif
constexpr int SORT_SEG = GEV_SORT_SEG;
// the next segment's cell starts are loaded while a segment is ranked
// (1.56 against 1.64 ms at 512^3, z = 0)
#ifndef GEV_SORT_PF
#define GEV_SORT_PF 1
#endif
static_assert(TPB <= 256, "a segment's cell index fits 8 bits");
constexpr int SORT_LDS = 2048;  // particles of a listed cell k_sort_big ranks in LDS

__device__ __forceinline__ void insertion_sort(uint32_t *__restrict__ perm, const int64_t *__restrict__ id,
                                               uint32_t b0, uint32_t b1) {
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

__device__ __forceinline__ void cswap(uint32_t *v, int64_t *k, int a, int b) {
    if (k[a] > k[b]) {
        const int64_t tk = k[a];
        k[a] = k[b];
        k[b] = tk;
        const uint32_t tv = v[a];
        v[a] = v[b];
        v[b] = tv;
    }
}

// thread per cell: the tiny sorts; a segment with a larger cell is flagged
// (a flag per segment: one shared list counter would serialise the late,
// clustered cycles, which flag most segments, on one atomic)
// the ID array the filing sorts by: in a batch, the set the drift just wrote
// (the other one than the box's live set), at the box's slab
__device__ __forceinline__ const int64_t *sort_ids(const int64_t *id, const int64_t *id_alt, const gev::Batch &bt,
                                                   int b) {
    if (!bt.cyc) return id;
    return gev::box_ptr(bt.cyc[b].parity ? id : id_alt, bt.stride, b);
}

__global__ __launch_bounds__(TPB) void k_sort_cells(const uint32_t *__restrict__ start,
                                                    const int64_t *__restrict__ id, int64_t n3v,
                                                    uint32_t *__restrict__ perm, uint8_t *__restrict__ flag,
                                                    const int64_t *__restrict__ id_alt, gev::Batch bt) {
    if (!gev::box_on(bt, blockIdx.y)) return;
    id = sort_ids(id, id_alt, bt, blockIdx.y);
    start = gev::box_ptr(start, bt.stride, blockIdx.y);
    perm = gev::box_ptr(perm, bt.stride, blockIdx.y);
    flag = gev::box_ptr(flag, bt.stride, blockIdx.y);
    const int64_t c = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const uint32_t b0 = c < n3v ? start[c] : 0u;
    const uint32_t m = c < n3v ? start[c + 1] - b0 : 0u;
    if (m >= 2 && m <= SORT_TINY) {
        uint32_t v[SORT_TINY];
        int64_t k[SORT_TINY];
#pragma unroll
        for (int i = 0; i < (int)SORT_TINY; i++) v[i] = (uint32_t)i < m ? perm[b0 + i] : 0u;
#pragma unroll
        for (int i = 0; i < (int)SORT_TINY; i++) k[i] = (uint32_t)i < m ? id[v[i]] : INT64_MAX;
        // bubble network: stable (exchanges only on strictly greater keys)
        cswap(v, k, 0, 1);
        cswap(v, k, 1, 2);
        cswap(v, k, 2, 3);
        cswap(v, k, 0, 1);
        cswap(v, k, 1, 2);
        cswap(v, k, 0, 1);
#pragma unroll
        for (int i = 0; i < (int)SORT_TINY; i++)
            if ((uint32_t)i < m) perm[b0 + i] = v[i];
    }
    const int big = __syncthreads_or(m > SORT_TINY);
    if (threadIdx.x == 0) flag[blockIdx.x] = big ? 1 : 0;
}

// Workgroup b takes the flagged ones among segments b + k * gridDim.x,
// k < SORT_SEG_BATCH (its first wave reads their flags at once): the larger
// cells' entries are ranked in LDS, or, in a segment of more than SORT_SEG
// entries, listed for k_sort_big.  The stride spreads neighbouring (and so
// similarly clustered) segments over the workgroups.
constexpr int SORT_SEG_BATCH = 32;
static_assert(SORT_SEG_BATCH <= 64, "one wavefront reads a batch's flags");

__global__ __launch_bounds__(TPB) void k_sort_segments(const uint32_t *__restrict__ start,
                                                       const int64_t *__restrict__ id, int64_t n3v,
                                                       uint32_t *__restrict__ perm, const uint8_t *__restrict__ flag,
                                                       uint32_t nsegs, uint32_t *__restrict__ big,
                                                       uint32_t *__restrict__ nbig, const int64_t *__restrict__ id_alt,
                                                       gev::Batch bt) {
    if (!gev::box_on(bt, blockIdx.y)) return;
    id = sort_ids(id, id_alt, bt, blockIdx.y);
    start = gev::box_ptr(start, bt.stride, blockIdx.y);
    perm = gev::box_ptr(perm, bt.stride, blockIdx.y);
    flag = gev::box_ptr(flag, bt.stride, blockIdx.y);
    big = gev::box_ptr(big, bt.stride, blockIdx.y);
    nbig = gev::box_ptr(nbig, bt.stride, blockIdx.y);
    __shared__ int64_t s_key[SORT_SEG];
    __shared__ uint32_t s_val[SORT_SEG];
    __shared__ uint32_t s_out[SORT_SEG];
    __shared__ uint32_t s_cs[TPB + 1];  // cell starts relative to the segment
    __shared__ uint8_t s_cell[SORT_SEG];  // each entry's cell in the segment
    __shared__ unsigned long long s_mask;
    if (threadIdx.x < 64) {
        const uint32_t e = blockIdx.x + threadIdx.x * gridDim.x;
        const bool f = threadIdx.x < SORT_SEG_BATCH && e < nsegs && flag[e] != 0;
        const unsigned long long m = __ballot(f);
        if (threadIdx.x == 0) s_mask = m;
    }
    __syncthreads();
    // GEV_SORT_PF: the cell starts of the next flagged segment are loaded
    // while this one is ranked (thread t: cell t, thread 0 also the end;
    // a segment has at most TPB cells)
    uint32_t pf_s = 0, pf_e = 0;
    auto prefetch = [&](unsigned long long m) {
        if (m == 0ull) return;
        const int64_t c = (int64_t)(blockIdx.x + (uint32_t)(__ffsll((long long)m) - 1) * gridDim.x) * TPB;
        const int nc = (int)(n3v - c < TPB ? n3v - c : TPB);
        if ((int)threadIdx.x <= nc) pf_s = start[c + threadIdx.x];
        if (threadIdx.x == 0) pf_e = start[c + nc];
    };
    if (GEV_SORT_PF) prefetch(s_mask);
    for (unsigned long long mask = s_mask; mask != 0ull; mask &= mask - 1ull) {
        const uint32_t e = blockIdx.x + (uint32_t)(__ffsll((long long)mask) - 1) * gridDim.x;
        const int64_t c0 = (int64_t)e * TPB;
        const int ncell = (int)(n3v - c0 < TPB ? n3v - c0 : TPB);
        uint32_t s0 = 0;
        if (GEV_SORT_PF) {
            // absolute starts (start[c0] is thread 0's value, read back
            // after the barrier)
            if ((int)threadIdx.x <= ncell) s_cs[threadIdx.x] = pf_s;
            if (threadIdx.x == 0) s_cs[ncell] = pf_e;
            prefetch(mask & (mask - 1ull));
        } else {
            s0 = start[c0];
            for (int t = threadIdx.x; t <= ncell; t += TPB) s_cs[t] = start[c0 + t] - s0;
        }
        __syncthreads();
        if (GEV_SORT_PF) s0 = s_cs[0];
        const uint32_t sb = GEV_SORT_PF ? s0 : 0u;  // s_cs[t] - sb: cell t's start in the segment
        const uint32_t total = s_cs[ncell] - sb;
        if (total > (uint32_t)SORT_SEG) {
            // dense segment: its larger cells are ranked one workgroup each
            if (threadIdx.x < ncell && s_cs[threadIdx.x + 1] - s_cs[threadIdx.x] > SORT_TINY)
                big[atomicAdd(nbig, 1u)] = (uint32_t)(c0 + threadIdx.x);
            __syncthreads();  // s_cs is rewritten for the next segment
            continue;
        }
        auto cell_of = [&](uint32_t i) {
            int lo = 0, hi = ncell - 1;  // the last cell whose start is <= i
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_cs[mid] - sb <= i) lo = mid;
                else hi = mid - 1;
            }
            return lo;
        };
        for (uint32_t i = threadIdx.x; i < total; i += TPB) {
            const int c = cell_of(i);
            s_cell[i] = (uint8_t)c;
            if (s_cs[c + 1] - s_cs[c] > SORT_TINY) {
                const uint32_t v = perm[s0 + i];
                s_val[i] = v;
                s_key[i] = id[v];
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < total; i += TPB) {
            const int c = s_cell[i];
            const uint32_t b = s_cs[c] - sb, f = s_cs[c + 1] - sb;
            if (f - b <= SORT_TINY) continue;
            const int64_t k = s_key[i];
            uint32_t r = 0;
            for (uint32_t j = b; j < f; j++) {
                const int64_t kj = s_key[j];
                r += (kj < k || (kj == k && j < i)) ? 1u : 0u;
            }
            s_out[b + r] = s_val[i];
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < total; i += TPB) {
            const int c = s_cell[i];
            if (s_cs[c + 1] - s_cs[c] > SORT_TINY) perm[s0 + i] = s_out[i];
        }
        __syncthreads();  // the next segment rewrites s_cs and the copies
    }
}

// One workgroup per listed cell (grid-stride over the list): every particle's
// rank among the cell's IDs (ties by position) from an LDS copy, then one
// store each.  Cells above SORT_LDS particles fall back to one thread.
__global__ __launch_bounds__(TPB) void k_sort_big(const uint32_t *__restrict__ start, const int64_t *__restrict__ id,
                                                  uint32_t *__restrict__ perm, const uint32_t *__restrict__ big,
                                                  const uint32_t *__restrict__ nbig, const int64_t *__restrict__ id_alt,
                                                  gev::Batch bt) {
    // (nbig: the big list's count)
    if (!gev::box_on(bt, blockIdx.y)) return;
    id = sort_ids(id, id_alt, bt, blockIdx.y);
    start = gev::box_ptr(start, bt.stride, blockIdx.y);
    perm = gev::box_ptr(perm, bt.stride, blockIdx.y);
    big = gev::box_ptr(big, bt.stride, blockIdx.y);
    nbig = gev::box_ptr(nbig, bt.stride, blockIdx.y);
    __shared__ int64_t s_key[SORT_LDS];
    __shared__ uint32_t s_val[SORT_LDS];
    const uint32_t nb = *nbig;
    for (uint32_t e = blockIdx.x; e < nb; e += gridDim.x) {
        const uint32_t c = big[e];
        const uint32_t b0 = start[c], b1 = start[c + 1], m = b1 - b0;
        if (m > (uint32_t)SORT_LDS) {
            if (threadIdx.x == 0) insertion_sort(perm, id, b0, b1);
            __syncthreads();
            continue;
        }
        for (uint32_t i = threadIdx.x; i < m; i += TPB) {
            const uint32_t v = perm[b0 + i];
            s_val[i] = v;
            s_key[i] = id[v];
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < m; i += TPB) {
            const int64_t k = s_key[i];
            uint32_t r = 0;
            for (uint32_t j = 0; j < m; j++) {
                const int64_t kj = s_key[j];
                r += (kj < k || (kj == k && j < i)) ? 1u : 0u;
            }
            perm[b0 + r] = s_val[i];
        }
        __syncthreads();  // the next cell overwrites the LDS copy
    }
}

// ------------------------------------------------------------ kick / drift

// Particles are updated tile by tile: a workgroup takes the particles filed
// in a KTX x KTY x KTZ block of cells (default 32x4x4; runs of the filing
// permutation) and first stages phi, chi and B in LDS, each over the block
// plus the halo its stencils reach (Region below), so every gather of
// update_q / update_pos is an LDS read.  The long x extent keeps the halo's
// share of the staged cells down (16x4x4 and 16x8x4 measured slower).
#ifndef GEV_KICK_TX
#define GEV_KICK_TX 32
#endif
#ifndef GEV_KICK_TY
#define GEV_KICK_TY 4
#endif
#ifndef GEV_KICK_TZ
#define GEV_KICK_TZ 4
#endif
constexpr int KTX = GEV_KICK_TX, KTY = GEV_KICK_TY, KTZ = GEV_KICK_TZ;

// The staged region of one field: the tile's cells plus the cell above on
// every axis (the stencils reach offsets 0 and 1), and the cell below on
// axis LO only: the gradients of update_q reach offset -1 in x for B0, in y
// for B1 and in z for B2, never for phi or chi (LO = -1).
template <int LO> struct Region {
    static constexpr int OX = LO == 0, OY = LO == 1, OZ = LO == 2;  // lower halo per axis
    static constexpr int SX = KTX + 1 + OX, SY = KTY + 1 + OY, SZ = KTZ + 1 + OZ;
    static constexpr int SIZE = SX * SY * SZ;
};

// one field around the particle's cell: (dx, dy, dz) within the region
// (F: the field element, double, or float in the FP32 form; values are read
// as double)
template <int LO, typename F> struct FieldView {
    const F *f;  // LDS image of the field over its staged region
    int base;    // region index of the particle's cell
    __device__ __forceinline__ double operator()(int dx, int dy, int dz) const {
        return (double)f[base + dx + Region<LO>::SX * (dy + Region<LO>::SY * dz)];
    }
    __device__ __forceinline__ F raw(int dx, int dy, int dz) const {
        return f[base + dx + Region<LO>::SX * (dy + Region<LO>::SY * dz)];
    }
};
template <int LO, typename F>
__device__ __forceinline__ FieldView<LO, F> view(const F *f, const int c[3], int x0, int y0, int z0) {
    using R = Region<LO>;
    return FieldView<LO, F>{f, (c[0] - x0 + R::OX) + R::SX * ((c[1] - y0 + R::OY) + R::SY * (c[2] - z0 + R::OZ))};
}
template <typename F> using ViewS = FieldView<-1, F>;  // phi, chi
template <typename F> using View0 = FieldView<0, F>;
template <typename F> using View1 = FieldView<1, F>;
template <typename F> using View2 = FieldView<2, F>;

// Compiler-only memory fence between the phases of the particle update: the
// LDS field values of one phase are read again in the next rather than held
// in registers, which bounds the live register set (more waves per SIMD).
#define GEV_LDS_PHASE() asm volatile("" ::: "memory")

// update_q, gevolution.hpp:570-678 (GRADIENT_ORDER 1); returns v^2/a^2
template <bool POW2, typename F>
__device__ __forceinline__ double update_q(double q[3], const double rd[3], ViewS<F> phi,
                                           ViewS<F> chi, View0<F> B0, View1<F> B1, View2<F> B2,
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

    GEV_LDS_PHASE();
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

    GEV_LDS_PHASE();
    if (nfield >= 3) {
        pgradB[0] = ((1. - rd[2]) * (B1(1, 0, 0) - B1(0, 0, 0)) + rd[2] * (B1(1, 0, 1) - B1(0, 0, 1))) * q[1];
        pgradB[0] += ((1. - rd[1]) * (B2(1, 0, 0) - B2(0, 0, 0)) + rd[1] * (B2(1, 1, 0) - B2(0, 1, 0))) * q[2];
        pgradB[0] += (1. - rd[1]) * (1. - rd[2]) * ((rd[0] - 1.) * B0(-1, 0, 0) + (1. - 2. * rd[0]) * B0(0, 0, 0) + rd[0] * B0(1, 0, 0)) * q[0];
        pgradB[0] += rd[1] * (1. - rd[2]) * ((rd[0] - 1.) * B0(-1, 1, 0) + (1. - 2. * rd[0]) * B0(0, 1, 0) + rd[0] * B0(1, 1, 0)) * q[0];
        pgradB[0] += (1. - rd[1]) * rd[2] * ((rd[0] - 1.) * B0(-1, 0, 1) + (1. - 2. * rd[0]) * B0(0, 0, 1) + rd[0] * B0(1, 0, 1)) * q[0];
        pgradB[0] += rd[1] * rd[2] * ((rd[0] - 1.) * B0(-1, 1, 1) + (1. - 2. * rd[0]) * B0(0, 1, 1) + rd[0] * B0(1, 1, 1)) * q[0];

        GEV_LDS_PHASE();
        pgradB[1] = ((1. - rd[0]) * (B2(0, 1, 0) - B2(0, 0, 0)) + rd[0] * (B2(1, 1, 0) - B2(1, 0, 0))) * q[2];
        pgradB[1] += ((1. - rd[2]) * (B0(0, 1, 0) - B0(0, 0, 0)) + rd[2] * (B0(0, 1, 1) - B0(0, 0, 1))) * q[0];
        pgradB[1] += (1. - rd[0]) * (1. - rd[2]) * ((rd[1] - 1.) * B1(0, -1, 0) + (1. - 2. * rd[1]) * B1(0, 0, 0) + rd[1] * B1(0, 1, 0)) * q[1];
        pgradB[1] += rd[0] * (1. - rd[2]) * ((rd[1] - 1.) * B1(1, -1, 0) + (1. - 2. * rd[1]) * B1(1, 0, 0) + rd[1] * B1(1, 1, 0)) * q[1];
        pgradB[1] += (1. - rd[0]) * rd[2] * ((rd[1] - 1.) * B1(0, -1, 1) + (1. - 2. * rd[1]) * B1(0, 0, 1) + rd[1] * B1(0, 1, 1)) * q[1];
        pgradB[1] += rd[0] * rd[2] * ((rd[1] - 1.) * B1(1, -1, 1) + (1. - 2. * rd[1]) * B1(1, 0, 1) + rd[1] * B1(1, 1, 1)) * q[1];

        GEV_LDS_PHASE();
        pgradB[2] = ((1. - rd[1]) * (B0(0, 0, 1) - B0(0, 0, 0)) + rd[1] * (B0(0, 1, 1) - B0(0, 1, 0))) * q[0];
        pgradB[2] += ((1. - rd[0]) * (B1(0, 0, 1) - B1(0, 0, 0)) + rd[0] * (B1(1, 0, 1) - B1(1, 0, 0))) * q[1];
        pgradB[2] += (1. - rd[0]) * (1. - rd[1]) * ((rd[2] - 1.) * B2(0, 0, -1) + (1. - 2. * rd[2]) * B2(0, 0, 0) + rd[2] * B2(0, 0, 1)) * q[2];
        pgradB[2] += rd[0] * (1. - rd[1]) * ((rd[2] - 1.) * B2(1, 0, -1) + (1. - 2. * rd[2]) * B2(1, 0, 0) + rd[2] * B2(1, 0, 1)) * q[2];
        pgradB[2] += (1. - rd[0]) * rd[1] * ((rd[2] - 1.) * B2(0, 1, -1) + (1. - 2. * rd[2]) * B2(0, 1, 0) + rd[2] * B2(0, 1, 1)) * q[2];
        pgradB[2] += rd[0] * rd[1] * ((rd[2] - 1.) * B2(1, 1, -1) + (1. - 2. * rd[2]) * B2(1, 1, 0) + rd[2] * B2(1, 1, 1)) * q[2];

        GEV_LDS_PHASE();
        gradphi[0] += pgradB[0] / params[1] / e2;
        gradphi[1] += pgradB[1] / params[1] / e2;
        gradphi[2] += pgradB[2] / params[1] / e2;
    }

    v2 = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        q[i] -= over_dx<POW2>(dtau * e2 * gradphi[i], dx, 1.0 / dx);
        v2 += q[i] * q[i];
    }
    return v2 / params[0] / params[0];
}

// update_q_Newton, gevolution.hpp:709-776
template <bool POW2, typename F>
__device__ __forceinline__ double update_q_newton(double q[3], const double rd[3], ViewS<F> psi,
                                                  ViewS<F> chi, int nfield, const double params[2],
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
        q[i] -= over_dx<POW2>(dtau * params[0] * g[i], dx, 1.0 / dx);
        v2 += q[i] * q[i];
    }
    return v2 / params[0] / params[0];
}

// update_pos, gevolution.hpp:810-871
template <typename F>
__device__ __forceinline__ void update_pos(double pos[3], const double q[3], const double rd[3],
                                           ViewS<F> phi, ViewS<F> chi, View0<F> B0,
                                           View1<F> B1, View2<F> B2, int nfield,
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

// ---- The FP32 form's gathers.  The CIC gradient and interpolation stencils
// over the float fields run in float (the weights are the position's
// fractional parts rounded once); the momentum and position updates
// themselves, e = sqrt(q^2 + a^2) and every factor involving q, stay FP64.
// Same terms as update_q / update_q_newton / update_pos above.

// trilinear gradient of a float field at fractional position (r0, r1, r2)
template <int LO>
__device__ __forceinline__ void grad_f32(FieldView<LO, float> f, float r0, float r1, float r2, float g[3]) {
    const float s0 = 1.f - r0, s1 = 1.f - r1, s2 = 1.f - r2;
    g[0] = s1 * s2 * (f.raw(1, 0, 0) - f.raw(0, 0, 0)) + r1 * s2 * (f.raw(1, 1, 0) - f.raw(0, 1, 0)) +
           s1 * r2 * (f.raw(1, 0, 1) - f.raw(0, 0, 1)) + r1 * r2 * (f.raw(1, 1, 1) - f.raw(0, 1, 1));
    g[1] = s0 * s2 * (f.raw(0, 1, 0) - f.raw(0, 0, 0)) + r0 * s2 * (f.raw(1, 1, 0) - f.raw(1, 0, 0)) +
           s0 * r2 * (f.raw(0, 1, 1) - f.raw(0, 0, 1)) + r0 * r2 * (f.raw(1, 1, 1) - f.raw(1, 0, 1));
    g[2] = s0 * s1 * (f.raw(0, 0, 1) - f.raw(0, 0, 0)) + r0 * s1 * (f.raw(1, 0, 1) - f.raw(1, 0, 0)) +
           s0 * r1 * (f.raw(0, 1, 1) - f.raw(0, 1, 0)) + r0 * r1 * (f.raw(1, 1, 1) - f.raw(1, 1, 0));
}

// trilinear interpolation of a float field
__device__ __forceinline__ float interp_f32(ViewS<float> f, float r0, float r1, float r2) {
    const float s0 = 1.f - r0, s1 = 1.f - r1, s2 = 1.f - r2;
    return s2 * (s1 * (s0 * f.raw(0, 0, 0) + r0 * f.raw(1, 0, 0)) + r1 * (s0 * f.raw(0, 1, 0) + r0 * f.raw(1, 1, 0))) +
           r2 * (s1 * (s0 * f.raw(0, 0, 1) + r0 * f.raw(1, 0, 1)) + r1 * (s0 * f.raw(0, 1, 1) + r0 * f.raw(1, 1, 1)));
}

// update_q, gevolution.hpp:570-678, FP32 form
template <bool POW2>
__device__ __forceinline__ double update_q(double q[3], const double rd[3], ViewS<float> phi, ViewS<float> chi,
                                           View0<float> B0, View1<float> B1, View2<float> B2, int nfield,
                                           const double params[2], double dtau, double dx) {
    const float r0 = (float)rd[0], r1 = (float)rd[1], r2 = (float)rd[2];
    const float s0 = 1.f - r0, s1 = 1.f - r1, s2 = 1.f - r2;
    double v2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    double e2 = v2 + params[0] * params[0];
    // the force in float: (v^2 + e^2) / e^2 and the B term's 1 / (N a^2 e) are
    // evaluated in FP64 and rounded once
    float gradphi[3];
    grad_f32(phi, r0, r1, r2, gradphi);
    const float fac = (float)((v2 + e2) / e2);
    gradphi[0] *= fac;
    gradphi[1] *= fac;
    gradphi[2] *= fac;
    GEV_LDS_PHASE();
    if (nfield >= 2) {
        float g[3];
        grad_f32(chi, r0, r1, r2, g);
        gradphi[0] -= g[0];
        gradphi[1] -= g[1];
        gradphi[2] -= g[2];
    }
    e2 = sqrt(e2);
    GEV_LDS_PHASE();
    if (nfield >= 3) {
        // pgradB_i = sum_j A_ij q_j with the A_ij float stencils of B
        const float a01 = s2 * (B1.raw(1, 0, 0) - B1.raw(0, 0, 0)) + r2 * (B1.raw(1, 0, 1) - B1.raw(0, 0, 1));
        const float a02 = s1 * (B2.raw(1, 0, 0) - B2.raw(0, 0, 0)) + r1 * (B2.raw(1, 1, 0) - B2.raw(0, 1, 0));
        const float c0 = r0 - 1.f, c1 = 1.f - 2.f * r0;
        const float a00 = s1 * s2 * (c0 * B0.raw(-1, 0, 0) + c1 * B0.raw(0, 0, 0) + r0 * B0.raw(1, 0, 0)) +
                          r1 * s2 * (c0 * B0.raw(-1, 1, 0) + c1 * B0.raw(0, 1, 0) + r0 * B0.raw(1, 1, 0)) +
                          s1 * r2 * (c0 * B0.raw(-1, 0, 1) + c1 * B0.raw(0, 0, 1) + r0 * B0.raw(1, 0, 1)) +
                          r1 * r2 * (c0 * B0.raw(-1, 1, 1) + c1 * B0.raw(0, 1, 1) + r0 * B0.raw(1, 1, 1));
        GEV_LDS_PHASE();
        const float a12 = s0 * (B2.raw(0, 1, 0) - B2.raw(0, 0, 0)) + r0 * (B2.raw(1, 1, 0) - B2.raw(1, 0, 0));
        const float a10 = s2 * (B0.raw(0, 1, 0) - B0.raw(0, 0, 0)) + r2 * (B0.raw(0, 1, 1) - B0.raw(0, 0, 1));
        const float d0 = r1 - 1.f, d1 = 1.f - 2.f * r1;
        const float a11 = s0 * s2 * (d0 * B1.raw(0, -1, 0) + d1 * B1.raw(0, 0, 0) + r1 * B1.raw(0, 1, 0)) +
                          r0 * s2 * (d0 * B1.raw(1, -1, 0) + d1 * B1.raw(1, 0, 0) + r1 * B1.raw(1, 1, 0)) +
                          s0 * r2 * (d0 * B1.raw(0, -1, 1) + d1 * B1.raw(0, 0, 1) + r1 * B1.raw(0, 1, 1)) +
                          r0 * r2 * (d0 * B1.raw(1, -1, 1) + d1 * B1.raw(1, 0, 1) + r1 * B1.raw(1, 1, 1));
        GEV_LDS_PHASE();
        const float a20 = s1 * (B0.raw(0, 0, 1) - B0.raw(0, 0, 0)) + r1 * (B0.raw(0, 1, 1) - B0.raw(0, 1, 0));
        const float a21 = s0 * (B1.raw(0, 0, 1) - B1.raw(0, 0, 0)) + r0 * (B1.raw(1, 0, 1) - B1.raw(1, 0, 0));
        const float e0 = r2 - 1.f, e1 = 1.f - 2.f * r2;
        const float a22 = s0 * s1 * (e0 * B2.raw(0, 0, -1) + e1 * B2.raw(0, 0, 0) + r2 * B2.raw(0, 0, 1)) +
                          r0 * s1 * (e0 * B2.raw(1, 0, -1) + e1 * B2.raw(1, 0, 0) + r2 * B2.raw(1, 0, 1)) +
                          s0 * r1 * (e0 * B2.raw(0, 1, -1) + e1 * B2.raw(0, 1, 0) + r2 * B2.raw(0, 1, 1)) +
                          r0 * r1 * (e0 * B2.raw(1, 1, -1) + e1 * B2.raw(1, 1, 0) + r2 * B2.raw(1, 1, 1));
        GEV_LDS_PHASE();
        const float qf0 = (float)q[0], qf1 = (float)q[1], qf2 = (float)q[2];
        const float inv = (float)(1. / params[1] / e2);
        gradphi[0] += (a01 * qf1 + a02 * qf2 + a00 * qf0) * inv;
        gradphi[1] += (a12 * qf2 + a10 * qf0 + a11 * qf1) * inv;
        gradphi[2] += (a20 * qf0 + a21 * qf1 + a22 * qf2) * inv;
    }
    v2 = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        q[i] -= over_dx<POW2>(dtau * e2 * (double)gradphi[i], dx, 1.0 / dx);
        v2 += q[i] * q[i];
    }
    return v2 / params[0] / params[0];
}

// update_q_Newton, gevolution.hpp:709-776, FP32 form
template <bool POW2>
__device__ __forceinline__ double update_q_newton(double q[3], const double rd[3], ViewS<float> psi,
                                                  ViewS<float> chi, int nfield, const double params[2],
                                                  double dtau, double dx) {
    const float r0 = (float)rd[0], r1 = (float)rd[1], r2 = (float)rd[2];
    float g[3], c[3] = {0.f, 0.f, 0.f};
    grad_f32(psi, r0, r1, r2, g);
    if (nfield >= 2) grad_f32(chi, r0, r1, r2, c);
    double v2 = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        q[i] -= over_dx<POW2>(dtau * params[0] * (double)(g[i] - c[i]), dx, 1.0 / dx);
        v2 += q[i] * q[i];
    }
    return v2 / params[0] / params[0];
}

// update_pos, gevolution.hpp:810-871, FP32 form
__device__ __forceinline__ void update_pos(double pos[3], const double q[3], const double rd[3], ViewS<float> phi,
                                           ViewS<float> chi, View0<float> B0, View1<float> B1, View2<float> B2,
                                           int nfield, const double params[2], double dtau) {
    const float r0 = (float)rd[0], r1 = (float)rd[1], r2 = (float)rd[2];
    const float s0 = 1.f - r0, s1 = 1.f - r1, s2 = 1.f - r2;
    double v2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    const double e2 = v2 + params[0] * params[0];
    const double ph = nfield >= 1 ? (double)interp_f32(phi, r0, r1, r2) : 0.;
    const double ch = nfield >= 2 ? (double)interp_f32(chi, r0, r1, r2) : 0.;
    v2 = (1. + (3. - v2 / e2) * ph - ch) / sqrt(e2);
    double v[3] = {q[0] * v2, q[1] * v2, q[2] * v2};
    if (nfield >= 3) {
        const float b0 = s2 * (s1 * B0.raw(0, 0, 0) + r1 * B0.raw(0, 1, 0)) + r2 * (s1 * B0.raw(0, 0, 1) + r1 * B0.raw(0, 1, 1));
        const float b1 = s2 * (s0 * B1.raw(0, 0, 0) + r0 * B1.raw(1, 0, 0)) + r2 * (s0 * B1.raw(0, 0, 1) + r0 * B1.raw(1, 0, 1));
        const float b2 = s1 * (s0 * B2.raw(0, 0, 0) + r0 * B2.raw(1, 0, 0)) + r1 * (s0 * B2.raw(0, 1, 0) + r0 * B2.raw(1, 1, 0));
        pos[0] += dtau * (v[0] + b0 / params[1]);
        pos[1] += dtau * (v[1] + b1 / params[1]);
        pos[2] += dtau * (v[2] + b2 / params[1]);
    } else {
#pragma unroll
        for (int l = 0; l < 3; l++) pos[l] += dtau * v[l];
    }
}

// periodic wrap of the re-filing step; a coordinate whose cell would round
// to N is set to 0 so that 0 <= floor(p/dx) < N for every filed particle
template <bool POW2> __device__ __forceinline__ double wrap_pos(double p, double dx, int N) {
    if (p >= 1.0) p -= 1.0;
    if (p < 0.0) p += 1.0;
    // beyond one box length: reduce modulo 1; non-finite: file at 0
    if (!(p >= 0.0 && p < 1.0)) p = p - floor(p);
    if (!(p >= 0.0) || cell1<POW2>(p, dx, N) >= N) p = 0.0;
    return p;
}

struct StepArgs {
    int nfk, nfd;
    double dtk, dtd;
    double pk[2], pd[2];
};

// Per tile: stage the fields, then for every particle filed in the tile:
// locate, optional kick, optional drift (+ wrap + count into the new cell).
// QM / PM select the callbacks (0 = GR, 1 = Newton).  The filing
// (start / perm) must describe the current positions.
// The FP32 form's particle pass asks for 5 waves per SIMD (96 VGPRs, no
// spill; 5.2 against 5.5 ms at 512^3)
#ifndef GEV_KICK_NT
#define GEV_KICK_NT 1
#endif
// whole tiles away from the box faces compute their staging indices as one
// offset from the region's corner (no clamp, no wrap): 6.46 -> 5.90 ms per
// FP64 kick-drift at 512^3 early, 4.69 -> 4.49 ms in the FP32 form
#ifndef GEV_KICK_INNER
#define GEV_KICK_INNER 1
#endif
#ifndef GEV_KICK_WAVES_F32
#define GEV_KICK_WAVES_F32 5
#endif
// GEV_KICK_WAVES > 0 asks the compiler for that many waves per SIMD
#ifndef GEV_KICK_WAVES
#define GEV_KICK_WAVES 0
#endif
#if GEV_KICK_WAVES > 0
#define GEV_KICK_OCC __attribute__((amdgpu_waves_per_eu(GEV_KICK_WAVES)))
#else
#define GEV_KICK_OCC
#endif

// BATCH: the launch covers the boxes of a batch; INPLACE: a compact box, the
// drift writes each particle back at its own index (each its own
// instantiation, so the plain one-box kernel keeps its register budget).
template <bool KICK, int QM, bool DRIFT, int PM, bool POW2, typename F, bool BATCH = false, bool INPLACE = false>
__global__ __launch_bounds__(TPB, (sizeof(F) == 4 && !BATCH ? GEV_KICK_WAVES_F32 : 1)) GEV_KICK_OCC void k_update_tiled(
    double *__restrict__ pos, double *__restrict__ q, int64_t cap, int N, int ntx, int nty, int ntz,
    const uint32_t *__restrict__ start, const uint32_t *__restrict__ perm, const F *__restrict__ phi,
    const F *__restrict__ chi, const F *__restrict__ Bi, StepArgs sa,
    unsigned long long *__restrict__ vmax, uint32_t *__restrict__ count, uint32_t *__restrict__ cell,
    double *__restrict__ pos_out, double *__restrict__ q_out,
    const int64_t *__restrict__ id_in, int64_t *__restrict__ id_out, gev::Batch bt) {
    if constexpr (BATCH) {
        // box blockIdx.y of a batch: its arrays, the live set in and the other out, its kick and drift
        const int b = blockIdx.y;
        if (!gev::box_on(bt, b)) return;
        const gev::BoxCycle &bc = bt.cyc[b];
        if (bc.parity) {
            double *t = pos;
            pos = pos_out;
            pos_out = t;
            t = q;
            q = q_out;
            q_out = t;
            const int64_t *u = id_in;
            id_in = id_out;
            id_out = const_cast<int64_t *>(u);
        }
        pos = gev::box_ptr(pos, bt.stride, b);
        q = gev::box_ptr(q, bt.stride, b);
        pos_out = gev::box_ptr(pos_out, bt.stride, b);
        q_out = gev::box_ptr(q_out, bt.stride, b);
        id_in = gev::box_ptr(id_in, bt.stride, b);
        id_out = gev::box_ptr(id_out, bt.stride, b);
        start = gev::box_ptr(start, bt.stride, b);
        perm = gev::box_ptr(perm, bt.stride, b);
        phi = gev::box_ptr(phi, bt.stride, b);
        chi = gev::box_ptr(chi, bt.stride, b);
        Bi = gev::box_ptr(Bi, bt.stride, b);
        vmax = gev::box_ptr(vmax, bt.stride, b);
        count = gev::box_ptr(count, bt.stride, b);
        cell = gev::box_ptr(cell, bt.stride, b);
        sa.nfk = bc.nfk;
        sa.nfd = bc.nfd;
        sa.dtk = bc.dtk;
        sa.dtd = bc.dtd;
        sa.pk[0] = bc.pk0;
        sa.pk[1] = bc.pk1;
        sa.pd[0] = bc.pd0;
        sa.pd[1] = bc.pd1;
    }
    __shared__ F s_phi[Region<-1>::SIZE], s_chi[Region<-1>::SIZE];
    __shared__ F s_b0[Region<0>::SIZE], s_b1[Region<1>::SIZE], s_b2[Region<2>::SIZE];
    __shared__ unsigned long long sh[TPB / 64];

    const int64_t nblk = (int64_t)ntx * nty * ntz;
    int64_t tile = blockIdx.x;
    if (nblk % 8 == 0) tile = (int64_t)(blockIdx.x % 8) * (nblk / 8) + blockIdx.x / 8;  // XCD-aware
    const int x0 = (int)(tile % ntx) * KTX, y0 = (int)((tile / ntx) % nty) * KTY;
    const int z0 = (int)(tile / ((int64_t)ntx * nty)) * KTZ;
    const int nx = min(KTX, N - x0), ny = min(KTY, N - y0), nz = min(KTZ, N - z0);
    const int64_t n3v = (int64_t)N * N * N;
    const double dx = 1.0 / (double)N;

    const bool need_phi = KICK || (DRIFT && PM == 0 && sa.nfd >= 1);
    const bool need_chi = (KICK && sa.nfk >= 2) || (DRIFT && PM == 0 && sa.nfd >= 2);
    const bool need_B = (KICK && QM == 0 && sa.nfk >= 3) || (DRIFT && PM == 0 && sa.nfd >= 3);
    // rows of the tile: cells (x0 .. x0+nx-1, y, z) are one run of the
    // permutation.  Lane r of every wavefront reads row r's extent (one
    // vector load latency) and the extents are made wave-uniform, so every
    // wave finds its particles without waiting for the field staging.
    static_assert(KTY * KTZ <= 64, "one lane per row");
    uint32_t rb = 0, re = 0;
    {
        const int r = threadIdx.x & 63, ly = r % KTY, lz = r / KTY;
        if (r < KTY * KTZ && ly < ny && lz < nz) {
            const int64_t g0 = L(N, x0, y0 + ly, z0 + lz);
            rb = start[g0];
            re = start[g0 + nx];
        }
    }
    uint32_t rowbeg[KTY * KTZ], rowoff[KTY * KTZ + 1];
    rowoff[0] = 0;
#pragma unroll
    for (int r = 0; r < KTY * KTZ; r++) {
        const uint32_t b = __builtin_amdgcn_readlane(rb, r);
        rowbeg[r] = b;
        rowoff[r + 1] = rowoff[r] + (__builtin_amdgcn_readlane(re, r) - b);
    }
    const uint32_t total = rowoff[KTY * KTZ];
    // cell-sorted index of the tile's j-th particle
    auto sorted_index = [&](uint32_t jj) {
        uint32_t s = rowbeg[0] + jj;
#pragma unroll
        for (int r = 1; r < KTY * KTZ; r++)
            if (rowoff[r] <= jj) s = rowbeg[r] + (jj - rowoff[r]);
        return s;
    };

    // Loads in the order they are needed: this thread's first two
    // permutation entries, the field staging (no dependence), then the two
    // particles.  Loads complete in order, so the particle loads wait only
    // for the permutation entries and the LDS writes only for the staging.
    uint32_t j = threadIdx.x, js = 0, p = 0, jsn = 0, pn = 0;
    if (j < total) {
        js = sorted_index(j);
        p = perm[js];
    }
    if (j + TPB < total) {
        jsn = sorted_index(j + TPB);
        pn = perm[jsn];
    }
    // field staging: element k of this thread's share of field f's region.
    // A tile cut short by the box edge (nx < KTX, e.g. every tile of N < 32)
    // only reads slots up to its own extent plus the upper halo, so the
    // region is clamped there: the coordinate stays within [x0 - 1, x0 + nx]
    // (x0 + nx <= N), which one periodic step maps into [0, N).  The slots
    // beyond hold copies of the last staged cell and are never read.
    // A whole tile away from the box faces (no coordinate of its regions
    // wraps, nothing to clamp) takes its cells as one offset from the
    // region's corner.
    const bool inner = nx == KTX && ny == KTY && nz == KTZ && x0 > 0 && y0 > 0 && z0 > 0 && x0 + KTX < N &&
                       y0 + KTY < N && z0 + KTZ < N;
    auto cell_of = [&](auto reg, int i) {
        using R = decltype(reg);
        if (GEV_KICK_INNER && inner) {
            const uint32_t corner = (uint32_t)L(N, x0 - R::OX, y0 - R::OY, z0 - R::OZ);  // N^3 < 2^32
            return (int64_t)(corner + (uint32_t)(i % R::SX) +
                             (uint32_t)N * ((uint32_t)((i / R::SX) % R::SY) + (uint32_t)N * (uint32_t)(i / (R::SX * R::SY))));
        }
        const int rx = min(i % R::SX, nx + R::OX), ry = min((i / R::SX) % R::SY, ny + R::OY);
        const int rz = min(i / (R::SX * R::SY), nz + R::OZ);
        return L(N, wrapi(x0 - R::OX + rx, N), wrapi(y0 - R::OY + ry, N), wrapi(z0 - R::OZ + rz, N));
    };
    constexpr int PS = (Region<-1>::SIZE + TPB - 1) / TPB, P0 = (Region<0>::SIZE + TPB - 1) / TPB;
    constexpr int P1 = (Region<1>::SIZE + TPB - 1) / TPB, P2 = (Region<2>::SIZE + TPB - 1) / TPB;
    F vp[PS], vc[PS], v0[P0], v1[P1], v2[P2];
#pragma unroll
    for (int k = 0; k < PS; k++) {
        const int i = threadIdx.x + k * TPB;
        if (i < Region<-1>::SIZE) {
            const int64_t g = cell_of(Region<-1>{}, i);
            if (need_phi) vp[k] = phi[g];
            if (need_chi) vc[k] = chi[g];
        }
    }
    if (need_B) {
#pragma unroll
        for (int k = 0; k < P0; k++) {
            const int i = threadIdx.x + k * TPB;
            if (i < Region<0>::SIZE) v0[k] = Bi[cell_of(Region<0>{}, i)];
        }
#pragma unroll
        for (int k = 0; k < P1; k++) {
            const int i = threadIdx.x + k * TPB;
            if (i < Region<1>::SIZE) v1[k] = Bi[n3v + cell_of(Region<1>{}, i)];
        }
#pragma unroll
        for (int k = 0; k < P2; k++) {
            const int i = threadIdx.x + k * TPB;
            if (i < Region<2>::SIZE) v2[k] = Bi[2 * n3v + cell_of(Region<2>{}, i)];
        }
    }
    double x[3] = {0., 0., 0.}, qq[3] = {0., 0., 0.}, xn[3] = {0., 0., 0.}, qn[3] = {0., 0., 0.};
    int64_t pid = 0, pidn = 0;
    auto load_particle = [&](uint32_t pp, double *xo, double *qo, int64_t &io) {
        xo[0] = pos[pp];
        xo[1] = pos[cap + pp];
        xo[2] = pos[2 * cap + pp];
        qo[0] = q[pp];
        qo[1] = q[cap + pp];
        qo[2] = q[2 * cap + pp];
        if (DRIFT) io = id_in[pp];
    };
    if (j < total) load_particle(p, x, qq, pid);
    if (j + TPB < total) load_particle(pn, xn, qn, pidn);
#pragma unroll
    for (int k = 0; k < PS; k++) {
        const int i = threadIdx.x + k * TPB;
        if (i < Region<-1>::SIZE) {
            if (need_phi) s_phi[i] = vp[k];
            if (need_chi) s_chi[i] = vc[k];
        }
    }
    if (need_B) {
#pragma unroll
        for (int k = 0; k < P0; k++)
            if (threadIdx.x + k * TPB < Region<0>::SIZE) s_b0[threadIdx.x + k * TPB] = v0[k];
#pragma unroll
        for (int k = 0; k < P1; k++)
            if (threadIdx.x + k * TPB < Region<1>::SIZE) s_b1[threadIdx.x + k * TPB] = v1[k];
#pragma unroll
        for (int k = 0; k < P2; k++)
            if (threadIdx.x + k * TPB < Region<2>::SIZE) s_b2[threadIdx.x + k * TPB] = v2[k];
    }
    __syncthreads();

    double rmax = 0.;
    while (j < total) {
        // filed site and ref_dist = modf(pos/dx) (Particles_gevolution.hpp:170 idiom)
        double rd[3];
        int ci[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double t = over_dx<POW2>(x[i], dx, (double)N);
            const double fl = floor(t);
            ci[i] = (int)fl;
            rd[i] = t - fl;
        }
        const ViewS<F> fphi = view<-1>(s_phi, ci, x0, y0, z0), fchi = view<-1>(s_chi, ci, x0, y0, z0);
        const View0<F> fb0 = view<0>(s_b0, ci, x0, y0, z0);
        const View1<F> fb1 = view<1>(s_b1, ci, x0, y0, z0);
        const View2<F> fb2 = view<2>(s_b2, ci, x0, y0, z0);
        if (KICK) {
            double r;
            if (QM == 0)
                r = update_q<POW2>(qq, rd, fphi, fchi, fb0, fb1, fb2, sa.nfk, sa.pk, sa.dtk, dx);
            else
                r = update_q_newton<POW2>(qq, rd, fphi, fchi, sa.nfk, sa.pk, sa.dtk, dx);
            rmax = r > rmax ? r : rmax;
            if (!DRIFT) {
                q[p] = qq[0];
                q[cap + p] = qq[1];
                q[2 * cap + p] = qq[2];
            }
        }
        GEV_LDS_PHASE();
        if (DRIFT) {
            if (PM == 0) {
                update_pos(x, qq, rd, fphi, fchi, fb0, fb1, fb2, sa.nfd, sa.pd, sa.dtd);
            } else {
#pragma unroll
                for (int l = 0; l < 3; l++) x[l] += sa.dtd * qq[l] / sa.pd[0];  // gevolution.hpp:902
            }
#pragma unroll
            for (int l = 0; l < 3; l++) x[l] = wrap_pos<POW2>(x[l], dx, N);
            // stored at the sorted index in the other set (the storage order
            // becomes the cell order), or in place in a compact box
            const uint32_t w = INPLACE ? p : js;
            double *const po = INPLACE ? pos : pos_out;
            double *const qo = INPLACE ? q : q_out;
            // (the FP64 form streams a large box's particles past the caches:
            // 7.86 against 8.19 ms per kick-drift at 512^3, z = 0; no gain in
            // the FP32 form)
            const bool nt = GEV_KICK_NT && sizeof(F) == 8 && N >= gev::STREAM_MIN_N;
            gev::store_stream(po + w, x[0], nt);
            gev::store_stream(po + cap + w, x[1], nt);
            gev::store_stream(po + 2 * cap + w, x[2], nt);
            gev::store_stream(qo + w, qq[0], nt);
            gev::store_stream(qo + cap + w, qq[1], nt);
            gev::store_stream(qo + 2 * cap + w, qq[2], nt);
            if (!INPLACE) gev::store_stream(id_out + js, pid, nt);
            const uint32_t c = (uint32_t)L(N, cell1<POW2>(x[0], dx, N), cell1<POW2>(x[1], dx, N),
                                           cell1<POW2>(x[2], dx, N));
            cell[w] = c;
            // no return value: the wave does not wait for it (one atomic per
            // run of equal cells, as in the scatter, was slower: 6.49 against
            // 6.02 ms early at 512^3, faster only near z = 0)
            atomicAdd(count + c, 1u);
        }
        // the next particle was fetched one step ahead; fetch the one after
        j += TPB;
        js = jsn;
        p = pn;
        pid = pidn;
#pragma unroll
        for (int l = 0; l < 3; l++) {
            x[l] = xn[l];
            qq[l] = qn[l];
        }
        if (j + TPB < total) {
            jsn = sorted_index(j + TPB);
            pn = perm[jsn];
            load_particle(pn, xn, qn, pidn);
        }
    }
    if (KICK) {
        // max of the callback returns (>= 0, so the bit patterns order as integers)
        unsigned long long b = __double_as_longlong(rmax);
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(b, off);
            b = o > b ? o : b;
        }
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = b;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = sh[0];
            for (int w = 1; w < TPB / 64; w++) m = sh[w] > m ? sh[w] : m;
            atomicMax(vmax, m);
        }
    }
}

template <bool KICK, int QM, bool DRIFT, int PM, typename F>
void launch_update_f(gev_ctx *ctx, const StepArgs &sa, const gev::Batch &bt) {
    const int N = ctx->N;
    const int ntx = (N + KTX - 1) / KTX, nty = (N + KTY - 1) / KTY, ntz = (N + KTZ - 1) / KTZ;
    const int64_t nblk = (int64_t)ntx * nty * ntz;
    const unsigned nbox = bt.cyc ? (unsigned)ctx->nbox : 1u;
    const F *phi = gev::fld<F>(ctx->phi), *chi = gev::fld<F>(ctx->chi), *Bi = gev::fld<F>(ctx->Bi);
#define GEV_KICK_LAUNCH(P2, B, IP)                                                                             \
    hipLaunchKernelGGL((k_update_tiled<KICK, QM, DRIFT, PM, P2, F, B, IP>), dim3((unsigned)nblk, nbox), dim3(TPB),  \
                       0, ctx->stream, ctx->pos, ctx->q, ctx->cap, N, ntx, nty, ntz, ctx->start, ctx->perm, phi, chi, \
                       Bi, sa, ctx->red_u, ctx->count, ctx->cell, ctx->pos2, ctx->q2, ctx->id, ctx->id2, bt)
    const bool pow2 = (N & (N - 1)) == 0;
    if (bt.cyc) {
        // the batched cycle runs the kick and drift together only
        if constexpr (KICK && DRIFT) {
            if (pow2) GEV_KICK_LAUNCH(true, true, false);
            else GEV_KICK_LAUNCH(false, true, false);
        }
    } else if (ctx->compact && DRIFT) {
        if constexpr (DRIFT) {
            if (pow2) GEV_KICK_LAUNCH(true, false, true);
            else GEV_KICK_LAUNCH(false, false, true);
        }
    } else if (pow2) {
        GEV_KICK_LAUNCH(true, false, false);
    } else {
        GEV_KICK_LAUNCH(false, false, false);
    }
#undef GEV_KICK_LAUNCH
}

template <bool KICK, int QM, bool DRIFT, int PM>
void launch_update_t(gev_ctx *ctx, const StepArgs &sa, const gev::Batch &bt = gev::Batch{}) {
    if (ctx->fp32) launch_update_f<KICK, QM, DRIFT, PM, float>(ctx, sa, bt);
    else launch_update_f<KICK, QM, DRIFT, PM, double>(ctx, sa, bt);
}

}  // namespace

namespace gev {

// scan count[] into start[] and build perm[] from cell[] (count[] ends zero).
// Over a batch (bt.cyc set) every box is filed by the same launches; the
// list counter sort_n of each box was cleared before the particle stage.
static int finish_filing(gev_ctx *ctx, const Batch &bt = Batch{}) {
    const int64_t n3v = n3(ctx);
    const int64_t nblk = (n3v + SCAN_BLOCK - 1) / SCAN_BLOCK;
    const unsigned nb = bt.cyc ? (unsigned)ctx->nbox : 1u;
    {
        ScopedTimer t(ctx, "filing_scan");
        hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nblk, nb), dim3(TPB), 0, ctx->stream, ctx->count,
                           n3v, ctx->scan_part, bt);
        hipLaunchKernelGGL(k_scan_partials, dim3(1, nb), dim3(TPB), 0, ctx->stream, ctx->scan_part, nblk, bt);
        hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nblk, nb), dim3(TPB), 0, ctx->stream, ctx->count,
                           n3v, ctx->scan_part, ctx->start, bt);
    }
    if (ctx->n > 0) {
        ScopedTimer t(ctx, "filing_scatter");
        hipLaunchKernelGGL(k_scatter, dim3(grid_for(ctx->n, TPB), nb), dim3(TPB), 0, ctx->stream,
                           ctx->cell, ctx->start, ctx->count, ctx->n, ctx->perm, bt);
    }
    {
        ScopedTimer t(ctx, "filing_sort");
        // sort_n[1]: the number of listed cells
        if (!bt.cyc) {
            int rc = check_hip(ctx, hipMemsetAsync(ctx->sort_n, 0, 2 * sizeof(uint32_t), ctx->stream),
                               "sort list clear");
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_sort_cells, dim3(grid_for(n3v, TPB), nb), dim3(TPB), 0, ctx->stream,
                           ctx->start, ctx->id, n3v, ctx->perm, reinterpret_cast<uint8_t *>(ctx->sort_segs), ctx->id2,
                           bt);
    }
    {
        ScopedTimer t(ctx, "filing_sort_segments");
        // SORT_SEG_BATCH segments per workgroup: the late, clustered cycles flag
        // most segments and need the parallelism; early ones flag almost none
        const unsigned nsw = (unsigned)((grid_for(n3v, TPB) + SORT_SEG_BATCH - 1) / SORT_SEG_BATCH);
        hipLaunchKernelGGL(k_sort_segments, dim3(nsw, nb), dim3(TPB), 0, ctx->stream, ctx->start, ctx->id, n3v,
                           ctx->perm, reinterpret_cast<const uint8_t *>(ctx->sort_segs),
                           (uint32_t)grid_for(n3v, TPB), ctx->sort_big, ctx->sort_n + 1, ctx->id2, bt);
    }
    {
        ScopedTimer t(ctx, "filing_sort_dense_cells");
        hipLaunchKernelGGL(k_sort_big, dim3(1024, nb), dim3(TPB), 0, ctx->stream, ctx->start, ctx->id, ctx->perm,
                           ctx->sort_big, ctx->sort_n + 1, ctx->id2, bt);
    }
    const int rc = check_hip(ctx, hipGetLastError(), "filing launch");
    if (!rc) ctx->count_clean = true;  // the scatter counted every cell back to zero
    return rc;
}

int launch_refile(gev_ctx *ctx) {
    int rc = check_hip(ctx, hipMemsetAsync(ctx->count, 0, sizeof(uint32_t) * n3(ctx), ctx->stream),
                       "count clear");
    if (rc) return rc;
    ctx->count_clean = false;
    if (ctx->n > 0) {
        ScopedTimer t(ctx, "filing_count");
        hipLaunchKernelGGL(k_count_cells, dim3(grid_for(ctx->n, TPB)), dim3(TPB), 0, ctx->stream,
                           ctx->pos, ctx->cap, ctx->n, ctx->N, ctx->count, ctx->cell);
    }
    rc = check_hip(ctx, hipGetLastError(), "count launch");
    if (rc) return rc;
    return finish_filing(ctx);
}

// The particle stage over the boxes of a batch: kick + drift of every box
// (its live set in, the other set out, its coefficients from its BoxCycle),
// then the filing.  No clearing here: the batch clears vmax and the sort
// list counter of every box before this stage, and every completed filing
// leaves the counts at zero.
int launch_particles_batch(gev_ctx *ctx, bool gr, const Batch &bt) {
    StepArgs sa{};
    if (ctx->n > 0) {
        ScopedTimer t(ctx, "kick_drift");
        if (gr) launch_update_t<true, 0, true, 0>(ctx, sa, bt);
        else launch_update_t<true, 1, true, 1>(ctx, sa, bt);
    }
    int rc = check_hip(ctx, hipGetLastError(), "particle update launch");
    return rc ? rc : finish_filing(ctx, bt);
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
    if (drift && !ctx->count_clean) {
        // a filing that did not complete left counts behind
        rc = check_hip(ctx, hipMemsetAsync(ctx->count, 0, sizeof(uint32_t) * n3(ctx), ctx->stream),
                       "count clear");
        if (rc) return rc;
        ctx->count_clean = true;
    }
    if (drift) ctx->count_clean = false;
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
    if (drift && ctx->n > 0 && !ctx->compact) {
        // the drift wrote the particles in cell order into the second set
        std::swap(ctx->pos, ctx->pos2);
        std::swap(ctx->q, ctx->q2);
        std::swap(ctx->id, ctx->id2);
    }
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
