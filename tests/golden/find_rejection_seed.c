/* Finds seeds for which the initial-condition realisation on an N^3 grid hits
 * a zero PRNG output at an r1 position, i.e. exercises the redraw loop of
 * ic_basic.hpp:1149-1156.  Links the oracle for the PRNG (orc_prng_at).
 *
 * The nominal r1 positions follow from the stream layout of
 * generateDisplacementField (ic_basic.hpp:1089-1379): line (ky, kz) of
 * quadrant Q starts at ((Q S + kz') S + ky') S with S = 65536, each drawn mode
 * takes two outputs, and the kx = 0 modes of the kz > N/2 half reuse the
 * first two outputs of another quadrant's line.  Usage: find_rejection_seed N
 * ksphere first_seed count -> prints "seed kx ky kz" of the first hit.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include "gev_oracle.h"

static int drawn(int kx, int ky, int kz, int kmax, int ksphere)
{
    if (kx >= kmax || ky >= kmax || kz >= kmax) return 0;
    if (ksphere && (kx * kx + ky * ky + kz * kz) >= kmax * kmax) return 0;
    return 1;
}

int main(int argc, char **argv)
{
    if (argc < 5) return 2;
    const int N = atoi(argv[1]), ksphere = atoi(argv[2]);
    const uint64_t first = strtoull(argv[3], NULL, 10), count = strtoull(argv[4], NULL, 10);
    const int H = N / 2, kmax = H - 1;
    const uint64_t S = 65536;
    /* list the nominal r1 positions once */
    uint64_t *pos = malloc(sizeof(uint64_t) * (size_t)N * N * (H + 2));
    int *who = malloc(sizeof(int) * 3 * (size_t)N * N * (H + 2));
    int64_t np = 0;
    for (int kz = 0; kz < N; kz++)
        for (int ky = 0; ky < N; ky++) {
            const int zq = kz <= H ? kz : N - kz, yq = ky <= H ? ky : N - ky;
            const int zhi = kz > H, yhi = ky > H;
            uint64_t base, base0 = 0;
            int kx0;
            if (!yhi && !zhi) { base = ((uint64_t)kz * S + ky) * S; kx0 = (ky == 0 && kz == 0); }
            else if (yhi && !zhi) { base = ((S + kz) * S + (uint64_t)(N - ky)) * S; kx0 = 0; }
            else if (!yhi) { base = ((2 * S + (uint64_t)(N - kz)) * S + ky) * S; kx0 = 1;
                             base0 = ((uint64_t)(N - kz) * S + ky) * S; }
            else { base = ((3 * S + (uint64_t)(N - kz)) * S + (uint64_t)(N - ky)) * S; kx0 = 1;
                   base0 = ((2 * S + (uint64_t)(N - kz)) * S + (uint64_t)(N - ky)) * S; }
            for (int kx = kx0; kx <= H; kx++)
                if (drawn(kx, yq, zq, kmax, ksphere)) {
                    pos[np] = base + 2 * (uint64_t)(kx - kx0);
                    who[3 * np] = kx; who[3 * np + 1] = ky; who[3 * np + 2] = kz;
                    np++;
                }
            if (zhi && drawn(0, yq, zq, kmax, ksphere)) {
                pos[np] = base0;
                who[3 * np] = 0; who[3 * np + 1] = ky; who[3 * np + 2] = kz;
                np++;
            }
        }
    int64_t found = -1;
    int64_t fi = -1;
    #pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t s = 0; s < (int64_t)count; s++) {
        if (found >= 0) continue;
        const uint32_t seed = (uint32_t)(first + (uint64_t)s);
        for (int64_t i = 0; i < np; i++)
            if (orc_prng_at(seed, pos[i]) == 0) {
                #pragma omp critical
                {
                    if (found < 0 || s < found) { found = s; fi = i; }
                }
                break;
            }
    }
    if (found < 0) { printf("none\n"); return 1; }
    printf("%llu %d %d %d\n", (unsigned long long)(first + (uint64_t)found), who[3 * fi], who[3 * fi + 1],
           who[3 * fi + 2]);
    free(pos);
    free(who);
    return 0;
}
