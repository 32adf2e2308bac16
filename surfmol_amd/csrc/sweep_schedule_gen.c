/* sweep_schedule_gen.c -- writes sweep_schedules.h, the lane schedules of the
 * one-wave pair sweep of k_relax for systems of 1..43 atoms.
 *
 *   gcc -O2 -o /tmp/sweep_schedule_gen sweep_schedule_gen.c -lm
 *   /tmp/sweep_schedule_gen > sweep_schedules.h
 *
 * The sweep rotates 64 reaction accumulators by one lane per round: on round
 * t (1..T) lane v pairs its atom a(v) with the atom g(v+t mod 64) of LDS row
 * v+t, and accumulator "row x" collects that row's reactions.  Lanes and rows
 * below N are the atoms themselves (a(v) = v, g(x) = x); the other 64-N lanes
 * and rows hold copies of atoms, chosen here so that every unordered pair of
 * the N atoms is met within as few rounds T as possible (simulated annealing,
 * fixed seeds: the output is reproducible).  Where the search does not beat
 * the plain rotation (atoms 0..N-1, then copies of atoms 0..N/2-1, T = N/2)
 * the plain rotation is written.
 *
 * Per N the header holds T, a[64], g[64] (-1: none), keep[64] (bit t-1 of
 * lane v: the pair of round t is the one that sums this unordered pair; the
 * other meetings of a pair, self pairs and empty slots are not summed) and
 * src[64]: for atom i < N up to four further sums that make up its pair force
 * besides lane i's own sum, one byte each (lane << 1 | kind; kind 0: that
 * lane's own sum, 1: that lane's reaction accumulator after the sweep;
 * 0xff: none), in summation order.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NMAX 43
static int N, T;
static int a[64], g[64], cov[64][64], miss;
static unsigned long long rs;
static unsigned long long xr(void) {
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return rs;
}
static double ur(void) { return (xr() >> 11) * (1.0 / 9007199254740992.0); }
static void addp(int i, int j, int d) {
    if (i < 0 || j < 0 || i == j) return;
    if (i > j) { int t = i; i = j; j = t; }
    if (d > 0) { if (cov[i][j] == 0) miss--; cov[i][j]++; }
    else { cov[i][j]--; if (cov[i][j] == 0) miss++; }
}
static void lanep(int v, int d) { for (int t = 1; t <= T; t++) addp(a[v], g[(v + t) & 63], d); }
static void rowp(int x, int d) { for (int t = 1; t <= T; t++) addp(a[(x - t) & 63], g[x], d); }

static void plain(void) {
    const int nd = N / 2;
    T = nd;
    for (int v = 0; v < 64; v++) a[v] = v < N ? v : -1;
    for (int x = 0; x < 64; x++) g[x] = x < N ? x : (x < N + nd ? x - N : -1);
}

static int anneal(long iters, unsigned seed) {
    rs = 0x9E3779B97F4A7C15ull ^ ((unsigned long long)seed * 0x100000001B3ull) ^ ((unsigned long long)N << 32) ^ T;
    for (int k = 0; k < 8; k++) xr();
    for (int v = 0; v < 64; v++) {
        a[v] = v < N ? v : (int)(xr() % N);
        g[v] = v < N ? v : (int)(xr() % N);
    }
    memset(cov, 0, sizeof cov);
    miss = N * (N - 1) / 2;
    for (int v = 0; v < 64; v++) lanep(v, 1);
    double temp = 2.0;
    for (long it = 0; it < iters && miss > 0; it++) {
        const int which = (int)(xr() & 1), k = N + (int)(xr() % (64 - N)), nv = (int)(xr() % (N + 1)) - 1;
        const int old = which ? g[k] : a[k], m0 = miss;
        if (nv == old) continue;
        if (which) { rowp(k, -1); g[k] = nv; rowp(k, 1); }
        else { lanep(k, -1); a[k] = nv; lanep(k, 1); }
        const int dm = miss - m0;
        if (dm > 0 && ur() >= exp(-dm / temp)) {
            if (which) { rowp(k, -1); g[k] = old; rowp(k, 1); }
            else { lanep(k, -1); a[k] = old; lanep(k, 1); }
        }
        temp *= 0.9999995;
        if (temp < 0.05) temp = 0.05;
    }
    return miss;
}

/* copies that the coverage does not need are dropped (lanes and rows above
   N, highest first), then every atom must have at most 4 sums besides its
   own lane's: its rows and its extra lanes */
static int prune_and_check(void) {
    for (int k = 63; k >= N; k--) {
        if (g[k] >= 0) {
            const int old = g[k];
            rowp(k, -1); g[k] = -1; rowp(k, 1);
            if (miss) { rowp(k, -1); g[k] = old; rowp(k, 1); }
        }
        if (a[k] >= 0) {
            const int old = a[k];
            lanep(k, -1); a[k] = -1; lanep(k, 1);
            if (miss) { lanep(k, -1); a[k] = old; lanep(k, 1); }
        }
    }
    for (int i = 0; i < N; i++) {
        int k = 0;
        for (int x = 0; x < 64; x++) k += g[x] == i;
        for (int v = N; v < 64; v++) k += a[v] == i;
        if (k > 4) return 0;
    }
    return 1;
}

static void recount(void) {
    memset(cov, 0, sizeof cov);
    miss = N * (N - 1) / 2;
    for (int v = 0; v < 64; v++) lanep(v, 1);
}

int main(void) {
    int sT[NMAX + 1], sa[NMAX + 1][64], sg[NMAX + 1][64], ssrc[NMAX + 1][64];
    unsigned skeep[NMAX + 1][64];
    /* the fewest rounds the search finds */
    sT[0] = 0;
    for (N = 1; N <= NMAX; N++) {
        plain();
        int bT = T, ba[64], bg[64];
        memcpy(ba, a, sizeof a);
        memcpy(bg, g, sizeof g);
        if (N >= 4) {
            /* from one below the previous size's rounds (or the counting
               bound) up to the first T the search meets every pair in */
            int lb = (N * (N - 1) / 2 + 63) / 64;
            if (lb < 1) lb = 1;
            if (sT[N - 1] - 1 > lb) lb = sT[N - 1] - 1;
            for (T = lb; T < bT; T++) {
                int ok = 0;
                for (unsigned s = 0; s < 4 && !ok; s++) ok = anneal(10000000L, s) == 0 && prune_and_check();
                if (!ok) continue;
                bT = T;
                memcpy(ba, a, sizeof a);
                memcpy(bg, g, sizeof g);
                break;
            }
        }
        sT[N] = bT;
        memcpy(sa[N], ba, sizeof ba);
        memcpy(sg[N], bg, sizeof bg);
    }
    /* a schedule for more atoms serves fewer: its atoms >= N dropped */
    for (N = NMAX - 1; N >= 4; N--) {
        if (sT[N + 1] >= sT[N]) continue;
        T = sT[N + 1];
        for (int v = 0; v < 64; v++) {
            a[v] = sa[N + 1][v] < N ? sa[N + 1][v] : -1;
            g[v] = sg[N + 1][v] < N ? sg[N + 1][v] : -1;
        }
        recount();
        if (miss || !prune_and_check()) continue;
        sT[N] = T;
        memcpy(sa[N], a, sizeof a);
        memcpy(sg[N], g, sizeof g);
    }
    for (N = 1; N <= NMAX; N++) {
        T = sT[N];
        memcpy(a, sa[N], sizeof a);
        memcpy(g, sg[N], sizeof g);
        /* the round that sums each pair: the first meeting in (t, v) order */
        static int owner_t[64][64];
        for (int i = 0; i < 64; i++)
            for (int j = 0; j < 64; j++) owner_t[i][j] = -1;
        unsigned keep[64];
        memset(keep, 0, sizeof keep);
        for (int t = 1; t <= T; t++)
            for (int v = 0; v < 64; v++) {
                int i = a[v], j = g[(v + t) & 63];
                if (i < 0 || j < 0 || i == j) continue;
                if (i > j) { int q = i; i = j; j = q; }
                if (owner_t[i][j] >= 0) continue;
                owner_t[i][j] = t;
                keep[v] |= 1u << (t - 1);
            }
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++)
                if (owner_t[i][j] < 0) { fprintf(stderr, "N %d: pair %d %d not met\n", N, i, j); return 1; }
        /* sums making up atom i's pair force besides lane i's own sum: the
           reaction accumulators of its rows (row x ends in lane x - T), then
           the own sums of its extra lanes */
        int src[64];
        for (int i = 0; i < 64; i++) src[i] = -1;
        for (int i = 0; i < N; i++) {
            int k = 0, packed = 0;
            for (int x = 0; x < 64; x++)
                if (g[x] == i) {
                    if (k == 4) { fprintf(stderr, "N %d: atom %d has more than 4 sums\n", N, i); return 1; }
                    packed |= ((((x - T) & 63) << 1) | 1) << (8 * k);
                    k++;
                }
            for (int v = N; v < 64; v++)
                if (a[v] == i) {
                    if (k == 4) { fprintf(stderr, "N %d: atom %d has more than 4 sums\n", N, i); return 1; }
                    packed |= ((v << 1) | 0) << (8 * k);
                    k++;
                }
            for (; k < 4; k++) packed |= 0xff << (8 * k);
            src[i] = packed;
        }
        memcpy(skeep[N], keep, sizeof keep);
        memcpy(ssrc[N], src, sizeof src);
        fprintf(stderr, "N %2d: T %2d (plain %2d)\n", N, T, N / 2);
    }
    int tmax = 0;
    printf("// sweep_schedules.h -- generated by sweep_schedule_gen.c; do not edit.\n");
    printf("// Lane schedules of the one-wave pair sweep for 1..%d atoms (see the generator).\n", NMAX);
    printf("#pragma once\n\nnamespace smr {\n\nconstexpr int kSchedMaxAtoms = %d;\n\n", NMAX);
    printf("// most rounds of any schedule for 1..n atoms (LDS rows of a launch)\n");
    printf("constexpr signed char kSchedTmax[%d] = {0", NMAX + 1);
    for (int n = 1; n <= NMAX; n++) {
        if (sT[n] > tmax) tmax = sT[n];
        printf(", %d", tmax);
    }
    printf("};\n\n");
    printf("struct SweepSchedule {\n    int T;\n    signed char a[64], g[64];\n    unsigned keep[64];\n    unsigned src[64];\n};\n\n");
    printf("// the table itself, in the one translation unit that defines SMR_SWEEP_SCHEDULE_TABLE\n");
    printf("#ifdef SMR_SWEEP_SCHEDULE_TABLE\n");
    printf("__device__ const SweepSchedule kSched[%d] = {\n    {},\n", NMAX + 1);
    for (int n = 1; n <= NMAX; n++) {
        printf("    {%d,\n     {", sT[n]);
        for (int v = 0; v < 64; v++) printf("%s%d", v ? "," : "", sa[n][v]);
        printf("},\n     {");
        for (int v = 0; v < 64; v++) printf("%s%d", v ? "," : "", sg[n][v]);
        printf("},\n     {");
        for (int v = 0; v < 64; v++) printf("%s0x%xu", v ? "," : "", skeep[n][v]);
        printf("},\n     {");
        for (int v = 0; v < 64; v++) printf("%s0x%xu", v ? "," : "", (unsigned)ssrc[n][v]);
        printf("}},\n");
    }
    printf("};\n#endif\n\n}  // namespace smr\n");
    return 0;
}
