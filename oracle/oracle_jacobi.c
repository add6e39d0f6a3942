/*
 * oracle_jacobi.c -- host replay of the 3-D 7-point Jacobi workload on the
 * whole periodic G^3 domain (no decomposition, no halos).  TEST
 * INFRASTRUCTURE ONLY.  The decomposed GPU run must equal it bit for bit.
 *
 * v = 0.25 * u + 0.125 * (((((u[x-1] + u[x+1]) + u[y-1]) + u[y+1]) + u[z-1]) + u[z+1])
 * with periodic neighbours; compiled with -ffp-contract=off.
 * Initial value of cell g = (x * G + y) * G + z: PCG32 (seed, stream 54)
 * words 2g and 2g+1 as (w0 << 32 | w1) >> 11, times 2^-53.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>

void oracle_pcg32_fill(uint32_t *out, size_t n, uint64_t seed, uint64_t seq);

void oracle_jacobi_init(double *u, int G, uint64_t seed) {
    const size_t n = (size_t)G * G * G;
    uint32_t *w = (uint32_t *)u; /* 2n words fill exactly n doubles */
    oracle_pcg32_fill(w, 2 * n, seed, 54);
    for (size_t i = 0; i < n; ++i) {
        const uint64_t hi = w[2 * i], lo = w[2 * i + 1];
        u[i] = (double)(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
    }
}

/* the step for the planes x0 <= x < x1 */
static void step_planes(const double *u, double *v, int G, int x0, int x1) {
    const size_t sx = (size_t)G * G, sy = (size_t)G;
    for (int x = x0; x < x1; ++x) {
        const int xm = (x + G - 1) % G, xp = (x + 1) % G;
        for (int y = 0; y < G; ++y) {
            const int ym = (y + G - 1) % G, yp = (y + 1) % G;
            for (int z = 0; z < G; ++z) {
                const int zm = (z + G - 1) % G, zp = (z + 1) % G;
                const size_t c = x * sx + y * sy + z;
                const double s = ((((u[xm * sx + y * sy + z] + u[xp * sx + y * sy + z]) + u[x * sx + ym * sy + z]) +
                                   u[x * sx + yp * sy + z]) +
                                  u[x * sx + y * sy + zm]) +
                                 u[x * sx + y * sy + zp];
                v[c] = 0.25 * u[c] + 0.125 * s;
            }
        }
    }
}

void oracle_jacobi_step(const double *u, double *v, int G) { step_planes(u, v, G, 0, G); }

/* `steps` steps; returns 0 if the result is in u, 1 if in v */
int oracle_jacobi_run(double *u, double *v, int G, int steps) {
    for (int t = 0; t < steps; ++t) {
        if (t & 1) oracle_jacobi_step(v, u, G);
        else oracle_jacobi_step(u, v, G);
    }
    return steps & 1;
}

/* The same run with the planes of every step split between `threads`
 * threads (at most 64, at most G), joined before the next step starts.  Every
 * cell is computed by the expression of oracle_jacobi_step from the same
 * inputs, so the result is bit-identical to oracle_jacobi_run; only the wall
 * time changes.  A slab whose thread cannot start is computed by the caller's
 * thread.  Returns as oracle_jacobi_run. */
struct mt_slab {
    const double *u;
    double *v;
    int G, x0, x1;
};

static void *mt_slab_step(void *p) {
    const struct mt_slab *a = (const struct mt_slab *)p;
    step_planes(a->u, a->v, a->G, a->x0, a->x1);
    return NULL;
}

int oracle_jacobi_run_mt(double *u, double *v, int G, int steps, int threads) {
    if (threads > G) threads = G;
    if (threads > 64) threads = 64;
    if (threads < 1) threads = 1;
    pthread_t tid[64];
    int live[64];
    struct mt_slab slab[64];
    for (int t = 0; t < steps; ++t) {
        const double *src = (t & 1) ? v : u;
        double *dst = (t & 1) ? u : v;
        for (int k = 0; k < threads; ++k) {
            slab[k] = (struct mt_slab){src, dst, G, (int)((long)G * k / threads), (int)((long)G * (k + 1) / threads)};
            live[k] = k > 0 && pthread_create(&tid[k], NULL, mt_slab_step, &slab[k]) == 0;
            if (k > 0 && !live[k]) mt_slab_step(&slab[k]);
        }
        mt_slab_step(&slab[0]);
        for (int k = 1; k < threads; ++k)
            if (live[k]) pthread_join(tid[k], NULL);
    }
    return steps & 1;
}
