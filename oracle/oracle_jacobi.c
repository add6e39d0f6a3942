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

void oracle_jacobi_step(const double *u, double *v, int G) {
    const size_t sx = (size_t)G * G, sy = (size_t)G;
    for (int x = 0; x < G; ++x) {
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

/* `steps` steps; returns 0 if the result is in u, 1 if in v */
int oracle_jacobi_run(double *u, double *v, int G, int steps) {
    for (int t = 0; t < steps; ++t) {
        if (t & 1) oracle_jacobi_step(v, u, G);
        else oracle_jacobi_step(u, v, G);
    }
    return steps & 1;
}
