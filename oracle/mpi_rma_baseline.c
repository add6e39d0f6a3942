/*
 * mpi_rma_baseline.c -- the reference's put/get/p/wait semantics run as host
 * MPI-3 RMA over a shared-memory window: the CPU baseline reported beside the
 * GPU numbers.  TEST/BENCH INFRASTRUCTURE ONLY (never part of the product).
 *
 * Mapping (reference device API -> MPI-3 RMA, passive target, lock_all):
 *   nvshmem_T_p / nvshmem_T_put  (nvshmem_device.cu:438-512) -> MPI_Put + MPI_Win_flush
 *   nvshmem_T_get                (nvshmem_device.cu:514-540) -> MPI_Get + MPI_Win_flush
 *   nvshmem_quiet                (nvshmem_device.cu:431-436) -> MPI_Win_flush_all
 *   nvshmem_int_wait             (nvshmem_device.cu:542-562) -> poll local window memory, MPI_Win_sync between reads
 *   nvshmem_barrier_all          (nvshmem.cu:858-867)        -> MPI_Win_flush_all + MPI_Barrier
 *   nvshmem_barrier_all_offload  (nvshmem_kernels.cu:56-102) -> count on rank 0's window (MPI_Fetch_and_op),
 *                                  release by rank 0's puts of the episode number (barrier mode)
 * Every wait gives up after NVSHMEM_WAIT_TIMEOUT_MS (default 10000, the
 * library's budget): the rank reports the timeout and the job ends with a
 * non-zero status, where the reference's barrier would continue
 * (nvshmem_kernels.cu:65-68).
 *
 * Modes (2 ranks):
 *   pingpong ITERS                 p + wait round trips (PE0 p(flag, i); PE1
 *                                  wait(flag, i-1) and replies), then one 1 MiB
 *                                  f32 put of PCG32(seed 42) uniform[-1,1)
 *                                  values, checked bit-exact on PE1
 *   putget OP DTYPE MIN MAX REPS BUDGET_S
 *                                  OP put|get, DTYPE int|float|double, sizes in
 *                                  bytes MIN..MAX by powers of two; REPS
 *                                  operations per size or until BUDGET_S
 *                                  seconds; the last destination is checked
 *                                  bit-exact against the seeded source
 *   jacobi N STEPS CHECK (8 ranks)  the 2x2x2 periodic Jacobi application,
 *                                  N^3 float64 cells per rank; CHECK=1
 *                                  compares with oracle_jacobi_run
 *   batch NREQ SEED STEPS           the uneven-put workload, 64 logical PEs
 *                                  over the ranks (8 ranks x 8 PEs)
 *   barrier ROUNDS SKIP             ROUNDS barrier episodes on the window; rank
 *                                  SKIP (-1: none) leaves out the last one, so
 *                                  the others' waits must time out
 * Output: one JSON object per line on rank 0.
 */
#define _GNU_SOURCE
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int cmp_dbl(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

static double pct(double *v, size_t n, double p) {
    if (n == 0) return 0.0;
    size_t i = (size_t)(p * (double)(n - 1) + 0.5);
    return v[i];
}

static double g_wait_budget_s = 10.0; /* NVSHMEM_WAIT_TIMEOUT_MS */

/* a wait that outlived the budget: report it and end the whole job */
static void wait_timed_out(const char *what, long episode) {
    int rank = 0;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    fprintf(stderr, "[mpi_rma_baseline] [%d] %s %ld timed out after %.3f s\n", rank, what, episode, g_wait_budget_s);
    fflush(stderr);
    MPI_Abort(MPI_COMM_WORLD, 3);
}

static void wait_while_equal(MPI_Win win, volatile int *flag, int cmp) {
    /* nvshmem_int_wait: block while *flag == cmp, within the budget */
    const double t0 = now_s();
    for (unsigned long n = 0;; ++n) {
        MPI_Win_sync(win);
        if (oracle_int_wait_done(*flag, cmp)) return;
        if ((n & 255u) == 255u && now_s() - t0 > g_wait_budget_s) wait_timed_out("wait for a value other than", cmp);
    }
}

static void wait_until_at_least(MPI_Win win, volatile int *word, int value, const char *what, long episode) {
    const double t0 = now_s();
    for (unsigned long n = 0;; ++n) {
        MPI_Win_sync(win);
        if (*word >= value) return;
        if ((n & 255u) == 255u && now_s() - t0 > g_wait_budget_s) wait_timed_out(what, episode);
    }
}

/* Barrier episodes through the window, the offload barrier's protocol: every
 * rank adds one to rank 0's count word, rank 0 waits for all of them and
 * writes the episode number into every rank's release word.  Rank `skip`
 * leaves out the last episode (a PE that never reaches the barrier). */
static int do_barrier(int rank, int nranks, long rounds, int skip) {
    MPI_Win win;
    int *base = NULL;
    MPI_Win_allocate_shared(64, sizeof(int), MPI_INFO_NULL, MPI_COMM_WORLD, &base, &win);
    memset(base, 0, 64);
    MPI_Win_lock_all(MPI_MODE_NOCHECK, win);
    MPI_Win_sync(win);
    MPI_Barrier(MPI_COMM_WORLD);
    volatile int *count = base, *release = base + 8;
    const double t0 = now_s();
    for (long e = 1; e <= rounds; ++e) {
        if (rank == skip && e == rounds) break;
        const int one = 1;
        int old = 0;
        MPI_Fetch_and_op(&one, &old, MPI_INT, 0, 0, MPI_SUM, win);
        MPI_Win_flush(0, win);
        if (rank == 0) {
            wait_until_at_least(win, count, (int)(e * nranks), "barrier episode (arrivals)", e);
            const int v = (int)e;
            for (int r = 0; r < nranks; ++r) MPI_Put(&v, 1, MPI_INT, r, 8, 1, MPI_INT, win);
            MPI_Win_flush_all(win);
        }
        wait_until_at_least(win, release, (int)e, "barrier episode (release)", e);
    }
    const double per = (now_s() - t0) / (double)rounds;
    if (rank == 0) {
        printf("{\"mode\": \"barrier\", \"ranks\": %d, \"rounds\": %ld, \"episode_us\": %.2f}\n", nranks, rounds,
               per * 1e6);
        fflush(stdout);
    }
    MPI_Win_unlock_all(win);
    MPI_Win_free(&win);
    return 0;
}

static int do_pingpong(int rank, long iters) {
    const size_t payload_words = (1u << 20) / 4; /* 1 MiB of f32 */
    MPI_Win win;
    char *base = NULL;
    MPI_Aint bytes = 64 + (MPI_Aint)payload_words * 4;
    MPI_Win_allocate_shared(bytes, 1, MPI_INFO_NULL, MPI_COMM_WORLD, &base, &win);
    memset(base, 0, (size_t)bytes);
    MPI_Win_lock_all(MPI_MODE_NOCHECK, win);
    MPI_Win_sync(win);
    MPI_Barrier(MPI_COMM_WORLD);
    volatile int *flag = (volatile int *)base;
    const int peer = 1 - rank;
    double *samples = rank == 0 ? (double *)malloc(sizeof(double) * (size_t)iters) : NULL;

    for (long i = 1; i <= iters; ++i) {
        int v = (int)i;
        if (rank == 0) {
            double t0 = now_s();
            MPI_Put(&v, 1, MPI_INT, peer, 0, 1, MPI_INT, win);
            MPI_Win_flush(peer, win);
            wait_while_equal(win, flag, (int)(i - 1));
            samples[i - 1] = (now_s() - t0) * 1e9;
        } else {
            wait_while_equal(win, flag, (int)(i - 1));
            MPI_Put(&v, 1, MPI_INT, peer, 0, 1, MPI_INT, win);
            MPI_Win_flush(peer, win);
        }
    }
    MPI_Win_flush_all(win);
    MPI_Barrier(MPI_COMM_WORLD);

    /* 1 MiB float put of PCG32(seed 42) uniform[-1,1) values, checked bit-exact */
    float *src = (float *)malloc(payload_words * 4);
    oracle_pcg32_uniform_pm1(src, payload_words, 42, 54);
    int ok = 1;
    double put_s = 0.0;
    if (rank == 0) {
        double t0 = now_s();
        MPI_Put(src, (int)payload_words, MPI_FLOAT, 1, 64, (int)payload_words, MPI_FLOAT, win);
        MPI_Win_flush(1, win);
        put_s = now_s() - t0;
        int v = -1;
        MPI_Put(&v, 1, MPI_INT, 1, 0, 1, MPI_INT, win);
        MPI_Win_flush(1, win);
    } else {
        wait_while_equal(win, flag, (int)iters);
        ok = memcmp(base + 64, src, payload_words * 4) == 0;
    }
    MPI_Bcast(&ok, 1, MPI_INT, 1, MPI_COMM_WORLD);
    if (rank == 0) {
        qsort(samples, (size_t)iters, sizeof(double), cmp_dbl);
        double sum = 0;
        for (long i = 0; i < iters; ++i) sum += samples[i];
        printf("{\"mode\": \"pingpong\", \"iters\": %ld, \"rtt_p50_ns\": %.1f, \"rtt_p99_ns\": %.1f, "
               "\"rtt_mean_ns\": %.1f, \"put_1mib_s\": %.6e, \"put_1mib_bitexact\": %s}\n",
               iters, pct(samples, (size_t)iters, 0.5), pct(samples, (size_t)iters, 0.99), sum / (double)iters,
               put_s, ok ? "true" : "false");
        fflush(stdout);
    }
    free(src);
    free(samples);
    MPI_Win_unlock_all(win);
    MPI_Win_free(&win);
    return ok ? 0 : 1;
}

static int do_putget(int rank, const char *op, const char *dtype, size_t min_b, size_t max_b, long reps, double budget) {
    MPI_Datatype dt = MPI_INT;
    size_t w = 4;
    if (strcmp(dtype, "float") == 0) dt = MPI_FLOAT;
    else if (strcmp(dtype, "double") == 0) { dt = MPI_DOUBLE; w = 8; }
    const int is_put = strcmp(op, "put") == 0;
    MPI_Win win;
    char *base = NULL;
    MPI_Win_allocate_shared((MPI_Aint)max_b, 1, MPI_INFO_NULL, MPI_COMM_WORLD, &base, &win);
    char *local = (char *)malloc(max_b);
    MPI_Win_lock_all(MPI_MODE_NOCHECK, win);
    int all_ok = 1;
    for (size_t n = min_b; n <= max_b; n *= 2) {
        const size_t words = n / 4;
        uint32_t *ref = (uint32_t *)malloc(n);
        oracle_pcg32_fill(ref, words, 0x5eedULL + n, 7);
        /* put: rank 0's local source -> rank 1's window; get: rank 1's window -> rank 0's local target */
        if (is_put) {
            if (rank == 0) memcpy(local, ref, n);
        } else if (rank == 1) {
            memcpy(base, ref, n);
        }
        if (is_put && rank == 1) memset(base, 0, n);
        if (!is_put && rank == 0) memset(local, 0, n);
        MPI_Win_sync(win);
        MPI_Barrier(MPI_COMM_WORLD);
        double *samples = (double *)malloc(sizeof(double) * (size_t)reps);
        long done = 0;
        double t_begin = now_s();
        if (rank == 0) {
            for (long r = 0; r < reps; ++r) {
                double t0 = now_s();
                if (is_put) MPI_Put(local, (int)(n / w), dt, 1, 0, (int)(n / w), dt, win);
                else MPI_Get(local, (int)(n / w), dt, 1, 0, (int)(n / w), dt, win);
                MPI_Win_flush(1, win);
                samples[done++] = now_s() - t0;
                if (now_s() - t_begin > budget) break;
            }
        }
        MPI_Barrier(MPI_COMM_WORLD);
        MPI_Win_sync(win);
        int ok = 1;
        if (is_put && rank == 1) ok = memcmp(base, ref, n) == 0;
        if (!is_put && rank == 0) ok = memcmp(local, ref, n) == 0;
        int ok_all = 0;
        MPI_Allreduce(&ok, &ok_all, 1, MPI_INT, MPI_LAND, MPI_COMM_WORLD);
        all_ok &= ok_all;
        if (rank == 0) {
            double total = 0;
            for (long i = 0; i < done; ++i) total += samples[i];
            qsort(samples, (size_t)done, sizeof(double), cmp_dbl);
            double mean = total / (double)(done ? done : 1);
            printf("{\"mode\": \"%s\", \"dtype\": \"%s\", \"bytes\": %zu, \"reps\": %ld, \"mean_s\": %.6e, "
                   "\"p50_ns\": %.1f, \"p99_ns\": %.1f, \"gbps\": %.4f, \"bitexact\": %s}\n",
                   op, dtype, n, done, mean, pct(samples, (size_t)done, 0.5) * 1e9,
                   pct(samples, (size_t)done, 0.99) * 1e9, 2.0 * (double)n / mean / 1e9, ok_all ? "true" : "false");
            fflush(stdout);
        }
        free(samples);
        free(ref);
    }
    MPI_Win_unlock_all(win);
    MPI_Win_free(&win);
    free(local);
    return all_ok ? 0 : 1;
}

/* ---- config 3: the Jacobi application, 8 ranks in a periodic 2x2x2 grid ----
 * N^3 float64 cells per rank with a one-cell halo, initial values from the
 * global PCG32 stream exactly as oracle_jacobi_init lays them out.  A step:
 * pack the six boundary planes, MPI_Put each into the neighbour's receive
 * slot (slots alternate by step parity), MPI_Win_fence (the barrier),
 * unpack the own slots into the halo, update in the oracle's operation order
 * (compiled without contraction).  With `check`, rank 0 gathers the blocks
 * and compares them bit for bit with oracle_jacobi_run on the whole domain. */
static size_t jidx(size_t P, int i, int j, int k) { return ((size_t)i * P + (size_t)j) * P + (size_t)k; }

static void jface(int d, int a, int b, int n, int halo, int *i, int *j, int *k) {
    const int plane = (d & 1) ? (halo ? n + 1 : n) : (halo ? 0 : 1);
    if (d < 2) *i = plane, *j = a + 1, *k = b + 1;
    else if (d < 4) *i = a + 1, *j = plane, *k = b + 1;
    else *i = a + 1, *j = b + 1, *k = plane;
}

static int do_jacobi(int rank, int nranks, int n, int steps, int check) {
    if (nranks != 8) {
        if (rank == 0) fprintf(stderr, "jacobi needs 8 ranks\n");
        return 1;
    }
    const int cx = rank >> 2, cy = (rank >> 1) & 1, cz = rank & 1;
    const size_t P = (size_t)n + 2, cells = P * P * P, nn = (size_t)n * n;
    double *u = (double *)calloc(cells, sizeof(double)), *v = (double *)calloc(cells, sizeof(double));
    double *send = (double *)malloc(6 * nn * sizeof(double));
    const uint64_t G = 2 * (uint64_t)n;
    uint32_t *w = (uint32_t *)malloc(2 * (size_t)n * sizeof(uint32_t));
    for (int i = 1; i <= n; ++i)
        for (int j = 1; j <= n; ++j) {
            const uint64_t gx = (uint64_t)cx * n + (uint64_t)(i - 1), gy = (uint64_t)cy * n + (uint64_t)(j - 1);
            const uint64_t g = (gx * G + gy) * G + (uint64_t)cz * n;
            oracle_pcg32_fill_at(w, 2 * (size_t)n, 42, 54, 2 * g);
            for (int k = 1; k <= n; ++k) {
                const uint64_t hi = w[2 * (k - 1)], lo = w[2 * (k - 1) + 1];
                u[jidx(P, i, j, k)] = (double)(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
            }
        }
    free(w);
    double *recv = NULL;
    MPI_Win win;
    MPI_Win_allocate((MPI_Aint)(2 * 6 * nn * sizeof(double)), sizeof(double), MPI_INFO_NULL, MPI_COMM_WORLD, &recv,
                     &win);
    MPI_Win_fence(0, win);
    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = now_s();
    for (int t = 0; t < steps; ++t) {
        const int par = t & 1;
        for (int d = 0; d < 6; ++d) {
            for (int a = 0; a < n; ++a)
                for (int b = 0; b < n; ++b) {
                    int i, j, k;
                    jface(d, a, b, n, 0, &i, &j, &k);
                    send[(size_t)d * nn + (size_t)a * n + b] = u[jidx(P, i, j, k)];
                }
            int x = cx, y = cy, z = cz;
            if (d < 2) x ^= 1;
            else if (d < 4) y ^= 1;
            else z ^= 1;
            const int nb = (x << 2) | (y << 1) | z;
            MPI_Put(send + (size_t)d * nn, (int)nn, MPI_DOUBLE, nb, (MPI_Aint)(((size_t)par * 6 + (size_t)(d ^ 1)) * nn),
                    (int)nn, MPI_DOUBLE, win);
        }
        MPI_Win_fence(0, win);
        for (int sd = 0; sd < 6; ++sd)
            for (int a = 0; a < n; ++a)
                for (int b = 0; b < n; ++b) {
                    int i, j, k;
                    jface(sd, a, b, n, 1, &i, &j, &k);
                    u[jidx(P, i, j, k)] = recv[((size_t)par * 6 + (size_t)sd) * nn + (size_t)a * n + b];
                }
        for (int i = 1; i <= n; ++i)
            for (int j = 1; j <= n; ++j)
                for (int k = 1; k <= n; ++k) {
                    const double sum = ((((u[jidx(P, i - 1, j, k)] + u[jidx(P, i + 1, j, k)]) + u[jidx(P, i, j - 1, k)]) +
                                         u[jidx(P, i, j + 1, k)]) +
                                        u[jidx(P, i, j, k - 1)]) +
                                       u[jidx(P, i, j, k + 1)];
                    v[jidx(P, i, j, k)] = 0.25 * u[jidx(P, i, j, k)] + 0.125 * sum;
                }
        double *tmp = u;
        u = v;
        v = tmp;
    }
    double dt = now_s() - t0, dt_max = 0;
    MPI_Reduce(&dt, &dt_max, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    int ok = -1;
    if (check) {
        double *blk = (double *)malloc((size_t)n * nn * sizeof(double));
        for (int i = 1; i <= n; ++i)
            for (int j = 1; j <= n; ++j)
                memcpy(blk + ((size_t)(i - 1) * n + (size_t)(j - 1)) * n, u + jidx(P, i, j, 1), (size_t)n * sizeof(double));
        double *all = rank == 0 ? (double *)malloc(8 * (size_t)n * nn * sizeof(double)) : NULL;
        MPI_Gather(blk, (int)((size_t)n * nn), MPI_DOUBLE, all, (int)((size_t)n * nn), MPI_DOUBLE, 0, MPI_COMM_WORLD);
        if (rank == 0) {
            const size_t Gc = (size_t)G * G * G;
            double *gu = (double *)malloc(Gc * sizeof(double)), *gv = (double *)malloc(Gc * sizeof(double));
            oracle_jacobi_init(gu, (int)G, 42);
            double *res = oracle_jacobi_run(gu, gv, (int)G, steps) ? gv : gu;
            ok = 1;
            for (int r = 0; r < 8 && ok; ++r) {
                const int rx = r >> 2, ry = (r >> 1) & 1, rz = r & 1;
                for (int i = 0; i < n && ok; ++i)
                    for (int j = 0; j < n && ok; ++j) {
                        const size_t gofs = (((size_t)rx * n + i) * G + ((size_t)ry * n + j)) * G + (size_t)rz * n;
                        ok = memcmp(all + ((size_t)r * n * nn + ((size_t)i * n + j) * n), res + gofs,
                                    (size_t)n * sizeof(double)) == 0;
                    }
            }
            free(gu);
            free(gv);
            free(all);
        }
        free(blk);
    }
    if (rank == 0) {
        printf("{\"mode\": \"jacobi\", \"n\": %d, \"steps\": %d, \"ranks\": 8, \"ms_per_step\": %.4f, \"bitexact\": %s}\n",
               n, steps, dt_max * 1e3 / steps, ok < 0 ? "null" : (ok ? "true" : "false"));
        fflush(stdout);
    }
    MPI_Win_free(&win);
    free(u);
    free(v);
    free(send);
    return ok == 0 ? 1 : 0;
}

/* ---- config 4: the uneven-put workload, 64 logical PEs as 8 ranks x 8 ----
 * Each logical PE's heap (32 MiB sources, PCG32 seed 100 + pe, then 64 MiB
 * of targets) lives in its rank's part of one shared window.  A step: every
 * rank issues the requests whose source PE it owns (MPI_Put from its own
 * window memory), then MPI_Win_flush_all and MPI_Barrier.  Afterwards each
 * rank compares the targets it owns with the source bytes, read through the
 * shared window. */
static int do_batch(int rank, int nranks, int nreq, uint64_t seed, int steps) {
    const int npes = 64;
    if (npes % nranks) {
        if (rank == 0) fprintf(stderr, "batch needs a rank count dividing 64\n");
        return 1;
    }
    const int per = npes / nranks;
    const size_t SRC = (size_t)32 << 20, DST = (size_t)64 << 20, HEAP = SRC + DST;
    char *base = NULL;
    MPI_Win win;
    MPI_Win_allocate_shared((MPI_Aint)((size_t)per * HEAP), 1, MPI_INFO_NULL, MPI_COMM_WORLD, &base, &win);
    for (int p = 0; p < per; ++p) {
        oracle_pcg32_fill((uint32_t *)(base + (size_t)p * HEAP), SRC / 4, 100 + (uint64_t)(rank * per + p), 54);
        memset(base + (size_t)p * HEAP + SRC, 0, DST);
    }
    int *sp = (int *)malloc(sizeof(int) * (size_t)nreq), *dp = (int *)malloc(sizeof(int) * (size_t)nreq);
    int *dty = (int *)malloc(sizeof(int) * (size_t)nreq);
    uint64_t *so = (uint64_t *)malloc(8 * (size_t)nreq), *dof = (uint64_t *)malloc(8 * (size_t)nreq);
    uint64_t *nb = (uint64_t *)malloc(8 * (size_t)nreq);
    if (oracle_batch_generate(seed, nreq, npes, SRC, DST, 6, 22, sp, dp, so, dof, nb, dty) != 0) {
        if (rank == 0) fprintf(stderr, "batch generation failed\n");
        return 1;
    }
    uint64_t payload = 0;
    for (int i = 0; i < nreq; ++i) payload += nb[i];
    MPI_Win_lock_all(MPI_MODE_NOCHECK, win);
    MPI_Win_sync(win);
    MPI_Barrier(MPI_COMM_WORLD);
    double best = 1e30;
    for (int t = 0; t < steps; ++t) {
        MPI_Barrier(MPI_COMM_WORLD);
        const double t0 = now_s();
        for (int i = 0; i < nreq; ++i) {
            if (sp[i] / per != rank) continue;
            MPI_Put(base + (size_t)(sp[i] % per) * HEAP + so[i], (int)nb[i], MPI_BYTE, dp[i] / per,
                    (MPI_Aint)((size_t)(dp[i] % per) * HEAP + SRC + dof[i]), (int)nb[i], MPI_BYTE, win);
        }
        MPI_Win_flush_all(win);
        MPI_Barrier(MPI_COMM_WORLD);
        const double dt = now_s() - t0;
        if (dt < best) best = dt;
    }
    MPI_Win_sync(win);
    MPI_Barrier(MPI_COMM_WORLD);
    int ok = 1;
    for (int i = 0; i < nreq && ok; ++i) {
        if (dp[i] / per != rank) continue;
        MPI_Aint size;
        int disp;
        char *src_base = NULL;
        MPI_Win_shared_query(win, sp[i] / per, &size, &disp, &src_base);
        ok = memcmp(base + (size_t)(dp[i] % per) * HEAP + SRC + dof[i], src_base + (size_t)(sp[i] % per) * HEAP + so[i],
                    nb[i]) == 0;
    }
    int ok_all = 0;
    MPI_Allreduce(&ok, &ok_all, 1, MPI_INT, MPI_LAND, MPI_COMM_WORLD);
    if (rank == 0) {
        printf("{\"mode\": \"batch\", \"requests\": %d, \"pes\": %d, \"ranks\": %d, \"payload_bytes\": %llu, "
               "\"steps\": %d, \"best_ms\": %.4f, \"gbps\": %.3f, \"bitexact\": %s}\n",
               nreq, npes, nranks, (unsigned long long)payload, steps, best * 1e3, 2.0 * (double)payload / best / 1e9,
               ok_all ? "true" : "false");
        fflush(stdout);
    }
    MPI_Win_unlock_all(win);
    MPI_Win_free(&win);
    free(sp), free(dp), free(dty), free(so), free(dof), free(nb);
    return ok_all ? 0 : 1;
}

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const char *budget = getenv("NVSHMEM_WAIT_TIMEOUT_MS");
    if (budget && atof(budget) > 0) g_wait_budget_s = atof(budget) * 1e-3;
    int rc = 2;
    const int two_rank_mode = argc >= 2 && (strcmp(argv[1], "pingpong") == 0 || strcmp(argv[1], "putget") == 0);
    if (two_rank_mode && size != 2) {
        if (rank == 0) fprintf(stderr, "%s needs exactly 2 ranks\n", argv[1]);
    } else if (argc >= 3 && strcmp(argv[1], "pingpong") == 0) {
        rc = do_pingpong(rank, atol(argv[2]));
    } else if (argc >= 5 && strcmp(argv[1], "jacobi") == 0) {
        rc = do_jacobi(rank, size, atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
    } else if (argc >= 5 && strcmp(argv[1], "batch") == 0) {
        rc = do_batch(rank, size, atoi(argv[2]), (uint64_t)atoll(argv[3]), atoi(argv[4]));
    } else if (argc >= 4 && strcmp(argv[1], "barrier") == 0) {
        rc = do_barrier(rank, size, atol(argv[2]), atoi(argv[3]));
    } else if (argc >= 8 && strcmp(argv[1], "putget") == 0) {
        rc = do_putget(rank, argv[2], argv[3], (size_t)atoll(argv[4]), (size_t)atoll(argv[5]), atol(argv[6]),
                       atof(argv[7]));
    } else if (rank == 0) {
        fprintf(stderr,
                "usage: %s pingpong ITERS | putget put|get int|float|double MIN MAX REPS BUDGET_S | jacobi N STEPS CHECK "
                "| batch NREQ SEED STEPS | barrier ROUNDS SKIP\n",
                argv[0]);
    }
    MPI_Finalize();
    return rc;
}
