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

static void wait_while_equal(MPI_Win win, volatile int *flag, int cmp) {
    /* nvshmem_int_wait: block while *flag == cmp */
    for (;;) {
        MPI_Win_sync(win);
        if (oracle_int_wait_done(*flag, cmp)) return;
    }
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

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    int rc = 2;
    if (size != 2) {
        if (rank == 0) fprintf(stderr, "needs exactly 2 ranks\n");
    } else if (argc >= 3 && strcmp(argv[1], "pingpong") == 0) {
        rc = do_pingpong(rank, atol(argv[2]));
    } else if (argc >= 8 && strcmp(argv[1], "putget") == 0) {
        rc = do_putget(rank, argv[2], argv[3], (size_t)atoll(argv[4]), (size_t)atoll(argv[5]), atol(argv[6]),
                       atof(argv[7]));
    } else if (rank == 0) {
        fprintf(stderr, "usage: %s pingpong ITERS | putget put|get int|float|double MIN MAX REPS BUDGET_S\n", argv[0]);
    }
    MPI_Finalize();
    return rc;
}
