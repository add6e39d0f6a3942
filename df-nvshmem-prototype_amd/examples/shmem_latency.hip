/*
 * shmem_latency.hip -- put + flag ping-pong between two MPI processes: the
 * caller pattern of the reference's microbench/shmem_latency.cu:48-139, built
 * against the drop-in headers and libnvshmem.a the way a reference user builds
 * (hipcc -fgpu-rdc ... -lnvshmem -lmpi).
 *
 * Iteration i: one PE puts `len` ints into its peer, quiets, raises the
 * peer's flag to i with nvshmem_int_p and quiets; the peer, waiting, answers
 * the same way.  Waits are nvshmem_int_wait(flag, i - 1) ("while equal to the
 * previous value"): a flag advanced twice before the waiter looks still
 * releases it, which an equality wait on i does not guarantee.  Unlike the
 * reference, the payload is checked after every size: PE 0's pattern must
 * have reached both PEs intact.
 *
 * Output (PE 0): size in bytes, round-trip latency in ns (wall clock) and in
 * shader cycles, payload check.
 */
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <hip/hip_runtime.h>

#include "nvshmem.h"
#include "nvshmem_device.h"

#define CHECK_HIP(x)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) {                                                                         \
            fprintf(stderr, "[%s:%d] hip failed with %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            exit(-1);                                                                                   \
        }                                                                                               \
    } while (0)

__global__ void ping_pong(int *data, int *flag, int len, int pe, int iter, int skip, long long *out) {
    const int peer = 1 - pe;
    long long t0 = 0, c0 = 0;
    for (int i = 1; i <= iter + skip; ++i) {
        if (i == skip + 1) {
            t0 = wall_clock64();
            c0 = clock64();
        }
        const bool waits_first = (i & 1) == pe;
        if (waits_first) nvshmem_int_wait(flag, i - 1);
        nvshmem_int_put(data, data, len, peer);
        nvshmem_quiet();
        nvshmem_int_p(flag, i, peer);
        nvshmem_quiet();
        if (!waits_first) nvshmem_int_wait(flag, i - 1);
    }
    out[0] = wall_clock64() - t0;
    out[1] = clock64() - c0;
}

__global__ void fill(int *p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v + i;
}

__global__ void count_mismatch(const int *p, int n, int v, int *bad) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (p[i] != v + i) atomicAdd(bad, 1);
}

int main(int argc, char **argv) {
    const int iter = 200, skip = 20, max_bytes = 64 * 1024;
    int dev_count = 0;
    CHECK_HIP(hipGetDeviceCount(&dev_count));
    const char *lr = getenv("MPI_LOCALRANKID");
    CHECK_HIP(hipSetDevice((lr ? atoi(lr) : 0) % dev_count));
    MPI_Init(&argc, &argv);
    nvstart_pes();
    const int mype = nvshmem_my_pe(), npes = nvshmem_n_pes();
    int rc = 0;
    if (npes != 2) {
        if (mype == 0) fprintf(stderr, "this test needs exactly two processes\n");
        rc = 1;
    } else {
        int *data = (int *)nvshmalloc(max_bytes);
        int *flag = (int *)nvshmalloc(sizeof(int));
        long long *out = nullptr;
        int *bad = nullptr;
        CHECK_HIP(hipMalloc(&out, 2 * sizeof(long long)));
        CHECK_HIP(hipMalloc(&bad, sizeof(int)));
        int khz = 100000;
        CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
        if (mype == 0) printf("%10s %14s %14s %s\n", "size", "latency(ns)", "latency(cyc)", "payload");
        for (int size = (int)sizeof(int); size <= max_bytes; size *= 2) {
            const int len = size / (int)sizeof(int);
            CHECK_HIP(hipMemset(flag, 0, sizeof(int)));
            hipLaunchKernelGGL(fill, dim3(64), dim3(256), 0, 0, data, len, 1000000 * mype + size);
            CHECK_HIP(hipDeviceSynchronize());
            nvshmem_barrier_all();
            hipLaunchKernelGGL(ping_pong, dim3(1), dim3(1), 0, 0, data, flag, len, mype, iter, skip, out);
            CHECK_HIP(hipDeviceSynchronize());
            nvshmem_barrier_all();
            /* PE 0 sends first in iteration 1, so its pattern is what both hold */
            CHECK_HIP(hipMemset(bad, 0, sizeof(int)));
            hipLaunchKernelGGL(count_mismatch, dim3(64), dim3(256), 0, 0, data, len, size, bad);
            int h_bad = 0, bad_all = 0;
            long long h_out[2];
            CHECK_HIP(hipMemcpy(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost));
            CHECK_HIP(hipMemcpy(h_out, out, sizeof(h_out), hipMemcpyDeviceToHost));
            MPI_Allreduce(&h_bad, &bad_all, 1, MPI_INT, MPI_SUM, MPI_COMM_WORLD);
            if (bad_all) rc = 2;
            if (mype == 0)
                printf("%10d %14.1f %14lld %s\n", size, (double)h_out[0] * 1e6 / khz / iter, h_out[1] / iter,
                       bad_all ? "MISMATCH" : "ok");
            fflush(stdout);
        }
        CHECK_HIP(hipFree(out));
        CHECK_HIP(hipFree(bad));
        nvshmcleanup();
    }
    nvstop_pes();
    MPI_Finalize();
    return rc;
}
