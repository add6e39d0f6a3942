/*
 * shmem_sync_latency.hip -- p + wait ping-pong between two MPI processes, the
 * caller pattern of the reference's microbench/shmem_sync_latency.cu:48-127,
 * built against the drop-in headers and libnvshmem.a.  Same wait discipline
 * as shmem_latency.hip.  Output (PE 0): round-trip latency in ns and shader
 * cycles, and the final flag values (must equal the iteration count).
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

__global__ void ping_pong(int *flag, int pe, int iter, int skip, long long *out) {
    const int peer = 1 - pe;
    long long t0 = 0, c0 = 0;
    for (int i = 1; i <= iter + skip; ++i) {
        if (i == skip + 1) {
            t0 = wall_clock64();
            c0 = clock64();
        }
        const bool waits_first = (i & 1) == pe;
        if (waits_first) nvshmem_int_wait(flag, i - 1);
        nvshmem_int_p(flag, i, peer);
        nvshmem_quiet();
        if (!waits_first) nvshmem_int_wait(flag, i - 1);
    }
    out[0] = wall_clock64() - t0;
    out[1] = clock64() - c0;
}

int main(int argc, char **argv) {
    const int iter = 10000, skip = 100;
    CHECK_HIP(hipSetDevice(0));
    MPI_Init(&argc, &argv);
    nvstart_pes();
    const int mype = nvshmem_my_pe(), npes = nvshmem_n_pes();
    int rc = 0;
    if (npes != 2) {
        if (mype == 0) fprintf(stderr, "this test needs exactly two processes\n");
        rc = 1;
    } else {
        int *flag = (int *)nvshmalloc(sizeof(int));
        long long *out = nullptr;
        CHECK_HIP(hipMalloc(&out, 2 * sizeof(long long)));
        CHECK_HIP(hipMemset(flag, 0, sizeof(int)));
        CHECK_HIP(hipDeviceSynchronize());
        nvshmem_barrier_all();
        hipLaunchKernelGGL(ping_pong, dim3(1), dim3(1), 0, 0, flag, mype, iter, skip, out);
        CHECK_HIP(hipDeviceSynchronize());
        nvshmem_barrier_all();
        long long h_out[2];
        int h_flag = 0, flags[2];
        CHECK_HIP(hipMemcpy(h_out, out, sizeof(h_out), hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(&h_flag, flag, sizeof(int), hipMemcpyDeviceToHost));
        MPI_Gather(&h_flag, 1, MPI_INT, flags, 1, MPI_INT, 0, MPI_COMM_WORLD);
        int khz = 100000;
        CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
        if (mype == 0) {
            printf("sync latency: %.1f ns %lld cycles flags %d %d\n", (double)h_out[0] * 1e6 / khz / iter,
                   h_out[1] / iter, flags[0], flags[1]);
            if (flags[0] != iter + skip || flags[1] != iter + skip) rc = 2;
        }
        MPI_Bcast(&rc, 1, MPI_INT, 0, MPI_COMM_WORLD);
        CHECK_HIP(hipFree(out));
        nvshmcleanup();
    }
    nvstop_pes();
    MPI_Finalize();
    return rc;
}
