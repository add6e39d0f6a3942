/*
 * shmem_nbi_exchange.hip -- puts posted from device code between two MPI
 * processes, built against the drop-in headers and libnvshmem.a with
 * relocatable device code like the other examples.  Each PE's kernel posts
 * 903 puts of 64 B .. 4 KiB (cycling through the seven powers of two, laid
 * end to end over its 1 MiB source) into the other PE's heap with
 * nvshmemx_putmem_nbi; the host moves them with nvshmemx_nbi_flush_on_stream,
 * and after the barrier each PE checks every byte it received and that the
 * bytes past the last put stayed zero.  Output (each PE): a line ending in
 * "ok", or the first wrong byte.  No counterpart in the reference, whose
 * puts are blocking.
 */
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <vector>
#include <hip/hip_runtime.h>

#include "nvshmem.h"
#include "nvshmem_device.h"
#include "nvshmemx.h"
#include "nvshmemx_device.h"

#define CHECK_HIP(x)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) {                                                                         \
            fprintf(stderr, "[%s:%d] hip failed with %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            exit(-1);                                                                                   \
        }                                                                                               \
    } while (0)

constexpr size_t kBytes = 1 << 20;
constexpr int kCycle = 7;               /* piece t is 64 << (t % 7) bytes */
constexpr size_t kCycleBytes = 64 * 127; /* one cycle of the seven sizes */
constexpr int kPieces = (int)(kBytes / kCycleBytes) * kCycle;

__host__ __device__ inline unsigned char pattern(int pe, size_t i) {
    unsigned x = (unsigned)i * 2654435761u ^ (unsigned)(pe + 1) * 40503u;
    x ^= x >> 13;
    return (unsigned char)(x * 2246822519u >> 24);
}
__host__ __device__ inline size_t piece_off(int t) {
    return (size_t)(t / kCycle) * kCycleBytes + 64 * (((size_t)1 << (t % kCycle)) - 1);
}
__host__ __device__ inline size_t piece_len(int t) { return (size_t)64 << (t % kCycle); }

__global__ void fill(unsigned char *src, int pe) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < kBytes; i += (size_t)gridDim.x * blockDim.x)
        src[i] = pattern(pe, i);
}

/* every thread of the grid posts pieces t, t + threads, ... */
__global__ void post(unsigned char *dst_sym, const unsigned char *src, int peer) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < kPieces; t += gridDim.x * blockDim.x)
        nvshmemx_putmem_nbi(dst_sym + piece_off(t), src + piece_off(t), piece_len(t), peer);
}

int main(int argc, char **argv) {
    CHECK_HIP(hipSetDevice(0));
    MPI_Init(&argc, &argv);
    nvstart_pes();
    const int mype = nvshmem_my_pe(), npes = nvshmem_n_pes();
    int rc = 0;
    if (npes != 2) {
        if (mype == 0) fprintf(stderr, "shmem_nbi_exchange needs exactly two processes (got %d)\n", npes);
        rc = 1;
    } else {
        const int peer = 1 - mype;
        unsigned char *src = (unsigned char *)nvshmalloc(kBytes);
        unsigned char *dst = (unsigned char *)nvshmalloc(kBytes);
        hipLaunchKernelGGL(fill, dim3(256), dim3(256), 0, 0, src, mype);
        CHECK_HIP(hipMemset(dst, 0, kBytes));
        CHECK_HIP(hipDeviceSynchronize());
        nvshmem_barrier_all();
        hipLaunchKernelGGL(post, dim3(4), dim3(256), 0, 0, dst, src, peer);
        nvshmemx_nbi_flush_on_stream(0);
        CHECK_HIP(hipDeviceSynchronize());
        nvshmem_barrier_all();
        std::vector<unsigned char> got(kBytes);
        CHECK_HIP(hipMemcpy(got.data(), dst, kBytes, hipMemcpyDeviceToHost));
        const size_t covered = piece_off(kPieces - 1) + piece_len(kPieces - 1);
        size_t bad = kBytes;
        for (size_t i = 0; i < kBytes && bad == kBytes; ++i)
            if (got[i] != (i < covered ? pattern(peer, i) : 0)) bad = i;
        if (bad == kBytes) {
            printf("PE %d: nbi exchange: %d puts, %zu bytes received, ok\n", mype, kPieces, covered);
        } else {
            printf("PE %d: nbi exchange: byte %zu is %u, expected %u\n", mype, bad, got[bad],
                   bad < covered ? pattern(peer, bad) : 0u);
            rc = 1;
        }
        if (nvshmemx_get_last_error()) {
            printf("PE %d: device error bits %#x\n", mype, nvshmemx_get_last_error());
            rc = 1;
        }
    }
    fflush(stdout);
    nvstop_pes();
    MPI_Finalize();
    return rc;
}
