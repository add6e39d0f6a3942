/*
 * kernels.hip -- host-launched kernels of the library: the grid-wide copy
 * behind nvshmemx_{put,get}mem_on_stream and the stream-ordered offload
 * barrier behind nvshmem_barrier_all_offload.
 *
 * Copy: bulk put/get is HBM-bound (2 bytes of traffic per byte moved, no
 * arithmetic), so the kernel is a streaming copy: 16-byte accesses, UNROLL
 * independent loads in flight per lane, lanes interleaved so each wavefront
 * access covers 64 * 16 = 1 KiB of contiguous memory, and a grid-stride loop
 * over the whole range so every CU streams from the start.
 *
 * Barrier (reference nvshmem_kernels.cu:56-102 gathers arrivals on PE 0 with
 * one spinning thread per peer, epoch values mod 65536, double-buffered by
 * parity, and fails open after 10^6 cycles): here one lane per PE does one
 * agent-scope fetch-add on a word in PE 0's internal heap region that holds
 * the arrival count and the generation (the barrier's sense), and waits for
 * the generation to advance; the last arriver stores the next generation with
 * a zero count, with release.  A wait that outlives the wall-clock budget
 * records NVSHMEMI_ERR_BARRIER_TIMEOUT.
 */
#include <algorithm>

#include "nvshmemi_device.h"
#include "nvshmemi_host.h"

namespace nvshmemi {

/* MODE 0: all lanes of the grid interleave over the whole range (grid-stride);
 * MODE 1: each workgroup streams one contiguous slice;
 * MODE 2: tiles of U * blockDim vectors alternating between `streams` address
 * runs, normally one tile per workgroup. */
template <int W, int U, int NT, int MODE>
__global__ void __launch_bounds__(1024) k_copy(char *dst, const char *src, size_t nbytes, int streams) {
    if constexpr (MODE == 0) {
        const size_t lane = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        const size_t n = (size_t)gridDim.x * blockDim.x;
        copy<W, U, NT>(dst, src, nbytes, lane, n);
    } else if constexpr (MODE == 1) {
        copy_chunked<W, U, NT>(dst, src, nbytes, blockIdx.x, gridDim.x, threadIdx.x, blockDim.x);
    } else {
        copy_tiled<W, U, NT>(dst, src, nbytes, (size_t)streams, blockIdx.x, gridDim.x, threadIdx.x, blockDim.x);
    }
}

typedef void (*copy_kernel_t)(char *, const char *, size_t, int);

template <int W, int NT, int MODE> static copy_kernel_t pick_unroll(int u) {
    switch (u) {
    case 1: return k_copy<W, 1, NT, MODE>;
    case 2: return k_copy<W, 2, NT, MODE>;
    case 8: return k_copy<W, 8, NT, MODE>;
    default: return k_copy<W, 4, NT, MODE>;
    }
}

template <int W, int MODE> static copy_kernel_t pick_nt(int nt, int u) {
    if (nt & ST_SC1) return (nt & NT_LOAD) ? pick_unroll<W, NT_LOAD | ST_SC1, MODE>(u) : pick_unroll<W, ST_SC1, MODE>(u);
    switch (nt & 3) {
    case NT_LOAD: return pick_unroll<W, NT_LOAD, MODE>(u);
    case NT_STORE: return pick_unroll<W, NT_STORE, MODE>(u);
    case NT_BOTH: return pick_unroll<W, NT_BOTH, MODE>(u);
    default: return pick_unroll<W, NT_NONE, MODE>(u);
    }
}

template <int MODE> static copy_kernel_t pick_width(const CopyConfig &c) {
    if (c.vec_bytes >= 16) return pick_nt<16, MODE>(c.nontemporal, c.unroll);
    if (c.vec_bytes >= 8) return pick_nt<8, MODE>(c.nontemporal, c.unroll);
    return pick_nt<4, MODE>(c.nontemporal, c.unroll);
}

static copy_kernel_t pick_copy_kernel(const CopyConfig &c) {
    if (c.mode == 2) return pick_width<2>(c);
    return c.mode == 1 ? pick_width<1>(c) : pick_width<0>(c);
}

/* Launch shape by size, from sweeps on MI355X through this library's put
 * into a peer process's IPC-mapped heap (copies rotated over a 1 GiB
 * footprint so the 256 MB MALL cannot serve them; streams of back-to-back
 * launches).  From 32 MiB up the best shape is a tiled one with ONE 16-byte
 * access per lane: each workgroup of 256 threads copies one contiguous 4 KiB
 * tile and exits, so the dispatcher streams short-lived workgroups through a
 * narrow moving window.  Two runs, TB/s:
 *
 *   size      tiled 256x1  tiled 128x1  tiled 512x1  tiled 256x2  tiled 512x2  interleaved 8192x512x2
 *   64 MiB    5.51-6.03    5.51-5.88    5.49-5.83    5.78-5.82    5.73-5.76    5.98-6.02
 *   128 MiB   6.41-6.42    6.24-6.32    6.21-6.24    6.24-6.29    6.15-6.20    5.95-5.96
 *   256 MiB   6.59-6.60    6.44-6.49    6.43-6.44    6.45-6.48    6.43-6.45    6.03-6.06
 *   512 MiB   6.67-6.68    6.58-6.59    6.52         6.61         6.48-6.53    5.69-5.71
 *   1 GiB     6.70-6.72    6.60-6.65    6.57         6.61         6.35-6.39    5.42-5.45
 *
 * The same ranking held with both PEs in one process.  Two accesses per lane
 * cost 1-4% wherever they were measured; the interleaved shape with one
 * access per lane (8192 x 512) was 6.09 TB/s at 64 MiB and 5.5 at 1 GiB.
 * Below 32 MiB a grid sized to the copy (interleaved, several accesses in
 * flight per lane).  From 32 MiB up loads are non-temporal and stores are
 * written through the L2 (sc1): a kernel then leaves no dirty lines for its
 * end-of-kernel release, worth 1.5-3% over non-temporal stores.  No shape or
 * policy brings 64 MiB to 6.4 TB/s (80% of 8 TB/s): each launch pays about
 * 1.2 us between launches and 1 us of drain, and copies alternated over
 * several streams ran slower, not faster. */
CopyConfig default_copy_config(size_t nbytes) {
    CopyConfig c;
    c.streams = 1;
    c.vec_bytes = 16;
    c.nontemporal = NT_LOAD | ST_SC1;
    if (nbytes >= ((size_t)32 << 20)) {
        c.mode = 2, c.block = 256, c.workgroups = 0, c.unroll = 1, c.streams = 1;
        return c;
    }
    c.nontemporal = NT_BOTH;
    c.mode = 0;
    c.block = nbytes >= ((size_t)16 << 20) ? 512 : 256;
    c.unroll = nbytes >= ((size_t)16 << 20) ? 8 : 4;
    if (nbytes < ((size_t)8 << 20)) c.nontemporal = NT_NONE; /* small puts are read again soon */
    const size_t per_wg = (size_t)c.block * c.vec_bytes * c.unroll;
    size_t wgs = (nbytes + per_wg - 1) / per_wg;
    if (wgs < 1) wgs = 1;
    if (wgs > 4096) wgs = 4096;
    c.workgroups = (int)wgs;
    return c;
}

/* Mode 2 grid: one tile per workgroup at the widest access the relative
 * alignment allows (tiles are rounded up to whole runs), capped so the grid
 * stays a valid launch; the kernel's tile loop covers any remainder. */
static size_t tiled_grid(const void *dst, const void *src, size_t nbytes, const CopyConfig &c) {
    const uintptr_t rel = (uintptr_t)dst ^ (uintptr_t)src;
    size_t w = (size_t)(c.vec_bytes >= 16 ? 16 : (c.vec_bytes >= 8 ? 8 : 4));
    while (w > 1 && (rel & (w - 1))) w >>= 1;
    const size_t tile_bytes = w * (size_t)c.unroll * (size_t)c.block;
    const size_t tiles = (nbytes + tile_bytes - 1) / tile_bytes;
    const size_t streams = (size_t)c.streams;
    size_t g = (tiles + streams - 1) / streams * streams;
    if (g < 1) g = 1;
    const size_t most = 0xFFFFFFFFu / (size_t)c.block; /* the grid's threads within 32 bits */
    if (g > most) g = most / streams * streams;
    return g;
}

void launch_copy(void *dst, const void *src, size_t nbytes, const CopyConfig &cfg, hipStream_t s) {
    if (cfg.block < 64 || cfg.block > 1024 || cfg.block % 64 != 0)
        NVSHMEMI_FATAL("copy block size %d must be a multiple of 64 in [64, 1024]", cfg.block);
    if (cfg.mode < 0 || cfg.mode > 2) NVSHMEMI_FATAL("copy mode %d must be 0, 1 or 2", cfg.mode);
    if (cfg.mode == 2 && (cfg.streams < 1 || cfg.streams > 64))
        NVSHMEMI_FATAL("copy streams %d must be in [1, 64]", cfg.streams);
    size_t grid = (size_t)(cfg.workgroups > 0 ? cfg.workgroups : 0);
    if (grid == 0) {
        if (cfg.mode == 2) {
            grid = tiled_grid(dst, src, nbytes, cfg);
        } else { /* one pass of `unroll` accesses per lane, at most 16384 workgroups */
            const size_t per_wg = (size_t)cfg.block * 16 * (size_t)(cfg.unroll > 0 ? cfg.unroll : 1);
            grid = std::min<size_t>(16384, std::max<size_t>(1, (nbytes + per_wg - 1) / per_wg));
        }
    }
    copy_kernel_t k = pick_copy_kernel(cfg);
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(cfg.block), 0, s, (char *)dst, (const char *)src, nbytes,
                       cfg.streams);
    NVSHMEMI_HIP(hipGetLastError());
}

/* 2-D copy: rows go to wavefronts (narrow rows) or to workgroups (rows of at
 * least 4 KiB), grid-stride over the rows. */
template <bool WG_ROWS>
__global__ void __launch_bounds__(256) k_copy_2d(char *dst, size_t dpitch, const char *src, size_t spitch, size_t width,
                                                 size_t height) {
    if constexpr (WG_ROWS) {
        for (size_t r = blockIdx.x; r < height; r += gridDim.x)
            copy<16, 4>(dst + r * dpitch, src + r * spitch, width, threadIdx.x, blockDim.x);
    } else {
        const size_t waves = (size_t)gridDim.x * (blockDim.x / 64), w = (size_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
        for (size_t r = w; r < height; r += waves)
            copy<16, 4>(dst + r * dpitch, src + r * spitch, width, threadIdx.x & 63u, 64);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_strided(T *dst, const T *src, ptrdiff_t dst_st, ptrdiff_t src_st, size_t n) {
    const size_t lanes = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += lanes)
        dst[(ptrdiff_t)i * dst_st] = src[(ptrdiff_t)i * src_st];
}

void launch_copy_2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                    hipStream_t s) {
    if (!width || !height) return;
    const bool wg_rows = width >= 4096;
    const size_t per_wg = wg_rows ? 1 : 4; /* rows per workgroup per pass */
    const unsigned grid = (unsigned)std::min<size_t>(8192, (height + per_wg - 1) / per_wg);
    if (wg_rows)
        hipLaunchKernelGGL(k_copy_2d<true>, dim3(grid), dim3(256), 0, s, (char *)dst, dpitch, (const char *)src, spitch,
                           width, height);
    else
        hipLaunchKernelGGL(k_copy_2d<false>, dim3(grid), dim3(256), 0, s, (char *)dst, dpitch, (const char *)src,
                           spitch, width, height);
    NVSHMEMI_HIP(hipGetLastError());
}

template <typename T>
void launch_strided(T *dst, const T *src, ptrdiff_t dst_st, ptrdiff_t src_st, size_t n, hipStream_t s) {
    if (!n) return;
    const unsigned grid = (unsigned)std::min<size_t>(8192, (n + 255) / 256);
    hipLaunchKernelGGL(k_strided<T>, dim3(grid), dim3(256), 0, s, dst, src, dst_st, src_st, n);
    NVSHMEMI_HIP(hipGetLastError());
}
template void launch_strided<int>(int *, const int *, ptrdiff_t, ptrdiff_t, size_t, hipStream_t);
template void launch_strided<float>(float *, const float *, ptrdiff_t, ptrdiff_t, size_t, hipStream_t);
template void launch_strided<double>(double *, const double *, ptrdiff_t, ptrdiff_t, size_t, hipStream_t);

__global__ void k_offload_barrier(unsigned int *word, int npes) {
    if (threadIdx.x != 0) return;
    gen_barrier(word, npes, NVSHMEMI_ERR_BARRIER_TIMEOUT);
}

void launch_offload_barrier(hipStream_t s) {
    char *root = g.heap_base[0];
    hipLaunchKernelGGL(k_offload_barrier, dim3(1), dim3(64), 0, s, (unsigned int *)(root + NVSHMEMI_OFF_BAR_WORD),
                       g.npes);
    NVSHMEMI_HIP(hipGetLastError());
}

} // namespace nvshmemi
