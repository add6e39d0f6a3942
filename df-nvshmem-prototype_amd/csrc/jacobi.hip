/*
 * jacobi.hip -- 3-D 7-point Jacobi with halo exchange through the symmetric
 * heap: the application the PE layer is measured on (8 PEs in a periodic
 * 2x2x2 grid, N^3 float64 cells per PE).  A caller of the library, exported
 * with C entry points (nvshmemh_jacobi_*) for the test and benchmark drivers.
 *
 * Per PE (symmetric allocations): u and v, each (N+2)^3 doubles with a
 * one-cell halo, index ((i * (N+2)) + j) * (N+2) + k, k fastest; and a receive
 * area recv[2][6][N*N] (parity x face slot).
 *
 * One step:
 *   1. pack+put: every boundary plane of u is written straight into the
 *      neighbour's receive slot for that side (translated symmetric address),
 *      so packing and the put are one pass; faces are strided in x and y, so
 *      the reference's contiguous put would need a separate pack;
 *   2. barrier: nvshmem_barrier_all_offload (separate processes) or the
 *      kernel boundary (all PEs in one launch, local mode);
 *   3. unpack: receive slots into u's halo;
 *   4. stencil: v = 0.25 * u + 0.125 * ((((( u[x-1] + u[x+1]) + u[y-1]) +
 *      u[y+1]) + u[z-1]) + u[z+1]) in this exact order (the host replay uses
 *      the same order; multiplications by powers of two are exact, so the
 *      result is bit-exact with or without FMA contraction);
 *   5. swap u, v.
 * Receive slots alternate by step parity, so one barrier per step suffices: a
 * PE can run at most one step ahead of a neighbour and then writes the other
 * parity.
 *
 * Initial values: global cell (x, y, z) of the 2N^3 domain gets the f64
 * uniform[0,1) value built from PCG32 words 2g and 2g+1, g = (x * 2N + y) * 2N
 * + z, i.e. (w0 << 32 | w1) >> 11 scaled by 2^-53.
 */
#include <stdint.h>

#include "nvshmem.h"
#include "nvshmem_device.h"
#include "nvshmemx.h"
#include "nvshmemx_device.h"
#include "nvshmemi_device.h"
#include "nvshmemi_host.h"

namespace {

constexpr uint64_t kMult = 6364136223846793005ULL;

__device__ __forceinline__ uint32_t out32(uint64_t old) {
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

__device__ __forceinline__ uint64_t advance(uint64_t state, uint64_t delta, uint64_t inc) {
    uint64_t am = 1, ap = 0, cm = kMult, cp = inc;
    while (delta) {
        if (delta & 1) {
            am *= cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm *= cm;
        delta >>= 1;
    }
    return am * state + ap;
}

struct geom {
    int n;  /* interior cells per side */
    int pe; /* this PE */
    int cx, cy, cz;
    __device__ __forceinline__ geom(int n_, int pe_) : n(n_), pe(pe_) {
        cx = pe >> 2;
        cy = (pe >> 1) & 1;
        cz = pe & 1;
    }
    __device__ __forceinline__ size_t idx(int i, int j, int k) const {
        return ((size_t)i * (n + 2) + j) * (size_t)(n + 2) + k;
    }
    /* neighbour PE across side d (0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z), periodic 2x2x2 */
    __device__ __forceinline__ int neighbour(int d) const {
        int x = cx, y = cy, z = cz;
        if (d < 2) x ^= 1;
        else if (d < 4) y ^= 1;
        else z ^= 1;
        return (x << 2) | (y << 1) | z;
    }
};

/* cell (a, b) of face d: interior plane on side d (send) or halo plane (recv) */
__device__ __forceinline__ void face_cell(const geom &g, int d, int a, int b, bool halo, int &i, int &j, int &k) {
    const int n = g.n;
    const int lo = halo ? 0 : 1, hi = halo ? n + 1 : n;
    const int plane = (d & 1) ? hi : lo;
    if (d < 2) i = plane, j = a + 1, k = b + 1;
    else if (d < 4) i = a + 1, j = plane, k = b + 1;
    else i = a + 1, j = b + 1, k = plane;
}

/* blockIdx.z = PE offset from pe0; blockIdx.y = face; x covers N*N cells */
__global__ void k_pack_put(double *u_sym, double *recv_sym, int n, int pe0, int parity) {
    const geom g(n, pe0 + (int)blockIdx.z);
    const int d = blockIdx.y;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= n * n) return;
    const int a = cell / n, b = cell % n;
    int i, j, k;
    face_cell(g, d, a, b, false, i, j, k);
    const double *u = (const double *)nvshmemx_ptr_d(u_sym, g.pe);
    /* side d of this PE is side d^1 of the neighbour */
    double *dst = (double *)nvshmemx_ptr_d(recv_sym, g.neighbour(d));
    dst[((size_t)parity * 6 + (size_t)(d ^ 1)) * (size_t)n * n + cell] = u[g.idx(i, j, k)];
}

__global__ void k_unpack(double *u_sym, const double *recv_sym, int n, int pe0, int parity) {
    const geom g(n, pe0 + (int)blockIdx.z);
    const int s = blockIdx.y;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= n * n) return;
    const int a = cell / n, b = cell % n;
    int i, j, k;
    face_cell(g, s, a, b, true, i, j, k);
    double *u = (double *)nvshmemx_ptr_d(u_sym, g.pe);
    const double *r = (const double *)nvshmemx_ptr_d(recv_sym, g.pe);
    u[g.idx(i, j, k)] = r[((size_t)parity * 6 + (size_t)s) * (size_t)n * n + cell];
}

/* thread (k, j) walks planes i in [i0, i0 + chunk) keeping x-1, x, x+1 in registers */
__global__ void __launch_bounds__(256) k_stencil(const double *u_sym, double *v_sym, int n, int pe0, int chunk) {
    const geom g(n, pe0 + (int)blockIdx.z);
    const int k = blockIdx.x * 64 + (threadIdx.x & 63) + 1;
    const int jb = (int)(gridDim.y / ((n + chunk - 1) / chunk));
    const int j = (blockIdx.y % jb) * 4 + (threadIdx.x >> 6) + 1;
    const int i0 = (blockIdx.y / jb) * chunk + 1;
    if (k > n || j > n) return;
    const double *__restrict__ u = (const double *)nvshmemx_ptr_d(u_sym, g.pe);
    double *__restrict__ v = (double *)nvshmemx_ptr_d(v_sym, g.pe);
    const size_t sx = (size_t)(n + 2) * (n + 2), sy = (size_t)(n + 2);
    size_t c = g.idx(i0, j, k);
    double xm = u[c - sx], x0 = u[c];
    const int i1 = i0 + chunk - 1 < n ? i0 + chunk - 1 : n;
#pragma unroll 4
    for (int i = i0; i <= i1; ++i, c += sx) {
        const double xp = u[c + sx];
        const double s = ((((xm + xp) + u[c - sy]) + u[c + sy]) + u[c - 1]) + u[c + 1];
        v[c] = 0.25 * x0 + 0.125 * s;
        xm = x0;
        x0 = xp;
    }
}

__global__ void k_init(double *u_sym, int n, int pe0, uint64_t s0, uint64_t inc) {
    const geom g(n, pe0 + (int)blockIdx.z);
    const size_t cells = (size_t)n * n * n;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells) return;
    const int i = (int)(t / ((size_t)n * n)) + 1, j = (int)((t / n) % n) + 1, k = (int)(t % n) + 1;
    const uint64_t G = 2ull * (uint64_t)n;
    const uint64_t gx = (uint64_t)g.cx * n + (i - 1), gy = (uint64_t)g.cy * n + (j - 1), gz = (uint64_t)g.cz * n + (k - 1);
    const uint64_t lin = (gx * G + gy) * G + gz;
    uint64_t st = advance(s0, 2 * lin, inc);
    const uint64_t hi = out32(st);
    st = st * kMult + inc;
    const uint64_t lo = out32(st);
    double *u = (double *)nvshmemx_ptr_d(u_sym, g.pe);
    u[g.idx(i, j, k)] = (double)(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

void check_n(int n) {
    if (n < 2 || n > 1024) NVSHMEMI_FATAL("jacobi: N = %d outside [2, 1024]", n);
}

} // namespace

extern "C" {

/* bytes of each symmetric allocation: u / v and the receive area */
size_t nvshmemh_jacobi_field_bytes(int n) { return (size_t)(n + 2) * (n + 2) * (n + 2) * sizeof(double); }
size_t nvshmemh_jacobi_recv_bytes(int n) { return (size_t)2 * 6 * n * n * sizeof(double); }

/* pe0..pe0+npes_here-1 are the PEs this call works for (1 in process mode,
 * all 8 in local mode) */
void nvshmemh_jacobi_init(double *u_sym, int n, int pe0, int npes_here, unsigned long long seed, hipStream_t s) {
    check_n(n);
    uint64_t inc = (54ull << 1u) | 1u, st = 0;
    st = st * kMult + inc;
    st += seed;
    st = st * kMult + inc;
    const size_t cells = (size_t)n * n * n;
    hipLaunchKernelGGL(k_init, dim3((unsigned)((cells + 255) / 256), 1, npes_here), dim3(256), 0, s, u_sym, n, pe0, st,
                       inc);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_jacobi_pack_put(double *u_sym, double *recv_sym, int n, int pe0, int npes_here, int parity,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_pack_put, dim3((unsigned)((n * n + 255) / 256), 6, npes_here), dim3(256), 0, s, u_sym,
                       recv_sym, n, pe0, parity);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_jacobi_unpack(double *u_sym, double *recv_sym, int n, int pe0, int npes_here, int parity,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_unpack, dim3((unsigned)((n * n + 255) / 256), 6, npes_here), dim3(256), 0, s, u_sym,
                       recv_sym, n, pe0, parity);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_jacobi_stencil(double *u_sym, double *v_sym, int n, int pe0, int npes_here, int chunk, hipStream_t s) {
    if (chunk <= 0) chunk = n;
    const int jb = (n + 3) / 4, ic = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_stencil, dim3((unsigned)((n + 63) / 64), (unsigned)(jb * ic), npes_here), dim3(256), 0, s,
                       u_sym, v_sym, n, pe0, chunk);
    NVSHMEMI_HIP(hipGetLastError());
}

/* `steps` full steps; returns the buffer holding the result (u or v).  In
 * process mode each step's barrier is nvshmem_barrier_all_offload; in local
 * mode the launches for all PEs are ordered on the stream. */
double *nvshmemh_jacobi_run(double *u_sym, double *v_sym, double *recv_sym, int n, int steps, int chunk, int first_step,
                            hipStream_t s) {
    const bool local = nvshmemx_mode() == NVSHMEMX_MODE_LOCAL;
    const int pe0 = local ? 0 : nvshmem_my_pe();
    const int here = local ? nvshmem_n_pes() : 1;
    if (nvshmem_n_pes() != 8) NVSHMEMI_FATAL("jacobi needs exactly 8 PEs (2x2x2), found %d", nvshmem_n_pes());
    double *a = u_sym, *b = v_sym;
    for (int t = 0; t < steps; ++t) {
        const int parity = (first_step + t) & 1;
        nvshmemh_jacobi_pack_put(a, recv_sym, n, pe0, here, parity, s);
        if (!local) nvshmemx_barrier_all_on_stream(s);
        nvshmemh_jacobi_unpack(a, recv_sym, n, pe0, here, parity, s);
        nvshmemh_jacobi_stencil(a, b, n, pe0, here, chunk, s);
        double *tmp = a;
        a = b;
        b = tmp;
    }
    return a;
}

} /* extern "C" */
