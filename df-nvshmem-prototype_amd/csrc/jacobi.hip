/*
 * jacobi.hip -- 3-D 7-point Jacobi with halo exchange through the symmetric
 * heap: the application the PE layer is measured on (8 PEs in a periodic
 * 2x2x2 grid, N^3 float64 cells per PE).  A caller of the library, exported
 * with C entry points (nvshmemh_jacobi_*) for the test and benchmark drivers.
 *
 * Per PE (symmetric allocations): u and v, each (N+2)^2 rows of N+32 doubles
 * holding a one-cell halo (cell (i, j, k), 0 <= i, j, k <= N+1, at
 * ((i * (N+2)) + j) * (N+32) + k + 15, so interior cell k = 1 starts a
 * 128-byte line when N is a multiple of 16); and a receive area
 * recv[2][6][N*N] (parity x face slot).
 *
 * The update, in this exact order (the host replay uses the same order;
 * multiplications by powers of two are exact, so the result is bit-exact with
 * or without FMA contraction):
 *     v = 0.25 * u + 0.125 * ((((( u[x-1] + u[x+1]) + u[y-1]) + u[y+1]) + u[z-1]) + u[z+1])
 *
 * A step of nvshmemh_jacobi_run is one fused launch for every PE of the
 * process: the halo is read straight from this PE's receive slots of the
 * step's parity, and every boundary value of v is also written into the
 * receive slot of the next parity on the neighbour across that face (a
 * translated symmetric address), so the next step's halo exchange rides on
 * this step's stores; then the barrier (nvshmem_barrier_all_offload between
 * processes, the kernel boundary in local mode).  A run starts with one
 * pack+put of u's faces.  The separate kernels (pack+put, unpack into u's
 * halo, stencil reading the halo from u) stay exported.  Faces are strided in
 * x and y, so the reference's contiguous put would need a separate pack.
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

/* Row layout: a row of n + 2 cells (k = 0 .. n+1) sits at offset kKOff in a
 * row of n + kRowPad doubles, so interior cell k = 1 starts on a 128-byte
 * line whenever n is a multiple of 16 (rows are then whole lines long). */
constexpr int kKOff = 15;
constexpr int kRowPad = 32;

constexpr unsigned kXcds = 8; /* MI355X: eight accelerator complex dies */

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
    __device__ __forceinline__ size_t row_stride() const { return (size_t)n + kRowPad; }
    __device__ __forceinline__ size_t idx(int i, int j, int k) const {
        return ((size_t)i * (n + 2) + j) * row_stride() + (size_t)(k + kKOff);
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

/* Workgroup barrier over LDS only: waits for this wavefront's LDS (and
 * scalar) accesses, then s_barrier; the "memory" clobber keeps the compiler
 * from moving any access across it.  Global loads and stores stay in flight
 * across it (__syncthreads() would wait for them too), so a plane's loads
 * overlap the barriers of the plane before. */
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

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

/* The same faces sent through the library's device put API instead of
 * direct stores: blockIdx.z = PE offset, blockIdx.y = face, blockIdx.x = a
 * band of kApiBand rows (x and y faces) or planes (z faces), dealt to the
 * workgroup's wavefronts in turn.  A row of an x or y face is n contiguous
 * doubles: one contiguous wavefront put (nvshmemx_putmem_wave) into the
 * neighbour's receive slot; a plane of a z face is one column of n doubles a
 * row stride apart: one strided wavefront put (nvshmemx_double_iput_wave).
 * The receive slot layout is k_pack_put's.  At 8 PEs x 256^3 this pack takes
 * 27.5 us against 27.8 us for the direct stores; with one pitched 2-D
 * workgroup put per band and one strided workgroup put per plane it took 86
 * us -- every workgroup-collective call ends with a release, which costs more
 * than the puts when a face is cut into many calls. */
constexpr int kApiBand = 16;
__global__ void __launch_bounds__(256) k_pack_put_api(double *u_sym, double *recv_sym, int n, int pe0, int parity) {
    const geom g(n, pe0 + (int)blockIdx.z);
    const int d = blockIdx.y;
    const int r0 = (int)blockIdx.x * kApiBand;
    if (r0 >= n) return; /* the whole workgroup: the device calls below are workgroup-collective */
    const int rows = kApiBand < n - r0 ? kApiBand : n - r0;
    const double *u = (const double *)nvshmemx_ptr_d(u_sym, g.pe);
    const int plane = (d & 1) ? n : 1;
    double *slot = recv_sym + ((size_t)parity * 6 + (size_t)(d ^ 1)) * (size_t)n * n; /* symmetric */
    const int nb = g.neighbour(d);
    const size_t row_bytes = (size_t)n * sizeof(double);
    const int waves = (int)(blockDim.x >> 6), w = (int)(threadIdx.x >> 6);
    if (d < 4) { /* the band's rows go to the workgroup's wavefronts in turn */
        for (int r = r0 + 1 + w; r <= r0 + rows; r += waves) {
            const double *src = d < 2 ? u + g.idx(plane, r, 1) : u + g.idx(r, plane, 1);
            nvshmemx_putmem_wave(slot + (size_t)(r - 1) * n, src, row_bytes, nb);
        }
    } else { /* the band's planes go to the workgroup's wavefronts in turn */
        for (int i = r0 + 1 + w; i <= r0 + rows; i += waves)
            nvshmemx_double_iput_wave(slot + (size_t)(i - 1) * n, u + g.idx(i, 1, plane), 1,
                                      (ptrdiff_t)g.row_stride(), (size_t)n, nb);
    }
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

/* Workgroup (k tile, j tile x i chunk, PE): walks planes i0..i1 of a
 * TJ x TK tile.  Each plane (tile plus a one-cell ring in j and k) is read
 * from memory once into one of two LDS buffers; plane i+2 is fetched into
 * registers while plane i is consumed.  Thread (wave w, lane l) owns columns
 * k0 + l + 64c (c < TK / 64) of rows j0 + w TJ/4 .. j0 + (w+1) TJ/4 - 1 and
 * loads exactly those cells of every plane (waves 0 and 3 also load the j
 * halo rows), so the x-1 and x+1 values it needs stay in its registers and
 * LDS only serves the in-plane neighbours: two buffers and one barrier per
 * plane suffice.
 *
 * FUSED: the halo is not in u but in this PE's receive slots of parity
 * `pin` (a halo cell is read straight from the slot of its face), and every
 * boundary value of v is also written into the receive slot of parity `pout`
 * of the neighbour across that face -- the next step's halo exchange rides on
 * this step's stores, so a step is one kernel plus the barrier. */
template <bool FUSED, int TJ, int TK>
__global__ void __launch_bounds__(256) k_stencil_t(const double *u_sym, double *v_sym, double *recv_sym, int n, int pe0,
                                                   int chunk, int pin, int pout) {
    constexpr int KC = TK / 64; /* columns per lane */
    /* XCD-aware tile order: the dispatcher deals workgroup ids round-robin
     * over the eight XCDs, each with its own L2.  Renumbering so that each XCD
     * gets a contiguous run of tiles (k fastest, then j) lets tiles that share
     * halo cache lines run on the same L2 instead of fetching them again. */
    const unsigned gx = gridDim.x, gy = gridDim.y, nwg = gx * gy * gridDim.z;
    unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (nwg % kXcds == 0) lin = (lin % kXcds) * (nwg / kXcds) + lin / kXcds;
    const unsigned bx = lin % gx, by = (lin / gx) % gy, bz = lin / (gx * gy);
    const geom g(n, pe0 + (int)bz);
    const int jtiles = (n + TJ - 1) / TJ;
    const int k0 = (int)bx * TK + 1;
    const int j0 = (int)(by % (unsigned)jtiles) * TJ + 1;
    const int i0 = (int)(by / (unsigned)jtiles) * chunk + 1;
    const int i1 = i0 + chunk - 1 < n ? i0 + chunk - 1 : n;
    const double *__restrict__ u = (const double *)nvshmemx_ptr_d(u_sym, g.pe);
    double *__restrict__ v = (double *)nvshmemx_ptr_d(v_sym, g.pe);
    const double *__restrict__ rin = FUSED ? (const double *)nvshmemx_ptr_d(recv_sym, g.pe) : nullptr;
    const size_t sy = g.row_stride(), sx = (size_t)(n + 2) * sy, nn = (size_t)n * n;
    __shared__ double tile[2][TJ + 2][TK + 2];
    const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
    const int jlim = j0 + TJ < n + 1 ? j0 + TJ : n + 1; /* last row to load (halo included) */
    const int klim = k0 + TK < n + 1 ? k0 + TK : n + 1; /* last column to load */
    constexpr int RPW = TJ / 4;                         /* rows per wavefront */
    const bool has_halo = threadIdx.x < 2 * (TJ + 2);
    const int hr = (int)(threadIdx.x >> 1), hhi = (int)(threadIdx.x & 1u);
    const int hk = hhi ? k0 + TK : k0 - 1;

    /* cell (i, j, k) of the (n+2)^3 block, 0 <= i, j, k <= n+1 */
    auto at = [&](int i, int j, int k) -> double {
        if constexpr (FUSED) {
            const bool hx = i == 0 || i == n + 1, hy = j == 0 || j == n + 1, hz = k == 0 || k == n + 1;
            if (hx || hy || hz) {
                if ((int)hx + (int)hy + (int)hz > 1) return 0.0; /* edges and corners: never used */
                if (hx) return rin[((size_t)pin * 6 + (i ? 1 : 0)) * nn + (size_t)(j - 1) * n + (k - 1)];
                if (hy) return rin[((size_t)pin * 6 + (j ? 3 : 2)) * nn + (size_t)(i - 1) * n + (k - 1)];
                return rin[((size_t)pin * 6 + (k ? 5 : 4)) * nn + (size_t)(i - 1) * n + (j - 1)];
            }
        }
        return u[(size_t)i * sx + (size_t)j * sy + (size_t)(k + kKOff)];
    };
    /* value of interior boundary cell (i, j, k) into the neighbour across face d */
    auto send = [&](int d, size_t idx, double val) {
        double *dst = (double *)nvshmemx_ptr_d(recv_sym, g.neighbour(d));
        dst[((size_t)pout * 6 + (size_t)(d ^ 1)) * nn + idx] = val;
    };

    /* plane i of the tile into registers (issued early), then into LDS.
     * Entry q < RPW is own row j0 + wv RPW + q; entry RPW the j halo row of
     * waves 0 (row j0 - 1) and 3 (row j0 + TJ). */
    const int hrow = wv == 0 ? 0 : (wv == 3 ? TJ + 1 : -1); /* tile row of the halo entry */
    double pre[RPW + 1][KC], preh = 0.0;
    auto issue = [&](int i) {
#pragma unroll
        for (int q = 0; q <= RPW; ++q) {
            const int r = q < RPW ? wv * RPW + q + 1 : hrow, j = j0 - 1 + r;
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                const int kc = k0 + lane + 64 * c;
                pre[q][c] = (r >= 0 && j <= jlim && kc <= klim) ? at(i, j, kc) : 0.0;
            }
        }
        if (has_halo) {
            const int j = j0 - 1 + hr;
            preh = (j <= jlim && hk <= n + 1) ? at(i, j, hk) : 0.0;
        }
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int q = 0; q <= RPW; ++q) {
            const int r = q < RPW ? wv * RPW + q + 1 : hrow;
            if (r >= 0) {
#pragma unroll
                for (int c = 0; c < KC; ++c) tile[buf][r][lane + 1 + 64 * c] = pre[q][c];
            }
        }
        if (has_halo) tile[buf][hr][hhi ? TK + 1 : 0] = preh;
    };

    double xm[RPW][KC], xp[RPW][KC]; /* own cells of planes i - 1 and i + 1 */
    bool ok[RPW][KC];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int j = j0 + wv * RPW + q;
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            const int kc = k0 + lane + 64 * c;
            ok[q][c] = kc <= n && j <= n;
            xm[q][c] = ok[q][c] ? at(i0 - 1, j, kc) : 0.0;
        }
    }
    issue(i0);
    commit(0);
    issue(i0 + 1);
    commit(1);
#pragma unroll
    for (int q = 0; q < RPW; ++q)
#pragma unroll
        for (int c = 0; c < KC; ++c) xp[q][c] = pre[q][c];
    __syncthreads();
    int bc = 0; /* buffer of plane i; plane i + 1 is in the other */
    for (int i = i0; i <= i1; ++i) {
        const bool fetch = i + 2 <= i1 + 1;
        if (fetch) issue(i + 2);
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int r = wv * RPW + q + 1;
            const int j = j0 + r - 1;
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                const int cc = lane + 1 + 64 * c, kc = k0 + lane + 64 * c;
                const double x0 = tile[bc][r][cc];
                const double s = ((((xm[q][c] + xp[q][c]) + tile[bc][r - 1][cc]) + tile[bc][r + 1][cc]) +
                                  tile[bc][r][cc - 1]) +
                                 tile[bc][r][cc + 1];
                const double val = 0.25 * x0 + 0.125 * s;
                if (ok[q][c]) {
                    __builtin_nontemporal_store(val, &v[(size_t)i * sx + (size_t)j * sy + (size_t)(kc + kKOff)]);
                    if constexpr (FUSED) {
                        if (i == 1) send(0, (size_t)(j - 1) * n + (kc - 1), val);
                        if (i == n) send(1, (size_t)(j - 1) * n + (kc - 1), val);
                        if (j == 1) send(2, (size_t)(i - 1) * n + (kc - 1), val);
                        if (j == n) send(3, (size_t)(i - 1) * n + (kc - 1), val);
                        if (kc == 1) send(4, (size_t)(i - 1) * n + (j - 1), val);
                        if (kc == n) send(5, (size_t)(i - 1) * n + (j - 1), val);
                    }
                }
                xm[q][c] = x0;
            }
        }
        /* every wave is done with buffer bc (plane i): refill it with plane
         * i + 2, which is read two iterations on, past the next barrier */
        __syncthreads();
        if (fetch) {
            commit(bc);
#pragma unroll
            for (int q = 0; q < RPW; ++q)
#pragma unroll
                for (int c = 0; c < KC; ++c) xp[q][c] = pre[q][c];
        }
        bc ^= 1;
    }
}

typedef double dbl2 __attribute__((ext_vector_type(2)));

/* The same step for n a multiple of TK (and so of TJ): the tiles cover the
 * block exactly, so which rows a wavefront reads depends only on the plane
 * and its row index, never on the lane.  Row addresses are then wave-uniform
 * (held in scalar registers), and lane l owns the adjacent columns
 * k0 + 2l, k0 + 2l + 1 of its rows: one 16-byte load and one 16-byte
 * nontemporal store per row (cell k = 1 sits at an even double offset and the
 * row stride, the face slots and n are even, so every access is 16-byte
 * aligned).  Same sums in the same order as k_stencil_t, so the results are
 * bit-identical; at 8 PEs x 256^3 a step takes 0.39 ms against 0.43. */
template <bool FUSED, int TJ, int TK>
__global__ void __launch_bounds__(256) k_stencil_v(const double *u_sym, double *v_sym, double *recv_sym, int n,
                                                   int pe0, int chunk, int pin, int pout) {
    static_assert(TK == 128, "64 lanes x 2 columns");
    constexpr int RPW = TJ / 4; /* rows per wavefront */
    const unsigned gx = gridDim.x, gy = gridDim.y, nwg = gx * gy * gridDim.z;
    unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (nwg % kXcds == 0) lin = (lin % kXcds) * (nwg / kXcds) + lin / kXcds;
    const unsigned bx = lin % gx, by = (lin / gx) % gy, bz = lin / (gx * gy);
    const geom g(n, pe0 + (int)bz);
    const int jtiles = n / TJ;
    const int k0 = (int)bx * TK + 1;
    const int j0 = (int)(by % (unsigned)jtiles) * TJ + 1;
    const int i0 = (int)(by / (unsigned)jtiles) * chunk + 1;
    const int i1 = i0 + chunk - 1 < n ? i0 + chunk - 1 : n;
    const double *__restrict__ u = (const double *)nvshmemx_ptr_d(u_sym, g.pe);
    double *__restrict__ v = (double *)nvshmemx_ptr_d(v_sym, g.pe);
    const double *__restrict__ rin = FUSED ? (const double *)nvshmemx_ptr_d(recv_sym, g.pe) : nullptr;
    const size_t nn = (size_t)n * n;
    __shared__ double tile[2][TJ + 2][TK + 2];
    const int lane = (int)(threadIdx.x & 63u);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int hrow = wv == 0 ? 0 : (wv == 3 ? TJ + 1 : -1); /* tile row of the j halo a wave loads */
    const bool has_halo = threadIdx.x < 2 * (TJ + 2);
    const int hr = (int)(threadIdx.x >> 1), hhi = (int)(threadIdx.x & 1u);
    const int hk = hhi ? k0 + TK : k0 - 1;
    const int kc = k0 + 2 * lane; /* the lane's first column */

    /* cell k of row (i, j) is rowp(i, j)[k]; in the fused step a halo row is
     * a row of a receive slot, and the rows of an edge (two halo
     * coordinates, never used by the stencil) read some valid row instead */
    auto rowp = [&](int i, int j) -> const double * {
        if constexpr (FUSED) {
            const bool hx = i == 0 || i == n + 1, hy = j == 0 || j == n + 1;
            if (hx) return rin + ((size_t)pin * 6 + (i ? 1 : 0)) * nn + (size_t)((hy ? 1 : j) - 1) * n - 1;
            if (hy) return rin + ((size_t)pin * 6 + (j ? 3 : 2)) * nn + (size_t)(i - 1) * n - 1;
        }
        return u + g.idx(i, j, 0);
    };
    auto halo_col = [&](int i, int j) -> double {
        if constexpr (FUSED) {
            if (hk == 0 || hk == n + 1) {
                const int ii = (i == 0 || i == n + 1) ? 1 : i, jj = (j == 0 || j == n + 1) ? 1 : j;
                return rin[((size_t)pin * 6 + (hk ? 5 : 4)) * nn + (size_t)(ii - 1) * n + (jj - 1)];
            }
        }
        return rowp(i, j)[hk];
    };
    auto slot = [&](int d) -> double * { /* receive slot of face d at the neighbour across it */
        return (double *)nvshmemx_ptr_d(recv_sym, g.neighbour(d)) + ((size_t)pout * 6 + (size_t)(d ^ 1)) * nn;
    };

    dbl2 pre[RPW + 1];
    double preh = 0.0;
    auto issue = [&](int i) {
#pragma unroll
        for (int q = 0; q <= RPW; ++q) {
            const int r = q < RPW ? wv * RPW + q + 1 : hrow;
            if (r >= 0) pre[q] = *(const dbl2 *)(rowp(i, j0 - 1 + r) + kc);
        }
        if (has_halo) preh = halo_col(i, j0 - 1 + hr);
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int q = 0; q <= RPW; ++q) {
            const int r = q < RPW ? wv * RPW + q + 1 : hrow;
            if (r >= 0) {
                tile[buf][r][2 * lane + 1] = pre[q].x;
                tile[buf][r][2 * lane + 2] = pre[q].y;
            }
        }
        if (has_halo) tile[buf][hr][hhi ? TK + 1 : 0] = preh;
    };

    dbl2 xm[RPW], xp[RPW]; /* own cells of planes i - 1 and i + 1 */
#pragma unroll
    for (int q = 0; q < RPW; ++q) xm[q] = *(const dbl2 *)(rowp(i0 - 1, j0 + wv * RPW + q) + kc);
    issue(i0);
    commit(0);
    issue(i0 + 1);
    commit(1);
#pragma unroll
    for (int q = 0; q < RPW; ++q) xp[q] = pre[q];
    __syncthreads();
    int bc = 0;
    for (int i = i0; i <= i1; ++i) {
        const bool fetch = i + 2 <= i1 + 1;
        if (fetch) issue(i + 2);
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int r = wv * RPW + q + 1, j = j0 + r - 1, cc = 2 * lane + 1;
            const double x0 = tile[bc][r][cc], x1 = tile[bc][r][cc + 1];
            const double s0 =
                ((((xm[q].x + xp[q].x) + tile[bc][r - 1][cc]) + tile[bc][r + 1][cc]) + tile[bc][r][cc - 1]) + x1;
            const double s1 =
                ((((xm[q].y + xp[q].y) + tile[bc][r - 1][cc + 1]) + tile[bc][r + 1][cc + 1]) + x0) + tile[bc][r][cc + 2];
            dbl2 val;
            val.x = 0.25 * x0 + 0.125 * s0;
            val.y = 0.25 * x1 + 0.125 * s1;
            __builtin_nontemporal_store(val, (dbl2 *)(v + g.idx(i, j, kc)));
            if constexpr (FUSED) {
                if (i == 1) *(dbl2 *)(slot(0) + (size_t)(j - 1) * n + (kc - 1)) = val;
                if (i == n) *(dbl2 *)(slot(1) + (size_t)(j - 1) * n + (kc - 1)) = val;
                if (j == 1) *(dbl2 *)(slot(2) + (size_t)(i - 1) * n + (kc - 1)) = val;
                if (j == n) *(dbl2 *)(slot(3) + (size_t)(i - 1) * n + (kc - 1)) = val;
                if (kc == 1) slot(4)[(size_t)(i - 1) * n + (j - 1)] = val.x;
                if (kc + 1 == n) slot(5)[(size_t)(i - 1) * n + (j - 1)] = val.y;
            }
            xm[q].x = x0;
            xm[q].y = x1;
        }
        /* every wave is done with buffer bc (plane i): refill it with plane
         * i + 2, which is read two iterations on, past the next barrier */
        __syncthreads();
        if (fetch) {
            commit(bc);
#pragma unroll
            for (int q = 0; q < RPW; ++q) xp[q] = pre[q];
        }
        bc ^= 1;
    }
}


/* a coordinate within two cells of an n-cell block (n >= 2), folded onto the
 * PE that holds it: returns the local coordinate and sets `flip` when that is
 * the other PE of the axis; coordinates further out are clamped first */
__device__ __forceinline__ int fold2(int c, int n, int &flip) {
    c = c < -1 ? -1 : (c > n + 2 ? n + 2 : c);
    flip = 0;
    if (c < 1) {
        c += n;
        flip = 1;
    } else if (c > n) {
        c -= n;
        flip = 1;
    }
    return c;
}

/* one cell of either step: the update in its fixed order from the cell's
 * x (planes), y (rows) and z (columns) neighbours */
__device__ __forceinline__ double upd(double c, double xm, double xp, double ym, double yp, double zm, double zp) {
    return 0.25 * c + 0.125 * (((((xm + xp) + ym) + yp) + zm) + zp);
}

/* Two steps per launch (temporal blocking).  A workgroup owns a TJ x 128
 * tile of rows j and columns k over a chunk of planes i and walks the planes.
 * u's cells within two of the tile (rows j0-2 .. j0+TJ+1, columns
 * k0-2 .. k0+129) stream in one plane at a time; the first step is computed
 * on the tile widened by one cell, the second on the tile itself, stored into
 * v.  Cells outside this PE's block are read straight from the neighbour's u
 * through the symmetric heap (in the periodic 2x2x2 grid a coordinate past
 * either end of an axis belongs to the other PE of that axis), so there is no
 * halo exchange, and one barrier per launch (two steps) suffices: a launch
 * only reads u and only writes v, on every PE.  The sums are those of the
 * one-step kernels in the same order, so a launch gives the bits of two steps
 * there.  Coordinates more than two cells past the block are clamped to valid
 * addresses; they feed only cells that are not stored.
 *
 * Each lane keeps its own cells in registers across planes: the x (plane)
 * neighbours of both steps come from registers, and two LDS rings of two
 * slots hold the input plane and the first-step plane whose in-plane (y and
 * z) neighbours are read (47.9 KB: three workgroups per CU).  Wavefront w
 * loads input rows (w + 1 + 4q) mod IR (tile rows - 2), computes the first
 * step on the rows just above them and the second step on the rows above
 * those, so every row's history stays with the lanes that loaded it; lanes
 * of wavefront 3 load the side columns and compute the first step on
 * columns -1 and 128.
 *
 * Per step, HBM traffic is half that of the one-step kernels (u read once and
 * v written once per two steps); each chunk re-reads two planes on either
 * side and each tile two rows and two columns on either side, mostly from L2
 * (tiles adjacent in j run on the same XCD).  At 8 PEs x 256^3 this took
 * 0.210-0.226 ms per step against 0.233-0.238 for the same walk with
 * three-plane LDS rings and every neighbour read from LDS (two workgroups
 * per CU), and 0.284-0.37 with 4-row tiles (three per CU, more re-read
 * rows). */
template <int TJ>
__global__ void __launch_bounds__(256) k_stencil_2s(const double *u_sym, double *v_sym, int n, int pe0, int chunk) {
    constexpr int TK = 128;
    constexpr int IR = TJ + 4; /* input rows: tile rows -2 .. TJ+1 */
    constexpr int MR = TJ + 2; /* first-step rows: tile rows -1 .. TJ */
    constexpr int LW = TK + 8; /* LDS row: tile column c at index c + 4 */
    constexpr int RQ = IR / 4; /* input rows per wavefront */
    static_assert(IR % 4 == 0 && TJ % 4 == 0 && 4 * IR <= 64, "tile rows over 4 wavefronts");
    __shared__ double lin[2][IR][LW];
    __shared__ double lmid[2][MR][LW];
    const unsigned gx = gridDim.x, gy = gridDim.y, nwg = gx * gy * gridDim.z;
    unsigned wl = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (nwg % kXcds == 0) wl = (wl % kXcds) * (nwg / kXcds) + wl / kXcds;
    const unsigned bx = wl % gx, by = (wl / gx) % gy, bz = wl / (gx * gy);
    const geom g(n, pe0 + (int)bz);
    const int jtiles = (n + TJ - 1) / TJ;
    const int k0 = (int)bx * TK + 1;
    const int j0 = (int)(by % (unsigned)jtiles) * TJ + 1;
    const int i0 = (int)(by / (unsigned)jtiles) * chunk + 1;
    const int i1 = i0 + chunk - 1 < n ? i0 + chunk - 1 : n;
    double *__restrict__ v = (double *)nvshmemi::translate(v_sym, g.pe);
    const size_t rs = g.row_stride(), ps = (size_t)(n + 2) * rs;
    const int lane = (int)(threadIdx.x & 63u);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const bool vec = (n & 1) == 0;
    const int kc = k0 + 2 * lane;
    const int ci = 2 * lane + 4;
    int kf0, kf1;
    const int kk0 = fold2(kc, n, kf0), kk1 = fold2(kc + 1, n, kf1);
    const bool pair = vec && kc + 1 <= n;

    int ir[RQ], rj[RQ], rfj[RQ];
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
        ir[q] = (wv + 1 + 4 * q) % IR;
        rj[q] = fold2(j0 - 2 + ir[q], n, rfj[q]);
    }
    /* side input cells: lanes 0..4*IR-1 of wavefront 3, columns -2, -1, 128,
     * 129 (lane & 3) of input row lane >> 2; the lanes on columns -1 and 128
     * of rows 1..IR-2 also compute the first step there */
    const bool side_ld = wv == 3 && lane < 4 * IR;
    const int s_r = lane >> 2, s_c = (lane & 3) < 2 ? (lane & 3) - 2 : TK + (lane & 3) - 2;
    const bool side_mid = side_ld && ((lane & 3) == 1 || (lane & 3) == 2) && s_r >= 1 && s_r <= IR - 2;
    const double *sb0 = u_sym, *sb1 = u_sym;
    if (side_ld) {
        int fj, fk;
        const int jj = fold2(j0 - 2 + s_r, n, fj), kk = fold2(k0 + s_c, n, fk);
        const int pyz = ((g.cy ^ fj) << 1) | (g.cz ^ fk);
        const size_t off = g.idx(0, jj, kk);
        sb0 = (const double *)nvshmemi::translate(u_sym, (g.cx << 2) | pyz) + off;
        sb1 = (const double *)nvshmemi::translate(u_sym, ((g.cx ^ 1) << 2) | pyz) + off;
    }

    /* histories: input planes m-1, m, m+1 (xm, x0, xp) of the lane's rows,
     * first-step planes m-2, m-1, m (ym, y0, yp) */
    dbl2 pre[RQ], xm[RQ], x0[RQ], xp[RQ], ym[RQ], y0[RQ], yp[RQ];
    double pres = 0.0, sxm = 0.0, sx0 = 0.0, sxp = 0.0;
#pragma unroll
    for (int q = 0; q < RQ; ++q) pre[q] = xm[q] = x0[q] = xp[q] = ym[q] = y0[q] = yp[q] = dbl2{0.0, 0.0};
    for (int m = i0 - 4; m <= i1 + 1; ++m) {
        /* the one barrier of an iteration: past it, slot (m+1)&1 of the input
         * ring (last read for plane m-1) and slot m&1 of the first-step ring
         * (last read for plane m-2) may be rewritten, and what the previous
         * iteration wrote (input plane m, first-step plane m-1) is visible */
        lds_barrier();
        if (m >= i0 - 3) { /* input plane m+1 arrived: into the histories and LDS */
            const int sl = (m + 1) & 1;
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                xm[q] = x0[q];
                x0[q] = xp[q];
                xp[q] = pre[q];
                lin[sl][ir[q]][ci] = pre[q].x;
                lin[sl][ir[q]][ci + 1] = pre[q].y;
            }
            if (side_ld) {
                sxm = sx0;
                sx0 = sxp;
                sxp = pres;
                lin[sl][s_r][s_c + 4] = pres;
            }
        }
        if (m + 2 <= i1 + 2) {
            int fi;
            const int ii = fold2(m + 2, n, fi);
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int pxy = ((g.cx ^ fi) << 2) | ((g.cy ^ rfj[q]) << 1);
                const size_t off = (size_t)ii * ps + (size_t)rj[q] * rs + kKOff;
                const double *p0 = (const double *)nvshmemi::translate(u_sym, pxy | g.cz) + off;
                if (pair) {
                    pre[q] = *(const dbl2 *)(p0 + kc);
                } else {
                    const double *p1 = (const double *)nvshmemi::translate(u_sym, pxy | (g.cz ^ 1)) + off;
                    pre[q].x = (kf0 ? p1 : p0)[kk0];
                    pre[q].y = (kf1 ? p1 : p0)[kk1];
                }
            }
            if (side_ld) pres = (fi ? sb1 : sb0)[(size_t)ii * ps];
        }
        if (m >= i0 - 1) { /* first step on plane m */
            const int sl = m & 1;
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int r = ir[q];
                ym[q] = y0[q];
                y0[q] = yp[q];
                if (r >= 1 && r <= IR - 2) {
                    const double a = x0[q].x, b = x0[q].y;
                    yp[q].x = upd(a, xm[q].x, xp[q].x, lin[sl][r - 1][ci], lin[sl][r + 1][ci], lin[sl][r][ci - 1], b);
                    yp[q].y = upd(b, xm[q].y, xp[q].y, lin[sl][r - 1][ci + 1], lin[sl][r + 1][ci + 1], a,
                                  lin[sl][r][ci + 2]);
                    lmid[sl][r - 1][ci] = yp[q].x;
                    lmid[sl][r - 1][ci + 1] = yp[q].y;
                }
            }
            if (side_mid) {
                const int c = s_c + 4;
                lmid[sl][s_r - 1][c] = upd(sx0, sxm, sxp, lin[sl][s_r - 1][c], lin[sl][s_r + 1][c], lin[sl][s_r][c - 1],
                                           lin[sl][s_r][c + 1]);
            }
        }
        if (m >= i0 + 1) { /* second step on plane m-1 */
            const int sl = (m - 1) & 1;
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int r = ir[q] - 1, ro = r - 1; /* first-step row, tile row */
                if (ro >= 0 && ro < TJ) {
                    const double a = y0[q].x, b = y0[q].y;
                    dbl2 val;
                    val.x = upd(a, ym[q].x, yp[q].x, lmid[sl][r - 1][ci], lmid[sl][r + 1][ci], lmid[sl][r][ci - 1], b);
                    val.y = upd(b, ym[q].y, yp[q].y, lmid[sl][r - 1][ci + 1], lmid[sl][r + 1][ci + 1], a,
                                lmid[sl][r][ci + 2]);
                    const int j = j0 + ro;
                    if (j <= n) {
                        double *dst = v + g.idx(m - 1, j, kc);
                        if (pair) {
                            __builtin_nontemporal_store(val, (dbl2 *)dst);
                        } else {
                            if (kc <= n) dst[0] = val.x;
                            if (kc + 1 <= n) dst[1] = val.y;
                        }
                    }
                }
            }
        }
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

/* Stencil tile: 8 rows x 128 columns per workgroup of 256 threads (two
 * columns and two rows per lane).  Against 16 x 64 it halves the j halo rows
 * a plane re-reads per output row; 4 x 256 re-reads more of them again. */
constexpr int kTJ = 8;
constexpr int kTK = 128;

/* Planes per workgroup when the caller passes chunk <= 0.  Measured at 8 PEs
 * x 256^3 (16-byte kernel, ms per fused step): chunk 4 0.404, 6 0.397, 8
 * 0.400, 16 0.407, 64 0.419, 256 0.401 with all PEs in one launch; 8
 * processes, one PE per launch: 4 0.476, 6 0.471, 16 0.477, 256 0.649.
 * Short chunks give many short-lived workgroups, which stream best here; a
 * chunk re-reads two planes per chunk, mostly from L2. */
constexpr int kDefaultChunk = 6;

template <bool FUSED>
static void launch_stencil(double *u_sym, double *v_sym, double *recv_sym, int n, int pe0, int npes_here, int chunk,
                           int pin, int pout, hipStream_t s) {
    if (chunk <= 0) chunk = kDefaultChunk;
    const int jt = (n + kTJ - 1) / kTJ, ic = (n + chunk - 1) / chunk;
    const dim3 grid((unsigned)((n + kTK - 1) / kTK), (unsigned)(jt * ic), (unsigned)npes_here);
    if (n % kTK == 0) /* whole tiles: the 16-byte kernel */
        hipLaunchKernelGGL((k_stencil_v<FUSED, kTJ, kTK>), grid, dim3(256), 0, s, u_sym, v_sym, recv_sym, n, pe0, chunk,
                           pin, pout);
    else
        hipLaunchKernelGGL((k_stencil_t<FUSED, kTJ, kTK>), grid, dim3(256), 0, s, u_sym, v_sym, recv_sym, n, pe0, chunk,
                           pin, pout);
    NVSHMEMI_HIP(hipGetLastError());
}

/* Planes per workgroup of the two-step kernel when the caller passes
 * chunk <= 0: a third of the column.  A chunk recomputes the first step on
 * two planes past its ends and re-reads two input planes on either side, so
 * it wants longer chunks than the one-step kernel, and the workgroup count
 * should fill whole rounds of the 768 resident slots (three per CU).
 * Measured at 8 PEs x 256^3 in one launch (ms per step): chunk 43 0.218,
 * 52 0.238, 64 0.224, 86 0.210, 96 0.225, 128 0.283 (1024 workgroups: one
 * round and a third), 172 0.213, 256 0.219; as 8 processes 64 0.271, 86
 * 0.268, 256 0.325. */
static int default_chunk2(int n) { return (n + 2) / 3; }

/* Tile rows of the two-step kernel (the tile is 128 columns wide). */
constexpr int kTJ2 = 8;

static void launch_stencil2(double *u_sym, double *v_sym, int n, int pe0, int npes_here, int chunk, hipStream_t s) {
    if (chunk <= 0) chunk = default_chunk2(n);
    const int jt = (n + kTJ2 - 1) / kTJ2, ic = (n + chunk - 1) / chunk;
    const dim3 grid((unsigned)((n + kTK - 1) / kTK), (unsigned)(jt * ic), (unsigned)npes_here);
    hipLaunchKernelGGL((k_stencil_2s<kTJ2>), grid, dim3(256), 0, s, u_sym, v_sym, n, pe0, chunk);
    NVSHMEMI_HIP(hipGetLastError());
}

} // namespace

extern "C" {

/* bytes of each symmetric allocation: u / v and the receive area */
size_t nvshmemh_jacobi_field_bytes(int n) { return (size_t)(n + 2) * (n + 2) * (size_t)(n + kRowPad) * sizeof(double); }

/* field layout: cell (i, j, k) is double number ((i * (n+2) + j) * row_stride + k + koff) */
void nvshmemh_jacobi_layout(int n, int *row_stride, int *koff) {
    *row_stride = n + kRowPad;
    *koff = kKOff;
}
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

/* the faces through the library's contiguous and strided device puts */
void nvshmemh_jacobi_pack_put_api(double *u_sym, double *recv_sym, int n, int pe0, int npes_here, int parity,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_pack_put_api, dim3((unsigned)((n + kApiBand - 1) / kApiBand), 6, npes_here), dim3(256), 0, s,
                       u_sym, recv_sym, n, pe0, parity);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_jacobi_unpack(double *u_sym, double *recv_sym, int n, int pe0, int npes_here, int parity,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_unpack, dim3((unsigned)((n * n + 255) / 256), 6, npes_here), dim3(256), 0, s, u_sym,
                       recv_sym, n, pe0, parity);
    NVSHMEMI_HIP(hipGetLastError());
}

void nvshmemh_jacobi_stencil(double *u_sym, double *v_sym, int n, int pe0, int npes_here, int chunk, hipStream_t s) {
    launch_stencil<false>(u_sym, v_sym, nullptr, n, pe0, npes_here, chunk, 0, 0, s);
}

/* One fused step: halo from this PE's receive slots of parity `pin`, v's
 * boundary into the neighbours' slots of parity `pout`. */
void nvshmemh_jacobi_stencil_fused(double *u_sym, double *v_sym, double *recv_sym, int n, int pe0, int npes_here,
                                   int chunk, int pin, int pout, hipStream_t s) {
    launch_stencil<true>(u_sym, v_sym, recv_sym, n, pe0, npes_here, chunk, pin & 1, pout & 1, s);
}

/* `steps` full steps; returns the buffer holding the result (u or v).  The
 * current field's faces are put once into the neighbours' slots of the first
 * step's parity; then every step is one fused stencil (halo in from the slots
 * of its parity, faces out to the slots of the next) followed by the barrier:
 * nvshmem_barrier_all_offload between processes, the kernel boundary in
 * local mode (every PE's work is in the same launch). */
double *nvshmemh_jacobi_run(double *u_sym, double *v_sym, double *recv_sym, int n, int steps, int chunk, int first_step,
                            hipStream_t s) {
    const bool local = nvshmemx_mode() == NVSHMEMX_MODE_LOCAL;
    const int pe0 = local ? 0 : nvshmem_my_pe();
    const int here = local ? nvshmem_n_pes() : 1;
    if (nvshmem_n_pes() != 8) NVSHMEMI_FATAL("jacobi needs exactly 8 PEs (2x2x2), found %d", nvshmem_n_pes());
    check_n(n);
    double *a = u_sym, *b = v_sym;
    if (steps <= 0) return a;
    nvshmemh_jacobi_pack_put(a, recv_sym, n, pe0, here, first_step & 1, s);
    if (!local) nvshmemx_barrier_all_on_stream(s);
    for (int t = 0; t < steps; ++t) {
        const int parity = (first_step + t) & 1;
        nvshmemh_jacobi_stencil_fused(a, b, recv_sym, n, pe0, here, chunk, parity, parity ^ 1, s);
        if (!local) nvshmemx_barrier_all_on_stream(s);
        double *tmp = a;
        a = b;
        b = tmp;
    }
    return a;
}

/* Two steps in one launch (u -> v) for every PE of this process, reading
 * the cells past this PE's block straight from the neighbours' u; every PE's
 * u must be complete before it starts (a barrier between processes). */
void nvshmemh_jacobi_stencil2(double *u_sym, double *v_sym, int n, int pe0, int npes_here, int chunk, hipStream_t s) {
    if (nvshmem_n_pes() != 8) NVSHMEMI_FATAL("jacobi needs exactly 8 PEs (2x2x2), found %d", nvshmem_n_pes());
    if (pe0 < 0 || npes_here < 1 || pe0 + npes_here > 8) NVSHMEMI_FATAL("jacobi: PEs %d..%d outside 0..7", pe0, pe0 + npes_here - 1);
    check_n(n);
    launch_stencil2(u_sym, v_sym, n, pe0, npes_here, chunk, s);
}

/* `steps` steps, two per launch, each launch followed by the barrier
 * (nvshmem_barrier_all_offload between processes, the kernel boundary in
 * local mode); an odd last step is one fused step (pack+put of the faces,
 * barrier, fused stencil, barrier).  No halo is exchanged for the two-step
 * launches.  Returns the buffer holding the result. */
double *nvshmemh_jacobi_run2(double *u_sym, double *v_sym, double *recv_sym, int n, int steps, int chunk,
                             hipStream_t s) {
    const bool local = nvshmemx_mode() == NVSHMEMX_MODE_LOCAL;
    const int pe0 = local ? 0 : nvshmem_my_pe();
    const int here = local ? nvshmem_n_pes() : 1;
    if (nvshmem_n_pes() != 8) NVSHMEMI_FATAL("jacobi needs exactly 8 PEs (2x2x2), found %d", nvshmem_n_pes());
    check_n(n);
    double *a = u_sym, *b = v_sym;
    if (steps <= 0) return a;
    if (steps >= 2 && !local) nvshmemx_barrier_all_on_stream(s);
    for (int t = 0; t + 1 < steps; t += 2) {
        launch_stencil2(a, b, n, pe0, here, chunk, s);
        if (!local) nvshmemx_barrier_all_on_stream(s);
        double *tmp = a;
        a = b;
        b = tmp;
    }
    if (steps & 1) a = nvshmemh_jacobi_run(a, b, recv_sym, n, 1, 0, 0, s);
    return a;
}

/* `steps` unfused steps: the faces of the current field put into the
 * neighbours' receive slots of the step's parity (use_api 0: the direct-store
 * pack kernel; 1: the library's contiguous and strided device puts), the barrier
 * (nvshmem_barrier_all_offload between processes, the kernel boundary in
 * local mode), the halo unpacked into the field, the stencil.  Returns the
 * buffer holding the result. */
double *nvshmemh_jacobi_run_separate(double *u_sym, double *v_sym, double *recv_sym, int n, int steps, int chunk,
                                     int use_api, hipStream_t s) {
    const bool local = nvshmemx_mode() == NVSHMEMX_MODE_LOCAL;
    const int pe0 = local ? 0 : nvshmem_my_pe();
    const int here = local ? nvshmem_n_pes() : 1;
    if (nvshmem_n_pes() != 8) NVSHMEMI_FATAL("jacobi needs exactly 8 PEs (2x2x2), found %d", nvshmem_n_pes());
    check_n(n);
    double *a = u_sym, *b = v_sym;
    for (int t = 0; t < steps; ++t) {
        if (use_api) nvshmemh_jacobi_pack_put_api(a, recv_sym, n, pe0, here, t & 1, s);
        else nvshmemh_jacobi_pack_put(a, recv_sym, n, pe0, here, t & 1, s);
        if (!local) nvshmemx_barrier_all_on_stream(s);
        nvshmemh_jacobi_unpack(a, recv_sym, n, pe0, here, t & 1, s);
        nvshmemh_jacobi_stencil(a, b, n, pe0, here, chunk, s);
        double *tmp = a;
        a = b;
        b = tmp;
    }
    return a;
}

} /* extern "C" */
