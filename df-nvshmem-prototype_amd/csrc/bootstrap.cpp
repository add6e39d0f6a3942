/*
 * bootstrap.cpp -- how PEs find each other at start-up.
 *
 * The reference bootstraps through MPI only (nvshmem.cu:480-481, 509-510,
 * 551-553, 863), or through MVAPICH2-X OpenSHMEM when compiled with
 * _USE_OPENSHMEM_.  Three bootstraps are offered here:
 *
 *  - MPI: used when the program has called MPI_Init (the reference's contract,
 *    README:23-26).  The MPI entry points are weak references, so the library
 *    itself carries no MPI dependency: an MPI program supplies them at link
 *    time, a non-MPI program (or a ctypes loader) leaves them unresolved.
 *  - shared memory: NVSHMEM_BOOTSTRAP=shm with NVSHMEM_BOOTSTRAP_RANK,
 *    NVSHMEM_BOOTSTRAP_NRANKS and NVSHMEM_BOOTSTRAP_NAME (a POSIX shm name
 *    unique to the job).  All PEs share one MI355X, hence one host, so a
 *    POSIX shared-memory segment with a counter/generation barrier replaces the
 *    MPI control plane.
 *  - single process: one PE (or, with NVSHMEM_LOCAL_PES, all PEs in this
 *    process; see host_runtime.cpp).
 */
#define MPICH_SKIP_MPICXX 1
#define OMPI_SKIP_MPICXX 1
#include <mpi.h>

#include <atomic>
#include <errno.h>
#include <fcntl.h>
#include <sched.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include "nvshmemi_host.h"

#pragma weak MPI_Initialized
#pragma weak MPI_Finalized
#pragma weak MPI_Comm_rank
#pragma weak MPI_Comm_size
#pragma weak MPI_Allgather
#pragma weak MPI_Barrier
#pragma weak MPI_Bcast

namespace nvshmemi {

bool mpi_available() {
    if (&MPI_Initialized == nullptr || &MPI_Comm_rank == nullptr) return false;
    int flag = 0;
    if (MPI_Initialized(&flag) != MPI_SUCCESS || !flag) return false;
    int fin = 0;
    if (&MPI_Finalized != nullptr && MPI_Finalized(&fin) == MPI_SUCCESS && fin) return false;
    return true;
}

namespace {

#define NVSHMEMI_MPI(stmt)                                                          \
    do {                                                                            \
        int rc_ = (stmt);                                                           \
        if (rc_ != MPI_SUCCESS) NVSHMEMI_FATAL("mpi call failed with error: %d (%s)", rc_, #stmt); \
    } while (0)

struct MpiBootstrap : Bootstrap {
    int r = 0, n = 1;
    MpiBootstrap() {
        NVSHMEMI_MPI(MPI_Comm_rank(MPI_COMM_WORLD, &r));
        NVSHMEMI_MPI(MPI_Comm_size(MPI_COMM_WORLD, &n));
    }
    const char *name() const override { return "mpi"; }
    int rank() const override { return r; }
    int size() const override { return n; }
    void allgather(const void *in, void *out, size_t bytes) override {
        NVSHMEMI_MPI(MPI_Allgather(in, (int)bytes, MPI_BYTE, out, (int)bytes, MPI_BYTE, MPI_COMM_WORLD));
    }
    void barrier() override { NVSHMEMI_MPI(MPI_Barrier(MPI_COMM_WORLD)); }
    void bcast(void *buf, size_t bytes, int root) override {
        NVSHMEMI_MPI(MPI_Bcast(buf, (int)bytes, MPI_BYTE, root, MPI_COMM_WORLD));
    }
};

struct SingleBootstrap : Bootstrap {
    const char *name() const override { return "single"; }
    int rank() const override { return 0; }
    int size() const override { return 1; }
    void allgather(const void *in, void *out, size_t bytes) override { memcpy(out, in, bytes); }
    void barrier() override {}
    void bcast(void *, size_t, int) override {}
};

/* Shared-memory segment: a barrier (arrival counter + generation) followed by
 * one exchange slot per rank. */
struct ShmHeader {
    std::atomic<uint32_t> count;
    char pad0[60];
    std::atomic<uint32_t> gen;
    char pad1[60];
    std::atomic<uint32_t> finished;
    char pad2[124];
};
static_assert(sizeof(ShmHeader) == 256, "header layout");

constexpr size_t kSlotBytes = 4096;

struct ShmBootstrap : Bootstrap {
    int r = 0, n = 1;
    std::string shm_name;
    size_t bytes = 0;
    char *base = nullptr;
    ShmHeader *hdr = nullptr;
    double timeout_s = 120.0;

    ShmBootstrap() {
        const char *rs = getenv("NVSHMEM_BOOTSTRAP_RANK");
        const char *ns = getenv("NVSHMEM_BOOTSTRAP_NRANKS");
        const char *nm = getenv("NVSHMEM_BOOTSTRAP_NAME");
        if (!rs || !ns || !nm)
            NVSHMEMI_FATAL("NVSHMEM_BOOTSTRAP=shm needs NVSHMEM_BOOTSTRAP_RANK, _NRANKS and _NAME");
        r = atoi(rs);
        n = atoi(ns);
        if (n < 1 || n > NVSHMEMI_MAX_PES || r < 0 || r >= n)
            NVSHMEMI_FATAL("bad shm bootstrap rank %d of %d (at most %d ranks)", r, n, NVSHMEMI_MAX_PES);
        if (const char *t = getenv("NVSHMEM_BOOTSTRAP_TIMEOUT_S")) timeout_s = atof(t);
        shm_name = nm[0] == '/' ? nm : std::string("/") + nm;
        bytes = sizeof(ShmHeader) + (size_t)n * kSlotBytes;
        int fd = shm_open(shm_name.c_str(), O_CREAT | O_RDWR, 0600);
        if (fd < 0) NVSHMEMI_FATAL("shm_open(%s) failed: %s", shm_name.c_str(), strerror(errno));
        struct stat st;
        if (fstat(fd, &st) != 0) NVSHMEMI_FATAL("fstat failed: %s", strerror(errno));
        if ((size_t)st.st_size < bytes && ftruncate(fd, (off_t)bytes) != 0)
            NVSHMEMI_FATAL("ftruncate(%s) failed: %s", shm_name.c_str(), strerror(errno));
        base = (char *)mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
        close(fd);
        if (base == MAP_FAILED) NVSHMEMI_FATAL("mmap of bootstrap segment failed: %s", strerror(errno));
        hdr = (ShmHeader *)base;
    }
    ~ShmBootstrap() override {
        if (base) munmap(base, bytes);
    }
    const char *name() const override { return "shm"; }
    int rank() const override { return r; }
    int size() const override { return n; }
    char *slot(int i) { return base + sizeof(ShmHeader) + (size_t)i * kSlotBytes; }

    void barrier() override {
        const uint32_t g0 = hdr->gen.load(std::memory_order_acquire);
        const uint32_t old = hdr->count.fetch_add(1, std::memory_order_acq_rel);
        if (old == (uint32_t)n - 1) {
            hdr->count.store(0, std::memory_order_relaxed);
            hdr->gen.store(g0 + 1, std::memory_order_release);
            return;
        }
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        unsigned spins = 0;
        while (hdr->gen.load(std::memory_order_acquire) == g0) {
            if (++spins < 1000) {
                sched_yield();
                continue;
            }
            usleep(50);
            if ((spins & 1023) == 0) {
                clock_gettime(CLOCK_MONOTONIC, &t1);
                double dt = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
                if (dt > timeout_s)
                    NVSHMEMI_FATAL("shm bootstrap barrier timed out after %.0f s (%u of %d ranks arrived)", dt,
                                   hdr->count.load() , n);
            }
        }
    }
    void allgather(const void *in, void *out, size_t nbytes) override {
        if (nbytes > kSlotBytes) NVSHMEMI_FATAL("shm allgather of %zu bytes exceeds the %zu-byte slot", nbytes, kSlotBytes);
        memcpy(slot(r), in, nbytes);
        barrier();
        for (int i = 0; i < n; ++i) memcpy((char *)out + (size_t)i * nbytes, slot(i), nbytes);
        barrier();
    }
    void bcast(void *buf, size_t nbytes, int root) override {
        if (nbytes > kSlotBytes) NVSHMEMI_FATAL("shm bcast of %zu bytes exceeds the %zu-byte slot", nbytes, kSlotBytes);
        if (r == root) memcpy(slot(root), buf, nbytes);
        barrier();
        if (r != root) memcpy(buf, slot(root), nbytes);
        barrier();
    }
    void finalize() override {
        barrier();
        if (r == 0) shm_unlink(shm_name.c_str());
    }
};

} // namespace

Bootstrap *bootstrap_create() {
    if (mpi_available()) return new MpiBootstrap();
    const char *b = getenv("NVSHMEM_BOOTSTRAP");
    if (b && strcmp(b, "shm") == 0) return new ShmBootstrap();
    if (b && strcmp(b, "mpi") == 0)
        NVSHMEMI_FATAL("NVSHMEM_BOOTSTRAP=mpi but MPI is not initialised (call MPI_Init before nvstart_pes)");
    return new SingleBootstrap();
}

} // namespace nvshmemi
