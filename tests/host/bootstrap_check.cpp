/*
 * bootstrap_check.cpp -- the shared-memory bootstrap (csrc/bootstrap.cpp)
 * exercised by real processes, built with AddressSanitizer and
 * UndefinedBehaviorSanitizer (host code only; no GPU is touched).
 *
 * Usage: bootstrap_check NRANKS ROUNDS.  Forks NRANKS ranks sharing one
 * segment; each runs ROUNDS of allgather (rank-dependent payloads), bcast
 * from a rotating root and barrier, checks every byte it receives, then
 * finalizes.  Exit status 0 when every rank saw the right data.
 */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/wait.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "nvshmemi_host.h"

namespace nvshmemi {
/* the library's fatal() lives in host_runtime.hip; this checker is host-only */
void fatal(const char *file, int line, const char *fmt, ...) {
    fprintf(stderr, "[%s:%d] ", file, line);
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
    _exit(3);
}
} // namespace nvshmemi

static unsigned char byte_of(int rank, int round, size_t i) { return (unsigned char)(rank * 131 + round * 7 + i * 13); }

static int run_rank(int rank, int nranks, int rounds, const char *name) {
    setenv("NVSHMEM_BOOTSTRAP", "shm", 1);
    setenv("NVSHMEM_BOOTSTRAP_RANK", std::to_string(rank).c_str(), 1);
    setenv("NVSHMEM_BOOTSTRAP_NRANKS", std::to_string(nranks).c_str(), 1);
    setenv("NVSHMEM_BOOTSTRAP_NAME", name, 1);
    setenv("NVSHMEM_BOOTSTRAP_TIMEOUT_S", "30", 1);
    nvshmemi::Bootstrap *b = nvshmemi::bootstrap_create();
    if (b->rank() != rank || b->size() != nranks || strcmp(b->name(), "shm") != 0) return 4;
    int bad = 0;
    for (int r = 0; r < rounds; ++r) {
        const size_t n = 1 + (size_t)(r * 97) % 4000; /* up to the 4096-byte slot */
        std::vector<unsigned char> mine(n), all(n * (size_t)nranks);
        for (size_t i = 0; i < n; ++i) mine[i] = byte_of(rank, r, i);
        b->allgather(mine.data(), all.data(), n);
        for (int q = 0; q < nranks; ++q)
            for (size_t i = 0; i < n; ++i) bad += all[(size_t)q * n + i] != byte_of(q, r, i);
        const int root = r % nranks;
        std::vector<unsigned char> buf(n, 0);
        if (rank == root)
            for (size_t i = 0; i < n; ++i) buf[i] = byte_of(root, r + 1000, i);
        b->bcast(buf.data(), n, root);
        for (size_t i = 0; i < n; ++i) bad += buf[i] != byte_of(root, r + 1000, i);
        b->barrier();
    }
    b->finalize();
    delete b;
    return bad ? 5 : 0;
}

int main(int argc, char **argv) {
    const int nranks = argc > 1 ? atoi(argv[1]) : 4;
    const int rounds = argc > 2 ? atoi(argv[2]) : 200;
    const std::string name = "/nvshmem_bootcheck_" + std::to_string((long)getpid());
    std::vector<pid_t> kids;
    for (int r = 0; r < nranks; ++r) {
        const pid_t p = fork();
        if (p == 0) _exit(run_rank(r, nranks, rounds, name.c_str()));
        kids.push_back(p);
    }
    int failed = 0;
    for (pid_t p : kids) {
        int st = 0;
        waitpid(p, &st, 0);
        if (!WIFEXITED(st) || WEXITSTATUS(st) != 0) {
            fprintf(stderr, "rank process %d ended with status 0x%x\n", (int)p, st);
            failed = 1;
        }
    }
    printf("%s: %d ranks x %d rounds of allgather / bcast / barrier\n", failed ? "FAILED" : "ok", nranks, rounds);
    return failed;
}
