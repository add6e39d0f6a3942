/*
 * costmodel.hip -- the learned part of the batched put: a small fp32 MLP
 * that predicts, for one request, the time of each of the 54 autotuner
 * candidates, so a batch can be planned without measuring (no counterpart in
 * the reference).
 *
 *   inputs  (4): log2 bytes, element width, concurrent requests, alignment
 *                class, normalised by nvshmemx_tune_features
 *   hidden     : 32 ReLU -> 32 ReLU
 *   outputs (C): log(time in microseconds) of each candidate
 *
 * Parameters, in one device buffer: W1[32][4] b1[32] W2[32][32] b2[32]
 * W3[C][32] b3[C].  Loss: mean over examples and outputs of the squared error
 * of the log-times.  Training: full-batch Adam on the GPU; each step is two
 * kernels (gradient, update) and runs of steps are captured once into a
 * hipGraph and replayed.  The step counter for Adam's bias correction lives in
 * device memory and is advanced by the gradient kernel, so a replayed graph
 * needs no host arguments; the update kernel clears the gradient buffer, so the
 * graph holds kernel nodes only.
 *
 * Work per example is about 2 * (4*32 + 32*32 + 32*C) = 5.8 kFLOP (C = 54): the
 * network is latency-bound, not MFMA-bound, so the kernels are plain fp32 with
 * the weights staged in LDS.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "nvshmem.h"
#include "nvshmemx.h"
#include "nvshmemi_device.h"
#include "nvshmemi_host.h"

using namespace nvshmemi;

namespace {

constexpr int IN = 4, HID = 32, MAXC = 64;
constexpr int EX_PER_WG = 64; /* examples per workgroup in the gradient kernel */

__host__ __device__ constexpr int n_params(int c) { return HID * IN + HID + HID * HID + HID + c * HID + c; }

struct offsets {
    int w1, b1, w2, b2, w3, b3, n;
    __host__ __device__ explicit offsets(int c)
        : w1(0), b1(HID * IN), w2(HID * IN + HID), b2(HID * IN + HID + HID * HID), w3(b2 + HID), b3(w3 + c * HID),
          n(n_params(c)) {}
};

/* forward for one example with parameters in LDS; h1, h2, y in registers */
__device__ __forceinline__ void forward_one(const float *P, const offsets &o, int C, const float *x, float *h1,
                                            float *h2, float *y) {
#pragma unroll
    for (int k = 0; k < HID; ++k) {
        float s = P[o.b1 + k];
#pragma unroll
        for (int i = 0; i < IN; ++i) s += P[o.w1 + k * IN + i] * x[i];
        h1[k] = s > 0.f ? s : 0.f;
    }
#pragma unroll
    for (int m = 0; m < HID; ++m) {
        float s = P[o.b2 + m];
#pragma unroll
        for (int k = 0; k < HID; ++k) s += P[o.w2 + m * HID + k] * h1[k];
        h2[m] = s > 0.f ? s : 0.f;
    }
    for (int j = 0; j < C; ++j) {
        float s = P[o.b3 + j];
#pragma unroll
        for (int m = 0; m < HID; ++m) s += P[o.w3 + j * HID + m] * h2[m];
        y[j] = s;
    }
}

__global__ void __launch_bounds__(256) k_mlp_forward(const float *params, int C, const float *X, int N, float *Y) {
    __shared__ float P[n_params(MAXC)];
    const offsets o(C);
    for (int i = threadIdx.x; i < o.n; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    float x[IN], h1[HID], h2[HID], y[MAXC];
#pragma unroll
    for (int i = 0; i < IN; ++i) x[i] = X[(size_t)e * IN + i];
    forward_one(P, o, C, x, h1, h2, y);
    for (int j = 0; j < C; ++j) Y[(size_t)e * C + j] = y[j];
}

/* Gradient of the mean squared error over all N examples, accumulated into
 * `grad` (which the update kernel clears).  Threads 0..63 each own one
 * example for the per-example forward/backward; then all 256 threads reduce
 * the weight gradients over the workgroup's examples from LDS and add them to
 * `grad` with one atomic per parameter per workgroup. */
__global__ void __launch_bounds__(256) k_mlp_grad(const float *params, int C, const float *X, const float *T, int N,
                                                  float *grad, int *step) {
    __shared__ float P[n_params(MAXC)];
    __shared__ float sx[EX_PER_WG][IN];
    __shared__ float sh1[EX_PER_WG][HID];
    __shared__ float sdh1[EX_PER_WG][HID];
    __shared__ float sh2[EX_PER_WG][HID];
    __shared__ float sdh2[EX_PER_WG][HID];
    __shared__ float sdy[EX_PER_WG][MAXC];
    const offsets o(C);
    if (blockIdx.x == 0 && threadIdx.x == 0 && step) step[0] += 1;
    for (int i = threadIdx.x; i < o.n; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const int t = threadIdx.x;
    const int e = blockIdx.x * EX_PER_WG + t;
    const float scale = 2.0f / ((float)N * (float)C);
    if (t < EX_PER_WG) {
        float x[IN], h1[HID], h2[HID], y[MAXC];
        if (e < N) {
#pragma unroll
            for (int i = 0; i < IN; ++i) x[i] = X[(size_t)e * IN + i];
            forward_one(P, o, C, x, h1, h2, y);
            for (int j = 0; j < C; ++j) sdy[t][j] = scale * (y[j] - T[(size_t)e * C + j]);
        } else {
#pragma unroll
            for (int i = 0; i < IN; ++i) x[i] = 0.f;
#pragma unroll
            for (int k = 0; k < HID; ++k) h1[k] = h2[k] = 0.f;
            for (int j = 0; j < C; ++j) sdy[t][j] = 0.f;
        }
#pragma unroll
        for (int i = 0; i < IN; ++i) sx[t][i] = x[i];
#pragma unroll
        for (int k = 0; k < HID; ++k) sh1[t][k] = h1[k], sh2[t][k] = h2[k];
        /* backward through layer 3 and the second ReLU */
        float dh2[HID];
#pragma unroll
        for (int m = 0; m < HID; ++m) {
            float s = 0.f;
            for (int j = 0; j < C; ++j) s += sdy[t][j] * P[o.w3 + j * HID + m];
            dh2[m] = h2[m] > 0.f ? s : 0.f;
            sdh2[t][m] = dh2[m];
        }
        /* through layer 2 and the first ReLU */
#pragma unroll
        for (int k = 0; k < HID; ++k) {
            float s = 0.f;
#pragma unroll
            for (int m = 0; m < HID; ++m) s += dh2[m] * P[o.w2 + m * HID + k];
            sdh1[t][k] = h1[k] > 0.f ? s : 0.f;
        }
    }
    __syncthreads();
    /* weight gradients: sum over this workgroup's examples */
    for (int idx = threadIdx.x; idx < o.n; idx += blockDim.x) {
        float s = 0.f;
        if (idx < o.b1) { /* W1[k][i] */
            const int k = idx / IN, i = idx % IN;
            for (int q = 0; q < EX_PER_WG; ++q) s += sdh1[q][k] * sx[q][i];
        } else if (idx < o.w2) { /* b1[k] */
            const int k = idx - o.b1;
            for (int q = 0; q < EX_PER_WG; ++q) s += sdh1[q][k];
        } else if (idx < o.b2) { /* W2[m][k] */
            const int m = (idx - o.w2) / HID, k = (idx - o.w2) % HID;
            for (int q = 0; q < EX_PER_WG; ++q) s += sdh2[q][m] * sh1[q][k];
        } else if (idx < o.w3) { /* b2[m] */
            const int m = idx - o.b2;
            for (int q = 0; q < EX_PER_WG; ++q) s += sdh2[q][m];
        } else if (idx < o.b3) { /* W3[j][m] */
            const int j = (idx - o.w3) / HID, m = (idx - o.w3) % HID;
            for (int q = 0; q < EX_PER_WG; ++q) s += sdy[q][j] * sh2[q][m];
        } else { /* b3[j] */
            const int j = idx - o.b3;
            for (int q = 0; q < EX_PER_WG; ++q) s += sdy[q][j];
        }
        atomicAdd(&grad[idx], s);
    }
}

/* Adam (Kingma & Ba, 2015) with bias correction; clears the gradient. */
__global__ void k_adam(float *params, float *grad, float *m, float *v, int n, float lr, const int *step) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int t = step[0];
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)t));
    const float vh = vi / (1.f - powf(b2, (float)t));
    params[i] -= lr * mh / (sqrtf(vh) + eps);
    grad[i] = 0.f;
}

uint32_t init_rng(uint64_t &st) {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    const uint32_t xs = (uint32_t)(((st >> 18u) ^ st) >> 27u), rot = (uint32_t)(st >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

} // namespace

struct nvshmemx_costmodel {
    int C = 54;
    int np = 0;
    float *d_params = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
    int *d_step = nullptr;
    float *d_X = nullptr, *d_T = nullptr, *d_Y = nullptr;
    int cap = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0, graph_n = 0;
    float graph_lr = 0;
};

static void ensure_data(nvshmemx_costmodel *m, int N) {
    if (N <= m->cap) return;
    if (m->d_X) NVSHMEMI_HIP(hipFree(m->d_X));
    if (m->d_T) NVSHMEMI_HIP(hipFree(m->d_T));
    if (m->d_Y) NVSHMEMI_HIP(hipFree(m->d_Y));
    m->cap = N;
    NVSHMEMI_HIP(hipMalloc((void **)&m->d_X, (size_t)N * IN * sizeof(float)));
    NVSHMEMI_HIP(hipMalloc((void **)&m->d_T, (size_t)N * m->C * sizeof(float)));
    NVSHMEMI_HIP(hipMalloc((void **)&m->d_Y, (size_t)N * m->C * sizeof(float)));
    if (m->exec) { /* the captured graph refers to the old buffers */
        NVSHMEMI_HIP(hipGraphExecDestroy(m->exec));
        NVSHMEMI_HIP(hipGraphDestroy(m->graph));
        m->exec = nullptr;
        m->graph = nullptr;
    }
}

extern "C" int nvshmemx_costmodel_create(int ncand, unsigned long long seed, nvshmemx_costmodel_t *out) {
    if (ncand < 1 || ncand > MAXC) NVSHMEMI_FATAL("cost model outputs %d not in [1, %d]", ncand, MAXC);
    nvshmemx_costmodel *m = new nvshmemx_costmodel();
    m->C = ncand;
    m->np = n_params(ncand);
    NVSHMEMI_HIP(hipMalloc((void **)&m->d_params, m->np * sizeof(float)));
    NVSHMEMI_HIP(hipMalloc((void **)&m->d_grad, m->np * sizeof(float)));
    NVSHMEMI_HIP(hipMalloc((void **)&m->d_m, m->np * sizeof(float)));
    NVSHMEMI_HIP(hipMalloc((void **)&m->d_v, m->np * sizeof(float)));
    NVSHMEMI_HIP(hipMalloc((void **)&m->d_step, sizeof(int)));
    NVSHMEMI_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    /* He-uniform weights, zero biases, from a seeded PCG stream */
    const offsets o(ncand);
    std::vector<float> p((size_t)m->np, 0.f);
    uint64_t st = seed * 2654435761ULL + 1;
    auto fill = [&](int off, int count, int fan_in) {
        const float lim = sqrtf(6.0f / (float)fan_in);
        for (int i = 0; i < count; ++i) p[(size_t)(off + i)] = lim * (2.0f * (float)init_rng(st) / 4294967296.0f - 1.0f);
    };
    fill(o.w1, HID * IN, IN);
    fill(o.w2, HID * HID, HID);
    fill(o.w3, ncand * HID, HID);
    NVSHMEMI_HIP(hipMemcpy(m->d_params, p.data(), m->np * sizeof(float), hipMemcpyHostToDevice));
    NVSHMEMI_HIP(hipMemset(m->d_grad, 0, m->np * sizeof(float)));
    NVSHMEMI_HIP(hipMemset(m->d_m, 0, m->np * sizeof(float)));
    NVSHMEMI_HIP(hipMemset(m->d_v, 0, m->np * sizeof(float)));
    NVSHMEMI_HIP(hipMemset(m->d_step, 0, sizeof(int)));
    *out = m;
    return 0;
}

extern "C" int nvshmemx_costmodel_num_params(nvshmemx_costmodel_t m) { return m->np; }
extern "C" int nvshmemx_costmodel_num_outputs(nvshmemx_costmodel_t m) { return m->C; }

extern "C" void nvshmemx_costmodel_get_params(nvshmemx_costmodel_t m, float *out) {
    NVSHMEMI_HIP(hipMemcpy(out, m->d_params, m->np * sizeof(float), hipMemcpyDeviceToHost));
}

extern "C" void nvshmemx_costmodel_set_params(nvshmemx_costmodel_t m, const float *in) {
    NVSHMEMI_HIP(hipMemcpy(m->d_params, in, m->np * sizeof(float), hipMemcpyHostToDevice));
}

/* One full-batch gradient at the current parameters (no update). */
extern "C" void nvshmemx_costmodel_gradient(nvshmemx_costmodel_t m, const float *X, const float *T, int N,
                                            float *grad_out) {
    ensure_data(m, N);
    NVSHMEMI_HIP(hipMemcpy(m->d_X, X, (size_t)N * IN * sizeof(float), hipMemcpyHostToDevice));
    NVSHMEMI_HIP(hipMemcpy(m->d_T, T, (size_t)N * m->C * sizeof(float), hipMemcpyHostToDevice));
    NVSHMEMI_HIP(hipMemset(m->d_grad, 0, m->np * sizeof(float)));
    hipLaunchKernelGGL(k_mlp_grad, dim3((N + EX_PER_WG - 1) / EX_PER_WG), dim3(256), 0, nullptr, m->d_params, m->C,
                       m->d_X, m->d_T, N, m->d_grad, (int *)nullptr);
    NVSHMEMI_HIP(hipGetLastError());
    NVSHMEMI_HIP(hipMemcpy(grad_out, m->d_grad, m->np * sizeof(float), hipMemcpyDeviceToHost));
    NVSHMEMI_HIP(hipMemset(m->d_grad, 0, m->np * sizeof(float)));
}

static void enqueue_step(nvshmemx_costmodel *m, int N, float lr, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_grad, dim3((N + EX_PER_WG - 1) / EX_PER_WG), dim3(256), 0, s, m->d_params, m->C, m->d_X,
                       m->d_T, N, m->d_grad, m->d_step);
    hipLaunchKernelGGL(k_adam, dim3((m->np + 255) / 256), dim3(256), 0, s, m->d_params, m->d_grad, m->d_m, m->d_v,
                       m->np, lr, m->d_step);
}

/* `steps` Adam steps on (X, T).  Steps run from a captured graph of up to
 * 100 steps; the graph is re-captured only when N or lr changes. */
extern "C" void nvshmemx_costmodel_train(nvshmemx_costmodel_t m, const float *X, const float *T, int N, int steps,
                                         float lr) {
    if (N < 1 || steps < 0) NVSHMEMI_FATAL("nvshmemx_costmodel_train: bad arguments");
    ensure_data(m, N);
    NVSHMEMI_HIP(hipMemcpy(m->d_X, X, (size_t)N * IN * sizeof(float), hipMemcpyHostToDevice));
    NVSHMEMI_HIP(hipMemcpy(m->d_T, T, (size_t)N * m->C * sizeof(float), hipMemcpyHostToDevice));
    const int chunk = 100;
    if (!m->exec || m->graph_n != N || m->graph_lr != lr) {
        if (m->exec) {
            NVSHMEMI_HIP(hipGraphExecDestroy(m->exec));
            NVSHMEMI_HIP(hipGraphDestroy(m->graph));
        }
        NVSHMEMI_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < chunk; ++i) enqueue_step(m, N, lr, m->stream);
        NVSHMEMI_HIP(hipStreamEndCapture(m->stream, &m->graph));
        NVSHMEMI_HIP(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
        m->graph_n = N;
        m->graph_lr = lr;
    }
    int done = 0;
    for (; done + chunk <= steps; done += chunk) NVSHMEMI_HIP(hipGraphLaunch(m->exec, m->stream));
    for (; done < steps; ++done) enqueue_step(m, N, lr, m->stream);
    NVSHMEMI_HIP(hipStreamSynchronize(m->stream));
    NVSHMEMI_HIP(hipGetLastError());
}

extern "C" void nvshmemx_costmodel_predict(nvshmemx_costmodel_t m, const float *X, int N, float *Y) {
    if (N < 1) return;
    ensure_data(m, N);
    NVSHMEMI_HIP(hipMemcpy(m->d_X, X, (size_t)N * IN * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_forward, dim3((N + 255) / 256), dim3(256), 0, nullptr, m->d_params, m->C, m->d_X, N,
                       m->d_Y);
    NVSHMEMI_HIP(hipGetLastError());
    NVSHMEMI_HIP(hipMemcpy(Y, m->d_Y, (size_t)N * m->C * sizeof(float), hipMemcpyDeviceToHost));
}

/* Per-request choice: the candidate with the smallest predicted time. */
extern "C" int nvshmemx_costmodel_choose(nvshmemx_costmodel_t m, const nvshmemx_put_request_t *reqs, int n,
                                         int concurrency, nvshmemx_batch_choice_t *out) {
    if (n < 1) return 0;
    std::vector<float> X((size_t)n * IN), Y((size_t)n * m->C);
    size_t total = 0;
    for (int i = 0; i < n; ++i) total += reqs[i].nbytes;
    for (int i = 0; i < n; ++i) {
        const void *dst = g.initialized ? nvshmem_ptr(reqs[i].target, reqs[i].pe) : reqs[i].target;
        nvshmemx_tune_features(log2((double)std::max<size_t>(1, reqs[i].nbytes)), reqs[i].elem_bytes,
                               nvshmemx_effective_concurrency(reqs[i].nbytes, total, concurrency),
                               nvshmemx_align_class(reqs[i].source, dst), &X[(size_t)i * IN]);
    }
    nvshmemx_costmodel_predict(m, X.data(), n, Y.data());
    const size_t cap = balance_cap(total);
    for (int i = 0; i < n; ++i) {
        const float *y = &Y[(size_t)i * m->C];
        int best = 0;
        if (m->C == nvshmemx_tune_candidate_count()) {
            best = pick_candidate(y, reqs[i].nbytes, cap, true);
        } else {
            for (int j = 1; j < m->C; ++j)
                if (y[j] < y[best]) best = j;
        }
        nvshmemx_tune_candidate(best, &out[i]);
    }
    return 0;
}

extern "C" int nvshmemx_costmodel_save(nvshmemx_costmodel_t m, const char *path) {
    std::vector<float> p((size_t)m->np);
    nvshmemx_costmodel_get_params(m, p.data());
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const int hdr[2] = {m->C, m->np};
    const bool ok = fwrite(hdr, sizeof(hdr), 1, f) == 1 && fwrite(p.data(), sizeof(float), p.size(), f) == p.size();
    fclose(f);
    return ok ? 0 : -1;
}

extern "C" int nvshmemx_costmodel_load(nvshmemx_costmodel_t m, const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int hdr[2] = {0, 0};
    std::vector<float> p((size_t)m->np);
    const bool ok = fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == m->C && hdr[1] == m->np &&
                    fread(p.data(), sizeof(float), p.size(), f) == p.size();
    fclose(f);
    if (!ok) return -1;
    nvshmemx_costmodel_set_params(m, p.data());
    return 0;
}

extern "C" void nvshmemx_costmodel_destroy(nvshmemx_costmodel_t m) {
    if (!m) return;
    NVSHMEMI_HIP(hipDeviceSynchronize());
    if (m->exec) NVSHMEMI_HIP(hipGraphExecDestroy(m->exec));
    if (m->graph) NVSHMEMI_HIP(hipGraphDestroy(m->graph));
    NVSHMEMI_HIP(hipStreamDestroy(m->stream));
    for (float *p : {m->d_params, m->d_grad, m->d_m, m->d_v, m->d_X, m->d_T, m->d_Y})
        if (p) NVSHMEMI_HIP(hipFree(p));
    NVSHMEMI_HIP(hipFree(m->d_step));
    delete m;
}
