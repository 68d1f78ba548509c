// galn_membench.hip -- random-access gather ceiling of the device (pointer chase).
//
// The search's cost is one random 64-byte index block per occurrence lookup,
// dependent on the previous one. This measures the chip's ceiling for exactly
// that access: many independent chains, each loading a whole 64-byte (or
// 128-byte) line whose first word names the next line. The lines form one
// cycle in a pseudo-random order (a 4-round Feistel permutation of the line
// indices, cycle-walked into [0, n_lines)), so no chain revisits a line until
// the whole buffer has been walked.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "galn_hip.h"

namespace {

struct Perm {
    uint64_t n;      // lines
    uint32_t half;   // bits per Feistel half
    uint32_t key[4];
};

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint64_t feistel(const Perm &p, uint64_t v)
{
    const uint64_t mask = (1ull << p.half) - 1;
    uint64_t L = v >> p.half, R = v & mask;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint64_t F = (uint64_t)mix32((uint32_t)R ^ p.key[r]) & mask;
        const uint64_t nl = R;
        R = L ^ F;
        L = nl;
    }
    return (L << p.half) | R;
}

// a bijection on [0, n): walk the Feistel permutation of [0, 4^half) until inside
__device__ __forceinline__ uint64_t perm(const Perm &p, uint64_t j)
{
    uint64_t y = feistel(p, j);
    while (y >= p.n) y = feistel(p, y);
    return y;
}

// line perm(j) points at line perm(j + 1): one cycle through every line
__global__ void k_fill(uint64_t *buf, uint32_t words_per_line, Perm p)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p.n) return;
    const uint64_t here = perm(p, j), next = perm(p, j + 1 == p.n ? 0 : j + 1);
    uint64_t *line = buf + here * words_per_line;
    line[0] = next;
    for (uint32_t w = 1; w < words_per_line; ++w) line[w] = here ^ w;
}

template <int LINE>
__global__ void __launch_bounds__(256) k_chase(const uint64_t *__restrict__ buf, Perm p, uint32_t n_chains, int iters,
                                              unsigned long long *sink)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_chains) return;
    uint64_t cur = perm(p, (uint64_t)c * (p.n / n_chains));
    uint64_t acc = 0;
    constexpr int V = LINE / 16;
    for (int it = 0; it < iters; ++it) {
        const uint4 *q = reinterpret_cast<const uint4 *>(buf + cur * (LINE / 8));
        uint4 v[V];
#pragma unroll
        for (int k = 0; k < V; ++k) v[k] = q[k];
#pragma unroll
        for (int k = 1; k < V; ++k) acc += v[k].x ^ v[k].w;
        cur = (uint64_t)v[0].y << 32 | v[0].x;
    }
    if (acc == 0x123456789abcdefull) atomicAdd(sink, 1ull);  // keeps the loads alive
    if (cur == ~0ull) atomicAdd(sink, 2ull);
}

thread_local char g_err[256];

}  // namespace

extern "C" int galn_hip_gather_ceiling(int device, uint64_t ws_bytes, int line_bytes, int waves_per_cu, int iters,
                                       double *lines_per_s, double *ms)
{
    if ((line_bytes != 64 && line_bytes != 128) || ws_bytes < (uint64_t)line_bytes * 1024 || waves_per_cu < 1 || iters < 1)
        return GALN_HIP_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return GALN_HIP_E_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return GALN_HIP_E_DEVICE;
    Perm p{};
    p.n = ws_bytes / (uint64_t)line_bytes;
    p.half = 1;
    while ((1ull << (2 * p.half)) < p.n) ++p.half;
    p.key[0] = 0x9e3779b9u;
    p.key[1] = 0x7f4a7c15u;
    p.key[2] = 0x85ebca6bu;
    p.key[3] = 0xc2b2ae35u;
    uint64_t *buf = nullptr;
    unsigned long long *sink = nullptr;
    if (hipMalloc(&buf, p.n * (uint64_t)line_bytes) != hipSuccess) return GALN_HIP_E_NOMEM;
    if (hipMalloc(&sink, 8) != hipSuccess) {
        (void)hipFree(buf);
        return GALN_HIP_E_NOMEM;
    }
    int rc = 0;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const uint32_t wpl = (uint32_t)(line_bytes / 8);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((p.n + 255) / 256)), dim3(256), 0, 0, buf, wpl, p);
    const uint64_t want = (uint64_t)prop.multiProcessorCount * (uint64_t)waves_per_cu * 64ull;
    const uint32_t chains = (uint32_t)(want < p.n ? want : p.n);
    const dim3 grid((chains + 255) / 256);
    // warm-up pass, then the timed pass
    for (int pass = 0; pass < 2 && !rc; ++pass) {
        const int it = pass == 0 ? (iters < 64 ? iters : 64) : iters;
        if (pass == 1) (void)hipEventRecord(e0, 0);
        if (line_bytes == 64)
            hipLaunchKernelGGL(k_chase<64>, grid, dim3(256), 0, 0, buf, p, chains, it, sink);
        else
            hipLaunchKernelGGL(k_chase<128>, grid, dim3(256), 0, 0, buf, p, chains, it, sink);
        if (hipGetLastError() != hipSuccess) rc = GALN_HIP_E_DEVICE;
        if (pass == 1) (void)hipEventRecord(e1, 0);
    }
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = GALN_HIP_E_DEVICE;
    float t = 0;
    if (!rc) (void)hipEventElapsedTime(&t, e0, e1);
    if (!rc && lines_per_s) *lines_per_s = t > 0 ? (double)chains * (double)iters / (t * 1e-3) : 0.0;
    if (!rc && ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    (void)hipFree(sink);
    (void)g_err;
    return rc;
}
