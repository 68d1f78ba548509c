// galn_index.hip -- BWA .bwt (and .sa samples) of T$ with T = F . revcomp(F), built on one
// MI355X (gfx950) for texts of any length that fit its HBM, 2^32 symbols and beyond.
//
// Reference behaviour: src/index/cpu_ref.c:56-106 (BWT of T$ from a suffix array, $ removed,
// primary = row of the suffix T[0..]), src/index/idxcommon.c:72-135 (the 64-byte block layout),
// upstream bntseq.c (T = F . revcomp(F)); the GPU builder src/index/gpu_sa.cu does the same job
// with its own chunked sort. Any correct suffix order gives the same bytes.
//
// Design (memory-lean, one pass over the suffixes per chunk):
//   * T is packed 2 bits per symbol, 32 per u64 word, first symbol in the top bits: n/4 bytes.
//   * A suffix's sort key is its next 27 symbols in base 5, digit = symbol + 1, and 0 past the
//     end of T. Because 0 sorts below every symbol, comparing padded keys orders suffixes exactly
//     as T$ does ($ smallest), and two different suffixes never tie once the keys reach past the
//     shorter one's end. 5^27 < 2^63: one 64-bit radix key per 27 symbols.
//   * Suffixes are split by their first 10 key digits (5^10 buckets, one histogram pass) into
//     chunks of consecutive buckets that fit the free memory (at most 2^31 suffixes). A chunk's
//     suffixes are collected by a scan of T, sorted by their first key (rocPRIM radix sort), and
//     the runs of equal keys are refined 27 symbols deeper each round: the unresolved elements
//     are re-keyed at the new depth and put in (group, key) order by two stable radix sorts
//     (key, then group), until every run is a single suffix. A chunk's rows are contiguous, so
//     its BWT symbols and SA samples are written once it is done.
//   * The BWT is kept by row, the $ row included (2 bits per row), and the .bwt body is packed
//     from it at the end with the $ row left out: per 128-symbol block, the exclusive counts
//     (prefix scans over per-block counts) and 8 words of 16 symbols, first symbol in bits 31-30.
// Device memory: n/4 (text) + n/4 (BWT by row) + n/4 with SA samples + ~90 bytes per suffix of
// the largest chunk, then n/2 for the packed body.

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "galn_hip.h"

__attribute__((visibility("hidden"))) void galn_hip_set_error(const char *msg);

namespace {

constexpr int KD = 27;                         // base-5 digits per sort key (5^27 < 2^63)
constexpr int BD = 10;                         // digits of the bucket id
constexpr uint32_t NBUCKET = 9765625;          // 5^10
constexpr uint64_t CHUNK_MAX = 1ull << 31;     // suffixes per chunk
constexpr uint64_t PIECE = 256ull << 20;       // forward bases uploaded per copy

int ifail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    galn_hip_set_error(buf);
    return code;
}

#define IHIP(call)                                                                                     \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return ifail(e_ == hipErrorOutOfMemory ? GALN_HIP_E_NOMEM : GALN_HIP_E_DEVICE, "%s:%d %s: %s", \
                         __FILE__, __LINE__, #call, hipGetErrorString(e_));                            \
    } while (0)

unsigned grid_for(uint64_t n, unsigned block = 256)
{
    const uint64_t g = (n + block - 1) / block;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, 1u << 20));
}

// symbol of T at position p (p < n)
__device__ __forceinline__ uint32_t tsym(const uint64_t *P, uint64_t p)
{
    return (uint32_t)(P[p >> 5] >> (62 - 2 * (p & 31))) & 3u;
}

// `digits` base-5 digits of the suffix at p: T[p + j] + 1 while p + j < n, then 0. P carries two
// zero words past its end, so the two-word window load stays in bounds.
__device__ __forceinline__ uint64_t key_at(const uint64_t *P, uint64_t n, uint64_t p, int digits)
{
    if (p >= n) return 0;
    const uint64_t wi = p >> 5;
    const int off = (int)(p & 31);
    const uint64_t hi = P[wi], lo = P[wi + 1];
    const uint64_t win = off ? (hi << (2 * off)) | (lo >> (64 - 2 * off)) : hi;
    const uint64_t rem = n - p;
    const int valid = rem < (uint64_t)digits ? (int)rem : digits;
    uint64_t key = 0;
    for (int j = 0; j < digits; ++j) {
        const uint64_t d = j < valid ? ((win >> (62 - 2 * j)) & 3u) + 1 : 0;
        key = key * 5 + d;
    }
    return key;
}

// T's words from one piece F[a, a + cnt) of the forward codes: the forward image T[a, a + cnt)
// and the reverse-complement image T[n - a - cnt, n - a) (T[n - 1 - p] = 3 - F[p]). Words the
// piece covers whole are stored, the others OR-ed in (T starts zeroed).
__global__ void k_pack_text(uint64_t *P, uint64_t n, const uint8_t *F, uint64_t a, uint64_t cnt)
{
    const uint64_t fw0 = a >> 5, fw1 = (a + cnt - 1) >> 5;
    const uint64_t rlo = n - a - cnt, rhi = n - a;  // [rlo, rhi)
    const uint64_t rw0 = rlo >> 5, rw1 = (rhi - 1) >> 5;
    const uint64_t nf = fw1 - fw0 + 1, nr = rw1 - rw0 + 1;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < nf + nr; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t v = 0;
        uint64_t w, lo, hi;
        if (t < nf) {
            w = fw0 + t;
            lo = std::max(a, w * 32);
            hi = std::min(a + cnt, w * 32 + 32);
            for (uint64_t p = lo; p < hi; ++p) v |= (uint64_t)(F[p - a] & 3u) << (62 - 2 * (p & 31));
        } else {
            w = rw0 + (t - nf);
            lo = std::max(rlo, w * 32);
            hi = std::min(rhi, w * 32 + 32);
            for (uint64_t q = lo; q < hi; ++q) v |= (uint64_t)(3u - (F[(n - 1 - q) - a] & 3u)) << (62 - 2 * (q & 31));
        }
        if (hi - lo == 32)
            P[w] = v;
        else if (v)
            atomicOr(reinterpret_cast<unsigned long long *>(P + w), (unsigned long long)v);
    }
}

// suffixes per bucket (first BD key digits)
__global__ void k_hist(const uint64_t *P, uint64_t n, unsigned long long *hist)
{
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(hist + key_at(P, n, p, BD), 1ull);
}

// positions whose bucket lies in [b0, b1), appended (one counter update per wave); all lanes of a
// wave run the same iterations, so the wave-wide ballot sees every lane
__global__ void k_collect(const uint64_t *P, uint64_t n, uint32_t b0, uint32_t b1, uint64_t *out,
                          unsigned long long *cnt)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    for (uint64_t base = blockIdx.x * (uint64_t)blockDim.x; base < n; base += stride) {
        const uint64_t p = base + threadIdx.x;
        bool take = false;
        if (p < n) {
            const uint64_t b = key_at(P, n, p, BD);
            take = b >= b0 && b < b1;
        }
        const uint64_t m = __ballot(take);
        if (m == 0) continue;
        unsigned long long at = 0;
        if (lane == __ffsll((long long)m) - 1) at = atomicAdd(cnt, (unsigned long long)__popcll(m));
        at = __shfl(at, __ffsll((long long)m) - 1);
        if (take) out[at + __popcll(m & ((1ull << lane) - 1))] = p;
    }
}

__global__ void k_keys(const uint64_t *P, uint64_t n, const uint64_t *pos, uint64_t m, uint64_t depth, uint64_t *key)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m; j += (uint64_t)gridDim.x * blockDim.x)
        key[j] = key_at(P, n, pos[j] + depth, KD);
}

// after the first sort: run heads, then the unresolved flags (members of runs of two or more)
__global__ void k_heads(const uint64_t *K, uint64_t m, uint8_t *H)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m; j += (uint64_t)gridDim.x * blockDim.x)
        H[j] = (j == 0 || K[j] != K[j - 1]) ? 1 : 0;
}

__global__ void k_unres_all(const uint8_t *H, uint64_t m, uint8_t *U)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m; j += (uint64_t)gridDim.x * blockDim.x)
        U[j] = (H[j] && (j + 1 == m || H[j + 1])) ? 0 : 1;
}

// refinement: element t of the unresolved list (chunk slot idx[t]) gets its key at `depth`
__global__ void k_rkeys(const uint64_t *P, uint64_t n, const uint64_t *V, const uint32_t *idx, uint32_t u,
                        uint64_t depth, uint64_t *nk, uint32_t *tt)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < u; t += gridDim.x * blockDim.x) {
        nk[t] = key_at(P, n, V[idx[t]] + depth, KD);
        tt[t] = t;
    }
}

__global__ void k_gather_u32(const uint32_t *src, const uint32_t *by, uint32_t u, uint32_t *dst)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < u; s += gridDim.x * blockDim.x) dst[s] = src[by[s]];
}

// new order of the unresolved slots: element s of the (group, key) order is the suffix of slot
// idx[tB[s]], and goes to slot idx[s]
__global__ void k_gather_pos(const uint64_t *V, const uint32_t *idx, const uint32_t *tB, uint32_t u, uint64_t *nv)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < u; s += gridDim.x * blockDim.x) nv[s] = V[idx[tB[s]]];
}

__global__ void k_scatter_pos(uint64_t *V, const uint32_t *idx, const uint64_t *nv, const uint32_t *gB,
                              const uint64_t *nk, const uint32_t *tB, uint32_t u, uint8_t *H)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < u; s += gridDim.x * blockDim.x) {
        const uint32_t j = idx[s];
        V[j] = nv[s];
        H[j] = (s == 0 || gB[s] != gB[s - 1] || nk[tB[s]] != nk[tB[s - 1]]) ? 1 : 0;
    }
}

__global__ void k_unres(const uint8_t *H, uint64_t m, const uint32_t *idx, uint32_t u, uint8_t *F)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < u; s += gridDim.x * blockDim.x) {
        const uint64_t j = idx[s];
        F[s] = (H[j] && (j + 1 == m || H[j + 1])) ? 0 : 1;
    }
}

// group of an unresolved element: the slot of its run's head (heads are the only elements with
// H set inside a run, and slots increase along the list)
struct HeadSlot {
    const uint8_t *H;
    const uint32_t *idx;
    __device__ uint32_t operator()(uint32_t t) const { return H[idx[t]] ? idx[t] : 0u; }
};

// BWT symbols by row for a finished chunk (rows r0 .. r0 + m - 1): the symbol before each
// suffix, 0 as a stand-in for the $ at the row of suffix 0 (whose row becomes `primary`), and the
// SA samples of rows 32, 64, ...
__global__ void k_emit(const uint64_t *P, const uint64_t *V, uint64_t m, uint64_t r0, uint64_t *R,
                       unsigned long long *primary, uint64_t *samples, uint64_t n_samples)
{
    const uint64_t w0 = r0 >> 5, w1 = (r0 + m - 1) >> 5;
    for (uint64_t w = w0 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w <= w1; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t lo = std::max(r0, w * 32), hi = std::min(r0 + m, w * 32 + 32);
        uint64_t v = 0;
        for (uint64_t r = lo; r < hi; ++r) {
            const uint64_t pos = V[r - r0];
            if (pos == 0)
                *primary = r;
            else
                v |= (uint64_t)tsym(P, pos - 1) << (62 - 2 * (r & 31));
            if (samples && (r & 31) == 0 && r >= 32 && (r >> 5) - 1 < n_samples) samples[(r >> 5) - 1] = pos;
        }
        if (hi - lo == 32)
            R[w] = v;
        else if (v)
            atomicOr(reinterpret_cast<unsigned long long *>(R + w), (unsigned long long)v);
    }
}

// row 0 (the suffix $): the last symbol of T
__global__ void k_row0(const uint64_t *P, uint64_t n, uint64_t *R)
{
    atomicOr(reinterpret_cast<unsigned long long *>(R), (unsigned long long)tsym(P, n - 1) << 62);
}

// symbol x of the $-less BWT: row x, or x + 1 from the primary row on; 0 beyond n
__device__ __forceinline__ uint32_t lsym(const uint64_t *R, uint64_t x, uint64_t n, uint64_t primary)
{
    if (x >= n) return 0;
    const uint64_t r = x + (x >= primary ? 1 : 0);
    return (uint32_t)(R[r >> 5] >> (62 - 2 * (r & 31))) & 3u;
}

// per 128-symbol block: symbol counts (four 8-bit fields, each at most 128)
__global__ void k_block_counts(const uint64_t *R, uint64_t n, uint64_t primary, uint64_t nb, uint32_t *bc)
{
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t c[4] = {0, 0, 0, 0};
        const uint64_t end = std::min(n, b * 128 + 128);
        for (uint64_t x = b * 128; x < end; ++x) {
            const uint32_t s = lsym(R, x, n, primary);
            c[0] += s == 0;
            c[1] += s == 1;
            c[2] += s == 2;
            c[3] += s == 3;
        }
        bc[b] = c[0] | c[1] << 8 | c[2] << 16 | c[3] << 24;
    }
}

struct Field {
    int c;
    __device__ uint64_t operator()(uint32_t v) const { return (v >> (8 * c)) & 0xffu; }
};

// the .bwt body: block b at word 16 b = [4 x u64 exclusive counts][symbol words], the last block's
// symbol words only as far as n reaches, then the 4 x u64 totals
__global__ void k_body(const uint64_t *R, uint64_t n, uint64_t primary, uint64_t nb, const uint64_t *ex0,
                       const uint64_t *ex1, const uint64_t *ex2, const uint64_t *ex3, uint32_t *body)
{
    const uint64_t total = nb * 16;  // slots; the last block's unused symbol words are skipped
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = t >> 4;
        const int k = (int)(t & 15);
        if (k < 8) {
            const uint64_t *ex = (k >> 1) == 0 ? ex0 : (k >> 1) == 1 ? ex1 : (k >> 1) == 2 ? ex2 : ex3;
            const uint64_t v = ex[b];
            body[t] = (k & 1) ? (uint32_t)(v >> 32) : (uint32_t)v;
        } else {
            const uint64_t x0 = b * 128 + (uint64_t)(k - 8) * 16;
            if (x0 >= n) continue;
            uint32_t v = 0;
            for (int i = 0; i < 16; ++i) v |= lsym(R, x0 + i, n, primary) << (30 - 2 * i);
            body[t] = v;
        }
    }
}

struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    template <class T>
    T *as() const
    {
        return static_cast<T *>(p);
    }
};

int dalloc(DevBuf &b, size_t bytes)
{
    if (b.p) IHIP(hipFree(b.p));
    b.p = nullptr;
    IHIP(hipMalloc(&b.p, std::max<size_t>(bytes, 16)));
    return 0;
}

struct Work {
    hipStream_t st = nullptr;
    DevBuf tmp;
    size_t tmp_bytes = 0;
    int need(size_t bytes)
    {
        if (bytes <= tmp_bytes) return 0;
        IHIP(hipStreamSynchronize(st));  // kernels in flight may still use the old buffer
        const int r = dalloc(tmp, bytes);
        tmp_bytes = r ? 0 : bytes;
        return r;
    }
};

// Sorts one chunk's suffixes (pos[0..m), collected) into suffix order in V.
struct ChunkBufs {
    DevBuf pos_in, key_in, key_out, V, nv, H, idx, idx2, gid, tA, tAo, gA, gB, tB, flag, cnt;
    uint64_t cap = 0;
};

int chunk_alloc(ChunkBufs &c, uint64_t m)
{
    if (m <= c.cap) return 0;
    int r;
    if ((r = dalloc(c.pos_in, m * 8)) || (r = dalloc(c.key_in, m * 8)) || (r = dalloc(c.key_out, m * 8)) ||
        (r = dalloc(c.V, m * 8)) || (r = dalloc(c.nv, m * 8)) || (r = dalloc(c.H, m)) || (r = dalloc(c.idx, m * 4)) ||
        (r = dalloc(c.idx2, m * 4)) || (r = dalloc(c.gid, m * 4)) || (r = dalloc(c.tA, m * 4)) ||
        (r = dalloc(c.tAo, m * 4)) || (r = dalloc(c.gA, m * 4)) || (r = dalloc(c.gB, m * 4)) || (r = dalloc(c.tB, m * 4)) ||
        (r = dalloc(c.flag, m)) || (r = dalloc(c.cnt, 16)))
        return r;
    c.cap = m;
    return 0;
}
// the chunk buffers, plus the radix sort's own copy of keys and values
constexpr uint64_t CHUNK_BYTES_PER_SUFFIX = 8 * 5 + 1 + 4 * 8 + 1 + 16;

int sort_chunk(Work &W, ChunkBufs &c, const uint64_t *P, uint64_t n, uint64_t m, int *rounds)
{
    hipStream_t st = W.st;
    uint64_t *V = c.V.as<uint64_t>();
    // first key, first sort
    hipLaunchKernelGGL(k_keys, dim3(grid_for(m)), dim3(256), 0, st, P, n, c.pos_in.as<uint64_t>(), m, (uint64_t)0,
                       c.key_in.as<uint64_t>());
    IHIP(hipGetLastError());
    size_t tb = 0;
    IHIP(rocprim::radix_sort_pairs(nullptr, tb, c.key_in.as<uint64_t>(), c.key_out.as<uint64_t>(), c.pos_in.as<uint64_t>(),
                                   V, (size_t)m, 0, 63, st));
    if (int r = W.need(tb)) return r;
    IHIP(rocprim::radix_sort_pairs(W.tmp.p, tb, c.key_in.as<uint64_t>(), c.key_out.as<uint64_t>(), c.pos_in.as<uint64_t>(),
                                   V, (size_t)m, 0, 63, st));
    uint8_t *H = c.H.as<uint8_t>();
    hipLaunchKernelGGL(k_heads, dim3(grid_for(m)), dim3(256), 0, st, c.key_out.as<const uint64_t>(), m, H);
    hipLaunchKernelGGL(k_unres_all, dim3(grid_for(m)), dim3(256), 0, st, (const uint8_t *)H, m, c.flag.as<uint8_t>());
    IHIP(hipGetLastError());
    uint32_t *idx = c.idx.as<uint32_t>(), *idx2 = c.idx2.as<uint32_t>();
    uint64_t *cnt = c.cnt.as<uint64_t>();
    rocprim::counting_iterator<uint32_t> ci(0u);
    IHIP(rocprim::select(nullptr, tb, ci, c.flag.as<uint8_t>(), idx, cnt, (size_t)m, st));
    if (int r = W.need(tb)) return r;
    IHIP(rocprim::select(W.tmp.p, tb, ci, c.flag.as<uint8_t>(), idx, cnt, (size_t)m, st));
    uint64_t u = 0;
    IHIP(hipMemcpyAsync(&u, cnt, 8, hipMemcpyDeviceToHost, st));
    IHIP(hipStreamSynchronize(st));
    uint64_t depth = 0;
    *rounds = 0;
    while (u > 0) {
        depth += KD;
        if (depth > n + KD) return ifail(GALN_HIP_E_DEVICE, "build_bwt: suffixes still tied past the text's end");
        ++*rounds;
        const uint32_t uu = (uint32_t)u;
        const unsigned g = grid_for(u);
        uint64_t *nk = c.key_in.as<uint64_t>(), *nkA = c.key_out.as<uint64_t>();
        uint32_t *tA = c.tA.as<uint32_t>(), *tAo = c.tAo.as<uint32_t>(), *gid = c.gid.as<uint32_t>();
        uint32_t *gA = c.gA.as<uint32_t>(), *gB = c.gB.as<uint32_t>(), *tB = c.tB.as<uint32_t>();
        // groups: the slot of each element's run head (max scan over head slots)
        rocprim::transform_iterator<rocprim::counting_iterator<uint32_t>, HeadSlot, uint32_t> hs(ci, HeadSlot{H, idx});
        IHIP(rocprim::inclusive_scan(nullptr, tb, hs, gid, (size_t)u, rocprim::maximum<uint32_t>(), st));
        if (int r = W.need(tb)) return r;
        IHIP(rocprim::inclusive_scan(W.tmp.p, tb, hs, gid, (size_t)u, rocprim::maximum<uint32_t>(), st));
        hipLaunchKernelGGL(k_rkeys, dim3(g), dim3(256), 0, st, P, n, (const uint64_t *)V, (const uint32_t *)idx, uu, depth,
                           nk, tA);
        IHIP(hipGetLastError());
        // (key) order, then a stable sort by group: (group, key) order
        IHIP(rocprim::radix_sort_pairs(nullptr, tb, nk, nkA, tA, tAo, (size_t)u, 0, 63, st));
        if (int r = W.need(tb)) return r;
        IHIP(rocprim::radix_sort_pairs(W.tmp.p, tb, nk, nkA, tA, tAo, (size_t)u, 0, 63, st));
        hipLaunchKernelGGL(k_gather_u32, dim3(g), dim3(256), 0, st, (const uint32_t *)gid, (const uint32_t *)tAo, uu, gA);
        IHIP(hipGetLastError());
        IHIP(rocprim::radix_sort_pairs(nullptr, tb, gA, gB, tAo, tB, (size_t)u, 0, 32, st));
        if (int r = W.need(tb)) return r;
        IHIP(rocprim::radix_sort_pairs(W.tmp.p, tb, gA, gB, tAo, tB, (size_t)u, 0, 32, st));
        hipLaunchKernelGGL(k_gather_pos, dim3(g), dim3(256), 0, st, (const uint64_t *)V, (const uint32_t *)idx,
                           (const uint32_t *)tB, uu, c.nv.as<uint64_t>());
        hipLaunchKernelGGL(k_scatter_pos, dim3(g), dim3(256), 0, st, V, (const uint32_t *)idx, c.nv.as<const uint64_t>(),
                           (const uint32_t *)gB, (const uint64_t *)nk, (const uint32_t *)tB, uu, H);
        hipLaunchKernelGGL(k_unres, dim3(g), dim3(256), 0, st, (const uint8_t *)H, m, (const uint32_t *)idx, uu,
                           c.flag.as<uint8_t>());
        IHIP(hipGetLastError());
        IHIP(rocprim::select(nullptr, tb, idx, c.flag.as<uint8_t>(), idx2, cnt, (size_t)u, st));
        if (int r = W.need(tb)) return r;
        IHIP(rocprim::select(W.tmp.p, tb, idx, c.flag.as<uint8_t>(), idx2, cnt, (size_t)u, st));
        IHIP(hipMemcpyAsync(&u, cnt, 8, hipMemcpyDeviceToHost, st));
        IHIP(hipStreamSynchronize(st));
        std::swap(c.idx.p, c.idx2.p);
        idx = c.idx.as<uint32_t>();
        idx2 = c.idx2.as<uint32_t>();
    }
    return 0;
}

int build(const uint8_t *fwd, uint64_t len, int device, galn_bwt_meta_t *meta, uint32_t *bwt, uint64_t *samples_out,
          double *stats)
{
    const uint64_t n = 2 * len;
    IHIP(hipSetDevice(device));
    Work W;
    IHIP(hipStreamCreateWithFlags(&W.st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{W.st};
    hipStream_t st = W.st;
    const uint64_t nw = (n + 31) / 32;
    DevBuf P, R, S, hist;
    int r;
    if ((r = dalloc(P, (nw + 2) * 8)) || (r = dalloc(R, ((n + 1 + 31) / 32 + 2) * 8))) return r;
    const uint64_t n_samples = (n + 32) / 32 - 1;
    if (samples_out && n_samples && (r = dalloc(S, n_samples * 8))) return r;
    IHIP(hipMemsetAsync(P.p, 0, (nw + 2) * 8, st));
    IHIP(hipMemsetAsync(R.p, 0, ((n + 1 + 31) / 32 + 2) * 8, st));
    // text: forward codes uploaded in pieces, packed into both images
    {
        DevBuf stage;
        if ((r = dalloc(stage, std::min(len, PIECE)))) return r;
        for (uint64_t a = 0; a < len; a += PIECE) {
            const uint64_t cnt = std::min(PIECE, len - a);
            IHIP(hipMemcpyAsync(stage.p, fwd + a, cnt, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_pack_text, dim3(grid_for(cnt / 16 + 2)), dim3(256), 0, st, P.as<uint64_t>(), n,
                               stage.as<const uint8_t>(), a, cnt);
            IHIP(hipGetLastError());
            IHIP(hipStreamSynchronize(st));  // the stage is reused
        }
    }
    const uint64_t *Pc = P.as<const uint64_t>();
    if ((r = dalloc(hist, (size_t)NBUCKET * 8))) return r;
    IHIP(hipMemsetAsync(hist.p, 0, (size_t)NBUCKET * 8, st));
    hipLaunchKernelGGL(k_hist, dim3(grid_for(n)), dim3(256), 0, st, Pc, n, hist.as<unsigned long long>());
    IHIP(hipGetLastError());
    std::vector<uint64_t> h(NBUCKET);
    IHIP(hipMemcpyAsync(h.data(), hist.p, (size_t)NBUCKET * 8, hipMemcpyDeviceToHost, st));
    IHIP(hipStreamSynchronize(st));
    // chunk capacity from the free memory (the packed body, n/2 bytes, comes after the chunks)
    size_t freeb = 0, totalb = 0;
    IHIP(hipMemGetInfo(&freeb, &totalb));
    const uint64_t reserve = (uint64_t)1 << 30;
    uint64_t cap = freeb > reserve ? (freeb - reserve) / (CHUNK_BYTES_PER_SUFFIX + 8) : 0;
    cap = std::min(cap, CHUNK_MAX);
    if (const char *e = getenv("GALN_HIP_TEST_BUILD_CHUNK")) cap = std::min<uint64_t>(cap, std::max(1L, strtol(e, nullptr, 10)));
    // chunks: runs of consecutive buckets of at most `cap` suffixes
    struct Chunk {
        uint32_t b0, b1;
        uint64_t m, r0;
    };
    std::vector<Chunk> chunks;
    uint64_t row = 1, maxm = 0;
    for (uint32_t b = 0; b < NBUCKET;) {
        if (h[b] > cap)
            return ifail(GALN_HIP_E_NOMEM, "build_bwt: %llu suffixes share a 10-symbol prefix; at most %llu fit a chunk",
                         (unsigned long long)h[b], (unsigned long long)cap);
        Chunk c{b, b, 0, row};
        while (b < NBUCKET && c.m + h[b] <= cap) c.m += h[b++];
        c.b1 = b;
        row += c.m;
        maxm = std::max(maxm, c.m);
        if (c.m) chunks.push_back(c);
    }
    if (row != n + 1) return ifail(GALN_HIP_E_DEVICE, "build_bwt: bucket counts sum to %llu, not %llu",
                                   (unsigned long long)(row - 1), (unsigned long long)n);
    hist.release();
    DevBuf primary;
    if ((r = dalloc(primary, 8))) return r;
    IHIP(hipMemsetAsync(primary.p, 0xff, 8, st));
    hipLaunchKernelGGL(k_row0, dim3(1), dim3(1), 0, st, Pc, n, R.as<uint64_t>());
    IHIP(hipGetLastError());
    const char *vb = getenv("GALN_HIP_VERBOSE");
    const bool verbose = vb && *vb && strcmp(vb, "0") != 0;
    {
        ChunkBufs cb;
        if ((r = chunk_alloc(cb, maxm))) return r;
        int max_rounds = 0;
        for (const Chunk &c : chunks) {
            IHIP(hipMemsetAsync(cb.cnt.p, 0, 8, st));
            hipLaunchKernelGGL(k_collect, dim3(grid_for(n)), dim3(256), 0, st, Pc, n, c.b0, c.b1, cb.pos_in.as<uint64_t>(),
                               cb.cnt.as<unsigned long long>());
            IHIP(hipGetLastError());
            uint64_t got = 0;
            IHIP(hipMemcpyAsync(&got, cb.cnt.p, 8, hipMemcpyDeviceToHost, st));
            IHIP(hipStreamSynchronize(st));
            if (got != c.m) return ifail(GALN_HIP_E_DEVICE, "build_bwt: chunk collected %llu of %llu suffixes",
                                         (unsigned long long)got, (unsigned long long)c.m);
            int rounds = 0;
            if ((r = sort_chunk(W, cb, Pc, n, c.m, &rounds))) return r;
            max_rounds = std::max(max_rounds, rounds);
            if (verbose)
                fprintf(stderr, "[galn_hip] build_bwt: chunk %zu of %zu: %llu suffixes sorted, %d refinement rounds\n",
                        (size_t)(&c - chunks.data()) + 1, chunks.size(), (unsigned long long)c.m, rounds);
            const uint64_t words = ((c.r0 + c.m - 1) >> 5) - (c.r0 >> 5) + 1;
            hipLaunchKernelGGL(k_emit, dim3(grid_for(words)), dim3(256), 0, st, Pc, cb.V.as<const uint64_t>(), c.m, c.r0,
                               R.as<uint64_t>(), primary.as<unsigned long long>(), samples_out ? S.as<uint64_t>() : nullptr,
                               n_samples);
            IHIP(hipGetLastError());
        }
        if (stats) {
            stats[0] = (double)chunks.size();
            stats[1] = (double)max_rounds;
            stats[2] = (double)maxm;
        }
    }
    P.release();
    uint64_t prim = 0;
    IHIP(hipMemcpyAsync(&prim, primary.p, 8, hipMemcpyDeviceToHost, st));
    IHIP(hipStreamSynchronize(st));
    if (prim == ~0ull || prim == 0 || prim > n) return ifail(GALN_HIP_E_DEVICE, "build_bwt: no primary row");
    // the .bwt body
    const uint64_t nb = (n + 127) / 128;
    const uint64_t words = galn_hip_bwt_words(n);
    DevBuf bc, ex[4], body;
    if ((r = dalloc(bc, nb * 4))) return r;
    for (auto &e : ex)
        if ((r = dalloc(e, nb * 8))) return r;
    if ((r = dalloc(body, words * 4))) return r;
    hipLaunchKernelGGL(k_block_counts, dim3(grid_for(nb)), dim3(256), 0, st, R.as<const uint64_t>(), n, prim, nb,
                       bc.as<uint32_t>());
    IHIP(hipGetLastError());
    uint64_t totals[4] = {0, 0, 0, 0};
    for (int c = 0; c < 4; ++c) {
        rocprim::transform_iterator<const uint32_t *, Field, uint64_t> it(bc.as<const uint32_t>(), Field{c});
        size_t tb = 0;
        IHIP(rocprim::exclusive_scan(nullptr, tb, it, ex[c].as<uint64_t>(), (uint64_t)0, (size_t)nb,
                                     rocprim::plus<uint64_t>(), st));
        if ((r = W.need(tb))) return r;
        IHIP(rocprim::exclusive_scan(W.tmp.p, tb, it, ex[c].as<uint64_t>(), (uint64_t)0, (size_t)nb,
                                     rocprim::plus<uint64_t>(), st));
        uint64_t last = 0;
        uint32_t lastc = 0;
        IHIP(hipMemcpyAsync(&last, ex[c].as<uint64_t>() + nb - 1, 8, hipMemcpyDeviceToHost, st));
        IHIP(hipMemcpyAsync(&lastc, bc.as<uint32_t>() + nb - 1, 4, hipMemcpyDeviceToHost, st));
        IHIP(hipStreamSynchronize(st));
        totals[c] = last + ((lastc >> (8 * c)) & 0xffu);
    }
    hipLaunchKernelGGL(k_body, dim3(grid_for(nb * 16)), dim3(256), 0, st, R.as<const uint64_t>(), n, prim, nb,
                       ex[0].as<const uint64_t>(), ex[1].as<const uint64_t>(), ex[2].as<const uint64_t>(),
                       ex[3].as<const uint64_t>(), body.as<uint32_t>());
    IHIP(hipGetLastError());
    const uint64_t last_syms = (n - 128 * (nb - 1) + 15) / 16;
    const uint64_t head = 16 * (nb - 1) + 8 + last_syms;
    IHIP(hipMemcpyAsync(bwt, body.p, head * 4, hipMemcpyDeviceToHost, st));
    if (samples_out && n_samples) IHIP(hipMemcpyAsync(samples_out, S.p, n_samples * 8, hipMemcpyDeviceToHost, st));
    IHIP(hipStreamSynchronize(st));
    memcpy(bwt + head, totals, sizeof totals);
    meta->primary = prim;
    meta->L2[0] = 0;
    for (int c = 0; c < 4; ++c) meta->L2[c + 1] = meta->L2[c] + totals[c];
    meta->seq_len = n;
    meta->bwt_size = words;
    if (meta->L2[4] != n) return ifail(GALN_HIP_E_DEVICE, "build_bwt: symbol counts sum to %llu, not %llu",
                                       (unsigned long long)meta->L2[4], (unsigned long long)n);
    return 0;
}

}  // namespace

extern "C" {

uint64_t galn_hip_bwt_words(uint64_t n)
{
    const uint64_t nb = (n + 127) / 128;
    return 8 * (nb + 1) + (n + 15) / 16;
}

int galn_hip_build_bwt(const uint8_t *fwd, uint64_t len, int device, galn_bwt_meta_t *meta, uint32_t *bwt,
                       uint64_t *sa_samples, double stats[4])
{
    if (!fwd || !meta || !bwt || len == 0) return ifail(GALN_HIP_E_ARG, "build_bwt: null argument or empty text");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ifail(GALN_HIP_E_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return ifail(GALN_HIP_E_DEVICE, "no HIP device %d (%d visible)", device, ndev);
    return build(fwd, len, device, meta, bwt, sa_samples, stats);
}

}  // extern "C"
