// galn_search.hip -- BWA-aln backtracking search on gfx950 (MI355X), and its C ABI.
//
// One lane searches one read (thread-per-read). The lane runs the read as a
// state machine whose every iteration performs exactly one occurrence (LF)
// lookup, whatever the read is doing -- widths (bwt_cal_width), the seed
// widths, expanding a popped stack entry (bwt_2occ4), or an exact-match tail
// (bwt_match_exact_alt) -- so all 64 lanes of a wave issue their random index
// loads together and the memory system sees one independent gather per lane.
//
// Per-lane search state (reference src/bwtgap_hd.h:338-473):
//   * the priority stack: one LIFO per score bucket, built from 16-entry
//     chunks of 16-byte packed entries; the top of each bucket (chunk id,
//     fill) lives in LDS, the non-empty-bucket bitmask in registers, so a pop
//     is one dependent global load and a push is none. Chunks come from a
//     device-wide pool in granules of 64 chunks: a lane takes granules as its
//     stack grows (lock-free tagged free stack, then a bump counter), keeps a
//     few for its next read and hands the rest back, so a read's stack grows to
//     its true peak without re-runs and the pool holds only what is live;
//   * the exact child of an expansion (pushed last, same score, so always the
//     next pop) stays in registers instead of going through the arena;
//   * width[] / seed_width[] as packed u64 (w:48 | bid:16) in the lane's arena.
// Pops, pushes, cut-offs and records follow the reference exactly, so records,
// record order, pop counts and trace hashes are those of the CPU search.
//
// Forms:
//   plain    one launch per 0x40000 reads (one lane per read, input order);
//   batched  one persistent launch over the whole read list; lanes claim the
//            next read from a device counter as soon as theirs is done.
// Reads that outgrow their record slots, or find the pool exhausted, are re-run
// from scratch in a later launch with more room, so results never depend on
// the capacities chosen (the reference's contract, src/bwtgap_hd.h:10-22).

#include <hip/hip_runtime.h>
#include <cstdlib>
#include <cstring>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "galn_hip.h"

namespace {

constexpr int CH = 16;                       // entries per stack chunk
constexpr int GRAN = 64;                     // chunks per pool granule
constexpr uint32_t MAX_POOL_CHUNKS = 1u << 28;  // 28-bit chunk ids in the bucket tops
constexpr int MAXB = 256;                    // score buckets (bitmask width)
constexpr int MAXLEN = 1023;                 // read length limit (10-bit entry fields)
constexpr int MAXDIFF = 126;                 // 7-bit n_mm / n_gapo fields (n_mm may reach max_diff + 1)
constexpr uint64_t KMASK = (1ull << 38) - 1; // 38-bit k and l
constexpr uint64_t W48 = (1ull << 48) - 1;
constexpr uint32_t NIL = 0xffffffffu;

constexpr int ERR_STACK = -1;   // lane arena full: re-run in a larger tier
constexpr int ERR_ALN = -2;     // record slots full: re-run with more
constexpr int ERR_SCORE = -3;   // final, as the reference
constexpr int ERR_UNSUP = -5;
constexpr int ERR_OUT = -8;     // output arena full: re-run after it grows
constexpr int ERR_INTERNAL = -9; // stack bookkeeping inconsistent (final)
constexpr int ERR_HEAVY = -10;   // pop budget spent: the read moves to the wave-per-read engine
constexpr int NOFIN = 0x7fffffff;  // k_search helpers: the read goes on
constexpr int HITP = 0x7ffffffe;   // k_search pop: the popped entry is a hit
constexpr int STARTS = 0x7ffffffd; // k_search widths: done, the search is set up

constexpr int MODE_GAPE = 0x01, MODE_LOGGAP = 0x04, MODE_NONSTOP = 0x10;
constexpr int ST_M = 0, ST_I = 1, ST_D = 2;

constexpr uint64_t TR_INIT = 1469598103934665603ull;
constexpr uint64_t TR_PRIME = 1099511628211ull;

struct DMeta {
    uint64_t primary;
    uint64_t L2[5];
    uint64_t seq_len;
};

// The cumulative symbol counts held in registers: a lane's count for its own symbol is picked
// by selects. Indexing the kernel-argument copy by a per-lane symbol is a vector memory load,
// whose wait also waits for every store the wave still has in flight.
struct L2Sel {
    uint64_t v0, v1, v2, v3, v4;
    __device__ explicit L2Sel(const DMeta &m) : v0(m.L2[0]), v1(m.L2[1]), v2(m.L2[2]), v3(m.L2[3]), v4(m.L2[4]) {}
    __device__ __forceinline__ uint64_t operator[](int c) const
    {
        return c == 0 ? v0 : c == 1 ? v1 : c == 2 ? v2 : c == 3 ? v3 : v4;
    }
};

// packed stack entry (layout below)
struct __align__(16) Ent {
    uint64_t w0, w1;
};

struct KArgs {
    const uint32_t *bwt;
    DMeta meta;
    // reads
    const uint8_t *query;
    const uint64_t *qoff;
    const int32_t *qlen;
    const int32_t *maxdiff;
    const int32_t *batch;
    const galn_search_opt_t *bopts;
    const int32_t *nstacks;
    // work: list positions [0, n_list); read id = list ? list[pos] : list_base + pos
    const int32_t *list;
    int32_t list_base;
    int32_t n_list;
    int32_t *qnext;          // persistent claims
    int32_t n_slots;
    // per-lane region: widths, seed widths, records, owned granules
    uint8_t *scratch;
    uint64_t slot_bytes;
    uint32_t sw_off, rec_off, own_off;
    int32_t own_cap;         // granules a lane may hold (enough for max_entries)
    int32_t keep;            // granules a lane keeps between reads
    int32_t rec_cap;
    // stack pool
    Ent *pool_chk;           // n_gran * GRAN * CH entries
    uint32_t *pool_link;     // per chunk: previous chunk of its bucket, or next free chunk of its lane
    uint32_t *gnext;         // per granule: next granule in the free stack
    unsigned long long *ghead;  // [NSHARD] free stack heads: tag << 32 | granule (0xffffffff = empty)
    uint32_t *gbump;         // granules never handed out start here
    uint32_t n_gran;
    // results (indexed by read id)
    int32_t *out_n;
    uint64_t *out_off;
    uint64_t *out_pops;
    uint64_t *out_trace;
    uint64_t *out_bytes;     // algorithmic index bytes per read
    galn_aln_rec_t *out_rec;
    uint64_t out_cap;
    unsigned long long *out_top;
    int32_t *fail_list;
    int32_t *fail_n;
    int32_t *fail_kind;      // [0] output full, [1] record slots, [2] stack arena
    unsigned long long *algo_bytes;  // this launch's algorithmic index bytes (final results only)
    int32_t fuse;
    uint64_t pop_budget;     // > 0: reads popping more entries go to heavy_list
    int32_t *heavy_list;
    int32_t *heavy_n;
    // hand-off of a heavy read's widths: slot h (its place in heavy_list) holds
    // [bytes | 1 << 63][W[0..len]][SW[0..seed_len]], or 0 when the widths were not kept
    uint64_t *hand;
    uint32_t hand_stride;    // u64 words per slot
    int32_t hand_cap;        // slots
};

// The kernel argument segment as k_search reads it, and a load that the compiler must perform
// where it is written (not hoisted out of the loop into registers).
typedef const __attribute__((address_space(4))) KArgs KArgsC;
template <class T>
__device__ __forceinline__ T cold_ld(const __attribute__((address_space(4))) T *p)
{
    return *reinterpret_cast<const volatile __attribute__((address_space(4))) T *>(p);
}
#define CA(f) cold_ld(&A->f)

// ---------------------------------------------------------------------------
// Occurrence counts. Device index layout: 128-symbol blocks of 16 u32 words, 4 x u64
// exclusive counts, then the block's symbols as two bit planes: 4 words of high bits, then 4
// words of low bits, symbol s at bit 31 - (s & 31) of word s >> 5 of each plane. The .bwt body
// (BWA's layout: 8 words of 16 two-bit symbols, first symbol in bits 31-30) is uploaded as is and
// converted in place by k_planes. Position p of the $-inclusive BWT is at p - (p >= primary).

__device__ __forceinline__ void load_block(const uint32_t *__restrict__ bwt, uint64_t blk, uint4 (&b)[4])
{
    const uint4 *p = reinterpret_cast<const uint4 *>(bwt + (blk << 4));
    b[0] = p[0];
    b[1] = p[1];
    b[2] = p[2];
    b[3] = p[3];
}

// counts of all 4 bases among symbols 0..r of a loaded block, plus its prefix counts: per plane
// word, the high bits count bases 2 and 3, the low bits 1 and 3, both together base 3
__device__ __forceinline__ void block_occ4(const uint4 (&b)[4], int r, uint64_t c[4])
{
    const uint32_t hp[4] = {b[2].x, b[2].y, b[2].z, b[2].w};
    const uint32_t lp[4] = {b[3].x, b[3].y, b[3].z, b[3].w};
    uint32_t nh = 0, nl = 0, n3 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // the top t bits of word j: t = symbols of the word at or before r, 0..32
        const uint32_t t = (uint32_t)min(max(r + 1 - 32 * j, 0), 32);
        const uint32_t valid = (uint32_t)(0xffffffff00000000ull >> t);
        const uint32_t h = hp[j] & valid, l = lp[j] & valid;
        nh += __builtin_popcount(h);
        nl += __builtin_popcount(l);
        n3 += __builtin_popcount(h & l);
    }
    const uint32_t n2 = nh - n3, n1 = nl - n3;
    const uint32_t n0 = (uint32_t)(r + 1) - n1 - n2 - n3;
    c[0] = ((uint64_t)b[0].y << 32 | b[0].x) + n0;
    c[1] = ((uint64_t)b[0].w << 32 | b[0].z) + n1;
    c[2] = ((uint64_t)b[1].y << 32 | b[1].x) + n2;
    c[3] = ((uint64_t)b[1].w << 32 | b[1].z) + n3;
}

// In-place conversion of the uploaded .bwt body's symbol words to bit planes, one block per
// thread (a last, shorter block included: its missing words were zero-filled).
__global__ void __launch_bounds__(256) k_planes(uint32_t *bwt, uint64_t n_blocks)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_blocks) return;
    uint4 *const p = reinterpret_cast<uint4 *>(bwt + (i << 4));
    const uint4 s0 = p[2], s1 = p[3];
    const uint32_t w[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    uint32_t hp[4] = {0, 0, 0, 0}, lp[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const int sym = k * 16 + t;  // symbol index in the block
            const uint32_t c = (w[k] >> (30 - 2 * t)) & 3u;
            hp[sym >> 5] |= (c >> 1) << (31 - (sym & 31));
            lp[sym >> 5] |= (c & 1u) << (31 - (sym & 31));
        }
    }
    p[2] = make_uint4(hp[0], hp[1], hp[2], hp[3]);
    p[3] = make_uint4(lp[0], lp[1], lp[2], lp[3]);
}

// Occ of all four bases at positions a (may be -1) and b >= 0: bwt_2occ4, src/bwtgap_hd.h:158-192.
// One 64-byte block when both fall in it, else two independent loads issued together.
__device__ __forceinline__ uint32_t lf_pair(const uint32_t *__restrict__ bwt, const DMeta &m, uint64_t a, uint64_t b,
                                            uint64_t ca[4], uint64_t cb[4])
{
    const uint64_t bp = b - (b >= m.primary);
    const uint64_t bb = bp >> 7;
    const bool aneg = a == ~0ull;
    const uint64_t ap = a - (a >= m.primary);
    const uint64_t ab = ap >> 7;
    const bool same = !aneg && ab == bb;
    uint4 B[4], A[4];
    load_block(bwt, bb, B);
    if (!aneg && !same) load_block(bwt, ab, A);
    block_occ4(B, (int)(bp & 127), cb);
    // a's block, picked per lane, so that one count serves lanes of both kinds
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        A[t].x = same ? B[t].x : A[t].x;
        A[t].y = same ? B[t].y : A[t].y;
        A[t].z = same ? B[t].z : A[t].z;
        A[t].w = same ? B[t].w : A[t].w;
    }
    block_occ4(A, (int)(ap & 127), ca);
    if (aneg) ca[0] = ca[1] = ca[2] = ca[3] = 0;
    return (aneg || same) ? 64u : 128u;
}

// Occ of base c alone at positions a (may be -1) and b, for the steps that follow one symbol (an
// exact tail, a width): one popcount per plane word, the planes flipped so that base c reads as
// ones. Same block loads and byte count as lf_pair.
__device__ __forceinline__ uint64_t block_occ1(const uint4 (&b)[4], int r, int c)
{
    const uint32_t hp[4] = {b[2].x, b[2].y, b[2].z, b[2].w};
    const uint32_t lp[4] = {b[3].x, b[3].y, b[3].z, b[3].w};
    const uint32_t hx = (c & 2) ? 0u : ~0u, lx = (c & 1) ? 0u : ~0u;
    uint32_t n = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t t = (uint32_t)min(max(r + 1 - 32 * j, 0), 32);
        const uint32_t valid = (uint32_t)(0xffffffff00000000ull >> t);
        n += __builtin_popcount((hp[j] ^ hx) & (lp[j] ^ lx) & valid);
    }
    // base c's count at the block start, by masks (a select between the block's words would be
    // turned into an indexed load of the block, which then lives in scratch memory)
    const uint32_t m0 = 0u - (uint32_t)(c == 0), m1 = 0u - (uint32_t)(c == 1), m2 = 0u - (uint32_t)(c == 2),
                   m3 = 0u - (uint32_t)(c == 3);
    const uint32_t lo = (b[0].x & m0) | (b[0].z & m1) | (b[1].x & m2) | (b[1].z & m3);
    const uint32_t hi = (b[0].y & m0) | (b[0].w & m1) | (b[1].y & m2) | (b[1].w & m3);
    return ((uint64_t)hi << 32 | lo) + n;
}

__device__ __forceinline__ uint32_t lf_one(const uint32_t *__restrict__ bwt, const DMeta &m, uint64_t a, uint64_t b,
                                           int c, uint64_t &oa, uint64_t &ob)
{
    const uint64_t bp = b - (b >= m.primary);
    const uint64_t bb = bp >> 7;
    const bool aneg = a == ~0ull;
    const uint64_t ap = a - (a >= m.primary);
    const uint64_t ab = ap >> 7;
    const bool same = !aneg && ab == bb;
    uint4 B[4], A[4];
    load_block(bwt, bb, B);
    if (!aneg && !same) load_block(bwt, ab, A);
    ob = block_occ1(B, (int)(bp & 127), c);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        A[t].x = same ? B[t].x : A[t].x;
        A[t].y = same ? B[t].y : A[t].y;
        A[t].z = same ? B[t].z : A[t].z;
        A[t].w = same ? B[t].w : A[t].w;
    }
    oa = aneg ? 0 : block_occ1(A, (int)(ap & 127), c);
    return (aneg || same) ? 64u : 128u;
}

// ---------------------------------------------------------------------------
// Packed stack entry (16 bytes); the score is the bucket.
//   w0 = k[0:38) | i[38:48) | n_mm[48:55) | n_gapo[55:62) | state[62:64)
//   w1 = l[0:38) | last_diff_pos[38:48) | n_gape[48:56) | n_ins[56:64)
// n_del = n_gapo + n_gape - n_ins (every gap open or extension adds one
// inserted or one deleted base, src/bwtgap_hd.h:430-454).


__device__ __forceinline__ Ent mk_ent(uint64_t k, uint64_t l, int i, int ldp, int mm, int go, int ge, int ins, int st)
{
    // the fields first, so that children differing only in (k, l) share them
    const uint64_t h0 = ((uint64_t)i << 38) | ((uint64_t)mm << 48) | ((uint64_t)go << 55) | ((uint64_t)st << 62);
    const uint64_t h1 = ((uint64_t)ldp << 38) | ((uint64_t)ge << 48) | ((uint64_t)ins << 56);
    Ent e;
    e.w0 = h0 | (k & KMASK);
    e.w1 = h1 | (l & KMASK);
    return e;
}
__device__ __forceinline__ uint64_t e_k(const Ent &e) { return e.w0 & KMASK; }
__device__ __forceinline__ int e_i(const Ent &e) { return (int)((e.w0 >> 38) & 1023); }
__device__ __forceinline__ int e_mm(const Ent &e) { return (int)((e.w0 >> 48) & 127); }
__device__ __forceinline__ int e_go(const Ent &e) { return (int)((e.w0 >> 55) & 127); }
__device__ __forceinline__ int e_st(const Ent &e) { return (int)(e.w0 >> 62); }
__device__ __forceinline__ uint64_t e_l(const Ent &e) { return e.w1 & KMASK; }
__device__ __forceinline__ int e_ldp(const Ent &e) { return (int)((e.w1 >> 38) & 1023); }
__device__ __forceinline__ int e_ge(const Ent &e) { return (int)((e.w1 >> 48) & 255); }
__device__ __forceinline__ int e_ins(const Ent &e) { return (int)(e.w1 >> 56); }

__device__ __forceinline__ uint64_t tr_mix(uint64_t h, uint64_t v) { return (h ^ v) * TR_PRIME; }

__device__ __forceinline__ int int_log2(uint32_t v) { return v ? 31 - __builtin_clz(v) : 0; }

__device__ __forceinline__ int wbid(uint64_t w) { return (int)(w >> 48); }
__device__ __forceinline__ uint64_t wval(uint64_t w) { return w & W48; }
__device__ __forceinline__ uint64_t wpack(uint64_t w, int bid) { return w | ((uint64_t)bid << 48); }

// what a lane of k_search does with the loads of an iteration: claim a read (no loads), a
// width step, evaluate a popped entry, expand (or prune, or start the exact tail of) an entry,
// one exact-tail step
enum : int { PH_CLAIM = 0, PH_CW = 1, PH_POP = 2, PH_EXP = 3, PH_EXACT = 4 };

// Free granules: NSHARD lock-free stacks, each with a head that carries a tag bumped on every
// change (no ABA), links read and written at device scope; then never-used granules. A
// workgroup uses the shard of its index, so contention on one head stays low.
constexpr int NSHARD = 64;

__device__ __forceinline__ uint32_t shard_pop(unsigned long long *head, uint32_t *gnext)
{
    for (;;) {
        unsigned long long old = __hip_atomic_load(head, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t g = (uint32_t)old;
        if (g == NIL) return NIL;
        const uint32_t nxt = __hip_atomic_load(gnext + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long nw = ((old >> 32) + 1) << 32 | nxt;
        if (__hip_atomic_compare_exchange_strong(head, &old, nw, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return g;
    }
}

__device__ __forceinline__ uint32_t gran_pop(unsigned long long *ghead, uint32_t *gnext, uint32_t *gbump, uint32_t n_gran)
{
    const int sh = (int)(blockIdx.x & (NSHARD - 1));
    uint32_t g = shard_pop(ghead + sh, gnext);
    if (g != NIL) return g;
    if (__hip_atomic_load(gbump, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n_gran) {
        const uint32_t b = atomicAdd(gbump, 1u);
        if (b < n_gran) return b;
    }
    // the never-used granules are gone: look in the other shards
    for (int t = 1; t < NSHARD && g == NIL; ++t) g = shard_pop(ghead + ((sh + t) & (NSHARD - 1)), gnext);
    return g;
}

// push the chain own[from..to) (linked through gnext) onto this workgroup's shard in one exchange
__device__ __forceinline__ void gran_push(unsigned long long *ghead, uint32_t *gnext, const uint32_t *own, int from,
                                          int to)
{
    unsigned long long *const head = ghead + (blockIdx.x & (NSHARD - 1));
    for (int j = from; j + 1 < to; ++j)
        __hip_atomic_store(gnext + own[j], own[j + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t first = own[from], last = own[to - 1];
    unsigned long long old = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        __hip_atomic_store(gnext + last, (uint32_t)old, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long nw = ((old >> 32) + 1) << 32 | first;
        if (__hip_atomic_compare_exchange_strong(head, &old, nw, __ATOMIC_RELEASE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

template <bool TRACE, bool PERSIST>
__global__ void __launch_bounds__(128) k_search(KArgs a)
{
    using TopT = uint32_t;  // chunk id << 4 | fill - 1
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.n_slots) return;
    // arguments used once per read (claims, results, granules) are read from the kernel
    // argument segment where they are used, so they do not hold scalar registers for the loop
    const KArgsC *const A = (const KArgsC *)__builtin_amdgcn_kernarg_segment_ptr();
    TopT *const tops = reinterpret_cast<TopT *>(lds_raw) + threadIdx.x;
    const int tstride = blockDim.x;

    uint8_t *const base = a.scratch + (uint64_t)slot * a.slot_bytes;
    uint64_t *const W = reinterpret_cast<uint64_t *>(base);
    uint64_t *const SW = reinterpret_cast<uint64_t *>(base + a.sw_off);
    galn_aln_rec_t *const REC = reinterpret_cast<galn_aln_rec_t *>(base + a.rec_off);
    uint32_t *const OWN = reinterpret_cast<uint32_t *>(base + a.own_off);
    uint32_t *const LINK = a.pool_link;
    Ent *const CHK = a.pool_chk;
    const DMeta &M = a.meta;
    const uint64_t seq_len = M.seq_len;

    // ---- lane state ----
    int rid = 0, len = 0;
    const uint8_t *q = nullptr;
    int s_mm = 0, s_gapo = 0, s_gape = 0, mode = 0, ies = 0, mdo = 0, max_entries = 0, max_gapo = 0, max_gape = 0,
        msd = 0, max_top2 = 0, n_stacks = 0, md_read = 0, seed_len = 0;
    bool seeding = false;
    // widths in progress: position cj of `pstr` (the query, or its seed part), interval [ck, cl]
    int cj = 0, plen = 0, cbid = 0, cbase = 0, nN = 0;
    const uint8_t *pstr = nullptr;
    uint64_t *wout = nullptr;
    bool in_seed = false;
    uint64_t ck = 0, cl = 0;
    // search
    int best_score = 0, best_diff = 0, max_diff = 0, best_cnt = 0, n_aln = 0, n_entries = 0;
    uint64_t bm0 = 0, bm1 = 0, bm2 = 0, bm3 = 0;
    Ent cur{0, 0}, e{0, 0}, pf{0, 0};
    bool has_cur = false, pf_valid = false;
    int cur_score = 0, e_score = 0, m = 0, m_seed = 0;
    uint32_t pf_addr = NIL, ld_ntop = NIL;  // prefetched entry; lowest top after a pop being loaded
    uint64_t hk = 0, hl = 0;  // a hit waiting to be applied
    uint32_t spare = NIL, free_head = NIL;
    int n_own = 0, own_next = 0, gleft = 0;  // granules held, next one to use, chunks left in it
    uint32_t gbase = 0;
    uint64_t pops = 0, trace = TR_INIT, obytes = 0, wbytes = 0;  // wbytes: of the widths
    // what the next iteration loads, and what it then does
    int phase = PH_CLAIM;
    bool rq_lf = false, rq_ent = false, rq_pf = false, rq_w = false;
    uint64_t qa = 0, qb = 0;
    uint32_t ent_addr = 0;
    const uint8_t *c1p = nullptr, *c2p = nullptr;  // query bytes to load (null: none)
    bool claimed = false;

    // ---- helpers (all lane-local) ----
    auto bm_test = [&](int s) __attribute__((always_inline)) -> bool {
        const uint64_t w = s < 64 ? bm0 : s < 128 ? bm1 : s < 192 ? bm2 : bm3;
        return (w >> (s & 63)) & 1ull;
    };
    auto bm_set = [&](int s) __attribute__((always_inline)) {
        const uint64_t bit = 1ull << (s & 63);
        const int w = s >> 6;
        bm0 |= w == 0 ? bit : 0ull;
        bm1 |= w == 1 ? bit : 0ull;
        bm2 |= w == 2 ? bit : 0ull;
        bm3 |= w == 3 ? bit : 0ull;
    };
    auto bm_clear = [&](int s) __attribute__((always_inline)) {
        const uint64_t bit = ~(1ull << (s & 63));
        const int w = s >> 6;
        bm0 &= w == 0 ? bit : ~0ull;
        bm1 &= w == 1 ? bit : ~0ull;
        bm2 &= w == 2 ? bit : ~0ull;
        bm3 &= w == 3 ? bit : ~0ull;
    };
    auto bm_any = [&]() __attribute__((always_inline)) -> bool { return (bm0 | bm1 | bm2 | bm3) != 0; };
    auto bm_low = [&]() __attribute__((always_inline)) -> int {
        return bm0 ? __builtin_ctzll(bm0)
                   : bm1 ? 64 + __builtin_ctzll(bm1) : bm2 ? 128 + __builtin_ctzll(bm2) : 192 + __builtin_ctzll(bm3);
    };
    auto chunk_alloc = [&]() __attribute__((always_inline)) -> uint32_t {
        uint32_t c;
        if (spare != NIL) {
            c = spare;
            spare = NIL;
        } else if (free_head != NIL) {
            c = free_head;
            free_head = LINK[c];
        } else if (gleft > 0) {
            c = gbase + (uint32_t)(GRAN - gleft);
            --gleft;
        } else {
            uint32_t g;
            if (own_next < n_own) {
                g = OWN[own_next];
            } else {
                if (n_own >= CA(own_cap)) return NIL;
                g = gran_pop(CA(ghead), CA(gnext), CA(gbump), CA(n_gran));
                if (g == NIL) return NIL;
                OWN[n_own++] = g;
            }
            ++own_next;
            gbase = g * GRAN;
            gleft = GRAN - 1;
            c = gbase;
        }
        return c;
    };
    auto chunk_free = [&](uint32_t c) __attribute__((always_inline)) {
        if (spare == NIL) {
            spare = c;
        } else {
            LINK[c] = free_head;
            free_head = c;
        }
    };
    // gap_push into the arena: src/bwtgap_hd.h:264-286
    auto push_mem = [&](int s, const Ent &en) __attribute__((always_inline)) -> int {
        if (s < 0 || s >= n_stacks) return ERR_SCORE;
        TopT *tp = tops + s * tstride;
        uint32_t chunk, idx;
        if (bm_test(s)) {
            const uint32_t t = *tp;
            chunk = t >> 4;
            idx = (t & 15u) + 1;
            if (idx == CH) {
                const uint32_t nc = chunk_alloc();
                if (nc == NIL) return ERR_STACK;
                LINK[nc] = chunk;
                chunk = nc;
                idx = 0;
            }
        } else {
            const uint32_t nc = chunk_alloc();
            if (nc == NIL) return ERR_STACK;
            LINK[nc] = NIL;
            chunk = nc;
            idx = 0;
            bm_set(s);
        }
        *tp = (TopT)(chunk << 4 | idx);
        CHK[(uint64_t)chunk * CH + idx] = en;
        ++n_entries;
        return 0;
    };
    auto score_of = [&](int mm, int go, int ge) __attribute__((always_inline)) -> int { return mm * s_mm + go * s_gapo + ge * s_gape; };
    // n (<= 16) consecutive pushes into bucket s: the first n1 go to chunk c0 from slot s0,
    // the rest to chunk c1 from slot 0; the bucket's top moves to the last
    auto reserve = [&](int s, int n, uint32_t &c0, uint32_t &s0, uint32_t &n1, uint32_t &c1)
        __attribute__((always_inline)) -> int {
        if (s < 0 || s >= n_stacks) return ERR_SCORE;
        TopT *tp = tops + s * tstride;
        if (bm_test(s)) {
            const uint32_t t = *tp;
            c0 = t >> 4;
            s0 = (t & 15u) + 1;
        } else {
            const uint32_t nc = chunk_alloc();
            if (nc == NIL) return ERR_STACK;
            LINK[nc] = NIL;
            c0 = nc;
            s0 = 0;
            bm_set(s);
        }
        n1 = min((uint32_t)n, (uint32_t)CH - s0);
        if ((uint32_t)n > n1) {
            const uint32_t nc = chunk_alloc();
            if (nc == NIL) return ERR_STACK;
            LINK[nc] = c0;
            c1 = nc;
            *tp = (TopT)(nc << 4 | ((uint32_t)n - n1 - 1));
        } else {
            c1 = NIL;
            *tp = (TopT)(c0 << 4 | (s0 + n1 - 1));
        }
        n_entries += n;
        return 0;
    };

    // result of the current read; the lane then claims the next one
    auto finish = [&](int code) __attribute__((always_inline)) {
        int n = code;
        if (n == ERR_HEAVY) {
            const int h = atomicAdd(CA(heavy_n), 1);
            CA(heavy_list)[h] = rid;
            if (h < CA(hand_cap)) {
                // the widths go along while no hit has changed them (gap_shadow only runs with a
                // new record); the wave engine then skips computing them again
                uint64_t *const dst = CA(hand) + (uint64_t)h * CA(hand_stride);
                if (n_aln == 0) {
                    for (int j = 0; j <= len; ++j) dst[1 + j] = W[j];
                    if (seeding)
                        for (int j = 0; j <= seed_len; ++j) dst[2 + len + j] = SW[j];
                    dst[0] = wbytes | (1ull << 63);
                } else {
                    dst[0] = 0;
                }
            }
        } else {
            if (n > 0) {
                const unsigned long long b0 = atomicAdd(CA(out_top), (unsigned long long)n);
                if (b0 + (unsigned long long)n > CA(out_cap)) {
                    n = ERR_OUT;
                } else {
                    for (int j = 0; j < n; ++j) CA(out_rec)[b0 + j] = REC[j];
                    CA(out_off)[rid] = b0;
                }
            }
            if (n == ERR_STACK || n == ERR_ALN || n == ERR_OUT) {
                const int f = atomicAdd(CA(fail_n), 1);
                CA(fail_list)[f] = rid;
                atomicOr(CA(fail_kind) + (n == ERR_OUT ? 0 : n == ERR_ALN ? 1 : 2), 1);
            }
            CA(out_n)[rid] = n;
            CA(out_pops)[rid] = pops;
            if (TRACE) CA(out_trace)[rid] = trace;
            if (CA(out_bytes)) CA(out_bytes)[rid] = obytes;
            if (n != ERR_STACK && n != ERR_ALN && n != ERR_OUT) atomicAdd(CA(algo_bytes), (unsigned long long)obytes);
        }
        if (n_own > CA(keep)) {
            gran_push(CA(ghead), CA(gnext), OWN, CA(keep), n_own);
            n_own = CA(keep);
        }
        rq_lf = rq_ent = rq_pf = rq_w = false;
        c1p = c2p = nullptr;
        phase = PH_CLAIM;
    };

    // gap_shadow: src/bwtgap_hd.h:309-320 (widths read eight at a time)
    auto shadow = [&](int x, int ldp) __attribute__((always_inline)) {
        const uint64_t ux = (uint64_t)(int64_t)x;
        int j = 0;
        for (int i0 = 0; i0 < ldp; i0 += 8) {
            uint64_t wv[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) wv[t] = i0 + t < ldp ? W[i0 + t] : 0;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                if (i0 + t >= ldp) break;
                uint64_t w = wval(wv[t]);
                int bid = wbid(wv[t]);
                if (w > ux) {
                    w -= ux;
                } else if (w == ux) {
                    bid = 1;
                    w = seq_len - (uint64_t)(++j);
                } else {
                    continue;
                }
                W[i0 + t] = wpack(w, bid);
            }
        }
    };

    // a hit (src/bwtgap_hd.h:385-409): NOFIN, or the code the read finishes with
    auto hit = [&](uint64_t k, uint64_t l) __attribute__((always_inline)) -> int {
        const int mm = e_mm(e), go = e_go(e), ge = e_ge(e);
        const int score = e_score;
        bool do_add = true;
        if (n_aln == 0) {
            best_score = score;
            best_diff = mm + go;
            if (mode & MODE_GAPE) best_diff += ge;
            if (!(mode & MODE_NONSTOP)) max_diff = (best_diff + 1 > md_read) ? md_read : best_diff + 1;
        }
        if (score == best_score) {
            best_cnt = (int)(uint32_t)((uint64_t)(int64_t)best_cnt + (l - k + 1));
        } else if (best_cnt > max_top2) {
            return n_aln;
        }
        if (go) {
            for (int j = 0; j < n_aln; ++j)
                if (REC[j].k == k && REC[j].l == l) {
                    do_add = false;
                    break;
                }
        }
        if (do_add) {
            shadow((int)(uint32_t)(l - k + 1), e_ldp(e));
            if (n_aln == CA(rec_cap)) return ERR_ALN;
            const int ins = e_ins(e), del = go + ge - ins;
            galn_aln_rec_t r;
            r.packed = ((uint64_t)(mm & 0xff)) | ((uint64_t)(go & 0xff) << 8) | ((uint64_t)(ge & 0xff) << 16) |
                       ((uint64_t)(score & 0xfffff) << 24) | ((uint64_t)(ins & 0x3ff) << 44) |
                       ((uint64_t)(del & 0x3ff) << 54);
            r.k = k;
            r.l = l;
            REC[n_aln] = r;
            ++n_aln;
        }
        return NOFIN;
    };

    // The top of bwt_match_gap's loop (src/bwtgap_hd.h:357-383): pop entries until one needs
    // an LF lookup (its loads are requested for the next iteration: NOFIN), an entry has to be
    // loaded first (NOFIN), a popped entry is a hit (HITP, at (hk, hl)), or the search ends
    // (the finish code). have_e: `e` was just loaded and not yet evaluated.
    auto pop_loop = [&](bool have_e) __attribute__((always_inline)) -> int {
        uint32_t ntop = NIL;      // the lowest bucket's top after the last pop from memory
        bool popped_mem = have_e;  // ntop is known
        if (have_e) ntop = ld_ntop;
        for (;;) {
            if (!have_e) {
                if (n_entries == 0 || n_entries > max_entries) return n_aln;
                if (has_cur) {
                    e = cur;
                    e_score = cur_score;
                    has_cur = false;
                    popped_mem = false;
                } else {
                    popped_mem = true;
                    if (!bm_any()) return ERR_INTERNAL;  // entries counted but none held: never expected
                    const int s = bm_low();
                    TopT *tp = tops + s * tstride;
                    const uint32_t t = *tp;
                    const uint32_t chunk = t >> 4, idx = t & 15u;
                    // the bucket's new top: the lowest bucket's top unless the bucket empties
                    if (idx > 0) {
                        ntop = t - 1;
                        *tp = (TopT)ntop;
                    } else {
                        const uint32_t p = LINK[chunk];
                        chunk_free(chunk);  // its entries stay readable until a later push
                        if (p != NIL) {
                            ntop = p << 4 | (CH - 1);
                            *tp = (TopT)ntop;
                        } else {
                            bm_clear(s);
                            ntop = bm_any() ? tops[bm_low() * tstride] : NIL;
                        }
                    }
                    e_score = s;
                    if (pf_valid && pf_addr == t) {
                        e = pf;
                        pf_valid = false;
                    } else {
                        ent_addr = t;
                        rq_ent = true;
                        ld_ntop = ntop;
                        phase = PH_POP;
                        // and the entry now on top of the lowest bucket
                        if (ntop != NIL && (!pf_valid || pf_addr != ntop)) {
                            pf_addr = ntop;
                            pf_valid = false;
                            rq_pf = true;
                        }
                        return NOFIN;
                    }
                }
            }
            have_e = false;
            --n_entries;
            ++pops;
            if (a.pop_budget && pops > a.pop_budget) return ERR_HEAVY;
            const int i0 = e_i(e), mm = e_mm(e), go = e_go(e), ge = e_ge(e), st = e_st(e);
            if (TRACE) {
                const int ins = e_ins(e), del = go + ge - ins;
                trace = tr_mix(trace, (uint64_t)((uint32_t)e_score << 21 | (uint32_t)i0));
                trace = tr_mix(trace, e_k(e));
                trace = tr_mix(trace, e_l(e));
                trace = tr_mix(trace, (uint64_t)((uint32_t)mm | (uint32_t)go << 8 | (uint32_t)ge << 16 | (uint32_t)st << 24));
                trace = tr_mix(trace, (uint64_t)ins | (uint64_t)del << 16);
                trace = tr_mix(trace, (uint64_t)(uint32_t)e_ldp(e));
                trace = tr_mix(trace, (uint64_t)(uint32_t)n_entries);
            }
            if (!(mode & MODE_NONSTOP) && (uint32_t)e_score > (uint32_t)(best_score + s_mm)) return n_aln;
            m = max_diff - (mm + go);
            if (mode & MODE_GAPE) m -= ge;
            if (m < 0) continue;
            if (seeding) {
                m_seed = msd - (mm + go);
                if (mode & MODE_GAPE) m_seed -= ge;
            }
            if (i0 == 0) {
                hk = e_k(e);
                hl = e_l(e);
                return HITP;
            }
            // the LF lookup at (k-1, l), whether the entry expands or starts an exact tail,
            // with the widths and the query bytes it needs
            qa = e_k(e) - 1;
            qb = e_l(e);
            rq_lf = true;
            rq_w = true;
            c1p = q + (i0 - 1);
            c2p = i0 >= 2 ? q + (i0 - 2) : nullptr;
            phase = PH_EXP;
            // and the entry on top of the lowest bucket, usually the next pop (after a pop from
            // memory ntop is that entry; after the exact child in registers it is unchanged)
            if (!popped_mem && bm_any()) ntop = tops[bm_low() * tstride];
            if (ntop != NIL && (!pf_valid || pf_addr != ntop)) {
                pf_addr = ntop;
                pf_valid = false;
                rq_pf = true;
            }
            return NOFIN;
        }
    };

    auto start_search = [&]() __attribute__((always_inline)) -> int {
        wbytes = obytes;
        best_score = score_of(md_read + 1, max_gapo + 1, max_gape + 1);
        best_diff = md_read + 1;
        max_diff = md_read;
        best_cnt = 0;
        n_aln = 0;
        pops = 0;
        trace = TR_INIT;
        bm0 = bm1 = bm2 = bm3 = 0;
        n_entries = 0;
        spare = free_head = NIL;
        own_next = 0;
        gleft = 0;
        has_cur = false;
        pf_valid = false;
        if (nN > md_read) return 0;
        if (n_stacks <= 0) return ERR_SCORE;
        // root: (i = len, k = 0, l = seq_len), score 0
        const Ent root = mk_ent(0, seq_len, len, 0, 0, 0, 0, 0, ST_M);
        if (a.fuse) {
            cur = root;
            cur_score = 0;
            has_cur = true;
            ++n_entries;
            return NOFIN;
        }
        const int r = push_mem(0, root);
        return r ? r : NOFIN;
    };

    // widths (bwt_cal_width, src/bwtgap_hd.h:216-242): consume the query byte `c0` of position
    // cj, loaded by the previous iteration; request the next LF lookup or byte (NOFIN), or,
    // when the widths are done, set the search up (STARTS: pop next; or a finish code)
    auto width_char = [&](int c0) __attribute__((always_inline)) -> int {
        for (;;) {
            if (cj == plen) {
                wout[plen] = wpack(0, ++cbid);
                if (!in_seed && seeding) {
                    in_seed = true;
                    pstr = q + (len - seed_len);
                    plen = seed_len;
                    wout = SW;
                    cj = 0;
                    ck = 0;
                    cl = seq_len;
                    cbid = 0;
                    if (plen == 0) continue;
                    c1p = pstr;
                    return NOFIN;
                }
                const int r = start_search();
                return r == NOFIN ? STARTS : r;
            }
            if (!in_seed) nN += c0 > 3;
            if (c0 <= 3) {
                cbase = 3 - c0;
                qa = ck - 1;
                qb = cl;
                rq_lf = true;
                c1p = cj + 1 < plen ? pstr + (cj + 1) : nullptr;
                return NOFIN;
            }
            // N: restart without a lookup
            ck = 0;
            cl = seq_len;
            ++cbid;
            wout[cj] = wpack(seq_len + 1, cbid);
            ++cj;
            if (cj < plen) {
                c1p = pstr + cj;
                return NOFIN;
            }
            c0 = 4;
        }
    };

    auto init_read = [&](int pos) __attribute__((always_inline)) -> int {
        rid = CA(list) ? CA(list)[pos] : CA(list_base) + pos;
        len = CA(qlen)[rid];
        q = CA(query) + CA(qoff)[rid];
        const int b = CA(batch)[rid];
        const galn_search_opt_t o = CA(bopts)[b];
        s_mm = o.s_mm;
        s_gapo = o.s_gapo;
        s_gape = o.s_gape;
        mode = o.mode;
        ies = o.indel_end_skip;
        mdo = o.max_del_occ;
        max_entries = o.max_entries;
        max_gapo = o.max_gapo;
        max_gape = o.max_gape;
        msd = o.max_seed_diff;
        max_top2 = o.max_top2;
        n_stacks = CA(nstacks)[b];
        md_read = CA(maxdiff)[rid];
        seed_len = o.seed_len;
        seeding = seed_len < len;
        pops = 0;
        trace = TR_INIT;
        obytes = 0;
        pf_valid = false;
        if (len < 1 || len > MAXLEN || n_stacks > MAXB || md_read > MAXDIFF || max_gapo > MAXDIFF ||
            max_gapo + max_gape > 255 || max_gape < 0)
            return ERR_UNSUP;
        in_seed = false;
        pstr = q;
        plen = len;
        wout = W;
        cj = 0;
        ck = 0;
        cl = seq_len;
        cbid = 0;
        nN = 0;
        c1p = q;
        phase = PH_CW;
        return NOFIN;
    };

    // ---- main loop: every iteration loads what the lane asked for, then works on it ----
    for (;;) {
        int fin = NOFIN;
        if (phase == PH_CLAIM) {
            int pos;
            if (PERSIST) {
                pos = atomicAdd(CA(qnext), 1);
            } else {
                if (claimed) break;
                pos = slot;
            }
            claimed = true;
            if (pos >= CA(n_list)) break;
            fin = init_read(pos);
        } else {
            // ---- loads, all issued before any is used ----
            uint64_t wA = 0, wB = 0, sA = 0, sB = 0;
            int ch1 = 4, ch2 = 4;
            Ent ld{0, 0};
            uint64_t ca[4] = {0, 0, 0, 0}, cb[4] = {0, 0, 0, 0};
            uint32_t lf_bytes = 0;
            if (rq_w) {
                const int wi = e_i(e);
                wB = W[wi - 1];
                if (wi >= 2) wA = W[wi - 2];
                if (seeding) {
                    const int ii = wi - 1 - (len - seed_len);
                    if (ii > 0) {
                        sA = SW[ii - 1];
                        sB = SW[ii];
                    }
                }
            }
            if (c1p) ch1 = *c1p;
            if (c2p) ch2 = *c2p;
            if (rq_ent) ld = CHK[(uint64_t)(ent_addr >> 4) * CH + (ent_addr & 15u)];
            if (rq_pf) pf = CHK[(uint64_t)(pf_addr >> 4) * CH + (pf_addr & 15u)];
            if (rq_lf) lf_bytes = lf_pair(a.bwt, M, qa, qb, ca, cb);
            if (rq_pf) pf_valid = true;
            const bool had_lf = rq_lf;
            rq_lf = rq_ent = rq_pf = rq_w = false;
            c1p = c2p = nullptr;

            // ---- work: each phase ends in new loads, a hit and/or a pop, or a finish ----
            bool want_pop = false, have_e = false, want_hit = false;
            if (phase == PH_CW) {
                if (had_lf) {
                    obytes += lf_bytes;
                    const int c = cbase;
                    uint64_t k = M.L2[c] + ca[c] + 1, l = M.L2[c] + cb[c];
                    if (k > l) {
                        k = 0;
                        l = seq_len;
                        ++cbid;
                    }
                    wout[cj] = wpack(l - k + 1, cbid);
                    ck = k;
                    cl = l;
                    ++cj;
                }
                const int r = width_char(ch1);
                if (r == STARTS)
                    want_pop = true;
                else
                    fin = r;
            } else if (phase == PH_POP) {
                e = ld;
                want_pop = have_e = true;
            } else if (phase == PH_EXACT) {
                // one step of an exact tail (bwt_match_exact_alt, src/bwtgap_hd.h:201-208)
                obytes += lf_bytes;
                const int c = cbase;
                const uint64_t k = M.L2[c] + ca[c] + 1, l = M.L2[c] + cb[c];
                want_pop = true;
                if (k <= l) {
                    if (cj == 0) {
                        want_hit = (int)(uint32_t)(l - k + 1) != 0;
                        hk = k;
                        hl = l;
                    } else if (ch1 <= 3) {
                        --cj;
                        cbase = ch1;
                        qa = k - 1;
                        qb = l;
                        rq_lf = true;
                        c1p = cj >= 1 ? q + (cj - 1) : nullptr;
                        want_pop = false;
                    }
                }
            } else {
                // PH_EXP: the popped entry's lookup, widths and query bytes are here
                const int i0 = e_i(e);
                const int st = e_st(e), ge = e_ge(e);
                const int qbase = ch1;
                want_pop = true;
                if (m < wbid(wB)) {
                    // pruned by the widths
                } else if (m == 0 && (st == ST_M || (mode & MODE_GAPE) || ge == max_gape)) {
                    // exact tail: its first step is this lookup
                    if (qbase <= 3) {
                        obytes += lf_bytes;
                        const uint64_t k = M.L2[qbase] + ca[qbase] + 1, l = M.L2[qbase] + cb[qbase];
                        if (k <= l) {
                            if (i0 < 2) {
                                want_hit = (int)(uint32_t)(l - k + 1) != 0;
                                hk = k;
                                hl = l;
                            } else if (ch2 <= 3) {
                                cj = i0 - 2;
                                cbase = ch2;
                                qa = k - 1;
                                qb = l;
                                rq_lf = true;
                                c1p = cj >= 1 ? q + (cj - 1) : nullptr;
                                phase = PH_EXACT;
                                want_pop = false;
                            }
                        }
                    }
                } else {
                    // expansion: src/bwtgap_hd.h:412-470
                    obytes += lf_bytes;
                    const int i = i0 - 1;
                    const int mm = e_mm(e), go = e_go(e), ins = e_ins(e);
                    const uint64_t k = e_k(e), l = e_l(e), occ = l - k + 1;
                    bool allow_diff = true, allow_M = true;
                    if (i > 0) {
                        if (wbid(wA) > m - 1)
                            allow_diff = false;
                        else if (wbid(wA) == m - 1 && wbid(wB) == m - 1 && wval(wA) == wval(wB))
                            allow_M = false;
                        if (seeding) {
                            const int ii = i - (len - seed_len);
                            if (ii > 0) {
                                if (wbid(sA) > m_seed - 1)
                                    allow_diff = false;
                                else if (wbid(sA) == m_seed - 1 && wbid(sB) == m_seed - 1 && wval(sA) == wval(sB))
                                    allow_M = false;
                            }
                        }
                    }
                    const int tmp = (mode & MODE_LOGGAP) ? int_log2((uint32_t)(ge + go)) / 2 + 1 : go + ge;
                    const bool gap_ok = allow_diff && i >= ies + tmp && len - i >= ies + tmp;
                    // the children, in the order bwt_match_gap pushes them:
                    // bit 0 insertion, bits 1-4 deletions (bases 0..3), bits 5-8 the bases
                    // qbase+1 .. qbase+4 (the last one is the exact child), bit 9 the exact child
                    // alone (when differences are not allowed here)
                    // the children's intervals per base (constant indices: the counts come from
                    // scalar loads, not per-lane loads of the kernel arguments)
                    uint64_t KK[4], LL[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        KK[j] = M.L2[j] + ca[j] + 1;
                        LL[j] = M.L2[j] + cb[j];
                    }
                    auto pick = [](const uint64_t(&v)[4], int j) __attribute__((always_inline)) {
                        return j == 0 ? v[0] : j == 1 ? v[1] : j == 2 ? v[2] : v[3];
                    };
                    uint32_t mask = 0;
                    if (gap_ok) {
                        const bool open = st == ST_M && go < max_gapo;
                        if (open || (st == ST_I && ge < max_gape)) mask |= 1u;
                        if (open || (st == ST_D && ge < max_gape && (ge + go < max_diff || occ < (uint64_t)(int64_t)mdo))) {
    #pragma unroll
                            for (int j = 0; j < 4; ++j)
                                if (KK[j] <= LL[j]) mask |= 2u << j;
                        }
                    }
                    if (allow_diff && allow_M) {
    #pragma unroll
                        for (int j = 1; j <= 4; ++j) {
                            const int c = (qbase + j) & 3;
                            if (ca[c] + 1 <= cb[c]) mask |= 16u << j;
                        }
                    } else if (qbase < 4 && ca[qbase] + 1 <= cb[qbase]) {
                        mask |= 512u;
                    }
                    // Pushes, in bwt_match_gap's order: the gap children (bits 0-4) all score sg
                    // and the mismatches sm, so each group takes consecutive slots of one bucket
                    // (two groups with one score share it, gap children first); the exact child
                    // (bit 8 with qbase <= 3, or bit 9) comes last with the parent's score.
                    const uint32_t xbit = (qbase <= 3 ? 256u : 0u) | 512u;
                    const uint32_t mx = mask & xbit;
                    const uint32_t mg = mask & 31u, mmk = mask & 0x1e0u & ~xbit;
                    const int sg = st == ST_M ? score_of(mm, go + 1, ge) : score_of(mm, go, ge + 1);
                    const int sm = score_of(mm + 1, go, ge);
                    const int ng = __builtin_popcount(mg), nm = __builtin_popcount(mmk);
                    const bool merged = ng > 0 && nm > 0 && sg == sm;
                    uint32_t gc0 = 0, gs0 = 0, gn1 = 0, gc1 = 0, mc0 = 0, ms0 = 0, mn1 = 0, mc1 = 0;
                    int err = 0;
                    if (ng + (merged ? nm : 0) > 0) err = reserve(sg, ng + (merged ? nm : 0), gc0, gs0, gn1, gc1);
                    if (!err && nm > 0 && !merged) err = reserve(sm, nm, mc0, ms0, mn1, mc1);
                    if (!err) {
                        uint32_t rest = mg | mmk;
                        int rg = 0, rm = merged ? ng : 0;
                        while (rest) {
                            const int bit = __builtin_ctz(rest);
                            rest &= rest - 1;
                            Ent ch;
                            int r;
                            bool in_g = bit <= 4 || merged;
                            if (bit == 0) {
                                const int go2 = st == ST_M ? go + 1 : go, ge2 = st == ST_M ? ge : ge + 1;
                                ch = mk_ent(k, l, i, i, mm, go2, ge2, ins + 1, ST_I);
                            } else if (bit <= 4) {
                                const int j = bit - 1;
                                const uint64_t kk = pick(KK, j), ll = pick(LL, j);
                                const int go2 = st == ST_M ? go + 1 : go, ge2 = st == ST_M ? ge : ge + 1;
                                ch = mk_ent(kk, ll, i + 1, i + 1, mm, go2, ge2, ins, ST_D);
                            } else {
                                const int c = (qbase + bit - 4) & 3;
                                const uint64_t kk = pick(KK, c), ll = pick(LL, c);
                                ch = mk_ent(kk, ll, i, i, mm + 1, go, ge, ins, ST_M);
                            }
                            if (bit <= 4)
                                r = rg++;
                            else
                                r = rm++;
                            const uint32_t c0 = in_g ? gc0 : mc0, s0 = in_g ? gs0 : ms0, n1 = in_g ? gn1 : mn1,
                                           c1 = in_g ? gc1 : mc1;
                            const uint32_t slot_ = (uint32_t)r < n1 ? c0 * CH + s0 + (uint32_t)r : c1 * CH + ((uint32_t)r - n1);
                            CHK[slot_] = ch;
                        }
                        if (mx) {
                            const uint64_t kk = pick(KK, qbase & 3), ll = pick(LL, qbase & 3);  // (qbase <= 3 here)
                            const Ent ch = mk_ent(kk, ll, i, 0, mm, go, ge, ins, ST_M);
                            if (a.fuse) {
                                // pushed last with the parent's score: the next pop; kept in registers
                                cur = ch;
                                cur_score = e_score;
                                has_cur = true;
                                ++n_entries;
                            } else {
                                uint32_t c0, s0, n1, c1;
                                err = reserve(e_score, 1, c0, s0, n1, c1);
                                if (!err) CHK[n1 ? c0 * CH + s0 : c1 * CH] = ch;
                            }
                        }
                    }
                    if (err) {
                        fin = err;
                        want_pop = false;
                    }
                }
            }
            // ---- hits and pops, until new loads are requested or the read finishes ----
            for (;;) {
                if (want_hit) {
                    want_hit = false;
                    const int h = hit(hk, hl);
                    if (h != NOFIN) {
                        fin = h;
                        break;
                    }
                }
                if (!want_pop) break;
                const int r = pop_loop(have_e);
                have_e = false;
                if (r == HITP) {
                    want_hit = true;
                    continue;
                }
                fin = r;
                break;
            }
        }
        if (fin != NOFIN) finish(fin);
    }
}

// ---------------------------------------------------------------------------
// SA value of BWT rows (bwt_sa with bwt_invPsi, upstream bwt.c:53-59, :88-96): LF steps from each
// row to the nearest row divisible by the sample interval (fewer than intv steps, one 64-byte block
// each), then that row's sample plus the steps. Row 0, the $ row, counts as -1 as bwa sets it (a
// walk that ends there gives steps - 1). samples[j] is the suffix at row (j + 1) * intv; rows are
// at most seq_len (checked by the caller).
__global__ void __launch_bounds__(256) k_sa_walk(const uint32_t *__restrict__ bwt, DMeta m,
                                                  const uint64_t *__restrict__ samples, uint64_t intv,
                                                  const uint64_t *__restrict__ rows, uint64_t n, uint64_t *__restrict__ pos,
                                                  const int32_t *__restrict__ ref_len, int32_t *__restrict__ strand)
{
    const L2Sel L2(m);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k = rows[i], sa = 0;
        while (k & (intv - 1)) {
            ++sa;
            if (k == m.primary) {
                k = 0;
                break;
            }
            const uint64_t x = k - (k > m.primary);
            uint4 B[4];
            load_block(bwt, x >> 7, B);
            const int r = (int)(x & 127);
            const uint32_t hw = r < 32 ? B[2].x : r < 64 ? B[2].y : r < 96 ? B[2].z : B[2].w;
            const uint32_t lw = r < 32 ? B[3].x : r < 64 ? B[3].y : r < 96 ? B[3].z : B[3].w;
            const int sh = 31 - (r & 31);
            const int c = (int)(((hw >> sh) & 1u) << 1 | ((lw >> sh) & 1u));
            k = L2[c] + block_occ1(B, r, c);
        }
        uint64_t p = sa + (k == 0 ? ~0ull : samples[k / intv - 1]);
        if (ref_len) {
            // bwa_sa2pos (upstream bwase.c:112-123): a hit across the forward/reverse boundary is -1;
            // on the reverse half, the forward position of the hit's first base, strand 0
            const uint64_t l_pac = m.seq_len >> 1;
            const uint64_t rl = (uint64_t)(int64_t)ref_len[i];
            int st = 0;
            if (p < l_pac && l_pac < p + rl) {
                p = ~0ull;
            } else {
                const bool rev = p >= l_pac;
                if (rev) p = (l_pac << 1) - 1 - p;
                st = rev ? 0 : 1;
                if (rev) p = p + 1 < rl ? 0 : p - rl + 1;
            }
            strand[i] = st;
        }
        pos[i] = p;
    }
}

// ---------------------------------------------------------------------------
// Wave-per-read engine for heavy reads.
//
// Between two hits the search's state (max_diff, widths, best score) is fixed,
// and with positive penalties every non-exact child lands in a higher bucket
// than its parent. So the entries of the lowest bucket b each start an
// independent "chain" -- the entry, then its exact child, then that child's
// exact child, ... (the sequential search pops exactly this chain next, LIFO) --
// whose pushes all go to buckets above b. A wave takes the top K <= 64 entries
// of bucket b, runs the K chains in parallel (one lane each, pushes logged),
// then commits them in the sequential order: pops, pushes (placed by a
// per-bucket scan so every bucket keeps its LIFO order), removal of the K
// entries, and the first hit, which changes the state and ends the commit.
// Chains after that hit are discarded and re-run in the next round. The
// result, the pop count and the trace are those of the sequential search.
//
// Buckets are LIFO lists of segments (a 16-byte header naming the segment
// below, then the entries) in granules the wave takes from the stack pool.

constexpr int HW = 64;         // lanes, one wave per workgroup
constexpr int H_STEPS = 32;    // chain steps per lane per round
constexpr int H_PUSH = 128;    // logged pushes per lane per round
constexpr int H_CAP = 4096;    // pushes committed per round
static_assert(H_CAP < 65536, "a round's bucket totals are summed in 16-bit fields");
constexpr int H_KEEP = 64;     // granules a wave keeps between reads
constexpr int H_PIECES = 48;    // segment pieces one round may add (three buckets)
constexpr uint32_t GENT = (uint32_t)GRAN * CH;  // entries per granule

enum : int { K_END = 0, K_HIT = 1, K_TRUNC = 2, K_ESCORE = 3 };

struct HArgs {
    const uint32_t *bwt;
    DMeta meta;
    const uint8_t *query;
    const uint64_t *qoff;
    const int32_t *qlen;
    const int32_t *maxdiff;
    const int32_t *batch;
    const galn_search_opt_t *bopts;
    const int32_t *nstacks;
    const int32_t *list;
    int32_t n_list;
    const int32_t *n_list_dev;  // when set, the number of reads is read here (written by k_search)
    const uint64_t *hand;       // widths handed over by k_search (slot = list position), or null
    uint32_t hand_stride;
    int32_t hand_cap;
    int32_t *qnext;
    int32_t n_slots;
    uint8_t *scratch;
    uint64_t slot_bytes;
    uint32_t rec_off, own_off, plog_off, slog_off;
    int32_t own_cap, rec_cap;
    Ent *pool;
    uint32_t *gnext;
    unsigned long long *ghead;
    uint32_t *gbump;
    uint32_t n_gran;
    int32_t *out_n;
    uint64_t *out_off, *out_pops, *out_trace, *out_bytes;
    galn_aln_rec_t *out_rec;
    uint64_t out_cap;
    unsigned long long *out_top;
    int32_t *fail_list, *fail_n, *fail_kind;
    unsigned long long *algo_bytes;
    int32_t buckets;  // LDS buckets: the largest n_stacks of the launch
    int32_t wlen;     // LDS widths: longest read + 1
    int32_t swlen;    // LDS seed widths
    int32_t push_cap; // pushes committed per round, H_PUSH..H_CAP (H_CAP unless a test lowers it)
    unsigned long long *stats;  // [8] engine counters (rounds, chains, steps, cycles)
};

struct HShared {
    int rid, len, n_stacks, md_read, seeding, seed_len;
    int s_mm, s_gapo, s_gape, mode, ies, mdo, max_entries, max_gapo, max_gape, msd, max_top2;
    int best_score, best_diff, max_diff, best_cnt, n_aln, n_entries;
    unsigned long long pops, trace, obytes;
    unsigned long long bm[4];
    int has_pend, claim, status, result, K, b, J, jcap, jcut, tcut, n_piece, n_own, n_used, aleft;
    uint32_t abase;
    int nN;
    int l_full;             // the last committed chain's stop kind
    unsigned long long st[16];
    uint64_t hk, hl;
    Ent pend, he;
    int he_score;
};

__device__ __forceinline__ int hbm_low(const unsigned long long *bm)
{
    return bm[0] ? __builtin_ctzll(bm[0]) : bm[1] ? 64 + __builtin_ctzll(bm[1]) : bm[2] ? 128 + __builtin_ctzll(bm[2])
                 : bm[3] ? 192 + __builtin_ctzll(bm[3]) : -1;
}

// wave-wide (64-lane) scans and reductions; every lane of the wave must call them
__device__ __forceinline__ int wave_excl_scan(int v)
{
    const int lane = threadIdx.x & (HW - 1);
    int x = v;
    for (int d = 1; d < HW; d <<= 1) {
        const int t = __shfl_up(x, d, HW);
        if (lane >= d) x += t;
    }
    return x - v;
}


__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int d = HW / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, HW);
    return v;
}

__device__ __forceinline__ int wave_max_i32(int v)
{
    for (int d = HW / 2; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, HW));
    return v;
}

// bwt_cal_width by one lane (src/bwtgap_hd.h:216-242); returns the algorithmic bytes
__device__ uint64_t cal_width_one(const uint32_t *bwt, const DMeta &M, int len, const uint8_t *str, uint64_t *w)
{
    const L2Sel L2s(M);
    uint64_t k = 0, l = M.seq_len, bytes = 0;
    int bid = 0;
    for (int i = 0; i < len; ++i) {
        const int c0 = str[i];
        const int c = c0 > 3 ? 4 : 3 - c0;
        if (c < 4) {
            uint64_t oa, ob;
            bytes += lf_one(bwt, M, k - 1, l, c, oa, ob);
            k = L2s[c] + oa + 1;
            l = L2s[c] + ob;
        }
        if (k > l || c > 3) {
            k = 0;
            l = M.seq_len;
            ++bid;
        }
        w[i] = wpack(l - k + 1, bid);
    }
    w[len] = wpack(0, ++bid);
    return bytes;
}

// Widths of the reads given to the wave engine whose hand-off slot is empty (a record found by
// k_search had already changed them): one thread per read, all reads at once, so that the wave
// engine does not hold 64 lanes while one of them walks a read's sequential width pass.
struct WArgs {
    const uint32_t *bwt;
    DMeta meta;
    const uint8_t *query;
    const uint64_t *qoff;
    const int32_t *qlen;
    const int32_t *batch;
    const galn_search_opt_t *bopts;
    const int32_t *list;
    int32_t n_list;
    const int32_t *n_list_dev;
    uint64_t *hand;
    uint32_t hand_stride;
    int32_t hand_cap;
    int32_t wlen, swlen;
};

__global__ void __launch_bounds__(256) k_widths(WArgs a)
{
    const int n = min(a.n_list_dev ? *a.n_list_dev : a.n_list, a.hand_cap);
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < n; pos += gridDim.x * blockDim.x) {
        uint64_t *const dst = a.hand + (uint64_t)pos * a.hand_stride;
        if (dst[0] >> 63) continue;  // handed over by k_search
        const int rid = a.list[pos];
        const int len = a.qlen[rid];
        if (len < 1 || len + 1 > a.wlen) continue;  // the wave engine rejects it
        const int sl = a.bopts[a.batch[rid]].seed_len;
        if (sl < len && sl + 1 > a.swlen) continue;
        const uint8_t *const q = a.query + a.qoff[rid];
        uint64_t wb = cal_width_one(a.bwt, a.meta, len, q, dst + 1);
        if (sl < len) wb += cal_width_one(a.bwt, a.meta, sl, q + (len - sl), dst + 2 + len);
        dst[0] = wb | (1ull << 63);
    }
}

// WPE waves per SIMD: 4 (16 per CU) at 128 VGPRs, the compiler spilling a few values to scratch; the
// extra resident waves hide more of the index latency than the spills cost (measured 3.4% against 3
// per SIMD on the 100 Mbp workload, where 5 was 10% slower: that index sits in the MALL and the
// engine is bound by instruction issue). 5 (96 VGPRs, more spills) for indexes in HBM, where a wave
// waits on its loads: 20% faster on the 3.1 Gbp human-like workload, 25% slower on the 20 Gbp one.
template <bool TRACE, int WPE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) k_heavy(HArgs a)
{
    extern __shared__ __align__(16) uint8_t hlds[];
    HShared &F = *reinterpret_cast<HShared *>(hlds);
    uint8_t *lp = hlds + ((sizeof(HShared) + 15) & ~(size_t)15);
    const int NB = a.buckets;
    uint32_t *seg_base = reinterpret_cast<uint32_t *>(lp);
    uint32_t *seg_cnt = seg_base + NB;
    int32_t *first_piece = reinterpret_cast<int32_t *>(seg_cnt + NB);
    lp = reinterpret_cast<uint8_t *>(first_piece + NB);
    lp = hlds + (((size_t)(lp - hlds) + 15) & ~(size_t)15);
    uint4 *pieces = reinterpret_cast<uint4 *>(lp);  // {bucket, rank0, len, base}
    uint64_t *W = reinterpret_cast<uint64_t *>(pieces + H_PIECES);
    uint64_t *SW = W + a.wlen;
    int32_t *lcommit = reinterpret_cast<int32_t *>(SW + a.swlen);  // per chain, for the trace hash
    int32_t *lnstart = lcommit + HW;
    // per chain step: pushes (4 bits) | index bytes / 64 << 4 (a step's bytes are whole 64-byte
    // blocks, at most 2 per position of a read of at most 1023: under 4096 units)
    uint16_t *const SINFO_ALL = reinterpret_cast<uint16_t *>(lnstart + HW);
    uint8_t *const QL = reinterpret_cast<uint8_t *>(SINFO_ALL + HW * H_STEPS);  // the read's query

    const int lane = threadIdx.x;
    if ((int)blockIdx.x >= a.n_slots) return;
    uint8_t *const base = a.scratch + (uint64_t)blockIdx.x * a.slot_bytes;
    galn_aln_rec_t *const REC = reinterpret_cast<galn_aln_rec_t *>(base + a.rec_off);
    uint32_t *const OWN = reinterpret_cast<uint32_t *>(base + a.own_off);
    Ent *const PLOG = reinterpret_cast<Ent *>(base + a.plog_off) + lane * H_PUSH;
    Ent *const SLOG_ALL = reinterpret_cast<Ent *>(base + a.slog_off);
    Ent *const SLOG = SLOG_ALL + lane * H_STEPS;
    uint16_t *const SINFO = SINFO_ALL + lane * H_STEPS;
    static_assert(MAXLEN * 128 / 64 < 4096, "a step's index bytes fit the step info");
    Ent *const POOL = a.pool;
    const DMeta &M = a.meta;
    const L2Sel L2s(M);
    const uint64_t seq_len = M.seq_len;

    if (lane == 0) {
        for (int j = 0; j < 16; ++j) F.st[j] = 0;
        F.n_own = 0;  // granules this wave holds; a read reuses them before taking more
    }
    const int n_list = a.n_list_dev ? *a.n_list_dev : a.n_list;
    for (;;) {
        const long long t_read = a.stats ? wall_clock64() : 0;
        if (lane == 0) F.claim = atomicAdd(a.qnext, 1);
        __syncthreads();
        const int pos = F.claim;
        if (pos >= n_list) break;
        // ---- read set-up (lane 0), widths (lanes 0 and 1) ----
        const int rid = a.list[pos];
        const int len = a.qlen[rid];
        const uint8_t *const qg = a.query + a.qoff[rid];
        const uint8_t *const q = QL;
        for (int j = lane; j < len && j < a.wlen; j += HW) QL[j] = qg[j];
        if (lane == 0) {
            const int b0 = a.batch[rid];
            const galn_search_opt_t o = a.bopts[b0];
            F.rid = rid;
            F.len = len;
            F.s_mm = o.s_mm;
            F.s_gapo = o.s_gapo;
            F.s_gape = o.s_gape;
            F.mode = o.mode;
            F.ies = o.indel_end_skip;
            F.mdo = o.max_del_occ;
            F.max_entries = o.max_entries;
            F.max_gapo = o.max_gapo;
            F.max_gape = o.max_gape;
            F.msd = o.max_seed_diff;
            F.max_top2 = o.max_top2;
            F.n_stacks = a.nstacks[b0];
            F.md_read = a.maxdiff[rid];
            F.seed_len = o.seed_len;
            F.seeding = o.seed_len < len;
            F.nN = 0;
            F.obytes = 0;
            F.n_used = 0;
            F.aleft = 0;
            F.abase = 0;
            F.status = 0;
        }
        __syncthreads();
        const bool unsup = len < 1 || len > MAXLEN || len + 1 > a.wlen || F.n_stacks > NB || F.md_read > MAXDIFF ||
                           F.max_gapo > MAXDIFF || F.max_gapo + F.max_gape > 255;
        const uint64_t h0 = (!unsup && a.hand && pos < a.hand_cap) ? a.hand[(uint64_t)pos * a.hand_stride] : 0;
        if (!unsup && (h0 >> 63)) {
            // the widths k_search computed for this read
            const uint64_t *src = a.hand + (uint64_t)pos * a.hand_stride;
            for (int j = lane; j <= len; j += HW) W[j] = src[1 + j];
            if (F.seeding)
                for (int j = lane; j <= F.seed_len; j += HW) SW[j] = src[2 + len + j];
            if (lane == 0) F.obytes = h0 & ~(1ull << 63);
            int nN = 0;
            for (int j = lane; j < len; j += HW) nN += q[j] > 3;
            if (nN) atomicAdd(&F.nN, nN);
        } else if (!unsup) {
            uint64_t wb = 0;
            if (lane == 0) wb = cal_width_one(a.bwt, M, len, q, W);
            if (lane == 1 && F.seeding) wb = cal_width_one(a.bwt, M, F.seed_len, q + (len - F.seed_len), SW);
            int nN = 0;
            for (int j = lane; j < len; j += HW) nN += q[j] > 3;
            if (nN) atomicAdd(&F.nN, nN);
            if (wb) atomicAdd(&F.obytes, (unsigned long long)wb);
        }
        __syncthreads();
        // ---- search set-up ----
        if (lane == 0) {
            F.best_score = (F.md_read + 1) * F.s_mm + (F.max_gapo + 1) * F.s_gapo + (F.max_gape + 1) * F.s_gape;
            F.best_diff = F.md_read + 1;
            F.max_diff = F.md_read;
            F.best_cnt = 0;
            F.n_aln = 0;
            F.pops = 0;
            F.trace = TR_INIT;
            F.bm[0] = F.bm[1] = F.bm[2] = F.bm[3] = 0;
            F.has_pend = 0;
            F.n_entries = 0;
            if (unsup) {
                F.status = 1;
                F.result = ERR_UNSUP;
            } else if (F.nN > F.md_read) {
                F.status = 1;
                F.result = 0;
            } else if (F.n_stacks <= 0) {
                F.status = 1;
                F.result = ERR_SCORE;
            } else {
                F.has_pend = 1;
                F.pend = mk_ent(0, seq_len, len, 0, 0, 0, 0, 0, ST_M);
                F.b = 0;
                F.n_entries = 1;
            }
        }
        __syncthreads();
        if (a.stats && lane == 0) {
            F.st[12] += (unsigned long long)(wall_clock64() - t_read);
            F.st[15] += 1;
        }
        // ---- rounds ----
        long long t_round_end = 0;  // with stats: end of the last round's commit
        while (F.status == 0) {
            const long long t_round = a.stats ? wall_clock64() : 0;
            if (a.stats && lane == 0 && t_round_end) F.st[14] += (unsigned long long)(t_round - t_round_end);
            if (lane == 0) {
                // a pending exact child is the top of the lowest bucket (the bucket it came from)
                const int b = F.has_pend ? F.b : hbm_low(F.bm);
                if (F.has_pend == 0 && b < 0) {
                    F.status = 1;
                    F.result = F.n_aln;
                } else if (F.n_entries > F.max_entries) {
                    F.status = 1;
                    F.result = F.n_aln;
                } else {
                    F.b = b;
                    const bool bit = (F.bm[b >> 6] >> (b & 63)) & 1ull;
                    const uint32_t avail = (uint32_t)F.has_pend + (bit ? seg_cnt[b] : 0u);
                    if (!(F.mode & MODE_NONSTOP) && (uint32_t)b > (uint32_t)(F.best_score + F.s_mm)) {
                        // this pop ends the search (src/bwtgap_hd.h:366)
                        const Ent e = F.has_pend ? F.pend : POOL[seg_base[b] + seg_cnt[b] - 1];
                        F.pops += 1;
                        F.n_entries -= 1;
                        if (TRACE) {
                            const int i0 = e_i(e), mm = e_mm(e), go = e_go(e), ge = e_ge(e), st = e_st(e), ins = e_ins(e);
                            uint64_t t = F.trace;
                            t = tr_mix(t, (uint64_t)((uint32_t)b << 21 | (uint32_t)i0));
                            t = tr_mix(t, e_k(e));
                            t = tr_mix(t, e_l(e));
                            t = tr_mix(t, (uint64_t)((uint32_t)mm | (uint32_t)go << 8 | (uint32_t)ge << 16 | (uint32_t)st << 24));
                            t = tr_mix(t, (uint64_t)ins | (uint64_t)(go + ge - ins) << 16);
                            t = tr_mix(t, (uint64_t)(uint32_t)e_ldp(e));
                            t = tr_mix(t, (uint64_t)(uint32_t)F.n_entries);
                            F.trace = t;
                        }
                        F.status = 1;
                        F.result = F.n_aln;
                    } else {
                        F.K = (int)min(avail, (uint32_t)HW);
                    }
                }
            }
            __syncthreads();
            if (F.status) break;
            // ---- speculation: lane j runs the chain of the j-th entry from the top of bucket b ----
            const int b = F.b, K = F.K, hp = F.has_pend;
            const long long t_spec = a.stats ? wall_clock64() : 0;
            if (a.stats && lane == 0) F.st[13] += (unsigned long long)(t_spec - t_round);
            const int s_mm = F.s_mm, s_gapo = F.s_gapo, s_gape = F.s_gape, mode = F.mode, max_diff = F.max_diff;
            const int n_stacks = F.n_stacks, max_gapo = F.max_gapo, max_gape = F.max_gape, ies = F.ies, mdo = F.mdo;
            const int msd = F.msd, seed_len = F.seed_len;
            const bool seeding = F.seeding != 0;
            int steps = 0, np = 0, kind = K_END;
            uint32_t sby = 0;    // index bytes of the chain's steps
            uint32_t ccnt = 0;   // logged pushes per bucket code, 8 bits each
            uint64_t hk = 0, hl = 0;
            Ent nxt{0, 0};
            if (lane < K) {
                Ent e = (lane == 0 && hp) ? F.pend : POOL[seg_base[b] + seg_cnt[b] - 1 - (uint32_t)(lane - hp)];
                for (;;) {
                    if (steps == H_STEPS || np > H_PUSH - 9) {
                        kind = K_TRUNC;
                        nxt = e;
                        break;
                    }
                    if (TRACE) SLOG[steps] = e;  // the trace hash runs over every committed pop
                    uint32_t sp = 0, sbytes = 0;
                    bool cont = false;
                    Ent child{0, 0};
                    const int i0 = e_i(e), mm = e_mm(e), go = e_go(e), ge = e_ge(e), st = e_st(e), ins = e_ins(e);
                    int m = max_diff - (mm + go);
                    if (mode & MODE_GAPE) m -= ge;
                    if (m >= 0) {
                        int m_seed = 0;
                        if (seeding) {
                            m_seed = msd - (mm + go);
                            if (mode & MODE_GAPE) m_seed -= ge;
                        }
                        if (i0 == 0) {
                            kind = K_HIT;
                            hk = e_k(e);
                            hl = e_l(e);
                        } else {
                            const uint64_t wB = W[i0 - 1], wA = i0 >= 2 ? W[i0 - 2] : 0;
                            const int qbase = q[i0 - 1];
                            uint64_t ca[4], cb[4];
                            const uint32_t lfb = lf_pair(a.bwt, M, e_k(e) - 1, e_l(e), ca, cb);
                            if (m < wbid(wB)) {
                                // pruned
                            } else if (m == 0 && (st == ST_M || (mode & MODE_GAPE) || ge == max_gape)) {
                                // exact tail: bwt_match_exact_alt, src/bwtgap_hd.h:195-211
                                if (qbase <= 3) {
                                    sbytes += lfb;
                                    uint64_t k = L2s[qbase] + ca[qbase] + 1, l = L2s[qbase] + cb[qbase];
                                    bool ok = k <= l;
                                    for (int jn = i0 - 2; ok && jn >= 0; --jn) {
                                        const int c = q[jn];
                                        if (c > 3) {
                                            ok = false;
                                            break;
                                        }
                                        uint64_t oa, ob;
                                        sbytes += lf_one(a.bwt, M, k - 1, l, c, oa, ob);
                                        k = L2s[c] + oa + 1;
                                        l = L2s[c] + ob;
                                        ok = k <= l;
                                    }
                                    if (ok && (int)(uint32_t)(l - k + 1) != 0) {
                                        kind = K_HIT;
                                        hk = k;
                                        hl = l;
                                    }
                                }
                            } else {
                                // expansion: src/bwtgap_hd.h:412-470
                                sbytes += lfb;
                                const int i = i0 - 1;
                                const uint64_t k = e_k(e), l = e_l(e), occ = l - k + 1;
                                bool allow_diff = true, allow_M = true;
                                if (i > 0) {
                                    if (wbid(wA) > m - 1)
                                        allow_diff = false;
                                    else if (wbid(wA) == m - 1 && wbid(wB) == m - 1 && wval(wA) == wval(wB))
                                        allow_M = false;
                                    if (seeding) {
                                        const int ii = i - (len - seed_len);
                                        if (ii > 0) {
                                            const uint64_t sA = SW[ii - 1], sB = SW[ii];
                                            if (wbid(sA) > m_seed - 1)
                                                allow_diff = false;
                                            else if (wbid(sA) == m_seed - 1 && wbid(sB) == m_seed - 1 && wval(sA) == wval(sB))
                                                allow_M = false;
                                        }
                                    }
                                }
                                const int tmp = (mode & MODE_LOGGAP) ? int_log2((uint32_t)(ge + go)) / 2 + 1 : go + ge;
                                // the children's intervals per base
                                uint64_t KK[4], LL[4];
#pragma unroll
                                for (int j = 0; j < 4; ++j) {
                                    KK[j] = L2s[j] + ca[j] + 1;
                                    LL[j] = L2s[j] + cb[j];
                                }
                                auto logp = [&](int s, const Ent &en) __attribute__((always_inline)) {
                                    if (kind != K_END) return;
                                    if (s < 0 || s >= n_stacks) {
                                        kind = K_ESCORE;
                                        return;
                                    }
                                    PLOG[np] = en;
                                    // its bucket as the commit numbers them: 0 b+s_mm, 1 b+s_gapo,
                                    // 2 b+s_gape, 3 none of them (not possible with positive penalties)
                                    const int d = s - b;
                                    const int code = d == s_mm ? 0 : d == s_gapo ? 1 : d == s_gape ? 2 : 3;
                                    ccnt += 1u << (8 * code);  // at most H_PUSH of each
                                    ++np;
                                    ++sp;
                                };
                                if (allow_diff && i >= ies + tmp && len - i >= ies + tmp) {
                                    if (st == ST_M) {
                                        if (go < max_gapo) {
                                            const int sc = mm * s_mm + (go + 1) * s_gapo + ge * s_gape;
                                            logp(sc, mk_ent(k, l, i, i, mm, go + 1, ge, ins + 1, ST_I));
                                            for (int j = 0; j < 4; ++j)
                                                if (KK[j] <= LL[j]) logp(sc, mk_ent(KK[j], LL[j], i + 1, i + 1, mm, go + 1, ge, ins, ST_D));
                                        }
                                    } else if (st == ST_I) {
                                        if (ge < max_gape)
                                            logp(mm * s_mm + go * s_gapo + (ge + 1) * s_gape,
                                                 mk_ent(k, l, i, i, mm, go, ge + 1, ins + 1, ST_I));
                                    } else {
                                        if (ge < max_gape && (ge + go < max_diff || occ < (uint64_t)(int64_t)mdo)) {
                                            const int sc = mm * s_mm + go * s_gapo + (ge + 1) * s_gape;
                                            for (int j = 0; j < 4; ++j)
                                                if (KK[j] <= LL[j]) logp(sc, mk_ent(KK[j], LL[j], i + 1, i + 1, mm, go, ge + 1, ins, ST_D));
                                        }
                                    }
                                }
                                // the intervals rotated so that slot t holds base (qbase + t) & 3, the
                                // order of the mismatch loop (the deletions above are done with them)
                                uint64_t RK[4], RL[4];
                                {
                                    const bool r1 = (qbase & 1) != 0, r2 = (qbase & 2) != 0;
                                    uint64_t TK[4], TL[4];
#pragma unroll
                                    for (int t = 0; t < 4; ++t) {
                                        TK[t] = r1 ? KK[(t + 1) & 3] : KK[t];
                                        TL[t] = r1 ? LL[(t + 1) & 3] : LL[t];
                                    }
#pragma unroll
                                    for (int t = 0; t < 4; ++t) {
                                        RK[t] = r2 ? TK[(t + 2) & 3] : TK[t];
                                        RL[t] = r2 ? TL[(t + 2) & 3] : TL[t];
                                    }
                                }
                                if (allow_diff && allow_M) {
                                    for (int j = 1; j <= 4; ++j) {
                                        // base (qbase + j) & 3: slot j & 3 of the rotated intervals
                                        const int is_mm = (j != 4 || qbase > 3);
                                        const uint64_t kk = RK[j & 3], ll = RL[j & 3];
                                        if (kk > ll) continue;
                                        if (is_mm) {
                                            logp((mm + 1) * s_mm + go * s_gapo + ge * s_gape,
                                                 mk_ent(kk, ll, i, i, mm + 1, go, ge, ins, ST_M));
                                        } else if (kind == K_END) {
                                            child = mk_ent(kk, ll, i, 0, mm, go, ge, ins, ST_M);
                                            cont = true;
                                        }
                                    }
                                } else if (qbase < 4) {
                                    const uint64_t kk = RK[0], ll = RL[0];  // base qbase
                                    if (kk <= ll && kind == K_END) {
                                        child = mk_ent(kk, ll, i, 0, mm, go, ge, ins, ST_M);
                                        cont = true;
                                    }
                                }
                            }
                        }
                    }
                    if (!TRACE && (kind == K_HIT || kind == K_ESCORE)) SLOG[steps] = e;  // read by the commit
                    SINFO[steps] = (uint16_t)(sp | (sbytes >> 6) << 4);
                    sby += sbytes;
                    ++steps;
                    if (kind != K_END || !cont) break;
                    e = child;
                }
            }
            // ---- commit plan, by wave-wide scans over the chains ----
            if (lane >= K) {
                steps = 0;
                np = 0;
                sby = 0;
                kind = K_END;
            }
            const long long t_commit = a.stats ? wall_clock64() : 0;
            // J: the first chain that stops (a hit, a truncation, a score error); the chains after
            // it are re-run in the next round
            const uint64_t stopm = __ballot(kind != K_END);
            const int J = stopm ? __builtin_ctzll(stopm) : K - 1;
            const bool inJ = lane <= J;
            const int pex = wave_excl_scan(inJ ? np : 0);  // pushes of the chains before this one
            // entries before this chain's first pop: each chain pops its head, pushes its logged
            // children and, when truncated, leaves its next exact child pending
            const int nstart = F.n_entries + wave_excl_scan(inJ ? np - 1 + (kind == K_TRUNC ? 1 : 0) : 0);
            // the push cap falls in chain jcap, which waits for the next round with the chains
            // after it (chain 0 always fits: a chain logs at most H_PUSH pushes)
            const uint64_t capm = __ballot(inJ && pex + np > a.push_cap);
            const int jcap = capm ? max(1, __builtin_ctzll(capm)) : J + 1;
            const bool cand = inJ && lane < jcap;
            int c = 0, cut = -1;
            if (cand) {
                c = steps;
                // the max_entries check before each pop (src/bwtgap_hd.h:360): the entries before a
                // pop are nstart plus the chain's earlier pushes, so only a chain whose pushes could
                // cross the limit is walked
                if ((long)nstart + np > F.max_entries) {
                    long nb = nstart;
                    for (int t = 0; t < c; ++t) {
                        if (nb > F.max_entries) {
                            cut = t;
                            break;
                        }
                        nb += SINFO[t] & 0xf;
                    }
                }
                if (c > 0 && c == steps && (kind == K_HIT || kind == K_ESCORE)) {
                    F.hk = hk;
                    F.hl = hl;
                    F.he = SLOG[steps - 1];
                }
                if (c > 0 && c == steps && kind == K_TRUNC) F.pend = nxt;
            }
            // the search stops inside chain jcut (max_entries); otherwise chains 0..last commit
            const uint64_t cutm = __ballot(cand && cut >= 0);
            const int jcut = cutm ? __builtin_ctzll(cutm) : -1;
            const uint64_t posm = __ballot(cand && c > 0);
            const int last = jcut >= 0 ? jcut : (posm ? 63 - __builtin_clzll(posm) : -1);
            const int cj = lane <= last ? (lane == jcut ? cut : c) : 0;  // committed steps of this chain
            int ncp = 0;  // committed pushes of this chain
            uint64_t by = 0;
            if (cj == steps) {  // the whole chain
                ncp = np;
                by = sby;
            } else {
                for (int t = 0; t < cj; ++t) {
                    by += (uint64_t)(SINFO[t] >> 4) << 6;
                    ncp += SINFO[t] & 0xf;
                }
            }
            // pops and entry changes in one sum: pops in bits 0-31, entry change + 1 (>= 0) above
            const unsigned long long pe_sum =
                wave_sum_u64((unsigned long long)cj | (unsigned long long)(cj > 0 ? ncp : 1) << 32);
            const unsigned long long sum_pops = pe_sum & 0xffffffffull, sum_by = wave_sum_u64(by);
            const int sum_ne = (int)(pe_sum >> 32) - HW;
            if (TRACE) {
                lcommit[lane] = cj;
                lnstart[lane] = nstart;
            }
            // a committed chain commits whole (a cut at max_entries ends the search instead)
            if (lane == last) F.l_full = kind;
            __syncthreads();
            const long long t_c1 = a.stats ? wall_clock64() : 0;
            if (lane == 0) {
                if (TRACE) {
                    // the trace hash runs over the committed pops in order
                    uint64_t tr = F.trace;
                    for (int j = 0; j <= last; ++j) {
                        long nb = lnstart[j];
                        for (int t = 0; t < lcommit[j]; ++t) {
                            const uint32_t info = SINFO_ALL[j * H_STEPS + t];
                            nb -= 1;
                            const Ent e = SLOG_ALL[j * H_STEPS + t];
                            const int i0 = e_i(e), mm = e_mm(e), go = e_go(e), ge = e_ge(e), st = e_st(e), ins = e_ins(e);
                            tr = tr_mix(tr, (uint64_t)((uint32_t)b << 21 | (uint32_t)i0));
                            tr = tr_mix(tr, e_k(e));
                            tr = tr_mix(tr, e_l(e));
                            tr = tr_mix(tr, (uint64_t)((uint32_t)mm | (uint32_t)go << 8 | (uint32_t)ge << 16 | (uint32_t)st << 24));
                            tr = tr_mix(tr, (uint64_t)ins | (uint64_t)(go + ge - ins) << 16);
                            tr = tr_mix(tr, (uint64_t)(uint32_t)e_ldp(e));
                            tr = tr_mix(tr, (uint64_t)(uint32_t)nb);
                            nb += (info & 0xf) + 1;  // its pushes, and its exact child (popped next, if any)
                        }
                    }
                    F.trace = tr;
                }
                F.pops += sum_pops;
                F.obytes += sum_by;
                if (jcut >= 0) {
                    F.status = 1;
                    F.result = F.n_aln;
                }
            }
            __syncthreads();
            if (F.status) break;
            // ---- pushes: a committed push goes to bucket b + s_mm, b + s_gapo or b + s_gape
            // (a mismatch, a gap open, a gap extension); its rank in the bucket is its place in
            // the sequential push order, chain by chain ----
            const int tg0 = b + s_mm, tg1 = b + s_gapo, tg2 = b + s_gape;
            // pushes per bucket, counted while the log was written: a round commits a chain whole,
            // or ends the search (max_entries) before its pushes are placed
            const uint32_t cc = ncp > 0 ? ccnt : 0u;
            const int n0 = (int)(cc & 0xff), n1 = (int)((cc >> 8) & 0xff), n2 = (int)((cc >> 16) & 0xff);
            const bool bad = (cc >> 24) != 0 || (ncp != 0 && ncp != np);
            // the three totals in one sum (16-bit fields: a round commits at most H_CAP pushes)
            const unsigned long long tsum =
                wave_sum_u64((unsigned long long)n0 | (unsigned long long)n1 << 16 | (unsigned long long)n2 << 32);
            const int T0 = (int)(tsum & 0xffff), T1 = (int)((tsum >> 16) & 0xffff), T2 = (int)((tsum >> 32) & 0xffff);
            const bool badw = __ballot(bad) != 0;
            if (lane == 0) {
                // pieces of the wave's current granule for each bucket that receives pushes; a
                // bucket's new piece links to its previous top segment
                int np_ = 0;
                bool oom = false;
                for (int x = 0; x < 3 && !oom && !badw; ++x) {
                    // (selects: arrays indexed by x would live in scratch memory)
                    const int sb = x == 0 ? tg0 : x == 1 ? tg1 : tg2;
                    uint32_t rem = (uint32_t)(x == 0 ? T0 : x == 1 ? T1 : T2);
                    if (!rem) continue;
                    const bool bit = (F.bm[sb >> 6] >> (sb & 63)) & 1ull;
                    uint32_t prev_base = bit ? seg_base[sb] : NIL, prev_size = bit ? seg_cnt[sb] : 0u, rk = 0;
                    first_piece[sb] = np_;
                    while (rem > 0) {
                        if (F.aleft < 2) {
                            uint32_t g;
                            if (F.n_used < F.n_own) {
                                g = OWN[F.n_used];
                            } else {
                                if (F.n_own >= a.own_cap) {
                                    oom = true;
                                    break;
                                }
                                g = gran_pop(a.ghead, a.gnext, a.gbump, a.n_gran);
                                if (g == NIL) {
                                    oom = true;
                                    break;
                                }
                                if (a.stats) F.st[11] += 1;
                                OWN[F.n_own++] = g;
                            }
                            ++F.n_used;
                            F.abase = g * GENT;
                            F.aleft = (int)GENT;
                        }
                        const uint32_t plen = min(rem, (uint32_t)F.aleft - 1);
                        Ent hdr;
                        hdr.w0 = (uint64_t)prev_base | (uint64_t)prev_size << 32;
                        hdr.w1 = 0;
                        POOL[F.abase] = hdr;
                        pieces[np_++] = make_uint4((uint32_t)sb, rk, plen, F.abase + 1);
                        prev_base = F.abase + 1;
                        prev_size = plen;
                        F.abase += plen + 1;
                        F.aleft -= (int)plen + 1;
                        rem -= plen;
                        rk += plen;
                    }
                    seg_base[sb] = prev_base;
                    seg_cnt[sb] = prev_size;
                    F.bm[sb >> 6] |= 1ull << (sb & 63);
                }
                F.n_piece = np_;
                if (badw) {  // a push outside the three buckets: not possible with positive penalties
                    F.status = 1;
                    F.result = ERR_INTERNAL;
                } else if (oom) {
                    F.status = 1;
                    F.result = ERR_STACK;
                }
            }
            __syncthreads();
            if (F.status) break;
            const long long t_c2 = a.stats ? wall_clock64() : 0;
            {
                const int npieces = F.n_piece;
                const Ent *const PLOG_W = reinterpret_cast<const Ent *>(base + a.plog_off);  // lane j's log at j * H_PUSH
                bool bad_piece = false;
                // rank R of bucket tg's pushes this round to its pool slot, within this round's pieces
                auto place = [&](int tg, uint32_t R, const Ent &pe) __attribute__((always_inline)) {
                    int p = first_piece[tg];
                    while (p < npieces && R >= pieces[p].y + pieces[p].z) ++p;
                    if (p < npieces && (int)pieces[p].x == tg)
                        POOL[pieces[p].w + (R - pieces[p].y)] = pe;
                    else
                        bad_piece = true;
                };
                // every committed push, spread over the wave: push g of the round in the sequential
                // order (chain by chain, each in log order) is placed by lane g % 64. Its rank in
                // its bucket is the number of earlier pushes to that bucket in this order: a ballot
                // per bucket within a spread, plus the counts of the spreads before it.
                const int E = wave_excl_scan(ncp);
                const int M = __shfl(E + ncp, HW - 1);
                const uint64_t lt = (1ull << lane) - 1ull;  // the lanes before this one
                uint32_t c0 = 0, c1 = 0, c2 = 0;
                for (int g0 = 0; g0 < M; g0 += HW) {
                    const int g = g0 + lane;
                    int L = 0;  // the owner: the last lane whose first push is at or before g
#pragma unroll
                    for (int step = 32; step > 0; step >>= 1) {
                        const int el = __shfl(E, L + step);
                        if (el <= g) L += step;
                    }
                    const int u = g - __shfl(E, L);
                    const bool valid = g < M;
                    if (valid && (u < 0 || u >= H_PUSH)) bad_piece = true;  // (not possible: E is a scan)
                    Ent pe{0, 0};
                    int x = 3;
                    if (valid && u >= 0 && u < H_PUSH) {
                        pe = PLOG_W[L * H_PUSH + u];
                        const int sc = e_mm(pe) * s_mm + e_go(pe) * s_gapo + e_ge(pe) * s_gape;  // its bucket
                        x = sc == tg0 ? 0 : sc == tg1 ? 1 : sc == tg2 ? 2 : 3;
                    }
                    const uint64_t b0 = __ballot(x == 0), b1 = __ballot(x == 1), b2 = __ballot(x == 2);
                    if (valid && u >= 0 && u < H_PUSH) {
                        if (x == 3) {
                            bad_piece = true;  // (ruled out when the round was sized)
                        } else {
                            const uint32_t R = x == 0   ? c0 + (uint32_t)__builtin_popcountll(b0 & lt)
                                               : x == 1 ? c1 + (uint32_t)__builtin_popcountll(b1 & lt)
                                                        : c2 + (uint32_t)__builtin_popcountll(b2 & lt);
                            place(x == 0 ? tg0 : x == 1 ? tg1 : tg2, R, pe);
                        }
                    }
                    c0 += (uint32_t)__builtin_popcountll(b0);
                    c1 += (uint32_t)__builtin_popcountll(b1);
                    c2 += (uint32_t)__builtin_popcountll(b2);
                }
                if (bad_piece) {  // a rank outside this round's pieces: not possible
                    F.status = 1;
                    F.result = ERR_INTERNAL;
                }
            }
            const int mxs = a.stats ? wave_max_i32(steps) : 0;
            __syncthreads();
            if (F.status) break;
            const long long t_c3 = a.stats ? wall_clock64() : 0;
            // ---- bucket b: the committed start entries leave its top; the state moves on ----
            if (lane == 0) {
                const int taken = __builtin_popcountll(posm);
                const int from_seg = taken - (F.has_pend && taken > 0 ? 1 : 0);
                if (taken > 0) F.has_pend = 0;
                if (from_seg > 0) {
                    seg_cnt[b] -= (uint32_t)from_seg;
                    if (seg_cnt[b] == 0) {
                        const Ent hdr = POOL[seg_base[b] - 1];
                        const uint32_t pb = (uint32_t)hdr.w0, ps = (uint32_t)(hdr.w0 >> 32);
                        if (pb == NIL) {
                            F.bm[b >> 6] &= ~(1ull << (b & 63));
                        } else {
                            seg_base[b] = pb;
                            seg_cnt[b] = ps;
                        }
                    }
                }
                // entries: every committed pop -1, every committed push +1, a pending child +1
                long ne = F.n_entries + sum_ne;
                const int lfull = last >= 0 ? F.l_full : -1;
                if (last >= 0 && lfull == K_TRUNC) {
                    F.has_pend = 1;
                    ne += 1;
                }
                F.n_entries = (int)ne;
                if (a.stats) {
                    F.st[0] += 1;
                    F.st[1] += (unsigned long long)K;
                    F.st[2] += (unsigned long long)(last + 1);
                    F.st[3] += sum_pops;
                    F.st[4] += (unsigned long long)mxs;
                    F.st[5] += (unsigned long long)(t_commit - t_spec);
                    const long long t_end = wall_clock64();
                    t_round_end = t_end;
                    F.st[6] += (unsigned long long)(t_end - t_commit);
                    F.st[7] += (unsigned long long)(t_c1 - t_commit);
                    F.st[8] += (unsigned long long)(t_c2 - t_c1);
                    F.st[9] += (unsigned long long)(t_c3 - t_c2);
                    F.st[10] += (unsigned long long)(t_end - t_c3);
                }
                // a hit or a score error ends the last committed chain
                if (lfull == K_ESCORE) {
                    F.status = 1;
                    F.result = ERR_SCORE;
                } else if (lfull == K_HIT) {
                    const Ent e = F.he;
                    const uint64_t k = F.hk, l = F.hl;
                    const int mm = e_mm(e), go = e_go(e), ge = e_ge(e);
                    const int score = b;
                    bool do_add = true, stop = false;
                    if (F.n_aln == 0) {
                        F.best_score = score;
                        F.best_diff = mm + go + ((F.mode & MODE_GAPE) ? ge : 0);
                        if (!(F.mode & MODE_NONSTOP))
                            F.max_diff = (F.best_diff + 1 > F.md_read) ? F.md_read : F.best_diff + 1;
                    }
                    if (score == F.best_score) {
                        F.best_cnt = (int)(uint32_t)((uint64_t)(int64_t)F.best_cnt + (l - k + 1));
                    } else if (F.best_cnt > F.max_top2) {
                        stop = true;
                    }
                    if (!stop) {
                        if (go) {
                            for (int j = 0; j < F.n_aln; ++j)
                                if (REC[j].k == k && REC[j].l == l) {
                                    do_add = false;
                                    break;
                                }
                        }
                        if (do_add) {
                            const int x = (int)(uint32_t)(l - k + 1);
                            const uint64_t ux = (uint64_t)(int64_t)x;
                            int jj = 0;
                            for (int i = 0; i < e_ldp(e); ++i) {
                                uint64_t w = wval(W[i]);
                                int bid = wbid(W[i]);
                                if (w > ux) w -= ux;
                                else if (w == ux) {
                                    bid = 1;
                                    w = seq_len - (uint64_t)(++jj);
                                } else continue;
                                W[i] = wpack(w, bid);
                            }
                            if (F.n_aln == a.rec_cap) {
                                F.status = 1;
                                F.result = ERR_ALN;
                            } else {
                                const int ins = e_ins(e), del = go + ge - ins;
                                galn_aln_rec_t r;
                                r.packed = ((uint64_t)(mm & 0xff)) | ((uint64_t)(go & 0xff) << 8) |
                                           ((uint64_t)(ge & 0xff) << 16) | ((uint64_t)(score & 0xfffff) << 24) |
                                           ((uint64_t)(ins & 0x3ff) << 44) | ((uint64_t)(del & 0x3ff) << 54);
                                r.k = k;
                                r.l = l;
                                REC[F.n_aln] = r;
                                F.n_aln += 1;
                            }
                        }
                    } else {
                        F.status = 1;
                        F.result = F.n_aln;
                    }
                }
            }
            __syncthreads();
        }
        // ---- the read's result ----
        if (lane == 0) {
            int n = F.result;
            const int rd = F.rid;
            if (n > 0) {
                const unsigned long long b0 = atomicAdd(a.out_top, (unsigned long long)n);
                if (b0 + (unsigned long long)n > a.out_cap) {
                    n = ERR_OUT;
                } else {
                    for (int j = 0; j < n; ++j) a.out_rec[b0 + j] = REC[j];
                    a.out_off[rd] = b0;
                }
            }
            if (n == ERR_STACK || n == ERR_ALN || n == ERR_OUT) {
                const int f = atomicAdd(a.fail_n, 1);
                a.fail_list[f] = rd;
                atomicOr(a.fail_kind + (n == ERR_OUT ? 0 : n == ERR_ALN ? 1 : 2), 1);
            } else {
                atomicAdd(a.algo_bytes, F.obytes);
            }
            a.out_n[rd] = n;
            a.out_pops[rd] = F.pops;
            if (TRACE) a.out_trace[rd] = F.trace;
            if (a.out_bytes) a.out_bytes[rd] = F.obytes;
            // keep a few granules for the next read; the rest go back to the pool
            if (F.n_own > H_KEEP) {
                gran_push(a.ghead, a.gnext, OWN, H_KEEP, F.n_own);
                F.n_own = H_KEEP;
            }
        }
        __syncthreads();
    }
    if (a.stats && lane < 16) atomicAdd(a.stats + lane, F.st[lane]);
}

// A read's record count for the output offsets (error codes count none).
struct RecCount {
    __host__ __device__ uint64_t operator()(int32_t v) const { return v > 0 ? (uint64_t)v : 0; }
};

// Records of every read in read order: read i's n_aln[i] records move from the arena (at
// arena_off[i], where the search appended them) to incl[i] - n_aln[i] of the staging array.
__global__ void k_gather_records(const int32_t *n_aln, const uint64_t *arena_off, const uint64_t *incl,
                                 const galn_aln_rec_t *arena, galn_aln_rec_t *stage, uint64_t *rec_off, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int na = n_aln[i] > 0 ? n_aln[i] : 0;
    const uint64_t o = incl[i] - (uint64_t)na;
    rec_off[i] = o;
    for (int j = 0; j < na; ++j) stage[o + j] = arena[arena_off[i] + j];
}

// Fills n words with v: the search's counter and pool-head resets, as kernels so that a
// captured graph holds no memset nodes.
__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// host side

thread_local char g_err[512];

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

// progress lines on stderr for long calls (GALN_HIP_VERBOSE set and not "0")
bool galn_hip_verbose()
{
    const char *v = getenv("GALN_HIP_VERBOSE");
    return v && *v && strcmp(v, "0") != 0;
}

#define HIP_OK(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? GALN_HIP_E_NOMEM : GALN_HIP_E_DEVICE, "%s:%d %s: %s", \
                        __FILE__, __LINE__, #call, hipGetErrorString(e_));                             \
    } while (0)

struct Tier {
    int rec_cap;   // record slots per lane
    int keep;      // granules a lane keeps between reads
};

}  // namespace

// Where a search runs: a stream and a stack pool (grown on demand). The read sets of an index
// share its context; a read set of a server slot has one of its own, so that slots' launches run
// at the same time.
struct SearchCtx {
    hipStream_t stream = nullptr;
    Ent *d_pool_chk = nullptr;
    uint32_t *d_pool_link = nullptr, *d_gnext = nullptr, *d_gbump = nullptr;
    unsigned long long *d_ghead = nullptr;
    uint32_t n_gran = 0;
    bool grow = false;  // the last search used every never-used granule: double the pool next time
    // frees the pool (not the stream)
    void free_pool()
    {
        void *p[] = {d_pool_chk, d_pool_link, d_gnext, d_gbump, d_ghead};
        for (void *x : p)
            if (x) (void)hipFree(x);
        d_pool_chk = nullptr, d_pool_link = nullptr, d_gnext = nullptr, d_gbump = nullptr, d_ghead = nullptr;
        n_gran = 0;
    }
};

struct galn_hip_index {
    int device = 0;
    uint32_t *d_bwt = nullptr;
    DMeta meta{};
    galn_bwt_meta_t hmeta{};
    int n_cu = 0;
    SearchCtx ctx0;  // the stream and stack pool of the index's own read sets
    // SA samples (galn_hip_index_set_sa): the suffix at rows intv, 2 intv, ...
    uint64_t *d_sa = nullptr;
    uint64_t n_sa = 0;
    uint32_t sa_intv = 0;
    std::atomic<int> n_readsets{0};  // read sets not yet freed (they use the stream and pool)
    std::atomic<int> n_servers{0};   // open servers: the index then belongs to their workers
};

struct GraphKey {
    uint32_t flags;
    int32_t block, slots, hwaves, n;
    uint64_t budget;
    uint32_t rec_cap, keep;
    const int32_t *list;
    const void *ptr[20];
    uint64_t rec_cap_out, n_gran, scratch_bytes, hscratch_bytes;
    int32_t max_len, seed_len, max_stacks, max_entries;
    const void *hand;
    int32_t hand_cap;
    uint32_t hand_stride;
    int32_t push_cap;
    int32_t heavy_wpe;
};

struct galn_hip_readset {
    galn_hip_index *idx = nullptr;
    SearchCtx *ctx = nullptr;   // &idx->ctx0, or own_ctx
    SearchCtx *own_ctx = nullptr;
    int n = 0;
    int n_batches = 0;
    int seed_len = 0;
    int max_len = 0;
    int max_stacks = 0;
    int max_entries = 0;
    uint64_t qbytes = 0;
    // device inputs
    uint8_t *d_query = nullptr;
    uint64_t *d_qoff = nullptr;
    int32_t *d_qlen = nullptr, *d_maxdiff = nullptr, *d_batch = nullptr, *d_nstacks = nullptr;
    galn_search_opt_t *d_bopts = nullptr;
    int32_t *d_sorted = nullptr;   // read ids, longest first
    // device outputs
    int32_t *d_out_n = nullptr;
    uint64_t *d_out_off = nullptr, *d_pops = nullptr, *d_trace = nullptr, *d_bytes = nullptr;
    galn_aln_rec_t *d_rec = nullptr;
    uint64_t rec_cap = 0;
    unsigned long long *d_top = nullptr;
    int32_t *d_fail[2] = {nullptr, nullptr};
    int32_t *d_ctr = nullptr;      // [0],[1] failure counts, [2] queue head, [3] heavy count, [4..6] failure kinds
    int32_t *d_heavy = nullptr;    // reads handed to the wave-per-read engine
    unsigned long long *d_stats = nullptr;  // wave-per-read engine counters of the last search
    uint8_t *d_hscratch = nullptr; // per-wave regions of the wave-per-read engine
    size_t hscratch_bytes = 0;
    // per-lane regions (grown on demand)
    uint8_t *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    // host copies
    std::vector<galn_search_opt_t> h_bopts;
    // widths handed from k_search to k_heavy, one slot per heavy-list position
    uint64_t *d_hand = nullptr;
    size_t hand_bytes = 0;
    uint32_t hand_stride = 0;
    int32_t hand_cap = 0;
    // results in read order: inclusive prefix sums of record counts, output offsets, staging
    uint64_t *d_incl = nullptr, *d_roff = nullptr;
    galn_aln_rec_t *d_stage = nullptr;
    size_t stage_bytes = 0;
    void *d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    // capacities of the device arrays (reads, query bytes, batches)
    int cap_n = 0, cap_b = 0;
    uint64_t cap_q = 0;
    // last search: one record per kernel launch, timed with events on the search stream
    double first_ms = 0, total_ms = 0;
    // the batched form's tuning chosen by galn_hip_readset_autotune (used when a search passes none)
    bool has_tuned = false;
    galn_hip_tune_t tuned{};
    uint32_t last_flags = 0;
    int heavy_wpe = 4;  // waves per SIMD of the wave engine in the current search (4 or 5)
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    struct Launch {
        int tier;
        int slot;
        double ms;
        uint64_t bytes;
    };
    std::vector<Launch> launches;
    // the batched form's captured graph (GALN_HIP_GRAPH), its key and its launch records
    hipGraphExec_t gexec = nullptr, gexec2 = nullptr;  // prologue + first tier; wave engine
    bool capturing = false;  // launches skip their timing events (recorded around the graphs)
    GraphKey gkey{};
    std::vector<Launch> glaunches;
    std::vector<hipEvent_t> lev;           // 2 per launch, reused across searches
    unsigned long long *d_algo = nullptr;  // per-launch algorithmic byte counters
    static constexpr int MAX_LAUNCH = 256;
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Layout {
    uint64_t slot_bytes;
    uint32_t sw_off, rec_off, own_off;
    int own_cap;
};

Layout make_layout(int max_len, int seed_len, int rec_cap, int own_cap)
{
    Layout L{};
    size_t off = align_up((size_t)(max_len + 1) * 8, 256);
    L.sw_off = (uint32_t)off;
    const int sl = (seed_len > 0 && seed_len < max_len) ? seed_len : 0;
    off += align_up((size_t)(sl + 1) * 8, 256);
    L.rec_off = (uint32_t)off;
    off += align_up((size_t)rec_cap * sizeof(galn_aln_rec_t), 256);
    L.own_off = (uint32_t)off;
    off += align_up((size_t)own_cap * 4, 256);
    // an odd multiple of 256 bytes keeps lanes' regions off the same cache sets
    if ((off / 256) % 2 == 0) off += 256;
    L.slot_bytes = off;
    L.own_cap = own_cap;
    return L;
}

// A device buffer of at least `need` bytes (contents not kept). It grows to twice its size or
// more: hipFree waits for every stream of the device, so a read set whose launches vary in size
// (a server slot) reallocates a few times rather than every launch.
int grow_dev(void **p, size_t &cap, size_t need, hipStream_t s)
{
    if (*p && cap >= need) return 0;
    HIP_OK(hipStreamSynchronize(s));
    if (*p) HIP_OK(hipFree(*p));
    *p = nullptr;
    const size_t c = std::max(need, cap * 2);
    if (hipMalloc(p, c) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        cap = 0;
        if (hipMalloc(p, need) != hipSuccess) return fail(GALN_HIP_E_NOMEM, "device allocation of %zu bytes failed", need);
        cap = need;
        return 0;
    }
    cap = c;
    return 0;
}

int ensure_scratch(galn_hip_readset *rs, size_t bytes)
{
    return grow_dev((void **)&rs->d_scratch, rs->scratch_bytes, bytes, rs->ctx->stream);
}

// Granules the stack pool may hold at most: no limit, or the value of GALN_HIP_TEST_POOL_GRANULES,
// which the tests set to exhaust the pool and drive the re-runs with fewer lanes.
uint64_t pool_test_cap()
{
    const char *v = getenv("GALN_HIP_TEST_POOL_GRANULES");
    if (!v || !*v) return ~0ull;
    return (uint64_t)std::max(1L, strtol(v, nullptr, 10));
}

// Granules of a context's first pool: as asked, or the value of GALN_HIP_TEST_POOL_START, which the
// tests set low so that the pool's growth on demand runs.
uint64_t pool_test_start()
{
    const char *v = getenv("GALN_HIP_TEST_POOL_START");
    if (!v || !*v) return ~0ull;
    return (uint64_t)std::max(1024L, strtol(v, nullptr, 10));
}

// The stack pool: `want` granules of 64 chunks (16 KB each), within half the free memory and the
// 28-bit chunk ids of the bucket tops (2^22 granules, 64 GiB).
int ensure_pool(SearchCtx *idx, uint64_t want)
{
    // contexts of one process size their pools one at a time (server slots start together: each
    // would otherwise count the same free memory)
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const uint64_t gbytes = (uint64_t)GRAN * CH * sizeof(Ent);
    size_t freeb = 0, totalb = 0;
    HIP_OK(hipMemGetInfo(&freeb, &totalb));
    const uint64_t have = idx->n_gran;
    uint64_t cap = (uint64_t)((freeb + have * (gbytes + GRAN * 4 + 4)) / 2) / (gbytes + GRAN * 4 + 4);
    cap = std::min<uint64_t>(cap, MAX_POOL_CHUNKS / GRAN);
    cap = std::min<uint64_t>(cap, pool_test_cap());
    // a search that used up the never-used granules made every later allocation search the other
    // shards' free lists (one wave's read can hold millions of entries on a repeat-rich genome: on
    // the 3.1 Gbp human-like workload a 26 GB pool spent 3x the search time allocating)
    if (idx->grow) want = std::max<uint64_t>(want, 2 * have);
    idx->grow = false;
    if (have == 0) want = std::min(want, pool_test_start());
    want = std::min(std::max<uint64_t>(want, 1024), cap);
    if (have >= want) return 0;
    if (!idx->d_ghead) {
        unsigned long long *head = nullptr;
        uint32_t *bump = nullptr;
        HIP_OK(hipMalloc(&head, NSHARD * 8));
        if (hipMalloc(&bump, 4) != hipSuccess) {
            (void)hipFree(head);
            return fail(GALN_HIP_E_NOMEM, "stack pool: allocation of its heads failed");
        }
        idx->d_ghead = head;
        idx->d_gbump = bump;
    }
    // the old pool goes first (it is not copied: no search is running), then all three new
    // arrays are allocated before any of them is installed, so a failure leaves an empty pool
    // (null pointers, no granules) and nothing freed twice
    void *old[] = {idx->d_pool_chk, idx->d_pool_link, idx->d_gnext};
    idx->d_pool_chk = nullptr;
    idx->d_pool_link = nullptr;
    idx->d_gnext = nullptr;
    idx->n_gran = 0;
    for (void *p : old)
        if (p) HIP_OK(hipFree(p));
    Ent *chk = nullptr;
    uint32_t *link = nullptr, *gnext = nullptr;
    // half as many granules while the allocation fails (other allocations of the process, or of
    // other processes, since the free memory was read): a smaller pool only costs re-runs
    for (;;) {
        if (hipMalloc(&chk, want * gbytes) == hipSuccess && hipMalloc(&link, want * GRAN * 4) == hipSuccess &&
            hipMalloc(&gnext, want * 4) == hipSuccess)
            break;
        (void)hipGetLastError();
        for (void *p : {(void *)chk, (void *)link, (void *)gnext})
            if (p) (void)hipFree(p);
        chk = nullptr, link = nullptr, gnext = nullptr;
        if (want <= 1024)
            return fail(GALN_HIP_E_NOMEM, "stack pool of %llu granules does not fit in device memory",
                        (unsigned long long)want);
        want = std::max<uint64_t>(want / 2, 1024);
    }
    idx->d_pool_chk = chk;
    idx->d_pool_link = link;
    idx->d_gnext = gnext;
    idx->n_gran = (uint32_t)want;
    return 0;
}

template <bool TRACE, bool PERSIST>
int launch_one(hipStream_t st, const KArgs &a, int block, int grid, int buckets)
{
    const size_t lds = (size_t)block * buckets * sizeof(uint32_t);
    hipLaunchKernelGGL((k_search<TRACE, PERSIST>), dim3(grid), dim3(block), lds, st, a);
    HIP_OK(hipGetLastError());
    return 0;
}

int launch(hipStream_t st, const KArgs &a, int block, int grid, bool persist, bool trace, int buckets)
{
    if (trace) return persist ? launch_one<true, true>(st, a, block, grid, buckets)
                              : launch_one<true, false>(st, a, block, grid, buckets);
    return persist ? launch_one<false, true>(st, a, block, grid, buckets)
                   : launch_one<false, false>(st, a, block, grid, buckets);
}

int fill_u32(hipStream_t st, void *p, uint32_t v, size_t bytes)
{
    const uint32_t n = (uint32_t)(bytes / 4);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::max(1u, std::min(64u, (n + 255) / 256))), dim3(256), 0, st,
                       reinterpret_cast<uint32_t *>(p), v, n);
    HIP_OK(hipGetLastError());
    return 0;
}
#define FILL(p, v, bytes)                                                                              \
    do {                                                                                               \
        const int fr_ = fill_u32(fst, (p), (v), (bytes));                                              \
        if (fr_) return fr_;                                                                           \
    } while (0)

int buckets_of(const galn_hip_readset *rs) { return std::max(1, std::min(rs->max_stacks, MAXB)); }

// threads per workgroup: the bucket tops of a workgroup must fit 64 KB of LDS
int pick_block(const galn_hip_readset *rs) { return buckets_of(rs) * 128 * 4 <= 65536 ? 128 : 64; }

// persistent grid: the workgroups that are resident at once
int resident_slots(galn_hip_readset *rs, int block, bool trace, int waves_per_cu)
{
    galn_hip_index *idx = rs->idx;
    if (waves_per_cu > 0) return std::max(block, idx->n_cu * waves_per_cu * 64 / block * block);
    int nb = 0;
    const size_t lds = (size_t)block * buckets_of(rs) * 4;
    hipError_t e = trace ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_search<true, true>, block, lds)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_search<false, true>, block, lds);
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb * idx->n_cu * block;
}

struct HLayout {
    uint64_t slot_bytes;
    uint32_t rec_off, own_off, plog_off, slog_off;
};

HLayout make_hlayout(int rec_cap, int own_cap)
{
    HLayout L{};
    size_t off = 0;
    L.rec_off = 0;
    off += align_up((size_t)rec_cap * sizeof(galn_aln_rec_t), 256);
    L.own_off = (uint32_t)off;
    off += align_up((size_t)own_cap * 4, 256);
    L.plog_off = (uint32_t)off;
    off += (size_t)HW * H_PUSH * sizeof(Ent);
    L.slog_off = (uint32_t)off;
    off += (size_t)HW * H_STEPS * sizeof(Ent);
    if ((off / 256) % 2 == 0) off += 256;
    L.slot_bytes = off;
    return L;
}

size_t heavy_lds_bytes(int buckets, int wlen, int swlen)
{
    size_t b = (sizeof(HShared) + 15) & ~(size_t)15;
    b += (size_t)buckets * 12;
    b = (b + 15) & ~(size_t)15;
    b += (size_t)H_PIECES * 16 + (size_t)(wlen + swlen) * 8 + 2 * HW * 4 + (size_t)HW * H_STEPS * 2 + (size_t)wlen;
    return b;
}

// the wave-per-read engine keeps each chain's pops in order only when every child scores above its parent
bool heavy_ok(const galn_hip_readset *rs)
{
    for (const auto &o : rs->h_bopts)
        if (o.s_mm <= 0 || o.s_gapo <= 0 || o.s_gape <= 0) return false;
    return true;
}

KArgs base_args(galn_hip_readset *rs, uint32_t flags)
{
    KArgs a{};
    galn_hip_index *idx = rs->idx;
    a.bwt = idx->d_bwt;
    a.meta = idx->meta;
    a.query = rs->d_query;
    a.qoff = rs->d_qoff;
    a.qlen = rs->d_qlen;
    a.maxdiff = rs->d_maxdiff;
    a.batch = rs->d_batch;
    a.bopts = rs->d_bopts;
    a.nstacks = rs->d_nstacks;
    a.out_n = rs->d_out_n;
    a.out_off = rs->d_out_off;
    a.out_pops = rs->d_pops;
    a.out_trace = rs->d_trace;
    a.out_bytes = rs->d_bytes;
    a.out_rec = rs->d_rec;
    a.out_cap = rs->rec_cap;
    a.out_top = rs->d_top;
    a.pool_chk = rs->ctx->d_pool_chk;
    a.pool_link = rs->ctx->d_pool_link;
    a.gnext = rs->ctx->d_gnext;
    a.ghead = rs->ctx->d_ghead;
    a.gbump = rs->ctx->d_gbump;
    a.n_gran = rs->ctx->n_gran;
    // keeping the exact child in registers relies on children never scoring below their parent
    bool nonneg = true;
    for (const auto &o : rs->h_bopts) nonneg = nonneg && o.s_mm >= 0 && o.s_gapo >= 0 && o.s_gape >= 0;
    a.fuse = ((flags & GALN_HIP_NOFUSE) || !nonneg) ? 0 : 1;
    return a;
}

// One launch over list positions [0, n_list) (read id = list[pos], or list_base + pos);
// failures land in fail buffer `fo`.
// The allocations a first-tier launch over `slots` lanes needs (run before a graph capture).
int prepare_tier(galn_hip_readset *rs, const Tier &t, int slots, int pool_lanes)
{
    const int own_cap = (int)(((int64_t)(rs->max_entries + 16) / CH + MAXB + 2) / GRAN + 2);
    const Layout L = make_layout(rs->max_len, rs->seed_len, t.rec_cap, own_cap);
    int r = ensure_scratch(rs, (size_t)slots * L.slot_bytes);
    if (r) return r;
    // the pool for the lanes that get reads, rounded up to a power of two so that launches of
    // varying size regrow it a few times at most
    uint64_t want = (uint64_t)std::max(pool_lanes, 1) * (uint64_t)(t.keep + 12), p2 = 1024;
    while (p2 < want) p2 <<= 1;
    return ensure_pool(rs->ctx, p2);
}

// The allocations of a wave-engine launch of `waves` waves.
int prepare_heavy(galn_hip_readset *rs, int waves, int rec_cap)
{
    SearchCtx *cx = rs->ctx;
    // granules for the waves that can get reads, rounded up to a power of two
    uint64_t want = (uint64_t)std::max(1, std::min(waves, rs->n)) * 256, p2 = 1024;
    while (p2 < want) p2 <<= 1;
    int r = ensure_pool(cx, p2);
    if (r) return r;
    const int own_cap = (int)std::min<uint64_t>(cx->n_gran, 32768);
    const HLayout L = make_hlayout(rec_cap, own_cap);
    return grow_dev((void **)&rs->d_hscratch, rs->hscratch_bytes, (size_t)waves * L.slot_bytes, cx->stream);
}

// Hand-off slots for the widths of reads moving to the wave engine: one per read, up to
// HAND_BYTES or a quarter of the free device memory (reads beyond recompute their widths there,
// one lane walking the read while the wave waits: with every read sent to the wave engine, 10M
// reads of up to 150 bp need 12 GB).
constexpr uint64_t HAND_BYTES = 32ull << 30;

int prepare_hand(galn_hip_readset *rs)
{
    const int sl = (rs->seed_len > 0 && rs->seed_len < rs->max_len) ? rs->seed_len : 0;
    const uint32_t stride = (uint32_t)(1 + (rs->max_len + 1) + (sl + 1));
    size_t freeb = 0, totalb = 0;
    HIP_OK(hipMemGetInfo(&freeb, &totalb));
    const uint64_t limit = std::min<uint64_t>(HAND_BYTES, (uint64_t)(freeb + rs->hand_bytes) / 4);
    const int cap = (int)std::min<uint64_t>((uint64_t)rs->n, limit / (stride * 8ull));
    // slots of this stride in the bytes already held (another launch's stride may differ)
    const size_t need = (size_t)std::max(cap, 1) * stride * 8;
    if (!rs->d_hand || rs->hand_bytes < need) {
        if (grow_dev((void **)&rs->d_hand, rs->hand_bytes, need, rs->ctx->stream)) {
            (void)hipGetLastError();
            g_err[0] = 0;
            rs->d_hand = nullptr;  // no hand-off: the wave engine computes the widths itself
            rs->hand_bytes = 0;
            rs->hand_cap = 0;
            return 0;
        }
    }
    rs->hand_stride = stride;
    rs->hand_cap = (int)std::min<uint64_t>((uint64_t)cap, rs->hand_bytes / (stride * 8ull));
    return 0;
}

int heavy_waves(galn_hip_readset *rs, uint32_t flags);

// The batched form's buffers sized for launches of up to n reads of up to max_len bases: the
// hand-off slots, the stack pool and wave-engine scratch of min(grid, n) waves, the results' scan
// space and two staged records per read. Without it these grow by doubling as launches get larger
// (a server slot's launches vary in size), and each regrowth's hipFree waits for every stream of
// the device: on the 100 Mbp index a slot regrew ~30 times over its first 50 requests, frees of up
// to 190 ms, each holding the slot until the other slots' launches ended. Best effort: what does
// not fit is left to the searches, which allocate what they need.
int reserve_search(galn_hip_readset *rs, int n, int max_len, uint32_t flags)
{
    SearchCtx *cx = rs->ctx;
    hipStream_t s = cx->stream;
    n = std::max(n, 1);
    auto soft = [](int r) {
        if (r) (void)hipGetLastError(), g_err[0] = 0;
        return r;
    };
    // hand-off slots (seed widths of seeds up to 32 bases included), within prepare_hand's limit
    const uint64_t stride = (uint64_t)(1 + (max_len + 1) + 33);
    size_t freeb = 0, totalb = 0;
    HIP_OK(hipMemGetInfo(&freeb, &totalb));
    const uint64_t limit = std::min<uint64_t>(HAND_BYTES, (uint64_t)(freeb + rs->hand_bytes) / 4);
    const uint64_t hand = std::min<uint64_t>((uint64_t)n, limit / (stride * 8)) * stride * 8;
    if (hand > rs->hand_bytes && soft(grow_dev((void **)&rs->d_hand, rs->hand_bytes, hand, s))) {
        rs->d_hand = nullptr;
        rs->hand_bytes = 0;
        rs->hand_cap = 0;
    }
    // the stack pool and the wave engine's per-wave scratch, as prepare_heavy sizes them
    const int waves = std::max(1, std::min(heavy_waves(rs, flags), n));
    uint64_t p2 = 1024;
    while (p2 < (uint64_t)waves * 256) p2 <<= 1;
    if (!soft(ensure_pool(cx, p2))) {
        const HLayout L = make_hlayout(256, (int)std::min<uint64_t>(cx->n_gran, 32768));
        soft(grow_dev((void **)&rs->d_hscratch, rs->hscratch_bytes, (size_t)waves * L.slot_bytes, s));
    }
    // the results: the record counts' scan over n reads and the staged records
    size_t tmp = 0;
    const rocprim::transform_iterator<const int32_t *, RecCount, uint64_t> cnt(rs->d_out_n, RecCount());
    HIP_OK(rocprim::inclusive_scan(nullptr, tmp, cnt, rs->d_incl, (size_t)n, rocprim::plus<uint64_t>(), s));
    soft(grow_dev(&rs->d_scan_tmp, rs->scan_tmp_bytes, tmp, s));
    soft(grow_dev((void **)&rs->d_stage, rs->stage_bytes, (size_t)2 * n * sizeof(galn_aln_rec_t), s));
    return 0;
}

// What a captured search graph holds: when any of it changes, the graph is captured again.
void fill_key_ptrs(const galn_hip_readset *rs, GraphKey &k)
{
    const void *p[] = {rs->d_query, rs->d_qoff, rs->d_qlen, rs->d_maxdiff, rs->d_batch, rs->d_nstacks, rs->d_bopts,
                       rs->d_sorted, rs->d_out_n, rs->d_out_off, rs->d_pops, rs->d_trace, rs->d_bytes, rs->d_rec,
                       rs->d_fail[0], rs->d_fail[1], rs->d_heavy, rs->d_scratch, rs->d_hscratch, rs->ctx->d_pool_chk};
    static_assert(sizeof(p) == sizeof(k.ptr), "graph key pointers");
    memcpy(k.ptr, p, sizeof p);
    k.rec_cap_out = rs->rec_cap;
    k.n_gran = rs->ctx->n_gran;
    k.scratch_bytes = rs->scratch_bytes;
    k.hscratch_bytes = rs->hscratch_bytes;
    k.max_len = rs->max_len, k.seed_len = rs->seed_len, k.max_stacks = rs->max_stacks, k.max_entries = rs->max_entries;
    k.hand = rs->d_hand, k.hand_cap = rs->hand_cap, k.hand_stride = rs->hand_stride;
}

int run_tier(galn_hip_readset *rs, uint32_t flags, const Tier &t, const int32_t *list, int list_base, int n_list,
             bool persist, int slots, int fo, int block, int tier_id, uint64_t pop_budget = 0)
{
    SearchCtx *cx = rs->ctx;
    hipStream_t fst = cx->stream;
    if (n_list <= 0) return 0;
    if (!persist) slots = n_list;
    // scratch and pool for every lane of a full persistent grid, also when this launch has fewer
    // reads, so that launches of varying size reuse them
    const int alloc_lanes = std::max(1, slots);
    slots = std::max(1, std::min(slots, n_list));
    const int own_cap = (int)(((int64_t)(rs->max_entries + 16) / CH + MAXB + 2) / GRAN + 2);
    Layout L = make_layout(rs->max_len, rs->seed_len, t.rec_cap, own_cap);
    // scratch, and a pool for what the lanes keep plus room for the heavy reads in flight
    int r = prepare_tier(rs, t, alloc_lanes, slots);
    if (r) return r;
    KArgs a = base_args(rs, flags);
    a.list = list;
    a.list_base = list_base;
    a.n_list = n_list;
    a.qnext = rs->d_ctr + 2;
    a.n_slots = slots;
    a.scratch = rs->d_scratch;
    a.slot_bytes = L.slot_bytes;
    a.sw_off = L.sw_off;
    a.rec_off = L.rec_off;
    a.own_off = L.own_off;
    a.own_cap = L.own_cap;
    a.keep = t.keep;
    a.rec_cap = t.rec_cap;
    a.fail_list = rs->d_fail[fo];
    a.fail_n = rs->d_ctr + fo;
    a.fail_kind = rs->d_ctr + 4;
    a.pop_budget = pop_budget;
    a.heavy_list = rs->d_heavy;
    a.heavy_n = rs->d_ctr + 3;
    a.hand = pop_budget ? rs->d_hand : nullptr;
    a.hand_stride = rs->hand_stride;
    a.hand_cap = (pop_budget && rs->d_hand) ? rs->hand_cap : 0;
    if (persist) FILL(rs->d_ctr + 2, 0, sizeof(int32_t));
    // an empty free stack (tag and granule all ones) and no granule handed out yet
    FILL(cx->d_ghead, 0xffffffffu, NSHARD * 8);
    FILL(cx->d_gbump, 0, 4);
    const int li = (int)rs->launches.size();
    if (li >= galn_hip_readset::MAX_LAUNCH) return fail(GALN_HIP_E_DEVICE, "too many launches in one search");
    while ((int)rs->lev.size() < 2 * (li + 1)) {
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        rs->lev.push_back(e);
    }
    a.algo_bytes = rs->d_algo + li;
    rs->launches.push_back({tier_id, li, 0.0, 0});
    const int grid = (slots + block - 1) / block;
    if (!rs->capturing) HIP_OK(hipEventRecord(rs->lev[2 * li], fst));
    r = launch(fst, a, block, grid, persist, (flags & GALN_HIP_TRACE) != 0, buckets_of(rs));
    if (r) return r;
    if (!rs->capturing) HIP_OK(hipEventRecord(rs->lev[2 * li + 1], fst));
    return 0;
}

// Pushes the wave engine commits per round: H_CAP, or a lower value (at least H_PUSH, so that a
// round's first chain always fits) from GALN_HIP_TEST_PUSH_CAP, which the tests set to make the
// cap cut rounds short and check that the results do not change.
int heavy_push_cap()
{
    const char *v = getenv("GALN_HIP_TEST_PUSH_CAP");
    if (!v || !*v) return H_CAP;
    const long c = strtol(v, nullptr, 10);
    return (int)std::min<long>(H_CAP, std::max<long>(H_PUSH, c));
}

// The wave-per-read engine over list[0..n): one wave per read, `waves` resident waves
// claiming reads from a device counter; failures land in fail buffer `fo`.
int run_heavy(galn_hip_readset *rs, uint32_t flags, const int32_t *list, int n, int rec_cap, int fo, int tier_id,
              int waves)
{
    galn_hip_index *idx = rs->idx;
    SearchCtx *cx = rs->ctx;
    hipStream_t fst = cx->stream;
    const bool dev_count = n < 0;  // the count is the heavy-list length k_search leaves on the device
    if (n == 0) return 0;
    waves = std::max(1, dev_count ? waves : std::min(waves, n));
    int r = prepare_heavy(rs, waves, rec_cap);
    if (r) return r;
    const int own_cap = (int)std::min<uint64_t>(cx->n_gran, 32768);
    const HLayout L = make_hlayout(rec_cap, own_cap);
    const int wlen = rs->max_len + 1;
    const int swlen = (rs->seed_len > 0 && rs->seed_len < rs->max_len) ? rs->seed_len + 1 : 1;
    const size_t lds = heavy_lds_bytes(buckets_of(rs), wlen, swlen);
    if (lds > 65536) return fail(GALN_HIP_E_UNSUPPORTED, "wave engine: %zu bytes of LDS needed", lds);
    HArgs a{};
    a.bwt = idx->d_bwt;
    a.meta = idx->meta;
    a.query = rs->d_query;
    a.qoff = rs->d_qoff;
    a.qlen = rs->d_qlen;
    a.maxdiff = rs->d_maxdiff;
    a.batch = rs->d_batch;
    a.bopts = rs->d_bopts;
    a.nstacks = rs->d_nstacks;
    a.list = list;
    a.n_list = dev_count ? 0 : n;
    a.n_list_dev = dev_count ? rs->d_ctr + 3 : nullptr;
    a.hand = rs->d_hand;  // the list is the heavy list of the k_search launch just before
    a.hand_stride = rs->hand_stride;
    a.hand_cap = rs->d_hand ? rs->hand_cap : 0;
    a.qnext = rs->d_ctr + 2;
    a.n_slots = waves;
    a.scratch = rs->d_hscratch;
    a.slot_bytes = L.slot_bytes;
    a.rec_off = L.rec_off;
    a.own_off = L.own_off;
    a.plog_off = L.plog_off;
    a.slog_off = L.slog_off;
    a.own_cap = own_cap;
    a.rec_cap = rec_cap;
    a.pool = cx->d_pool_chk;
    a.gnext = cx->d_gnext;
    a.ghead = cx->d_ghead;
    a.gbump = cx->d_gbump;
    a.n_gran = cx->n_gran;
    a.out_n = rs->d_out_n;
    a.out_off = rs->d_out_off;
    a.out_pops = rs->d_pops;
    a.out_trace = rs->d_trace;
    a.out_bytes = rs->d_bytes;
    a.out_rec = rs->d_rec;
    a.out_cap = rs->rec_cap;
    a.out_top = rs->d_top;
    a.fail_list = rs->d_fail[fo];
    a.fail_n = rs->d_ctr + fo;
    a.fail_kind = rs->d_ctr + 4;
    a.buckets = buckets_of(rs);
    a.wlen = wlen;
    a.push_cap = heavy_push_cap();
    a.swlen = swlen;
    a.stats = (flags & GALN_HIP_STATS) ? rs->d_stats : nullptr;
    FILL(rs->d_ctr + 2, 0, sizeof(int32_t));
    FILL(cx->d_ghead, 0xffffffffu, NSHARD * 8);
    FILL(cx->d_gbump, 0, 4);
    const int li = (int)rs->launches.size();
    if (li >= galn_hip_readset::MAX_LAUNCH) return fail(GALN_HIP_E_DEVICE, "too many launches in one search");
    while ((int)rs->lev.size() < 2 * (li + 1)) {
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        rs->lev.push_back(e);
    }
    a.algo_bytes = rs->d_algo + li;
    rs->launches.push_back({tier_id, li, 0.0, 0});
    if (!rs->capturing) HIP_OK(hipEventRecord(rs->lev[2 * li], fst));
    if (a.hand && a.hand_cap > 0) {
        WArgs w{};
        w.bwt = a.bwt;
        w.meta = a.meta;
        w.query = a.query;
        w.qoff = a.qoff;
        w.qlen = a.qlen;
        w.batch = a.batch;
        w.bopts = a.bopts;
        w.list = list;
        w.n_list = a.n_list;
        w.n_list_dev = a.n_list_dev;
        w.hand = rs->d_hand;
        w.hand_stride = a.hand_stride;
        w.hand_cap = a.hand_cap;
        w.wlen = wlen;
        w.swlen = swlen;
        const int wgrid = dev_count ? idx->n_cu * 8 : std::max(1, std::min(idx->n_cu * 8, (n + 255) / 256));
        hipLaunchKernelGGL(k_widths, dim3(wgrid), dim3(256), 0, fst, w);
        HIP_OK(hipGetLastError());
    }
    const bool tr = (flags & GALN_HIP_TRACE) != 0;
    if (rs->heavy_wpe == 5) {
        if (tr)
            hipLaunchKernelGGL((k_heavy<true, 5>), dim3(waves), dim3(HW), lds, fst, a);
        else
            hipLaunchKernelGGL((k_heavy<false, 5>), dim3(waves), dim3(HW), lds, fst, a);
    } else {
        if (tr)
            hipLaunchKernelGGL((k_heavy<true, 4>), dim3(waves), dim3(HW), lds, fst, a);
        else
            hipLaunchKernelGGL((k_heavy<false, 4>), dim3(waves), dim3(HW), lds, fst, a);
    }
    HIP_OK(hipGetLastError());
    if (!rs->capturing) HIP_OK(hipEventRecord(rs->lev[2 * li + 1], fst));
    return 0;
}

int heavy_waves(galn_hip_readset *rs, uint32_t flags)
{
    const int wlen = rs->max_len + 1;
    const int swlen = (rs->seed_len > 0 && rs->seed_len < rs->max_len) ? rs->seed_len + 1 : 1;
    const size_t lds = heavy_lds_bytes(buckets_of(rs), wlen, swlen);
    int nb = 0;
    const bool tr = (flags & GALN_HIP_TRACE) != 0;
    hipError_t e;
    if (rs->heavy_wpe == 5)
        e = tr ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_heavy<true, 5>, HW, lds)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_heavy<false, 5>, HW, lds);
    else
        e = tr ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_heavy<true, 4>, HW, lds)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_heavy<false, 4>, HW, lds);
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb * rs->idx->n_cu;
}

int read_ctr(galn_hip_readset *rs, int which, int32_t *v)
{
    HIP_OK(hipMemcpyAsync(v, rs->d_ctr + which, sizeof(int32_t), hipMemcpyDeviceToHost, rs->ctx->stream));
    HIP_OK(hipStreamSynchronize(rs->ctx->stream));
    return 0;
}

int grow_records(galn_hip_readset *rs)
{
    hipStream_t st = rs->ctx->stream;
    unsigned long long top = 0;
    HIP_OK(hipMemcpyAsync(&top, rs->d_top, sizeof(top), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    const uint64_t used = std::min<uint64_t>(top, rs->rec_cap);
    const uint64_t ncap = std::max<uint64_t>(rs->rec_cap * 2, top + 4096);
    galn_aln_rec_t *nrec = nullptr;
    HIP_OK(hipMalloc(&nrec, ncap * sizeof(galn_aln_rec_t)));
    if (used) HIP_OK(hipMemcpyAsync(nrec, rs->d_rec, used * sizeof(galn_aln_rec_t), hipMemcpyDeviceToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    HIP_OK(hipFree(rs->d_rec));
    rs->d_rec = nrec;
    rs->rec_cap = ncap;
    unsigned long long u = used;
    HIP_OK(hipMemcpyAsync(rs->d_top, &u, sizeof(u), hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

// Re-run failed reads (fail buffer `fi`, count n) until none is left: a full output arena
// grows, full record slots get 16 times more, an exhausted pool gets fewer lanes sharing it.
int run_retries(galn_hip_readset *rs, uint32_t flags, int fi, int32_t n, Tier t, int slots, uint64_t budget,
                int hwaves)
{
    hipStream_t fst = rs->ctx->stream;
    const int block = pick_block(rs);
    for (int round = 0; n > 0; ++round) {
        if (round >= 24) return fail(GALN_HIP_E_NOMEM, "reads still failing after %d re-runs", round);
        int32_t kind[3] = {0, 0, 0};
        HIP_OK(hipMemcpy(kind, rs->d_ctr + 4, sizeof(kind), hipMemcpyDeviceToHost));
        int r;
        if (kind[0] && (r = grow_records(rs))) return r;
        if (kind[1]) t.rec_cap *= 16;
        if (kind[2]) {
            // an exhausted pool: twice the granules when there is room, else fewer lanes sharing it
            const uint32_t before = rs->ctx->n_gran;
            rs->ctx->grow = true;
            if ((r = ensure_pool(rs->ctx, before))) return r;
            if (rs->ctx->n_gran <= before) {
                slots = std::max(block, slots / 4);
                hwaves = std::max(1, hwaves / 4);
            }
        }
        slots = std::min(slots, std::max(block, (int)n));
        const int fo = fi ^ 1;
        FILL(rs->d_ctr + fo, 0, sizeof(int32_t));
        FILL(rs->d_ctr + 3, 0, 4 * sizeof(int32_t));
        if ((r = run_tier(rs, flags, t, rs->d_fail[fi], 0, n, true, slots, fo, block, 1 + round, budget))) return r;
        int32_t nh = 0;
        if ((r = read_ctr(rs, 3, &nh))) return r;
        if (nh > 0 && (r = run_heavy(rs, flags, rs->d_heavy, nh, 256 * (1 + round), fo, 11 + round, hwaves))) return r;
        int32_t nf = 0;
        if ((r = read_ctr(rs, fo, &nf))) return r;
        fi = fo;
        n = nf;
    }
    return 0;
}

// Pops after which the batched form hands a read to the wave-per-read engine when no tuning is
// given: measured best on the 10M-read aDNA workload (512: 2.67 s, 1024: 2.85 s, 4096: 3.34 s).
constexpr int DEFAULT_HEAVY_POPS = 512;

int readset_search(galn_hip_readset *rs, int form, uint32_t flags, const galn_hip_tune_t *tune)
{
    if (!tune && rs->has_tuned) tune = &rs->tuned;
    galn_hip_index *idx = rs->idx;
    HIP_OK(hipSetDevice(idx->device));
    hipStream_t fst = rs->ctx->stream;
    rs->last_flags = flags;
    rs->launches.clear();
    const bool trace = (flags & GALN_HIP_TRACE) != 0;
    const bool graph = form != 0 && (flags & GALN_HIP_GRAPH) != 0;
    auto prologue = [&]() -> int {
        FILL(rs->d_algo, 0, sizeof(unsigned long long) * galn_hip_readset::MAX_LAUNCH);
        FILL(rs->d_stats, 0, 16 * 8);
        FILL(rs->d_top, 0, sizeof(unsigned long long));
        FILL(rs->d_ctr, 0, 8 * sizeof(int32_t));
        return 0;
    };
    if (!graph) {
        const int pr = prologue();
        if (pr) return pr;
    }
    int block = pick_block(rs);
    Tier t0{64, 4};
    if (tune && tune->rec_per_lane > 0) t0.rec_cap = tune->rec_per_lane;
    if (tune && tune->chunks_per_lane > 0) t0.keep = std::max(1, tune->chunks_per_lane / GRAN);
    int r = 0, slots = 0, hwaves = 0;
    uint64_t budget = 0;
    if (form == 0) {
        // plain: one launch per 0x40000 reads, one lane per read, input order
        slots = 0x40000;
        for (int s = 0; s < rs->n && !r; s += 0x40000)
            r = run_tier(rs, flags, t0, nullptr, s, std::min(0x40000, rs->n - s), false, 0, 0, block, 0);
    } else {
        // batched: one persistent launch over the reads, longest first; reads that spend their
        // pop budget go to the wave-per-read engine
        // 64-thread workgroups unless tuned otherwise (the autotuner's choice on every workload
        // measured so far)
        block = 64;
        if (tune && tune->block_threads > 0 && buckets_of(rs) * tune->block_threads * 4 <= 65536)
            block = tune->block_threads;
        slots = resident_slots(rs, block, trace, tune ? tune->waves_per_cu : 0);
        const bool sorted = !tune || tune->sort_by_cost != 0;
        const int hp = tune ? tune->heavy_pops : 0;
        rs->heavy_wpe = tune && tune->heavy_waves_per_simd == 5 ? 5 : 4;
        budget = (hp < 0 || !heavy_ok(rs)) ? 0 : (uint64_t)(hp > 0 ? hp : DEFAULT_HEAVY_POPS);
        hwaves = heavy_waves(rs, flags);
        if (budget && (r = prepare_hand(rs))) return r;
        const int32_t *list = sorted ? rs->d_sorted : nullptr;
        if (graph) {
            // the prologue, the persistent launch and the wave-engine launch (its read count read
            // on the device) as one captured graph, replayed while nothing it holds has changed
            GraphKey key;
            memset(&key, 0, sizeof key);  // padding included: keys are compared bytewise
            key.flags = flags, key.block = block, key.slots = slots, key.hwaves = hwaves, key.n = rs->n;
            key.budget = budget, key.rec_cap = (uint32_t)t0.rec_cap, key.keep = (uint32_t)t0.keep, key.list = list;
            key.push_cap = heavy_push_cap();
            key.heavy_wpe = rs->heavy_wpe;
            if ((r = prepare_tier(rs, t0, slots, std::min(slots, rs->n))) ||
                (budget && (r = prepare_heavy(rs, hwaves, 256))))
                return r;
            while (rs->lev.size() < 4) {
                hipEvent_t e;
                HIP_OK(hipEventCreate(&e));
                rs->lev.push_back(e);
            }
            fill_key_ptrs(rs, key);
            if (!rs->gexec || memcmp(&key, &rs->gkey, sizeof key) != 0) {
                for (hipGraphExec_t *ge : {&rs->gexec, &rs->gexec2}) {
                    if (*ge) HIP_OK(hipGraphExecDestroy(*ge));
                    *ge = nullptr;
                }
                // two graphs, so that the timing events sit between them rather than inside
                auto capture = [&](int part, hipGraphExec_t *out) -> int {
                    HIP_OK(hipStreamBeginCapture(fst, hipStreamCaptureModeThreadLocal));
                    rs->capturing = true;
                    int rc = 0;
                    if (part == 0) {
                        rc = prologue();
                        if (!rc) rc = run_tier(rs, flags, t0, list, 0, rs->n, true, slots, 0, block, 0, budget);
                    } else {
                        rc = run_heavy(rs, flags, rs->d_heavy, -1, 256, 0, 10, hwaves);
                    }
                    rs->capturing = false;
                    hipGraph_t g = nullptr;
                    const hipError_t ec = hipStreamEndCapture(fst, &g);
                    if (rc) {
                        if (g) (void)hipGraphDestroy(g);
                        return rc;
                    }
                    if (ec != hipSuccess) return fail(GALN_HIP_E_DEVICE, "graph capture: %s", hipGetErrorString(ec));
                    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    if (ei != hipSuccess) {
                        *out = nullptr;
                        return fail(GALN_HIP_E_DEVICE, "graph instantiate: %s", hipGetErrorString(ei));
                    }
                    return 0;
                };
                if ((r = capture(0, &rs->gexec))) return r;
                if (budget && (r = capture(1, &rs->gexec2))) return r;
                rs->gkey = key;
                rs->glaunches = rs->launches;
            }
            rs->launches = rs->glaunches;
            HIP_OK(hipEventRecord(rs->lev[0], fst));
            HIP_OK(hipGraphLaunch(rs->gexec, fst));
            HIP_OK(hipEventRecord(rs->lev[1], fst));
            if (rs->gexec2) {
                HIP_OK(hipEventRecord(rs->lev[2], fst));
                HIP_OK(hipGraphLaunch(rs->gexec2, fst));
                HIP_OK(hipEventRecord(rs->lev[3], fst));
            }
        } else {
            r = run_tier(rs, flags, t0, list, 0, rs->n, true, slots, 0, block, 0, budget);
            int32_t nh = 0;
            if (!r && budget && !(r = read_ctr(rs, 3, &nh)) && nh > 0)
                r = run_heavy(rs, flags, rs->d_heavy, nh, 256, 0, 10, hwaves);
        }
    }
    if (r) return r;
    int32_t nf = 0;
    if ((r = read_ctr(rs, 0, &nf))) return r;
    if (nf > 0 && (r = run_retries(rs, flags, 0, nf, t0, slots, budget, std::max(1, hwaves)))) return r;
    HIP_OK(hipStreamSynchronize(fst));
    if (rs->ctx->d_gbump) {
        uint32_t used = 0;
        HIP_OK(hipMemcpyAsync(&used, rs->ctx->d_gbump, 4, hipMemcpyDeviceToHost, fst));
        HIP_OK(hipStreamSynchronize(fst));
        if (used >= rs->ctx->n_gran) rs->ctx->grow = true;
    }
    std::vector<unsigned long long> ab(rs->launches.size());
    if (!ab.empty())
        HIP_OK(hipMemcpy(ab.data(), rs->d_algo, ab.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    rs->first_ms = rs->total_ms = 0;
    for (size_t i = 0; i < rs->launches.size(); ++i) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, rs->lev[2 * i], rs->lev[2 * i + 1]));
        rs->launches[i].ms = ms;
        rs->launches[i].bytes = ab[i];
        if (rs->launches[i].tier == 0) rs->first_ms += ms;
        rs->total_ms += ms;
    }
    return 0;
}

int check_reads(const galn_hip_reads *r)
{
    if (!r || r->n_reads < 0) return fail(GALN_HIP_E_ARG, "reads: null or negative count");
    if (r->n_reads > 0 && (!r->query || !r->qoff || !r->qlen || !r->max_diff || !r->batch))
        return fail(GALN_HIP_E_ARG, "reads: null array");
    if (r->n_batches < 1 || !r->bopts || !r->n_stacks) return fail(GALN_HIP_E_ARG, "reads: no batch options");
    for (int b = 0; b < r->n_batches; ++b)
        if (r->bopts[b].seed_len < 0) return fail(GALN_HIP_E_UNSUPPORTED, "reads: negative seed length");
    for (int i = 0; i < r->n_reads; ++i)
        if (r->batch[i] < 0 || r->batch[i] >= r->n_batches) return fail(GALN_HIP_E_ARG, "reads: batch index out of range");
    return 0;
}

// Grow the per-read and per-batch device arrays to hold n reads, qbytes query bytes and nb
// batches (contents are not kept).
int readset_reserve(galn_hip_readset *rs, int n, uint64_t qbytes, int nb)
{
    const bool grow_n = n > rs->cap_n, grow_q = qbytes > rs->cap_q, grow_b = nb > rs->cap_b;
    if (!grow_n && !grow_q && !grow_b) return 0;
    HIP_OK(hipStreamSynchronize(rs->ctx->stream));
    auto regrow = [](void **p, size_t bytes) __attribute__((always_inline)) -> bool {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        return hipMalloc(p, std::max<size_t>(bytes, 16)) == hipSuccess;
    };
    bool ok = true;
    if (grow_q) {
        const uint64_t c = std::max<uint64_t>(qbytes, 2 * rs->cap_q);
        ok = regrow((void **)&rs->d_query, c + 16);
        rs->cap_q = ok ? c : 0;
    }
    if (ok && grow_b) {
        const int c = std::max(nb, 2 * rs->cap_b);
        ok = regrow((void **)&rs->d_nstacks, (size_t)c * 4) && regrow((void **)&rs->d_bopts, (size_t)c * sizeof(galn_search_opt_t));
        rs->cap_b = ok ? c : 0;
    }
    if (ok && grow_n) {
        const size_t c = (size_t)std::min<int64_t>(std::max<int64_t>(n, 2 * (int64_t)rs->cap_n), INT32_MAX);
        ok = regrow((void **)&rs->d_qoff, c * 8) && regrow((void **)&rs->d_qlen, c * 4) &&
             regrow((void **)&rs->d_maxdiff, c * 4) && regrow((void **)&rs->d_batch, c * 4) &&
             regrow((void **)&rs->d_sorted, c * 4) && regrow((void **)&rs->d_out_n, c * 4) &&
             regrow((void **)&rs->d_out_off, c * 8) && regrow((void **)&rs->d_pops, c * 8) &&
             regrow((void **)&rs->d_trace, c * 8) && regrow((void **)&rs->d_bytes, c * 8) &&
             regrow((void **)&rs->d_fail[0], c * 4) && regrow((void **)&rs->d_fail[1], c * 4) &&
             regrow((void **)&rs->d_heavy, c * 4) && regrow((void **)&rs->d_incl, c * 8) &&
             regrow((void **)&rs->d_roff, c * 8);
        rs->cap_n = ok ? (int)c : 0;
        if (ok && rs->rec_cap < 2 * c) {
            rs->rec_cap = std::max<uint64_t>(2 * c, 1 << 16);
            ok = regrow((void **)&rs->d_rec, rs->rec_cap * sizeof(galn_aln_rec_t));
            if (!ok) rs->rec_cap = 0;
        }
    }
    if (!ok) return fail(GALN_HIP_E_NOMEM, "readset: device allocation failed");
    return 0;
}

// Copy a read set (already checked) into rs's device arrays.
int readset_fill(galn_hip_readset *rs, const galn_hip_reads *reads)
{
    const int n = reads->n_reads;
    uint64_t qbytes = 0;
    int max_len = 1;
    for (int i = 0; i < n; ++i) {
        qbytes = std::max<uint64_t>(qbytes, reads->qoff[i] + (uint64_t)std::max(reads->qlen[i], 0));
        max_len = std::max(max_len, reads->qlen[i]);
    }
    int r = readset_reserve(rs, std::max(n, 1), qbytes, reads->n_batches);
    if (r) return r;
    rs->n = n;
    rs->n_batches = reads->n_batches;
    rs->max_len = std::min(max_len, MAXLEN);
    rs->qbytes = qbytes;
    // seed widths are kept for the longest seed that is shorter than some read
    rs->seed_len = 0;
    for (int b = 0; b < reads->n_batches; ++b)
        if (reads->bopts[b].seed_len < rs->max_len) rs->seed_len = std::max(rs->seed_len, reads->bopts[b].seed_len);
    rs->max_stacks = 1;
    for (int b = 0; b < reads->n_batches; ++b) rs->max_stacks = std::max(rs->max_stacks, reads->n_stacks[b]);
    rs->h_bopts.assign(reads->bopts, reads->bopts + reads->n_batches);
    rs->max_entries = 0;
    for (const auto &o : rs->h_bopts) rs->max_entries = std::max(rs->max_entries, std::max(o.max_entries, 0));
    // claim order of the batched form: longest reads first (their searches are the longest); a
    // counting sort by length (stable; lengths above the kernel limit share the top bin)
    std::vector<int32_t> order(n);
    {
        std::vector<int32_t> cnt(MAXLEN + 3, 0);
        auto bin = [&](int i) { return MAXLEN + 1 - std::min(std::max(reads->qlen[i], 0), MAXLEN + 1); };
        for (int i = 0; i < n; ++i) ++cnt[bin(i) + 1];
        for (int b = 1; b <= MAXLEN + 2; ++b) cnt[b] += cnt[b - 1];
        for (int i = 0; i < n; ++i) order[cnt[bin(i)]++] = i;
    }
    hipStream_t s = rs->ctx->stream;
    if (hipMemcpyAsync(rs->d_query, reads->query, qbytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_qoff, reads->qoff, (size_t)n * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_qlen, reads->qlen, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_maxdiff, reads->max_diff, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_batch, reads->batch, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_nstacks, reads->n_stacks, (size_t)reads->n_batches * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_bopts, reads->bopts, (size_t)reads->n_batches * sizeof(galn_search_opt_t),
                       hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_sorted, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(GALN_HIP_E_DEVICE, "readset: copy failed");
    return 0;
}

}  // namespace

extern "C" {

int galn_hip_abi_version(void) { return GALN_HIP_ABI_VERSION; }

void galn_hip_get_limits(galn_hip_limits_t *lim)
{
    if (!lim) return;
    lim->max_read_len = MAXLEN;
    lim->max_n_stacks = MAXB;
    lim->max_diff = MAXDIFF;
    lim->max_gapo_plus_gape = 255;
    lim->max_seq_len = KMASK;
}

const char *galn_hip_last_error(void) { return g_err; }

}  // extern "C"

// Used by the serving loop (galn_server.cpp), not exported: this thread's error text, and
// the count of handles that keep an index open.
__attribute__((visibility("hidden"))) void galn_hip_set_error(const char *msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }
__attribute__((visibility("hidden"))) void galn_hip_index_hold(galn_hip_index *idx, int delta) { idx->n_readsets += delta; }
// A server's worker threads mark themselves; while a server is open on an index, only they may
// search it (its stream and stack pool are theirs).
thread_local bool tl_server_worker = false;
__attribute__((visibility("hidden"))) void galn_hip_server_worker_mark() { tl_server_worker = true; }
__attribute__((visibility("hidden"))) void galn_hip_index_serve(galn_hip_index *idx, int delta) { idx->n_servers += delta; }

// A server slot's read set gets a stream and a stack pool of its own, so that the slots' launches
// run at the same time (the read set was made on the index's context).
__attribute__((visibility("hidden"))) int galn_hip_readset_make_private(galn_hip_readset *rs)
{
    if (rs->own_ctx) return 0;
    HIP_OK(hipSetDevice(rs->idx->device));
    HIP_OK(hipStreamSynchronize(rs->ctx->stream));
    SearchCtx *c = new SearchCtx;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(GALN_HIP_E_DEVICE, "readset: stream creation failed");
    }
    // graphs captured on the old stream hold the old pool: captured again on first use
    for (hipGraphExec_t *ge : {&rs->gexec, &rs->gexec2}) {
        if (*ge) (void)hipGraphExecDestroy(*ge);
        *ge = nullptr;
    }
    rs->own_ctx = c;
    rs->ctx = c;
    return 0;
}

// Capacity for launches of up to n reads, qbytes query bytes and nb batches, allocated now (a
// server slot's read set, before its first launch fills it): later launches of up to that size
// then allocate nothing per read. The read set's contents are not kept.
__attribute__((visibility("hidden"))) int galn_hip_readset_reserve(galn_hip_readset *rs, int n, uint64_t qbytes, int nb)
{
    HIP_OK(hipSetDevice(rs->idx->device));
    return readset_reserve(rs, std::max(n, 1), qbytes, std::max(nb, 1));
}

// The search buffers of launches of up to n reads of up to max_len bases (the read set's arrays
// must already hold n reads: galn_hip_readset_reserve).
__attribute__((visibility("hidden"))) int galn_hip_readset_reserve_search(galn_hip_readset *rs, int n, int max_len,
                                                                          uint32_t flags)
{
    HIP_OK(hipSetDevice(rs->idx->device));
    if (n > rs->cap_n) return fail(GALN_HIP_E_ARG, "readset_reserve_search: %d reads, the arrays hold %d", n, rs->cap_n);
    return reserve_search(rs, n, max_len, flags);
}

namespace {
int noalloc_setup(int device);

int busy_check(const galn_hip_index *idx, const char *what)
{
    if (idx && idx->n_servers.load() > 0 && !tl_server_worker)
        return fail(GALN_HIP_E_BUSY, "%s: a server holds this index until galn_hip_server_close", what);
    return 0;
}
}  // namespace

extern "C" {

int galn_hip_index_open(const uint32_t *bwt, const galn_bwt_meta_t *meta, int device, galn_hip_index **out)
{
    if (!bwt || !meta || !out) return fail(GALN_HIP_E_ARG, "index_open: null argument");
    *out = nullptr;
    if (meta->seq_len > KMASK) return fail(GALN_HIP_E_UNSUPPORTED, "index of %llu symbols exceeds 2^38-1",
                                           (unsigned long long)meta->seq_len);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GALN_HIP_E_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(GALN_HIP_E_DEVICE, "no HIP device %d (%d visible)", device, ndev);
    HIP_OK(hipSetDevice(device));
    galn_hip_index *idx = new galn_hip_index;
    idx->device = device;
    idx->hmeta = *meta;
    idx->meta.primary = meta->primary;
    for (int i = 0; i < 5; ++i) idx->meta.L2[i] = meta->L2[i];
    idx->meta.seq_len = meta->seq_len;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete idx;
        return fail(GALN_HIP_E_DEVICE, "hipGetDeviceProperties failed");
    }
    idx->n_cu = prop.multiProcessorCount;
    const size_t bytes = (size_t)meta->bwt_size * 4;
    // whole 64-byte blocks are loaded: pad so the last block's load stays in bounds
    const size_t alloc = align_up(bytes + 256, 256);
    if (hipMalloc(&idx->d_bwt, alloc) != hipSuccess) {
        delete idx;
        return fail(GALN_HIP_E_NOMEM, "index of %zu bytes does not fit in device memory", bytes);
    }
    if (hipStreamCreateWithFlags(&idx->ctx0.stream, hipStreamNonBlocking) != hipSuccess ||
        hipMemsetAsync(idx->d_bwt, 0, alloc, idx->ctx0.stream) != hipSuccess ||
        hipMemcpyAsync(idx->d_bwt, bwt, bytes, hipMemcpyHostToDevice, idx->ctx0.stream) != hipSuccess) {
        if (idx->ctx0.stream) (void)hipStreamDestroy(idx->ctx0.stream);
        (void)hipFree(idx->d_bwt);
        delete idx;
        return fail(GALN_HIP_E_DEVICE, "index upload failed");
    }
    const uint64_t n_blocks = ((uint64_t)meta->bwt_size + 15) / 16;
    if (n_blocks > 0)
        hipLaunchKernelGGL(k_planes, dim3((unsigned)((n_blocks + 255) / 256)), dim3(256), 0, idx->ctx0.stream, idx->d_bwt,
                           n_blocks);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(idx->ctx0.stream) != hipSuccess) {
        (void)hipStreamDestroy(idx->ctx0.stream);
        (void)hipFree(idx->d_bwt);
        delete idx;
        return fail(GALN_HIP_E_DEVICE, "index upload failed");
    }
    // device address of the no-allocation search's array on this device (once per device); a
    // failure here is reported by that entry point, which looks it up again
    if (noalloc_setup(device) != 0) g_err[0] = 0;
    *out = idx;
    return 0;
}

int galn_hip_index_close(galn_hip_index *idx)
{
    if (!idx) return 0;
    if (idx->n_readsets.load() > 0)
        return fail(GALN_HIP_E_BUSY, "index_close: %d read sets (or servers) of this index are still open",
                    idx->n_readsets.load());
    (void)hipSetDevice(idx->device);
    if (idx->ctx0.stream) (void)hipStreamSynchronize(idx->ctx0.stream);
    (void)hipFree(idx->d_bwt);
    if (idx->d_sa) (void)hipFree(idx->d_sa);
    idx->ctx0.free_pool();
    if (idx->ctx0.stream) (void)hipStreamDestroy(idx->ctx0.stream);
    delete idx;
    return 0;
}

int galn_hip_index_device(const galn_hip_index *idx) { return idx ? idx->device : -1; }

int galn_hip_readset_free(galn_hip_readset *rs)
{
    if (!rs) return 0;
    (void)hipSetDevice(rs->idx->device);
    (void)hipStreamSynchronize(rs->ctx->stream);
    void *ptrs[] = {rs->d_query, rs->d_qoff, rs->d_qlen, rs->d_maxdiff, rs->d_batch, rs->d_nstacks, rs->d_bopts,
                    rs->d_sorted, rs->d_out_n, rs->d_out_off, rs->d_pops, rs->d_trace, rs->d_bytes, rs->d_rec, rs->d_top,
                    rs->d_fail[0], rs->d_fail[1], rs->d_ctr, rs->d_scratch, rs->d_algo, rs->d_heavy, rs->d_hscratch,
                    rs->d_stats, rs->d_incl, rs->d_roff, rs->d_stage, rs->d_scan_tmp, rs->d_hand};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &e : rs->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : rs->lev) (void)hipEventDestroy(e);
    if (rs->gexec) (void)hipGraphExecDestroy(rs->gexec);
    if (rs->gexec2) (void)hipGraphExecDestroy(rs->gexec2);
    if (rs->own_ctx) {
        rs->own_ctx->free_pool();
        if (rs->own_ctx->stream) (void)hipStreamDestroy(rs->own_ctx->stream);
        delete rs->own_ctx;
    }
    rs->idx->n_readsets.fetch_sub(1);
    delete rs;
    return 0;
}

int galn_hip_readset_upload(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_readset **out)
{
    if (!idx || !out) return fail(GALN_HIP_E_ARG, "readset_upload: null argument");
    *out = nullptr;
    int r = busy_check(idx, "readset_upload");
    if (r) return r;
    if ((r = check_reads(reads))) return r;
    HIP_OK(hipSetDevice(idx->device));
    galn_hip_readset *rs = new galn_hip_readset;
    rs->idx = idx;
    rs->ctx = &idx->ctx0;
    idx->n_readsets.fetch_add(1);
    auto M = [&](void **p, size_t b) __attribute__((always_inline)) -> bool { return hipMalloc(p, b) == hipSuccess; };
    bool ok = M((void **)&rs->d_top, 8) && M((void **)&rs->d_ctr, 8 * sizeof(int32_t)) && M((void **)&rs->d_stats, 16 * 8) &&
              M((void **)&rs->d_algo, sizeof(unsigned long long) * galn_hip_readset::MAX_LAUNCH);
    if (!ok) {
        galn_hip_readset_free(rs);
        return fail(GALN_HIP_E_NOMEM, "readset_upload: device allocation failed");
    }
    for (auto &e : rs->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            galn_hip_readset_free(rs);
            return fail(GALN_HIP_E_DEVICE, "hipEventCreate failed");
        }
    if ((r = readset_fill(rs, reads))) {
        galn_hip_readset_free(rs);
        return r;
    }
    *out = rs;
    return 0;
}

int galn_hip_readset_reload(galn_hip_readset *rs, const galn_hip_reads *reads)
{
    if (!rs) return fail(GALN_HIP_E_ARG, "readset_reload: null handle");
    int r = busy_check(rs->idx, "readset_reload");
    if (r) return r;
    if ((r = check_reads(reads))) return r;
    HIP_OK(hipSetDevice(rs->idx->device));
    return readset_fill(rs, reads);
}

int galn_hip_readset_search(galn_hip_index *idx, galn_hip_readset *rs, int form, uint32_t flags,
                            const galn_hip_tune_t *tune)
{
    if (!idx || !rs || rs->idx != idx) return fail(GALN_HIP_E_ARG, "readset_search: bad handle");
    const int b = busy_check(idx, "readset_search");
    if (b) return b;
    if (rs->n == 0) return 0;
    return readset_search(rs, form, flags, tune);
}

int galn_hip_readset_autotune(galn_hip_index *idx, galn_hip_readset *rs, uint32_t flags, galn_hip_tune_t *out)
{
    if (!idx || !rs || rs->idx != idx) return fail(GALN_HIP_E_ARG, "readset_autotune: bad handle");
    const int b = busy_check(idx, "readset_autotune");
    if (b) return b;
    rs->has_tuned = false;
    galn_hip_tune_t best{};
    best.sort_by_cost = 1;
    const bool verbose = galn_hip_verbose();
    if (rs->n > 0) {
        double best_ms = 0;
        auto trial = [&](const galn_hip_tune_t &t) -> int {
            const auto t0 = std::chrono::steady_clock::now();
            const int r = readset_search(rs, 1, flags, &t);
            if (r) return r;
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (verbose)
                fprintf(stderr, "[galn_hip] autotune: cutoff %d, %d-thread groups, %d waves/CU, %d record slots, "
                                "%d chunks kept, wave engine at %d waves/SIMD: %.1f ms\n",
                        t.heavy_pops, t.block_threads, t.waves_per_cu, t.rec_per_lane, t.chunks_per_lane,
                        t.heavy_waves_per_simd == 5 ? 5 : 4, ms);
            if (best_ms == 0 || ms < best_ms) {
                best_ms = ms;
                best = t;
            }
            return 0;
        };
        // the thread-vs-wave cutoff first (when the wave engine may run), then the workgroup size,
        // then one change at a time from the best so far: the persistent grid's waves per CU, the
        // record slots per lane, and the stack chunks a lane keeps between reads
        galn_hip_tune_t t{};
        t.sort_by_cost = 1;
        t.block_threads = 128;
        // (1: every read that pops goes to the wave engine, k_search computing only its widths)
        const int cut[] = {1, 16, 64, 128, 256, 512, 1024, 4096};
        int r;
        if (heavy_ok(rs)) {
            for (int c : cut) {
                t.heavy_pops = c;
                if ((r = trial(t))) return r;
            }
        } else {
            t.heavy_pops = -1;
            if ((r = trial(t))) return r;
        }
        t = best;
        t.block_threads = 64;
        if ((r = trial(t))) return r;
        const int waves[] = {4, 6}, recs[] = {16, 256}, keeps[] = {GRAN, 16 * GRAN};
        for (int v : waves) {
            t = best;
            t.waves_per_cu = v;
            if ((r = trial(t))) return r;
        }
        for (int v : recs) {
            t = best;
            t.rec_per_lane = v;
            if ((r = trial(t))) return r;
        }
        for (int v : keeps) {
            t = best;
            t.chunks_per_lane = v;
            if ((r = trial(t))) return r;
        }
        if (heavy_ok(rs) && best.heavy_pops >= 0) {
            t = best;
            t.heavy_waves_per_simd = 5;
            if ((r = trial(t))) return r;
            // with more waves in flight the best cutoff can move: the small ones again
            if (best.heavy_waves_per_simd == 5) {
                const galn_hip_tune_t at5 = best;
                for (int c : {1, 16, 64, 128}) {
                    if (c == at5.heavy_pops) continue;
                    t = at5;
                    t.heavy_pops = c;
                    if ((r = trial(t))) return r;
                }
            }
        }
    }
    rs->tuned = best;
    rs->has_tuned = true;
    if (out) *out = best;
    return 0;
}

int galn_hip_readset_results(galn_hip_readset *rs, galn_hip_out *out)
{
    if (!rs || !out || (rs->n > 0 && (!out->n_aln || !out->rec_off)))
        return fail(GALN_HIP_E_ARG, "readset_results: null argument");
    galn_hip_index *idx = rs->idx;
    const int b = busy_check(idx, "readset_results");
    if (b) return b;
    HIP_OK(hipSetDevice(idx->device));
    const int n = rs->n;
    if (n == 0) {
        out->n_rec = 0;
        return 0;
    }
    hipStream_t s = rs->ctx->stream;
    // records in read order, arranged on the device: offsets = prefix sums of the record counts
    // (errors count 0), then every read's records gathered from the arena
    size_t tmp_bytes = 0;
    const rocprim::transform_iterator<const int32_t *, RecCount, uint64_t> cnt(rs->d_out_n, RecCount());
    HIP_OK(rocprim::inclusive_scan(nullptr, tmp_bytes, cnt, rs->d_incl, (size_t)n, rocprim::plus<uint64_t>(), s));
    if (int g = grow_dev(&rs->d_scan_tmp, rs->scan_tmp_bytes, tmp_bytes, s)) return g;
    HIP_OK(rocprim::inclusive_scan(rs->d_scan_tmp, tmp_bytes, cnt, rs->d_incl, (size_t)n, rocprim::plus<uint64_t>(), s));
    unsigned long long need = 0;
    HIP_OK(hipMemcpyAsync(&need, rs->d_incl + (n - 1), 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(out->n_aln, rs->d_out_n, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    if (out->n_pops) HIP_OK(hipMemcpyAsync(out->n_pops, rs->d_pops, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    if (out->trace) HIP_OK(hipMemcpyAsync(out->trace, rs->d_trace, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    if (out->occ_bytes) HIP_OK(hipMemcpyAsync(out->occ_bytes, rs->d_bytes, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    out->n_rec = need;
    if (need > out->rec_cap || (need > 0 && !out->rec))
        return fail(GALN_HIP_E_RECCAP, "record array holds %llu, %llu needed", (unsigned long long)out->rec_cap,
                    (unsigned long long)need);
    if (int g = grow_dev((void **)&rs->d_stage, rs->stage_bytes, std::max<uint64_t>(need, 1) * sizeof(galn_aln_rec_t), s))
        return g;
    hipLaunchKernelGGL(k_gather_records, dim3((n + 255) / 256), dim3(256), 0, s, rs->d_out_n, rs->d_out_off,
                       (const uint64_t *)rs->d_incl, (const galn_aln_rec_t *)rs->d_rec, rs->d_stage, rs->d_roff, n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out->rec_off, rs->d_roff, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    if (need) HIP_OK(hipMemcpyAsync(out->rec, rs->d_stage, need * sizeof(galn_aln_rec_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return 0;
}

int galn_hip_readset_kernel_ms(const galn_hip_readset *rs, double *first_tier_ms, double *total_ms)
{
    if (!rs) return fail(GALN_HIP_E_ARG, "readset_kernel_ms: null handle");
    if (first_tier_ms) *first_tier_ms = rs->first_ms;
    if (total_ms) *total_ms = rs->total_ms;
    return 0;
}

int galn_hip_readset_engine_stats(galn_hip_readset *rs, uint64_t out[16])
{
    if (!rs || !out) return fail(GALN_HIP_E_ARG, "readset_engine_stats: null argument");
    HIP_OK(hipMemcpy(out, rs->d_stats, 16 * 8, hipMemcpyDeviceToHost));
    return 0;
}

int galn_hip_readset_launches(const galn_hip_readset *rs, int cap, int32_t *tier, double *ms, uint64_t *algo_bytes)
{
    if (!rs) return fail(GALN_HIP_E_ARG, "readset_launches: null handle");
    const int n = (int)rs->launches.size();
    for (int i = 0; i < n && i < cap; ++i) {
        if (tier) tier[i] = rs->launches[i].tier;
        if (ms) ms[i] = rs->launches[i].ms;
        if (algo_bytes) algo_bytes[i] = rs->launches[i].bytes;
    }
    return n;
}

static int search_once(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_out *out, int form, uint32_t flags,
                       const galn_hip_tune_t *tune)
{
    galn_hip_readset *rs = nullptr;
    int r = galn_hip_readset_upload(idx, reads, &rs);
    if (r) return r;
    r = galn_hip_readset_search(idx, rs, form, flags, tune);
    if (!r) r = galn_hip_readset_results(rs, out);
    galn_hip_readset_free(rs);
    return r;
}

int galn_hip_search_plain(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_out *out, uint32_t flags)
{
    return search_once(idx, reads, out, 0, flags, nullptr);
}

int galn_hip_search_batched(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_out *out, uint32_t flags,
                            const galn_hip_tune_t *tune)
{
    return search_once(idx, reads, out, 1, flags, tune);
}

}  // extern "C"

extern "C" {

int galn_hip_index_set_sa(galn_hip_index *idx, const uint64_t *samples, uint64_t n_samples, uint32_t intv)
{
    if (!idx || (!samples && n_samples)) return fail(GALN_HIP_E_ARG, "index_set_sa: null argument");
    int r = busy_check(idx, "index_set_sa");
    if (r) return r;
    if (intv == 0 || (intv & (intv - 1)) != 0) return fail(GALN_HIP_E_ARG, "index_set_sa: interval %u is not a power of 2", intv);
    const uint64_t want = (idx->hmeta.seq_len + intv) / intv - 1;
    if (n_samples != want)
        return fail(GALN_HIP_E_ARG, "index_set_sa: %llu samples, %llu expected for interval %u",
                    (unsigned long long)n_samples, (unsigned long long)want, intv);
    HIP_OK(hipSetDevice(idx->device));
    if (idx->d_sa) HIP_OK(hipFree(idx->d_sa));
    idx->d_sa = nullptr;
    idx->n_sa = 0;
    HIP_OK(hipMalloc(&idx->d_sa, std::max<uint64_t>(n_samples, 1) * 8));
    if (n_samples) HIP_OK(hipMemcpy(idx->d_sa, samples, n_samples * 8, hipMemcpyHostToDevice));
    idx->n_sa = n_samples;
    idx->sa_intv = intv;
    return 0;
}

int galn_hip_sa_positions(galn_hip_index *idx, const uint64_t *rows, uint64_t n, uint64_t *pos, const int32_t *ref_len,
                          int32_t *strand)
{
    if (!idx || (n && (!rows || !pos)) || (!ref_len != !strand))
        return fail(GALN_HIP_E_ARG, "sa_positions: null argument (ref_len and strand come together)");
    int r = busy_check(idx, "sa_positions");
    if (r) return r;
    if (!idx->sa_intv) return fail(GALN_HIP_E_ARG, "sa_positions: no SA samples (galn_hip_index_set_sa)");
    for (uint64_t i = 0; i < n; ++i)
        if (rows[i] > idx->hmeta.seq_len)
            return fail(GALN_HIP_E_ARG, "sa_positions: row %llu beyond %llu", (unsigned long long)rows[i],
                        (unsigned long long)idx->hmeta.seq_len);
    if (n == 0) return 0;
    HIP_OK(hipSetDevice(idx->device));
    const uint64_t chunk = std::min<uint64_t>(n, 1ull << 24);
    uint64_t *d = nullptr;
    HIP_OK(hipMalloc(&d, chunk * (ref_len ? 24 : 16)));
    int32_t *const d_rl = ref_len ? reinterpret_cast<int32_t *>(d + 2 * chunk) : nullptr;
    int32_t *const d_st = ref_len ? d_rl + chunk : nullptr;
    hipStream_t s = idx->ctx0.stream;
    int rc = 0;
    for (uint64_t a = 0; a < n && !rc; a += chunk) {
        const uint64_t m = std::min(chunk, n - a);
        if (hipMemcpyAsync(d, rows + a, m * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
            (ref_len && hipMemcpyAsync(d_rl, ref_len + a, m * 4, hipMemcpyHostToDevice, s) != hipSuccess)) {
            rc = fail(GALN_HIP_E_DEVICE, "sa_positions: copy failed");
            break;
        }
        hipLaunchKernelGGL(k_sa_walk, dim3((unsigned)std::min<uint64_t>((m + 255) / 256, 65535)), dim3(256), 0, s,
                           (const uint32_t *)idx->d_bwt, idx->meta, (const uint64_t *)idx->d_sa, (uint64_t)idx->sa_intv,
                           (const uint64_t *)d, m, d + chunk, (const int32_t *)d_rl, d_st);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(pos + a, d + chunk, m * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
            (ref_len && hipMemcpyAsync(strand + a, d_st, m * 4, hipMemcpyDeviceToHost, s) != hipSuccess) ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = fail(GALN_HIP_E_DEVICE, "sa_positions: kernel or copy failed");
    }
    (void)hipFree(d);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Search without allocation (galn_hip_search_device_noalloc). Every buffer the search needs
// beyond the caller's inputs and outputs -- counters, failure lists, per-batch options, the
// lanes' regions (widths, record slots, granule lists) and the stack pool -- is carved out of
// one static device array, `scratch`, which each GPU holds from the library's first use on it
// until the process ends. Host code never passes the array's name to a kernel: its device
// address on each GPU comes from hipGetSymbolAddress, taken once per device (with that device
// current) when the first index is opened there.

namespace {
constexpr size_t NA_SCRATCH_BYTES = 256ull << 20;  // the array's size
constexpr int NA_MAX_READS = 1 << 20;              // reads per call
constexpr int NA_MAX_BATCHES = 4096;               // per-batch option sets per call
constexpr int NA_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

namespace {

uint8_t *g_na_ptr[NA_MAX_DEVICES];  // device address of `scratch` on each device
std::mutex g_na_setup_mu;
std::mutex g_na_mu[NA_MAX_DEVICES];  // calls on one device take turns (they share its array)

// Looks up the device address of `scratch` on `device`, which must be current. Idempotent.
int noalloc_setup(int device)
{
    if (device < 0 || device >= NA_MAX_DEVICES) return fail(GALN_HIP_E_DEVICE, "device %d beyond %d", device, NA_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_na_setup_mu);
    if (g_na_ptr[device]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return fail(GALN_HIP_E_DEVICE, "scratch array of device %d: hipGetSymbolAddress: %s", device,
                    hipGetErrorString(e));
    }
    g_na_ptr[device] = static_cast<uint8_t *>(p);
    return 0;
}

// Longest read (capped at MAXLEN + 1) and the number of reads whose batch index is out of range.
__global__ void k_na_scan(const int32_t *qlen, const int32_t *batch, int n, int n_batches, int32_t *res)
{
    int mx = 0, bad = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        mx = max(mx, min(qlen[i], MAXLEN + 1));
        bad += (batch[i] < 0 || batch[i] >= n_batches) ? 1 : 0;
    }
    if (mx > 0) atomicMax(res, mx);
    if (bad) atomicAdd(res + 1, bad);
}

// Offsets inside the array.
struct NaLayout {
    size_t ctr, top, algo, scan, ghead, gbump, bopts, nstacks, fail0, fail1, pops, lanes;
};

NaLayout na_layout()
{
    NaLayout L{};
    L.ctr = 0;      // 8 x i32: [0],[1] failure counts, [2] claim counter, [3] unused, [4..6] failure kinds
    L.top = 64;     // u64 records reserved in the caller's array
    L.algo = 72;    // u64 algorithmic index bytes
    L.scan = 80;    // 2 x i32 (k_na_scan)
    L.ghead = 256;  // NSHARD x u64 free-granule stack heads
    L.gbump = L.ghead + NSHARD * 8;
    size_t off = align_up(L.gbump + 4, 256);
    L.bopts = off;
    off = align_up(off + (size_t)NA_MAX_BATCHES * sizeof(galn_search_opt_t), 256);
    L.nstacks = off;
    off = align_up(off + (size_t)NA_MAX_BATCHES * 4, 256);
    L.fail0 = off;
    off += (size_t)NA_MAX_READS * 4;
    L.fail1 = off;
    off += (size_t)NA_MAX_READS * 4;
    L.pops = off;  // pop counts when the caller does not ask for them
    off += (size_t)NA_MAX_READS * 8;
    L.lanes = align_up(off, 256);
    return L;
}

// Bytes of one pool granule with its chunk links and free-stack link.
constexpr size_t NA_GRAN_BYTES = (size_t)GRAN * CH * sizeof(Ent) + (size_t)GRAN * 4 + 4;

// Lanes that fit the array beside the pool with their region of `slot_bytes`: at most `want`,
// each with at least two granules of pool. Returns 0 if not even one fits.
int na_lanes(size_t slot_bytes, int want, uint64_t *n_gran)
{
    const size_t avail = NA_SCRATCH_BYTES - na_layout().lanes - 3 * 256;
    for (int lanes = std::max(1, want);; lanes = lanes > 1 ? lanes / 2 : 0) {
        if (lanes == 0) return 0;
        const size_t region = align_up((size_t)lanes * slot_bytes, 256);
        if (region >= avail) continue;
        const uint64_t g = std::min<uint64_t>((avail - region) / NA_GRAN_BYTES, MAX_POOL_CHUNKS / GRAN);
        if (g >= 2ull * (uint64_t)lanes) {
            *n_gran = g;
            return lanes;
        }
    }
}

int noalloc_search(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_out *out, uint32_t flags)
{
    if (!idx || !reads || !out) return fail(GALN_HIP_E_ARG, "search_device_noalloc: null argument");
    const int n = reads->n_reads;
    if (n < 0 || n > NA_MAX_READS)
        return fail(GALN_HIP_E_ARG, "search_device_noalloc: %d reads (at most %d per call)", n, NA_MAX_READS);
    if (reads->n_batches < 1 || reads->n_batches > NA_MAX_BATCHES || !reads->bopts || !reads->n_stacks)
        return fail(GALN_HIP_E_ARG, "search_device_noalloc: %d batches (1 to %d, with host option arrays)",
                    reads->n_batches, NA_MAX_BATCHES);
    for (int b = 0; b < reads->n_batches; ++b)
        if (reads->bopts[b].seed_len < 0) return fail(GALN_HIP_E_UNSUPPORTED, "reads: negative seed length");
    const bool trace = (flags & GALN_HIP_TRACE) != 0;
    if (n > 0 && (!reads->query || !reads->qoff || !reads->qlen || !reads->max_diff || !reads->batch ||
                  !out->n_aln || !out->rec_off || (out->rec_cap > 0 && !out->rec) || (trace && !out->trace)))
        return fail(GALN_HIP_E_ARG, "search_device_noalloc: null array");
    int r = busy_check(idx, "search_device_noalloc");
    if (r) return r;
    HIP_OK(hipSetDevice(idx->device));
    if ((r = noalloc_setup(idx->device))) return r;
    std::lock_guard<std::mutex> lk(g_na_mu[idx->device]);
    out->n_rec = 0;
    if (n == 0) return 0;

    uint8_t *const S = g_na_ptr[idx->device];
    const NaLayout NL = na_layout();
    int32_t *const ctr = reinterpret_cast<int32_t *>(S + NL.ctr);
    unsigned long long *const top = reinterpret_cast<unsigned long long *>(S + NL.top);
    unsigned long long *const algo = reinterpret_cast<unsigned long long *>(S + NL.algo);
    int32_t *const scan = reinterpret_cast<int32_t *>(S + NL.scan);
    unsigned long long *const ghead = reinterpret_cast<unsigned long long *>(S + NL.ghead);
    uint32_t *const gbump = reinterpret_cast<uint32_t *>(S + NL.gbump);
    galn_search_opt_t *const d_bopts = reinterpret_cast<galn_search_opt_t *>(S + NL.bopts);
    int32_t *const d_nstacks = reinterpret_cast<int32_t *>(S + NL.nstacks);
    int32_t *const d_fail[2] = {reinterpret_cast<int32_t *>(S + NL.fail0), reinterpret_cast<int32_t *>(S + NL.fail1)};
    uint64_t *const d_pops = out->n_pops ? out->n_pops : reinterpret_cast<uint64_t *>(S + NL.pops);
    hipStream_t st = idx->ctx0.stream, fst = st;

    // per-batch options in, counters cleared (fill kernels: the array keeps the last call's contents)
    const int nb = reads->n_batches;
    HIP_OK(hipMemcpyAsync(d_bopts, reads->bopts, (size_t)nb * sizeof(galn_search_opt_t), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d_nstacks, reads->n_stacks, (size_t)nb * 4, hipMemcpyHostToDevice, st));
    FILL(S, 0, 256);  // counters, record top, byte count, scan results
    FILL(out->rec_off, 0, (size_t)n * 8);
    hipLaunchKernelGGL(k_na_scan, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, st, reads->qlen, reads->batch, n,
                       nb, scan);
    HIP_OK(hipGetLastError());
    int32_t hs[2] = {0, 0};
    HIP_OK(hipMemcpyAsync(hs, scan, sizeof hs, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (hs[1] > 0) return fail(GALN_HIP_E_ARG, "reads: %d batch indexes out of range", hs[1]);

    // what readset_fill derives from the reads
    const int max_len = std::max(1, std::min(hs[0], MAXLEN));
    int seed_len = 0, max_stacks = 1, max_entries = 0;
    bool nonneg = true;
    for (int b = 0; b < nb; ++b) {
        const galn_search_opt_t &o = reads->bopts[b];
        if (o.seed_len < max_len) seed_len = std::max(seed_len, o.seed_len);
        max_stacks = std::max(max_stacks, reads->n_stacks[b]);
        max_entries = std::max(max_entries, std::max(o.max_entries, 0));
        nonneg = nonneg && o.s_mm >= 0 && o.s_gapo >= 0 && o.s_gape >= 0;
    }
    const int buckets = std::max(1, std::min(max_stacks, MAXB));
    const int block = buckets * 128 * 4 <= 65536 ? 128 : 64;
    const int own_full = (int)(((int64_t)(max_entries + 16) / CH + MAXB + 2) / GRAN + 2);

    KArgs a{};
    a.bwt = idx->d_bwt;
    a.meta = idx->meta;
    a.query = reads->query;
    a.qoff = reads->qoff;
    a.qlen = reads->qlen;
    a.maxdiff = reads->max_diff;
    a.batch = reads->batch;
    a.bopts = d_bopts;
    a.nstacks = d_nstacks;
    a.qnext = ctr + 2;
    a.fail_kind = ctr + 4;
    a.algo_bytes = algo;
    a.out_n = out->n_aln;
    a.out_off = out->rec_off;
    a.out_pops = d_pops;
    a.out_trace = out->trace;
    a.out_bytes = out->occ_bytes;
    a.out_rec = out->rec;
    a.out_cap = out->rec_cap;
    a.out_top = top;
    a.ghead = ghead;
    a.gbump = gbump;
    a.fuse = ((flags & GALN_HIP_NOFUSE) || !nonneg) ? 0 : 1;
    a.heavy_n = ctr + 3;  // no pop budget: nothing is handed on

    int rec_cap = 64, want = std::min(n, 4096), fi = 0;
    int32_t n_todo = n;
    bool out_full = false;
    for (int round = 0; n_todo > 0; ++round) {
        if (round >= 24) return fail(GALN_HIP_E_NOMEM, "search_device_noalloc: reads still failing after %d re-runs", round);
        const Layout L0 = make_layout(max_len, seed_len, rec_cap, own_full);
        uint64_t n_gran = 0;
        const int lanes = na_lanes(L0.slot_bytes, std::min(want, (int)n_todo), &n_gran);
        if (lanes == 0)
            return fail(GALN_HIP_E_UNSUPPORTED, "search_device_noalloc: a lane's %llu bytes and its pool do not fit in %zu bytes",
                        (unsigned long long)L0.slot_bytes, NA_SCRATCH_BYTES);
        const int own_cap = (int)std::min<uint64_t>((uint64_t)own_full, n_gran);
        const Layout L = make_layout(max_len, seed_len, rec_cap, own_cap);
        uint8_t *const lanes_base = S + NL.lanes;
        uint8_t *const pool = S + align_up(NL.lanes + (size_t)lanes * L.slot_bytes, 256);
        a.scratch = lanes_base;
        a.slot_bytes = L.slot_bytes;
        a.sw_off = L.sw_off;
        a.rec_off = L.rec_off;
        a.own_off = L.own_off;
        a.own_cap = L.own_cap;
        a.keep = 1;
        a.rec_cap = rec_cap;
        a.pool_chk = reinterpret_cast<Ent *>(pool);
        a.pool_link = reinterpret_cast<uint32_t *>(pool + align_up(n_gran * GRAN * CH * sizeof(Ent), 256));
        a.gnext = reinterpret_cast<uint32_t *>(reinterpret_cast<uint8_t *>(a.pool_link) + align_up(n_gran * GRAN * 4, 256));
        a.n_gran = (uint32_t)n_gran;
        a.list = round == 0 ? nullptr : d_fail[fi];
        a.list_base = 0;
        a.n_list = n_todo;
        a.n_slots = lanes;
        const int fo = round == 0 ? 0 : fi ^ 1;
        a.fail_list = d_fail[fo];
        a.fail_n = ctr + fo;
        FILL(ctr + fo, 0, sizeof(int32_t));
        FILL(ctr + 2, 0, 2 * sizeof(int32_t));
        FILL(ctr + 4, 0, 3 * sizeof(int32_t));
        FILL(ghead, 0xffffffffu, NSHARD * 8);
        FILL(gbump, 0, 4);
        r = launch(st, a, block, (lanes + block - 1) / block, true, trace, buckets);
        if (r) return r;
        int32_t c[8];
        HIP_OK(hipMemcpyAsync(c, ctr, sizeof c, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (c[4]) {  // the caller's record array is full: stop, and report what was reserved
            out_full = true;
            break;
        }
        if (c[5]) rec_cap *= 16;
        if (c[6]) want = std::max(1, lanes / 4);
        fi = fo;
        n_todo = c[fo];
    }
    unsigned long long t = 0;
    HIP_OK(hipMemcpyAsync(&t, top, sizeof t, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    out->n_rec = t;
    if (out_full)
        return fail(GALN_HIP_E_RECCAP, "search_device_noalloc: record array holds %llu; at least %llu needed",
                    (unsigned long long)out->rec_cap, t);
    return 0;
}

}  // namespace

extern "C" int galn_hip_search_device_noalloc(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_out *out,
                                              uint32_t flags)
{
    return noalloc_search(idx, reads, out, flags);
}
