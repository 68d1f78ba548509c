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
//     chunks of 16-byte packed entries in the lane's global arena; the top
//     of each bucket (chunk id, fill) lives in LDS, the non-empty-bucket bitmask
//     in registers, so a pop is one dependent global load and a push is none;
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
// Reads that outgrow a lane's first-tier arena (stack or record space) are
// re-run from scratch in larger tiers on the device, so results never depend
// on the capacities chosen (the reference's contract, src/bwtgap_hd.h:10-22).

#include <hip/hip_runtime.h>

#include <algorithm>
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

constexpr int MODE_GAPE = 0x01, MODE_LOGGAP = 0x04, MODE_NONSTOP = 0x10;
constexpr int ST_M = 0, ST_I = 1, ST_D = 2;

constexpr uint64_t TR_INIT = 1469598103934665603ull;
constexpr uint64_t TR_PRIME = 1099511628211ull;

struct DMeta {
    uint64_t primary;
    uint64_t L2[5];
    uint64_t seq_len;
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
    int32_t seed_len;
    // work: list positions [0, n_list); read id = list ? list[pos] : list_base + pos
    const int32_t *list;
    int32_t list_base;
    int32_t n_list;
    int32_t *qnext;          // persistent claims
    int32_t n_slots;
    // per-lane arena
    uint8_t *scratch;
    uint64_t slot_bytes;
    uint32_t sw_off, rec_off, link_off, chunk_off;
    int32_t n_chunks;
    int32_t rec_cap;
    // results (indexed by read id)
    int32_t *out_n;
    uint64_t *out_off;
    uint64_t *out_pops;
    uint64_t *out_trace;
    galn_aln_rec_t *out_rec;
    uint64_t out_cap;
    unsigned long long *out_top;
    int32_t *fail_list;
    int32_t *fail_n;
    int32_t *fail_kind;      // [0] output full, [1] record slots, [2] stack arena
    int32_t fuse;
};

// ---------------------------------------------------------------------------
// Occurrence counts. Index layout: 128-symbol blocks of 16 u32 words, 4 x u64
// exclusive counts then 8 words of 16 two-bit symbols (first symbol in bits
// 31-30); position p of the $-inclusive BWT is at p - (p >= primary).

__device__ __forceinline__ void load_block(const uint32_t *__restrict__ bwt, uint64_t blk, uint4 (&b)[4])
{
    const uint4 *p = reinterpret_cast<const uint4 *>(bwt + (blk << 4));
    b[0] = p[0];
    b[1] = p[1];
    b[2] = p[2];
    b[3] = p[3];
}

// counts of all 4 bases among symbols 0..r of a loaded block, plus its prefix counts
__device__ __forceinline__ void block_occ4(const uint4 (&b)[4], int r, uint64_t c[4])
{
    const uint32_t w[8] = {b[2].x, b[2].y, b[2].z, b[2].w, b[3].x, b[3].y, b[3].z, b[3].w};
    const int last = r >> 4;
    const uint32_t part = 0x55555555u & ~((1u << (30 - 2 * (r & 15))) - 1u);
    uint32_t n1 = 0, n2 = 0, n3 = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t valid = j < last ? 0x55555555u : (j == last ? part : 0u);
        const uint32_t hi = (w[j] >> 1) & valid, lo = w[j] & valid;
        n3 += __builtin_popcount(hi & lo);
        n2 += __builtin_popcount(hi & ~lo);
        n1 += __builtin_popcount(lo & ~hi);
    }
    const uint32_t n0 = (uint32_t)(r + 1) - n1 - n2 - n3;
    c[0] = ((uint64_t)b[0].y << 32 | b[0].x) + n0;
    c[1] = ((uint64_t)b[0].w << 32 | b[0].z) + n1;
    c[2] = ((uint64_t)b[1].y << 32 | b[1].x) + n2;
    c[3] = ((uint64_t)b[1].w << 32 | b[1].z) + n3;
}

// Occ of all four bases at positions a (may be -1) and b >= 0: bwt_2occ4, src/bwtgap_hd.h:158-192.
// One 64-byte block when both fall in it, else two independent loads issued together.
__device__ __forceinline__ void lf_pair(const uint32_t *__restrict__ bwt, const DMeta &m, uint64_t a, uint64_t b,
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
    if (aneg) {
        ca[0] = ca[1] = ca[2] = ca[3] = 0;
    } else if (same) {
        block_occ4(B, (int)(ap & 127), ca);
    } else {
        block_occ4(A, (int)(ap & 127), ca);
    }
}

// ---------------------------------------------------------------------------
// Packed stack entry (16 bytes); the score is the bucket.
//   w0 = k[0:38) | i[38:48) | n_mm[48:55) | n_gapo[55:62) | state[62:64)
//   w1 = l[0:38) | last_diff_pos[38:48) | n_gape[48:56) | n_ins[56:64)
// n_del = n_gapo + n_gape - n_ins (every gap open or extension adds one
// inserted or one deleted base, src/bwtgap_hd.h:430-454).

struct __align__(16) Ent {
    uint64_t w0, w1;
};

__device__ __forceinline__ Ent mk_ent(uint64_t k, uint64_t l, int i, int ldp, int mm, int go, int ge, int ins, int st)
{
    Ent e;
    e.w0 = (k & KMASK) | ((uint64_t)i << 38) | ((uint64_t)mm << 48) | ((uint64_t)go << 55) | ((uint64_t)st << 62);
    e.w1 = (l & KMASK) | ((uint64_t)ldp << 38) | ((uint64_t)ge << 48) | ((uint64_t)ins << 56);
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

enum : int { PH_CW = 0, PH_POPPED = 1, PH_EXACT = 2 };

template <bool TRACE, bool PERSIST, typename TopT>
__global__ void __launch_bounds__(128) k_search(KArgs a)
{
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.n_slots) return;
    TopT *const tops = reinterpret_cast<TopT *>(lds_raw) + threadIdx.x;
    const int tstride = blockDim.x;

    uint8_t *const base = a.scratch + (uint64_t)slot * a.slot_bytes;
    uint64_t *const W = reinterpret_cast<uint64_t *>(base);
    uint64_t *const SW = reinterpret_cast<uint64_t *>(base + a.sw_off);
    galn_aln_rec_t *const REC = reinterpret_cast<galn_aln_rec_t *>(base + a.rec_off);
    uint32_t *const LINK = reinterpret_cast<uint32_t *>(base + a.link_off);
    Ent *const CHK = reinterpret_cast<Ent *>(base + a.chunk_off);
    const DMeta &M = a.meta;
    const uint64_t seq_len = M.seq_len;

    // ---- lane state ----
    int rid = 0, len = 0;
    const uint8_t *q = nullptr;
    int s_mm = 0, s_gapo = 0, s_gape = 0, mode = 0, ies = 0, mdo = 0, max_entries = 0, max_gapo = 0, max_gape = 0,
        msd = 0, max_top2 = 0, n_stacks = 0, md_read = 0;
    bool seeding = false;
    int seed_len = a.seed_len;
    // chain (widths / exact tail)
    int phase = PH_CW, cj = 0, plen = 0, cbid = 0, cbase = 0;
    const uint8_t *pstr = nullptr;
    uint64_t *wout = nullptr;
    bool in_seed = false;
    uint64_t ck = 0, cl = 0;
    // search
    int best_score = 0, best_diff = 0, max_diff = 0, best_cnt = 0, n_aln = 0, n_entries = 0;
    uint64_t bm0 = 0, bm1 = 0, bm2 = 0, bm3 = 0;
    bool has_cur = false;
    Ent cur{0, 0}, e{0, 0};
    int cur_score = 0, e_score = 0, m = 0, m_seed = 0;
    uint32_t spare = NIL, free_head = NIL, bump = 0;
    uint64_t pops = 0, trace = TR_INIT;
    // LF request
    bool need_lf = false;
    uint64_t qa = 0, qb = 0;
    bool need_claim = true, claimed = false;

    // ---- helpers (all lane-local) ----
    auto bm_test = [&](int s) -> bool {
        const uint64_t w = s < 64 ? bm0 : s < 128 ? bm1 : s < 192 ? bm2 : bm3;
        return (w >> (s & 63)) & 1ull;
    };
    auto bm_set = [&](int s) {
        const uint64_t bit = 1ull << (s & 63);
        const int w = s >> 6;
        bm0 |= w == 0 ? bit : 0ull;
        bm1 |= w == 1 ? bit : 0ull;
        bm2 |= w == 2 ? bit : 0ull;
        bm3 |= w == 3 ? bit : 0ull;
    };
    auto bm_clear = [&](int s) {
        const uint64_t bit = ~(1ull << (s & 63));
        const int w = s >> 6;
        bm0 &= w == 0 ? bit : ~0ull;
        bm1 &= w == 1 ? bit : ~0ull;
        bm2 &= w == 2 ? bit : ~0ull;
        bm3 &= w == 3 ? bit : ~0ull;
    };
    auto bm_low = [&]() -> int {
        return bm0 ? __builtin_ctzll(bm0)
                   : bm1 ? 64 + __builtin_ctzll(bm1) : bm2 ? 128 + __builtin_ctzll(bm2) : 192 + __builtin_ctzll(bm3);
    };
    auto chunk_alloc = [&]() -> uint32_t {
        uint32_t c;
        if (spare != NIL) {
            c = spare;
            spare = NIL;
        } else if (free_head != NIL) {
            c = free_head;
            free_head = LINK[c];
        } else if (bump < (uint32_t)a.n_chunks) {
            c = bump++;
        } else {
            c = NIL;
        }
        return c;
    };
    auto chunk_free = [&](uint32_t c) {
        if (spare == NIL) {
            spare = c;
        } else {
            LINK[c] = free_head;
            free_head = c;
        }
    };
    // gap_push into the arena: src/bwtgap_hd.h:264-286
    auto push_mem = [&](int s, const Ent &en) -> int {
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
    // a child: the exact one (same score, pushed last) is kept in registers
    auto push = [&](int s, const Ent &en, bool exact) -> int {
        if (exact && a.fuse) {
            cur = en;
            cur_score = s;
            has_cur = true;
            ++n_entries;
            return 0;
        }
        return push_mem(s, en);
    };
    auto score_of = [&](int mm, int go, int ge) -> int { return mm * s_mm + go * s_gapo + ge * s_gape; };

    // result of the current read; the lane then claims the next one
    auto finish = [&](int code) {
        int n = code;
        if (n > 0) {
            const unsigned long long b0 = atomicAdd(a.out_top, (unsigned long long)n);
            if (b0 + (unsigned long long)n > a.out_cap) {
                n = ERR_OUT;
            } else {
                for (int j = 0; j < n; ++j) a.out_rec[b0 + j] = REC[j];
                a.out_off[rid] = b0;
            }
        }
        if (n == ERR_STACK || n == ERR_ALN || n == ERR_OUT) {
            const int f = atomicAdd(a.fail_n, 1);
            a.fail_list[f] = rid;
            atomicOr(a.fail_kind + (n == ERR_OUT ? 0 : n == ERR_ALN ? 1 : 2), 1);
        }
        a.out_n[rid] = n;
        a.out_pops[rid] = pops;
        if (TRACE) a.out_trace[rid] = trace;
        need_lf = false;
        need_claim = true;
    };

    // gap_shadow: src/bwtgap_hd.h:309-320
    auto shadow = [&](int x, int ldp) {
        const uint64_t ux = (uint64_t)(int64_t)x;
        int j = 0;
        for (int i = 0; i < ldp; ++i) {
            const uint64_t wv = W[i];
            uint64_t w = wval(wv);
            int bid = wbid(wv);
            if (w > ux) {
                w -= ux;
            } else if (w == ux) {
                bid = 1;
                w = seq_len - (uint64_t)(++j);
            } else {
                continue;
            }
            W[i] = wpack(w, bid);
        }
    };

    // a hit (src/bwtgap_hd.h:385-409); returns true when the search stops here
    auto hit = [&](uint64_t k, uint64_t l) -> bool {
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
            finish(n_aln);
            return true;
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
            if (n_aln == a.rec_cap) {
                finish(ERR_ALN);
                return true;
            }
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
        return false;
    };

    // the top of bwt_match_gap's loop: pop entries until one needs an LF lookup
    // or the search ends (src/bwtgap_hd.h:357-383)
    auto advance = [&]() {
        for (;;) {
            if (n_entries == 0 || n_entries > max_entries) {
                finish(n_aln);
                return;
            }
            if (has_cur) {
                e = cur;
                e_score = cur_score;
                has_cur = false;
            } else {
                if (!(bm0 | bm1 | bm2 | bm3)) {  // entries counted but none held: never expected
                    finish(ERR_INTERNAL);
                    return;
                }
                const int s = bm_low();
                TopT *tp = tops + s * tstride;
                const uint32_t t = *tp;
                const uint32_t chunk = t >> 4, idx = t & 15u;
                e = CHK[(uint64_t)chunk * CH + idx];
                e_score = s;
                if (idx > 0) {
                    *tp = (TopT)(t - 1);
                } else {
                    const uint32_t p = LINK[chunk];
                    chunk_free(chunk);
                    if (p != NIL)
                        *tp = (TopT)(p << 4 | (CH - 1));
                    else
                        bm_clear(s);
                }
            }
            --n_entries;
            ++pops;
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
            if (!(mode & MODE_NONSTOP) && (uint32_t)e_score > (uint32_t)(best_score + s_mm)) {
                finish(n_aln);
                return;
            }
            m = max_diff - (mm + go);
            if (mode & MODE_GAPE) m -= ge;
            if (m < 0) continue;
            if (seeding) {
                m_seed = msd - (mm + go);
                if (mode & MODE_GAPE) m_seed -= ge;
            }
            if (i0 == 0) {
                if (hit(e_k(e), e_l(e))) return;
                continue;
            }
            // the next LF lookup is at (k-1, l) whether the entry expands or
            // starts an exact tail: issue it with the width loads
            qa = e_k(e) - 1;
            qb = e_l(e);
            need_lf = true;
            phase = PH_POPPED;
            return;
        }
    };

    auto start_search = [&]() {
        int nN = 0;
        for (int j = 0; j < len; ++j) nN += q[j] > 3;
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
        bump = 0;
        has_cur = false;
        if (nN > md_read) {
            finish(0);
            return;
        }
        // root: (i = len, k = 0, l = seq_len), score 0
        const int r = push(0, mk_ent(0, seq_len, len, 0, 0, 0, 0, 0, ST_M), n_stacks > 0);
        if (r) {
            finish(r);
            return;
        }
        advance();
    };

    // widths: step to the next position that needs an LF lookup (N restarts inline)
    auto prep_width = [&]() {
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
                    continue;
                }
                start_search();
                return;
            }
            const int c0 = pstr[cj];
            const int c = c0 > 3 ? 4 : 3 - c0;
            if (c < 4) {
                cbase = c;
                qa = ck - 1;
                qb = cl;
                need_lf = true;
                phase = PH_CW;
                return;
            }
            ck = 0;
            cl = seq_len;
            ++cbid;
            wout[cj] = wpack(seq_len + 1, cbid);
            ++cj;
        }
    };

    auto init_read = [&](int pos) {
        rid = a.list ? a.list[pos] : a.list_base + pos;
        len = a.qlen[rid];
        q = a.query + a.qoff[rid];
        const int b = a.batch[rid];
        const galn_search_opt_t o = a.bopts[b];
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
        n_stacks = a.nstacks[b];
        md_read = a.maxdiff[rid];
        seeding = seed_len < len;
        pops = 0;
        trace = TR_INIT;
        if (len < 1 || len > MAXLEN || n_stacks > MAXB || md_read > MAXDIFF || max_gapo > MAXDIFF ||
            max_gapo + max_gape > 255 || max_gape < 0) {
            finish(ERR_UNSUP);
            return;
        }
        in_seed = false;
        pstr = q;
        plen = len;
        wout = W;
        cj = 0;
        ck = 0;
        cl = seq_len;
        cbid = 0;
        prep_width();
    };

    // ---- main loop: one LF lookup per lane per iteration ----
    for (;;) {
        if (need_claim) {
            int pos;
            if (PERSIST) {
                pos = atomicAdd(a.qnext, 1);
            } else {
                if (claimed) break;
                pos = slot;
            }
            claimed = true;
            if (pos >= a.n_list) break;
            need_claim = false;
            init_read(pos);
            continue;
        }

        // the popped entry's widths and query base are loaded with its LF lookup
        uint64_t wA = 0, wB = 0, sA = 0, sB = 0;
        int qbase = 0;
        if (phase == PH_POPPED) {
            const int i0 = e_i(e);
            wB = W[i0 - 1];
            if (i0 >= 2) wA = W[i0 - 2];
            qbase = q[i0 - 1];
            if (seeding) {
                const int ii = i0 - 1 - (len - seed_len);
                if (ii > 0) {
                    sA = SW[ii - 1];
                    sB = SW[ii];
                }
            }
        }
        uint64_t ca[4], cb[4];
        if (need_lf) lf_pair(a.bwt, M, qa, qb, ca, cb);
        need_lf = false;

        if (phase == PH_CW) {
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
            prep_width();
        } else if (phase == PH_EXACT) {
            // bwt_match_exact_alt step: src/bwtgap_hd.h:201-208
            const int c = cbase;
            const uint64_t k = M.L2[c] + ca[c] + 1, l = M.L2[c] + cb[c];
            if (k > l) {
                advance();
                continue;
            }
            --cj;
            if (cj < 0) {
                if ((int)(uint32_t)(l - k + 1) != 0) {
                    if (hit(k, l)) continue;
                }
                advance();
                continue;
            }
            const int c2 = q[cj];
            if (c2 > 3) {
                advance();
                continue;
            }
            cbase = c2;
            ck = k;
            cl = l;
            qa = k - 1;
            qb = l;
            need_lf = true;
        } else {  // PH_POPPED: src/bwtgap_hd.h:375-470
            const int i0 = e_i(e);
            if (m < wbid(wB)) {
                advance();
                continue;
            }
            const int st = e_st(e), ge = e_ge(e);
            if (m == 0 && (st == ST_M || (mode & MODE_GAPE) || ge == max_gape)) {
                // exact tail, first step from this lookup
                if (qbase > 3) {
                    advance();
                    continue;
                }
                const uint64_t k = M.L2[qbase] + ca[qbase] + 1, l = M.L2[qbase] + cb[qbase];
                if (k > l) {
                    advance();
                    continue;
                }
                cj = i0 - 2;
                if (cj < 0) {
                    if ((int)(uint32_t)(l - k + 1) != 0) {
                        if (hit(k, l)) continue;
                    }
                    advance();
                    continue;
                }
                const int c2 = q[cj];
                if (c2 > 3) {
                    advance();
                    continue;
                }
                cbase = c2;
                ck = k;
                cl = l;
                qa = k - 1;
                qb = l;
                need_lf = true;
                phase = PH_EXACT;
                continue;
            }
            // expansion
            const int i = i0 - 1;
            const int mm = e_mm(e), go = e_go(e), ins = e_ins(e), del = go + ge - ins;
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
            int err = 0;
            if (allow_diff && i >= ies + tmp && len - i >= ies + tmp) {
                if (st == ST_M) {
                    if (go < max_gapo) {
                        err = push(score_of(mm, go + 1, ge), mk_ent(k, l, i, i, mm, go + 1, ge, ins + 1, ST_I), false);
                        for (int j = 0; j < 4 && !err; ++j) {
                            const uint64_t kk = M.L2[j] + ca[j] + 1, ll = M.L2[j] + cb[j];
                            if (kk <= ll)
                                err = push(score_of(mm, go + 1, ge), mk_ent(kk, ll, i + 1, i + 1, mm, go + 1, ge, ins, ST_D),
                                           false);
                        }
                    }
                } else if (st == ST_I) {
                    if (ge < max_gape)
                        err = push(score_of(mm, go, ge + 1), mk_ent(k, l, i, i, mm, go, ge + 1, ins + 1, ST_I), false);
                } else {
                    if (ge < max_gape && (ge + go < max_diff || occ < (uint64_t)(int64_t)mdo)) {
                        for (int j = 0; j < 4 && !err; ++j) {
                            const uint64_t kk = M.L2[j] + ca[j] + 1, ll = M.L2[j] + cb[j];
                            if (kk <= ll)
                                err = push(score_of(mm, go, ge + 1), mk_ent(kk, ll, i + 1, i + 1, mm, go, ge + 1, ins, ST_D),
                                           false);
                        }
                    }
                }
            }
            (void)del;
            if (!err) {
                if (allow_diff && allow_M) {
                    for (int j = 1; j <= 4 && !err; ++j) {
                        const int c = (qbase + j) & 3;
                        const int is_mm = (j != 4 || qbase > 3);
                        const uint64_t kk = M.L2[c] + ca[c] + 1, ll = M.L2[c] + cb[c];
                        if (kk <= ll)
                            err = push(score_of(mm + is_mm, go, ge), mk_ent(kk, ll, i, is_mm ? i : 0, mm + is_mm, go, ge, ins, ST_M),
                                       !is_mm);
                    }
                } else if (qbase < 4) {
                    const int c = qbase;
                    const uint64_t kk = M.L2[c] + ca[c] + 1, ll = M.L2[c] + cb[c];
                    if (kk <= ll) err = push(e_score, mk_ent(kk, ll, i, 0, mm, go, ge, ins, ST_M), true);
                }
            }
            if (err) {
                finish(err);
                continue;
            }
            advance();
        }
    }
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

#define HIP_OK(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? GALN_HIP_E_NOMEM : GALN_HIP_E_DEVICE, "%s:%d %s: %s", \
                        __FILE__, __LINE__, #call, hipGetErrorString(e_));                             \
    } while (0)

struct Tier {
    int n_chunks;
    int rec_cap;
    bool wide;  // u32 bucket tops
};

}  // namespace

struct galn_hip_index {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t *d_bwt = nullptr;
    DMeta meta{};
    galn_bwt_meta_t hmeta{};
    int n_cu = 0;
    size_t lds_per_block_max = 65536;
};

struct galn_hip_readset {
    galn_hip_index *idx = nullptr;
    int n = 0;
    int n_batches = 0;
    int seed_len = 0;
    int max_len = 0;
    int max_stacks = 0;
    uint64_t qbytes = 0;
    // device inputs
    uint8_t *d_query = nullptr;
    uint64_t *d_qoff = nullptr;
    int32_t *d_qlen = nullptr, *d_maxdiff = nullptr, *d_batch = nullptr, *d_nstacks = nullptr;
    galn_search_opt_t *d_bopts = nullptr;
    int32_t *d_sorted = nullptr;   // read ids, longest first
    // device outputs
    int32_t *d_out_n = nullptr;
    uint64_t *d_out_off = nullptr, *d_pops = nullptr, *d_trace = nullptr;
    galn_aln_rec_t *d_rec = nullptr;
    uint64_t rec_cap = 0;
    unsigned long long *d_top = nullptr;
    int32_t *d_fail[2] = {nullptr, nullptr};
    int32_t *d_ctr = nullptr;      // [0] fail count A, [1] fail count B, [2] queue head
    // scratch arena (grown on demand)
    uint8_t *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    // host copies
    std::vector<int32_t> h_qlen, h_nstacks, h_batch;
    std::vector<galn_search_opt_t> h_bopts;
    // last search
    double first_ms = 0, total_ms = 0;
    uint32_t last_flags = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Layout {
    uint64_t slot_bytes;
    uint32_t sw_off, rec_off, link_off, chunk_off;
};

Layout make_layout(int max_len, int seed_len, int n_chunks, int rec_cap)
{
    Layout L{};
    size_t off = align_up((size_t)(max_len + 1) * 8, 256);
    L.sw_off = (uint32_t)off;
    const int sl = (seed_len > 0 && seed_len < max_len) ? seed_len : 0;
    off += align_up((size_t)(sl + 1) * 8, 256);
    L.rec_off = (uint32_t)off;
    off += align_up((size_t)rec_cap * sizeof(galn_aln_rec_t), 256);
    L.link_off = (uint32_t)off;
    off += align_up((size_t)n_chunks * 4, 256);
    L.chunk_off = (uint32_t)off;
    off += (size_t)n_chunks * CH * sizeof(Ent);
    off = align_up(off, 256);
    // an odd multiple of 256 bytes keeps lanes' arenas off the same cache sets
    if ((off / 256) % 2 == 0) off += 256;
    L.slot_bytes = off;
    return L;
}

int ensure_scratch(galn_hip_readset *rs, size_t bytes)
{
    if (rs->scratch_bytes >= bytes) return 0;
    if (rs->d_scratch) HIP_OK(hipFree(rs->d_scratch));
    rs->d_scratch = nullptr;
    rs->scratch_bytes = 0;
    HIP_OK(hipMalloc(&rs->d_scratch, bytes));
    rs->scratch_bytes = bytes;
    return 0;
}

template <bool TRACE, bool PERSIST, typename TopT>
int launch_one(galn_hip_index *idx, const KArgs &a, int block, int grid, int buckets)
{
    const size_t lds = (size_t)block * buckets * sizeof(TopT);
    hipLaunchKernelGGL((k_search<TRACE, PERSIST, TopT>), dim3(grid), dim3(block), lds, idx->stream, a);
    HIP_OK(hipGetLastError());
    return 0;
}

template <bool TRACE>
int launch_t(galn_hip_index *idx, const KArgs &a, int block, int grid, bool persist, bool wide, int buckets)
{
    if (persist) {
        return wide ? launch_one<TRACE, true, uint32_t>(idx, a, block, grid, buckets)
                    : launch_one<TRACE, true, uint16_t>(idx, a, block, grid, buckets);
    }
    return wide ? launch_one<TRACE, false, uint32_t>(idx, a, block, grid, buckets)
                : launch_one<TRACE, false, uint16_t>(idx, a, block, grid, buckets);
}

int launch(galn_hip_index *idx, const KArgs &a, int block, int grid, bool persist, bool wide, bool trace, int buckets)
{
    return trace ? launch_t<true>(idx, a, block, grid, persist, wide, buckets)
                 : launch_t<false>(idx, a, block, grid, persist, wide, buckets);
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
    a.seed_len = rs->seed_len;
    a.out_n = rs->d_out_n;
    a.out_off = rs->d_out_off;
    a.out_pops = rs->d_pops;
    a.out_trace = rs->d_trace;
    a.out_rec = rs->d_rec;
    a.out_cap = rs->rec_cap;
    a.out_top = rs->d_top;
    // keeping the exact child in registers relies on children never scoring below their parent
    bool nonneg = true;
    for (const auto &o : rs->h_bopts) nonneg = nonneg && o.s_mm >= 0 && o.s_gapo >= 0 && o.s_gape >= 0;
    a.fuse = ((flags & GALN_HIP_NOFUSE) || !nonneg) ? 0 : 1;
    return a;
}

// Runs reads list[0..n) (or identity from list_base) in one tier; failures land in fail buffer `fo`.
int run_tier(galn_hip_readset *rs, uint32_t flags, const Tier &t, const int32_t *list, int list_base, int n_list,
             bool persist, int max_slots, int fo, int block)
{
    galn_hip_index *idx = rs->idx;
    if (n_list <= 0) return 0;
    Layout L = make_layout(rs->max_len, rs->seed_len, t.n_chunks, t.rec_cap);
    int slots = persist ? std::min(max_slots, n_list) : n_list;
    // keep the arena within a budget: at most 24 GB of lane arenas per launch
    const size_t budget = (size_t)24 << 30;
    if ((size_t)slots * L.slot_bytes > budget) {
        const int cap = (int)std::max<size_t>(block, budget / L.slot_bytes / block * block);
        if (!persist) {
            // split a plain launch into pieces that fit
            for (int s = 0; s < n_list; s += cap) {
                const int r = run_tier(rs, flags, t, list ? list + s : nullptr, list_base + (list ? 0 : s),
                                       std::min(cap, n_list - s), false, max_slots, fo, block);
                if (r) return r;
            }
            return 0;
        }
        slots = std::min(slots, cap);
    }
    int r = ensure_scratch(rs, (size_t)slots * L.slot_bytes);
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
    a.link_off = L.link_off;
    a.chunk_off = L.chunk_off;
    a.n_chunks = t.n_chunks;
    a.rec_cap = t.rec_cap;
    a.fail_list = rs->d_fail[fo];
    a.fail_n = rs->d_ctr + fo;
    a.fail_kind = rs->d_ctr + 4;
    if (persist) HIP_OK(hipMemsetAsync(rs->d_ctr + 2, 0, sizeof(int32_t), idx->stream));
    const int grid = (slots + block - 1) / block;
    return launch(idx, a, block, grid, persist, t.wide, (flags & GALN_HIP_TRACE) != 0,
                  std::max(1, std::min(rs->max_stacks, MAXB)));
}

int read_ctr(galn_hip_readset *rs, int which, int32_t *v)
{
    HIP_OK(hipMemcpyAsync(v, rs->d_ctr + which, sizeof(int32_t), hipMemcpyDeviceToHost, rs->idx->stream));
    HIP_OK(hipStreamSynchronize(rs->idx->stream));
    return 0;
}

int grow_records(galn_hip_readset *rs)
{
    galn_hip_index *idx = rs->idx;
    unsigned long long top = 0;
    HIP_OK(hipMemcpyAsync(&top, rs->d_top, sizeof(top), hipMemcpyDeviceToHost, idx->stream));
    HIP_OK(hipStreamSynchronize(idx->stream));
    const uint64_t used = std::min<uint64_t>(top, rs->rec_cap);
    const uint64_t ncap = std::max<uint64_t>(rs->rec_cap * 2, top + 4096);
    galn_aln_rec_t *nrec = nullptr;
    HIP_OK(hipMalloc(&nrec, ncap * sizeof(galn_aln_rec_t)));
    if (used) HIP_OK(hipMemcpyAsync(nrec, rs->d_rec, used * sizeof(galn_aln_rec_t), hipMemcpyDeviceToDevice, idx->stream));
    HIP_OK(hipStreamSynchronize(idx->stream));
    HIP_OK(hipFree(rs->d_rec));
    rs->d_rec = nrec;
    rs->rec_cap = ncap;
    unsigned long long u = used;
    HIP_OK(hipMemcpyAsync(rs->d_top, &u, sizeof(u), hipMemcpyHostToDevice, idx->stream));
    HIP_OK(hipStreamSynchronize(idx->stream));
    return 0;
}

int pick_block(int max_stacks, bool wide)
{
    (void)max_stacks;
    (void)wide;
    return 128;
}

// Overflow tiers: re-run failures (fail buffer `fi`, count n) until none is left.
int run_overflow(galn_hip_readset *rs, uint32_t flags, int fi, int32_t n)
{
    galn_hip_index *idx = rs->idx;
    int max_entries = 0;
    for (const auto &o : rs->h_bopts) max_entries = std::max(max_entries, o.max_entries);
    const int big_chunks = (int)std::min<long>((long)(max_entries + 16) / CH + MAXB + 2, (1l << 27) - 1);
    Tier tiers[3] = {{1024, 256, false}, {4095, 1024, false}, {big_chunks, 4096, true}};
    int level = 0;
    int rec_boost = 1;
    while (n > 0) {
        const int fo = fi ^ 1;
        HIP_OK(hipMemsetAsync(rs->d_ctr + fo, 0, sizeof(int32_t), idx->stream));
        HIP_OK(hipMemsetAsync(rs->d_ctr + 4, 0, 3 * sizeof(int32_t), idx->stream));
        Tier t = tiers[std::min(level, 2)];
        t.rec_cap *= rec_boost;
        const int slots = level == 0 ? idx->n_cu * 256 : level == 1 ? idx->n_cu * 64 : std::max(64, idx->n_cu * 2);
        int r = run_tier(rs, flags, t, rs->d_fail[fi], 0, n, true, slots, fo, 64);
        if (r) return r;
        int32_t nf = 0;
        if ((r = read_ctr(rs, fo, &nf))) return r;
        if (nf > 0) {
            // classify the failures: output arena full -> grow it; records -> more slots
            int32_t kind[3] = {0, 0, 0};
            HIP_OK(hipMemcpy(kind, rs->d_ctr + 4, sizeof(kind), hipMemcpyDeviceToHost));
            const bool any_out = kind[0] != 0, any_aln = kind[1] != 0, any_stack = kind[2] != 0;
            if (any_out && (r = grow_records(rs))) return r;
            if (any_stack) ++level;
            else if (any_aln) rec_boost *= 4;
            if (level > 2 && any_stack) return fail(GALN_HIP_E_NOMEM, "stack arena still too small at the last tier");
        }
        fi = fo;
        n = nf;
    }
    return 0;
}

int readset_search(galn_hip_readset *rs, int form, uint32_t flags, const galn_hip_tune_t *tune)
{
    galn_hip_index *idx = rs->idx;
    HIP_OK(hipSetDevice(idx->device));
    rs->last_flags = flags;
    HIP_OK(hipMemsetAsync(rs->d_top, 0, sizeof(unsigned long long), idx->stream));
    HIP_OK(hipMemsetAsync(rs->d_ctr, 0, 8 * sizeof(int32_t), idx->stream));
    const int block = pick_block(rs->max_stacks, false);
    Tier t0{64, 32, false};
    if (tune && tune->chunks_per_lane > 0) t0.n_chunks = std::min(tune->chunks_per_lane, 4095);
    if (tune && tune->rec_per_lane > 0) t0.rec_cap = tune->rec_per_lane;
    HIP_OK(hipEventRecord(rs->ev[0], idx->stream));
    int r = 0;
    if (form == 0) {
        // plain: one launch per 0x40000 reads, input order
        for (int s = 0; s < rs->n && !r; s += 0x40000)
            r = run_tier(rs, flags, t0, nullptr, s, std::min(0x40000, rs->n - s), false, 0, 0, block);
    } else {
        int waves = (tune && tune->waves_per_cu > 0) ? tune->waves_per_cu : 16;
        const int blk = (tune && tune->block_threads > 0) ? tune->block_threads : block;
        const int slots = idx->n_cu * waves * 64;
        const bool sorted = !tune || tune->sort_by_cost != 0;
        r = run_tier(rs, flags, t0, sorted ? rs->d_sorted : nullptr, 0, rs->n, true, slots, 0, blk);
    }
    if (r) return r;
    HIP_OK(hipEventRecord(rs->ev[1], idx->stream));
    int32_t nf = 0;
    if ((r = read_ctr(rs, 0, &nf))) return r;
    if (nf > 0) {
        // a first-tier failure may also be an output-arena overflow
        unsigned long long top = 0;
        HIP_OK(hipMemcpy(&top, rs->d_top, sizeof(top), hipMemcpyDeviceToHost));
        if (top > rs->rec_cap && (r = grow_records(rs))) return r;
        if ((r = run_overflow(rs, flags, 0, nf))) return r;
    }
    HIP_OK(hipEventRecord(rs->ev[2], idx->stream));
    HIP_OK(hipEventSynchronize(rs->ev[2]));
    float ms1 = 0, ms2 = 0;
    HIP_OK(hipEventElapsedTime(&ms1, rs->ev[0], rs->ev[1]));
    HIP_OK(hipEventElapsedTime(&ms2, rs->ev[0], rs->ev[2]));
    rs->first_ms = ms1;
    rs->total_ms = ms2;
    return 0;
}

int check_reads(const galn_hip_reads *r)
{
    if (!r || r->n_reads < 0) return fail(GALN_HIP_E_ARG, "reads: null or negative count");
    if (r->n_reads > 0 && (!r->query || !r->qoff || !r->qlen || !r->max_diff || !r->batch))
        return fail(GALN_HIP_E_ARG, "reads: null array");
    if (r->n_batches < 1 || !r->bopts || !r->n_stacks) return fail(GALN_HIP_E_ARG, "reads: no batch options");
    if (r->seed_len < 0) return fail(GALN_HIP_E_UNSUPPORTED, "reads: negative seed length");
    for (int i = 0; i < r->n_reads; ++i)
        if (r->batch[i] < 0 || r->batch[i] >= r->n_batches) return fail(GALN_HIP_E_ARG, "reads: batch index out of range");
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
    if (hipStreamCreateWithFlags(&idx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMemsetAsync(idx->d_bwt, 0, alloc, idx->stream) != hipSuccess ||
        hipMemcpyAsync(idx->d_bwt, bwt, bytes, hipMemcpyHostToDevice, idx->stream) != hipSuccess ||
        hipStreamSynchronize(idx->stream) != hipSuccess) {
        hipFree(idx->d_bwt);
        delete idx;
        return fail(GALN_HIP_E_DEVICE, "index upload failed");
    }
    *out = idx;
    return 0;
}

int galn_hip_index_close(galn_hip_index *idx)
{
    if (!idx) return 0;
    hipSetDevice(idx->device);
    if (idx->stream) hipStreamSynchronize(idx->stream);
    hipFree(idx->d_bwt);
    if (idx->stream) hipStreamDestroy(idx->stream);
    delete idx;
    return 0;
}

int galn_hip_index_device(const galn_hip_index *idx) { return idx ? idx->device : -1; }

int galn_hip_readset_free(galn_hip_readset *rs)
{
    if (!rs) return 0;
    hipSetDevice(rs->idx->device);
    hipStreamSynchronize(rs->idx->stream);
    void *ptrs[] = {rs->d_query, rs->d_qoff, rs->d_qlen, rs->d_maxdiff, rs->d_batch, rs->d_nstacks, rs->d_bopts,
                    rs->d_sorted, rs->d_out_n, rs->d_out_off, rs->d_pops, rs->d_trace, rs->d_rec, rs->d_top,
                    rs->d_fail[0], rs->d_fail[1], rs->d_ctr, rs->d_scratch};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (auto &e : rs->ev)
        if (e) hipEventDestroy(e);
    delete rs;
    return 0;
}

int galn_hip_readset_upload(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_readset **out)
{
    if (!idx || !out) return fail(GALN_HIP_E_ARG, "readset_upload: null argument");
    *out = nullptr;
    int r = check_reads(reads);
    if (r) return r;
    HIP_OK(hipSetDevice(idx->device));
    galn_hip_readset *rs = new galn_hip_readset;
    rs->idx = idx;
    const int n = reads->n_reads;
    rs->n = n;
    rs->n_batches = reads->n_batches;
    rs->seed_len = reads->seed_len;
    uint64_t qbytes = 0;
    int max_len = 1;
    for (int i = 0; i < n; ++i) {
        qbytes = std::max<uint64_t>(qbytes, reads->qoff[i] + (uint64_t)std::max(reads->qlen[i], 0));
        max_len = std::max(max_len, reads->qlen[i]);
    }
    rs->max_len = std::min(max_len, MAXLEN);
    rs->qbytes = qbytes;
    rs->max_stacks = 1;
    for (int b = 0; b < reads->n_batches; ++b) rs->max_stacks = std::max(rs->max_stacks, reads->n_stacks[b]);
    rs->h_qlen.assign(reads->qlen, reads->qlen + n);
    rs->h_batch.assign(reads->batch, reads->batch + n);
    rs->h_nstacks.assign(reads->n_stacks, reads->n_stacks + reads->n_batches);
    rs->h_bopts.assign(reads->bopts, reads->bopts + reads->n_batches);
    const size_t nn = (size_t)std::max(n, 1);
    auto M = [&](void **p, size_t b) -> bool { return hipMalloc(p, std::max<size_t>(b, 16)) == hipSuccess; };
    rs->rec_cap = std::max<uint64_t>((uint64_t)n * 2, 1 << 16);
    bool ok = M((void **)&rs->d_query, qbytes + 16) && M((void **)&rs->d_qoff, nn * 8) && M((void **)&rs->d_qlen, nn * 4) &&
              M((void **)&rs->d_maxdiff, nn * 4) && M((void **)&rs->d_batch, nn * 4) &&
              M((void **)&rs->d_nstacks, reads->n_batches * 4) &&
              M((void **)&rs->d_bopts, reads->n_batches * sizeof(galn_search_opt_t)) && M((void **)&rs->d_sorted, nn * 4) &&
              M((void **)&rs->d_out_n, nn * 4) && M((void **)&rs->d_out_off, nn * 8) && M((void **)&rs->d_pops, nn * 8) &&
              M((void **)&rs->d_trace, nn * 8) && M((void **)&rs->d_rec, rs->rec_cap * sizeof(galn_aln_rec_t)) &&
              M((void **)&rs->d_top, 8) && M((void **)&rs->d_fail[0], nn * 4) && M((void **)&rs->d_fail[1], nn * 4) &&
              M((void **)&rs->d_ctr, 8 * sizeof(int32_t));
    if (!ok) {
        galn_hip_readset_free(rs);
        return fail(GALN_HIP_E_NOMEM, "readset_upload: device allocation failed");
    }
    for (auto &e : rs->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            galn_hip_readset_free(rs);
            return fail(GALN_HIP_E_DEVICE, "hipEventCreate failed");
        }
    // claim order of the batched form: longest reads first (their searches are the longest)
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return reads->qlen[x] > reads->qlen[y]; });
    hipStream_t s = idx->stream;
    if (hipMemcpyAsync(rs->d_query, reads->query, qbytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_qoff, reads->qoff, (size_t)n * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_qlen, reads->qlen, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_maxdiff, reads->max_diff, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_batch, reads->batch, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_nstacks, reads->n_stacks, (size_t)reads->n_batches * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_bopts, reads->bopts, (size_t)reads->n_batches * sizeof(galn_search_opt_t),
                       hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(rs->d_sorted, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        galn_hip_readset_free(rs);
        return fail(GALN_HIP_E_DEVICE, "readset_upload: copy failed");
    }
    *out = rs;
    return 0;
}

int galn_hip_readset_search(galn_hip_index *idx, galn_hip_readset *rs, int form, uint32_t flags,
                            const galn_hip_tune_t *tune)
{
    if (!idx || !rs || rs->idx != idx) return fail(GALN_HIP_E_ARG, "readset_search: bad handle");
    if (rs->n == 0) return 0;
    return readset_search(rs, form, flags, tune);
}

int galn_hip_readset_results(galn_hip_readset *rs, galn_hip_out *out)
{
    if (!rs || !out || (rs->n > 0 && (!out->n_aln || !out->rec_off)))
        return fail(GALN_HIP_E_ARG, "readset_results: null argument");
    galn_hip_index *idx = rs->idx;
    HIP_OK(hipSetDevice(idx->device));
    const int n = rs->n;
    if (n == 0) {
        out->n_rec = 0;
        return 0;
    }
    hipStream_t s = idx->stream;
    std::vector<uint64_t> off(n);
    unsigned long long top = 0;
    HIP_OK(hipMemcpyAsync(out->n_aln, rs->d_out_n, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(off.data(), rs->d_out_off, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&top, rs->d_top, 8, hipMemcpyDeviceToHost, s));
    if (out->n_pops) HIP_OK(hipMemcpyAsync(out->n_pops, rs->d_pops, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    if (out->trace) HIP_OK(hipMemcpyAsync(out->trace, rs->d_trace, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    uint64_t need = 0;
    for (int i = 0; i < n; ++i)
        if (out->n_aln[i] > 0) need += (uint64_t)out->n_aln[i];
    out->n_rec = need;
    if (need > out->rec_cap || (need > 0 && !out->rec))
        return fail(GALN_HIP_E_RECCAP, "record array holds %llu, %llu needed", (unsigned long long)out->rec_cap,
                    (unsigned long long)need);
    const uint64_t used = std::min<uint64_t>(top, rs->rec_cap);
    std::vector<galn_aln_rec_t> arena(used);
    if (used) HIP_OK(hipMemcpyAsync(arena.data(), rs->d_rec, used * sizeof(galn_aln_rec_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    // records in read order
    uint64_t o = 0;
    for (int i = 0; i < n; ++i) {
        out->rec_off[i] = o;
        const int na = out->n_aln[i];
        if (na > 0) {
            memcpy(out->rec + o, arena.data() + off[i], (size_t)na * sizeof(galn_aln_rec_t));
            o += (uint64_t)na;
        }
    }
    return 0;
}

int galn_hip_readset_kernel_ms(const galn_hip_readset *rs, double *first_tier_ms, double *total_ms)
{
    if (!rs) return fail(GALN_HIP_E_ARG, "readset_kernel_ms: null handle");
    if (first_tier_ms) *first_tier_ms = rs->first_ms;
    if (total_ms) *total_ms = rs->total_ms;
    return 0;
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
