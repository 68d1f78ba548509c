// galn_cpu.cpp -- the CPU search engine (include/galn_cpu.h): BWA aln's bwt_match_gap on the
// host, for `galn aln --cpu-only` and for the reads the GPU kernels do not take (longer than
// their length limit, or options beyond their bucket and difference limits).
//
// Same results as the reference's CPU search (src/aln_host.c:87-147, src/bwtgap_hd.h:338-473):
// records in the order found, pop counts and trace hashes. Design:
//   * one read = one resumable search (class Search): it runs until it needs an occurrence
//     lookup, names the two positions whose 64-byte index blocks the lookup reads, and yields;
//   * one thread runs several searches in turn (the `interleave` slots): when it comes back to
//     a search its blocks have been prefetched while the other slots ran, so a core keeps that
//     many independent index loads in flight instead of one (the reference's aln_il does the
//     same with two hand-cut halves of its loop, src/aln_il.c);
//   * the priority stack is one growable LIFO array per score bucket plus a bitmask of the
//     non-empty buckets (upstream BWA's own layout), arrays kept from read to read;
//   * threads claim reads in chunks from an atomic counter, keep their records, and the records
//     are gathered in read order at the end.
// Entry fields keep the reference's widths (u8 counts, u16 indel lengths, a u32 info word of
// score << 21 | i, with i read back as info & 0xffff), so that even options or reads beyond
// everyday sizes take the reference's path.

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "galn_cpu.h"

namespace {

constexpr int MODE_GAPE = 0x01, MODE_LOGGAP = 0x04, MODE_NONSTOP = 0x10;
constexpr int ST_M = 0, ST_I = 1, ST_D = 2;
constexpr int E_STACK_SCORE = -3;
constexpr uint64_t TR_INIT = 1469598103934665603ull;
constexpr uint64_t TR_PRIME = 1099511628211ull;

inline uint64_t tr_mix(uint64_t h, uint64_t v) { return (h ^ v) * TR_PRIME; }

// ---- occurrence counts over the .bwt body (BWA layout: 128-symbol blocks of 16 u32 words,
// [4 x u64 counts before the block][8 words of 16 two-bit symbols, first in bits 31-30]) ----

struct Idx {
    const uint32_t *bwt;
    uint64_t primary, seq_len;
    uint64_t L2[5];
};

// The 8 symbol words of a block are read as 4 chunks of 64 bits (words 2q and 2q + 1, symbol j of
// the chunk in bits 63-2j and 62-2j): one popcount round counts 32 symbols.
constexpr uint64_t M55 = 0x5555555555555555ull;

inline uint64_t chunk_of(const uint32_t *w, int q) { return (uint64_t)w[2 * q] << 32 | w[2 * q + 1]; }

// the symbol fields of symbols [a, b) of a chunk (0 <= a < b <= 32), as the low bit of each field
inline uint64_t span_mask(int a, int b)
{
    uint64_t keep = M55;
    if (b < 32) keep &= ~(~0ull >> (2 * b));
    if (a > 0) keep &= ~0ull >> (2 * a);
    return keep;
}

// adds the counts of the four bases among symbols [s0, s1) of a block's words w (0 <= s0 <= s1 <= 128):
// popcounts of the high bits, the low bits and both, summed over the chunks (only the first and the
// last chunk of the range need a mask), then split into the four bases once
inline void range_count4(const uint32_t *w, int s0, int s1, uint64_t n[4])
{
    if (s0 >= s1) return;
    const int q0 = s0 >> 5, q1 = (s1 - 1) >> 5;
    uint64_t nh = 0, nl = 0, nb = 0;
    for (int q = q0; q <= q1; ++q) {
        uint64_t keep = M55;
        if (q == q0 || q == q1) keep = span_mask(q == q0 ? s0 - 32 * q : 0, q == q1 ? s1 - 32 * q : 32);
        const uint64_t x = chunk_of(w, q);
        const uint64_t hi = (x >> 1) & keep, lo = x & keep;
        nb += (uint64_t)__builtin_popcountll(hi & lo);
        nh += (uint64_t)__builtin_popcountll(hi);
        nl += (uint64_t)__builtin_popcountll(lo);
    }
    n[0] += (uint64_t)(s1 - s0) - nh - nl + nb;
    n[1] += nl - nb;
    n[2] += nh - nb;
    n[3] += nb;
}

// the count of base c among symbols [s0, s1) of a block's words w
inline uint64_t range_count1(const uint32_t *w, int s0, int s1, int c)
{
    uint64_t n = 0;
    if (s0 >= s1) return 0;
    const int q0 = s0 >> 5, q1 = (s1 - 1) >> 5;
    // the symbols equal to c, as the low bit of each field: x ^ ~c repeated has both bits set there
    const uint64_t pat = ~((uint64_t)(c & 3) * M55);
    for (int q = q0; q <= q1; ++q) {
        uint64_t keep = M55;
        if (q == q0 || q == q1) keep = span_mask(q == q0 ? s0 - 32 * q : 0, q == q1 ? s1 - 32 * q : 32);
        const uint64_t y = chunk_of(w, q) ^ pat;
        n += (uint64_t)__builtin_popcountll(y & (y >> 1) & keep);
    }
    return n;
}

inline const uint32_t *block_of(const Idx &x, uint64_t p) { return x.bwt + ((p >> 7) << 4); }

inline uint64_t count_before(const uint32_t *blk, int c)
{
    return (uint64_t)blk[2 * c] | (uint64_t)blk[2 * c + 1] << 32;
}

// Occ of the four bases at a and b (a < b; a = -1: none), bwt_2occ4: when both rows are in one
// block the second count continues from the first
void occ4_pair(const Idx &x, uint64_t a, uint64_t b, uint64_t ca[4], uint64_t cb[4])
{
    const uint64_t bp = b - (b >= x.primary);
    const uint32_t *bb = block_of(x, bp);
    const int rb = (int)(bp & 127) + 1;  // symbols of the block up to bp
    if (a == ~0ull) {
        ca[0] = ca[1] = ca[2] = ca[3] = 0;
    } else {
        const uint64_t ap = a - (a >= x.primary);
        const uint32_t *ab = block_of(x, ap);
        const int ra = (int)(ap & 127) + 1;
        uint64_t n[4] = {0, 0, 0, 0};
        range_count4(ab + 8, 0, ra, n);
        for (int c = 0; c < 4; ++c) ca[c] = count_before(ab, c) + n[c];
        if (ab == bb) {
            uint64_t d[4] = {0, 0, 0, 0};
            range_count4(bb + 8, ra, rb, d);
            for (int c = 0; c < 4; ++c) cb[c] = ca[c] + d[c];
            return;
        }
    }
    uint64_t n[4] = {0, 0, 0, 0};
    range_count4(bb + 8, 0, rb, n);
    for (int c = 0; c < 4; ++c) cb[c] = count_before(bb, c) + n[c];
}

// algorithmic bytes of a lookup at (a, b): 64 per distinct block; position -1 needs none
inline uint32_t lookup_bytes(const Idx &x, uint64_t a, uint64_t b)
{
    const uint64_t bp = b - (b >= x.primary);
    if (a == ~0ull) return 64;
    const uint64_t ap = a - (a >= x.primary);
    return (ap >> 7) == (bp >> 7) ? 64 : 128;
}

// Occ of one base at a and b (bwt_2occ: the k = seq_len total applies to the single-position
// form only, as in the reference's two-position count)
inline void occ1_pair(const Idx &x, uint64_t a, uint64_t b, int c, uint64_t &oa, uint64_t &ob)
{
    const uint64_t bp = b - (b >= x.primary);
    const uint32_t *bb = block_of(x, bp);
    const int rb = (int)(bp & 127) + 1;
    oa = 0;
    if (a != ~0ull) {
        const uint64_t ap = a - (a >= x.primary);
        const uint32_t *ab = block_of(x, ap);
        const int ra = (int)(ap & 127) + 1;
        oa = count_before(ab, c) + range_count1(ab + 8, 0, ra, c);
        if (ab == bb) {
            ob = oa + range_count1(bb + 8, ra, rb, c);
            return;
        }
    }
    ob = count_before(bb, c) + range_count1(bb + 8, 0, rb, c);
}

inline void prefetch_lookup(const Idx &x, uint64_t a, uint64_t b)
{
    __builtin_prefetch(block_of(x, b - (b >= x.primary)), 0, 0);
    if (a != ~0ull) __builtin_prefetch(block_of(x, a - (a >= x.primary)), 0, 0);
}

struct Width {
    uint64_t w;
    int bid;
};

struct Entry {
    uint64_t k, l;
    uint32_t info;  // score << 21 | i
    int32_t last_diff_pos;
    uint8_t n_mm, n_gapo, n_gape, state;
    uint16_t n_ins, n_del;
};

// ---- one read's search, resumable at every occurrence lookup ----

class Search {
public:
    // what run() stopped for
    enum { LOOKUP = 0, DONE = 1 };
    uint64_t qa = 0, qb = 0;  // the pending lookup's positions (prefetch their blocks)

    void begin(const Idx &x, const uint8_t *q, int len, const galn_search_opt_t &bo, int n_stacks, int md_read)
    {
        X = &x;
        seq = q;
        this->len = len;
        o = bo;
        this->n_stacks = n_stacks;
        max_diff_read = md_read;
        // aln_search_host's per-read options: src/aln_host.c:96-98
        o.max_diff = md_read;
        seeding = len > bo.seed_len;
        o.seed_len = seeding ? bo.seed_len : 0x7fffffff;
        W.resize((size_t)len + 1);
        if (seeding) SW.resize((size_t)bo.seed_len + 1);
        n_aln = 0;
        recs.clear();
        pops = 0;
        trace = TR_INIT;
        bytes = 0;
        result = 0;
        // widths of the query (reversed, complement undone), then of its seed part
        wpass = 0;
        wstr = seq;
        wlen = len;
        wout = W.data();
        wj = 0;
        wk = 0;
        wl = x.seq_len;
        wbid = 0;
        phase = PH_WIDTH;
        pending = false;
    }

    // Runs until the next occurrence lookup (LOOKUP: qa, qb set) or the end of the search (DONE).
    int run()
    {
        for (;;) {
            switch (phase) {
            case PH_WIDTH:
                if (width_step()) return LOOKUP;
                break;
            case PH_LOOP:
                if (loop_step()) return LOOKUP;
                break;
            case PH_EXACT:
                if (exact_step()) return LOOKUP;
                break;
            case PH_EXPAND:
                expand();
                break;
            default:
                return DONE;
            }
        }
    }

    int result = 0;  // n_aln, or a negative code
    std::vector<galn_aln_rec_t> recs;
    uint64_t pops = 0, trace = TR_INIT, bytes = 0;
    bool want_trace = true;  // hash the pops (the caller asked for traces)

private:
    enum { PH_WIDTH, PH_LOOP, PH_EXACT, PH_EXPAND, PH_DONE };
    const Idx *X = nullptr;
    const uint8_t *seq = nullptr;
    int len = 0, n_stacks = 0, max_diff_read = 0;
    galn_search_opt_t o{};
    bool seeding = false;
    std::vector<Width> W, SW;
    int phase = PH_DONE;
    bool pending = false;  // the lookup at (qa, qb) was requested and its blocks are on their way
    // widths in progress (bwt_cal_width, src/bwtgap_hd.h:216-242)
    int wpass = 0, wlen = 0, wj = 0, wbid = 0, wc = 0;
    const uint8_t *wstr = nullptr;
    Width *wout = nullptr;
    uint64_t wk = 0, wl = 0;
    // the search (bwt_match_gap, src/bwtgap_hd.h:338-473)
    std::vector<std::vector<Entry>> bucket;
    std::vector<uint64_t> nonempty;  // bit s: bucket s holds entries
    int n_entries = 0;
    int best_score = 0, best_diff = 0, max_diff = 0, best_cnt = 0, n_aln = 0;
    Entry e{};                       // the popped entry being worked on
    int ei = 0, m = 0, m_seed = 0;   // its position, allowed differences (all, seed)
    uint64_t ek = 0, el = 0;         // the interval an exact tail narrows
    int xj = 0;                      // exact tail: next position
    int xc = 0;

    static int score_of(int mm, int go, int ge, const galn_search_opt_t &p) { return mm * p.s_mm + go * p.s_gapo + ge * p.s_gape; }

    // ---- widths ----
    bool width_step()
    {
        for (;;) {
            if (wj == wlen) {
                wout[wlen].w = 0;
                wout[wlen].bid = ++wbid;
                if (wpass == 0 && seeding) {
                    wpass = 1;
                    wstr = seq + (len - o.seed_len);
                    wlen = o.seed_len;
                    wout = SW.data();
                    wj = 0;
                    wk = 0;
                    wl = X->seq_len;
                    wbid = 0;
                    continue;
                }
                start_search();
                return false;
            }
            const uint8_t q = wstr[wj];
            const int c = q > 3 ? 4 : 3 - q;
            if (c < 4) {
                if (!pending) {
                    qa = wk - 1;
                    qb = wl;
                    wc = c;
                    pending = true;
                    return true;
                }
                pending = false;
                uint64_t ok, ol;
                occ1_pair(*X, wk - 1, wl, c, ok, ol);
                bytes += lookup_bytes(*X, wk - 1, wl);
                wk = X->L2[c] + ok + 1;
                wl = X->L2[c] + ol;
            }
            if (wk > wl || c > 3) {
                wk = 0;
                wl = X->seq_len;
                ++wbid;
            }
            wout[wj].w = wl - wk + 1;
            wout[wj].bid = wbid;
            ++wj;
        }
    }

    // ---- the stack ----
    int push(int i, uint64_t k, uint64_t l, int mm, int go, int ge, int ins, int del, int state, int is_diff)
    {
        const int score = score_of(mm, go, ge, o);
        if (score < 0 || score >= n_stacks) return E_STACK_SCORE;
        Entry p;
        p.info = (uint32_t)score << 21 | (uint32_t)i;
        p.k = k;
        p.l = l;
        p.n_mm = (uint8_t)mm;
        p.n_gapo = (uint8_t)go;
        p.n_gape = (uint8_t)ge;
        p.n_ins = (uint16_t)ins;
        p.n_del = (uint16_t)del;
        p.state = (uint8_t)state;
        p.last_diff_pos = is_diff ? i : 0;
        bucket[(size_t)score].push_back(p);
        nonempty[(size_t)score >> 6] |= 1ull << (score & 63);
        ++n_entries;
        return 0;
    }

    void pop()
    {
        size_t wi = 0;
        while (!nonempty[wi]) ++wi;
        const int s = (int)(wi * 64 + (size_t)__builtin_ctzll(nonempty[wi]));
        std::vector<Entry> &b = bucket[(size_t)s];
        e = b.back();
        b.pop_back();
        if (b.empty()) nonempty[wi] &= ~(1ull << (s & 63));
        --n_entries;
        ++pops;
        if (!want_trace) return;
        trace = tr_mix(trace, e.info);
        trace = tr_mix(trace, e.k);
        trace = tr_mix(trace, e.l);
        trace = tr_mix(trace, (uint64_t)((uint32_t)e.n_mm | (uint32_t)e.n_gapo << 8 | (uint32_t)e.n_gape << 16 |
                                         (uint32_t)e.state << 24));
        trace = tr_mix(trace, (uint64_t)e.n_ins | (uint64_t)e.n_del << 16);
        trace = tr_mix(trace, (uint64_t)(uint32_t)e.last_diff_pos);
        trace = tr_mix(trace, (uint64_t)(uint32_t)n_entries);
    }

    void finish(int code)
    {
        result = code;
        phase = PH_DONE;
        // keep the arrays' capacity for the next read, not their entries
        for (size_t s = 0; s < nonempty.size() * 64 && s < bucket.size(); ++s)
            if (nonempty[s >> 6] >> (s & 63) & 1) bucket[s].clear();
    }

    void start_search()
    {
        best_score = score_of(o.max_diff + 1, o.max_gapo + 1, o.max_gape + 1, o);
        best_diff = o.max_diff + 1;
        max_diff = o.max_diff;
        best_cnt = 0;
        n_aln = 0;
        int nN = 0;
        for (int j = 0; j < len; ++j) nN += seq[j] > 3;
        if (nN > max_diff) {
            finish(0);
            return;
        }
        if ((int)bucket.size() < n_stacks) bucket.resize((size_t)n_stacks);
        nonempty.assign(((size_t)std::max(n_stacks, 1) + 63) / 64, 0);
        n_entries = 0;
        const int r = push(len, 0, X->seq_len, 0, 0, 0, 0, 0, ST_M, 0);
        if (r) {
            finish(r);
            return;
        }
        phase = PH_LOOP;
    }

    // gap_shadow, src/bwtgap_hd.h:309-320
    void shadow(int x, int last_diff_pos)
    {
        int j = 0;
        for (int i = 0; i < last_diff_pos; ++i) {
            if (W[(size_t)i].w > (uint64_t)(int64_t)x) {
                W[(size_t)i].w -= (uint64_t)(int64_t)x;
            } else if (W[(size_t)i].w == (uint64_t)(int64_t)x) {
                W[(size_t)i].bid = 1;
                W[(size_t)i].w = X->seq_len - (uint64_t)(++j);
            }
        }
    }

    // a hit at [k, l] by the popped entry: false when the search ends (top-2 cut-off)
    bool hit(uint64_t k, uint64_t l)
    {
        const int score = score_of(e.n_mm, e.n_gapo, e.n_gape, o);
        bool add = true;
        if (n_aln == 0) {
            best_score = score;
            best_diff = e.n_mm + e.n_gapo;
            if (o.mode & MODE_GAPE) best_diff += e.n_gape;
            if (!(o.mode & MODE_NONSTOP)) max_diff = (best_diff + 1 > o.max_diff) ? o.max_diff : best_diff + 1;
        }
        if (score == best_score)
            best_cnt = (int)(uint32_t)((uint64_t)(int64_t)best_cnt + (l - k + 1));
        else if (best_cnt > o.max_top2)
            return false;
        if (e.n_gapo) {
            for (int j = 0; j < n_aln; ++j)
                if (recs[(size_t)j].k == k && recs[(size_t)j].l == l) {
                    add = false;
                    break;
                }
        }
        if (add) {
            shadow((int)(uint32_t)(l - k + 1), e.last_diff_pos);
            galn_aln_rec_t r;
            r.packed = (uint64_t)(e.n_mm & 0xff) | (uint64_t)(e.n_gapo & 0xff) << 8 | (uint64_t)(e.n_gape & 0xff) << 16 |
                       (uint64_t)((uint32_t)score & 0xfffff) << 24 | (uint64_t)(e.n_ins & 0x3ff) << 44 |
                       (uint64_t)(e.n_del & 0x3ff) << 54;
            r.k = k;
            r.l = l;
            recs.push_back(r);
            ++n_aln;
        }
        return true;
    }

    // the top of the loop: pop entries until one needs a lookup
    bool loop_step()
    {
        for (;;) {
            if (n_entries == 0 || n_entries > o.max_entries) {
                finish(n_aln);
                return false;
            }
            pop();
            ei = (int)(e.info & 0xffff);
            if (!(o.mode & MODE_NONSTOP) && (e.info >> 21) > (uint32_t)(best_score + o.s_mm)) {
                finish(n_aln);
                return false;
            }
            m = max_diff - (e.n_mm + e.n_gapo);
            if (o.mode & MODE_GAPE) m -= e.n_gape;
            if (m < 0) continue;
            if (seeding) {
                m_seed = o.max_seed_diff - (e.n_mm + e.n_gapo);
                if (o.mode & MODE_GAPE) m_seed -= e.n_gape;
            }
            if (ei > 0 && m < W[(size_t)ei - 1].bid) continue;
            if (ei == 0) {
                if (!hit(e.k, e.l)) {
                    finish(n_aln);
                    return false;
                }
                continue;
            }
            if (m == 0 && (e.state == ST_M || (o.mode & MODE_GAPE) || e.n_gape == o.max_gape)) {
                // no difference allowed: the rest must match exactly (bwt_match_exact_alt)
                ek = e.k;
                el = e.l;
                xj = ei - 1;
                phase = PH_EXACT;
                return false;
            }
            qa = e.k - 1;
            qb = e.l;
            pending = true;
            phase = PH_EXPAND;
            return true;
        }
    }

    bool exact_step()
    {
        for (;;) {
            if (xj < 0) {
                phase = PH_LOOP;
                // bwt_match_exact_alt returns the interval size as an int: a size of 2^32 reads as no hit
                if ((int)(uint32_t)(el - ek + 1) != 0 && !hit(ek, el)) finish(n_aln);
                return false;
            }
            const uint8_t c = seq[xj];
            if (c > 3) {  // an N: no match
                phase = PH_LOOP;
                return false;
            }
            if (!pending) {
                qa = ek - 1;
                qb = el;
                xc = c;
                pending = true;
                return true;
            }
            pending = false;
            uint64_t ok, ol;
            occ1_pair(*X, ek - 1, el, c, ok, ol);
            bytes += lookup_bytes(*X, ek - 1, el);
            ek = X->L2[c] + ok + 1;
            el = X->L2[c] + ol;
            if (ek > el) {
                phase = PH_LOOP;
                return false;
            }
            --xj;
        }
    }

#define PUSH_OR_FINISH(...)                       \
    do {                                          \
        const int r_ = push(__VA_ARGS__);         \
        if (r_) {                                 \
            finish(r_);                           \
            return;                               \
        }                                         \
    } while (0)

    // the popped entry's expansion, its lookup done (src/bwtgap_hd.h:411-470)
    void expand()
    {
        pending = false;
        phase = PH_LOOP;
        const int i = ei - 1;
        uint64_t ck[4], cl[4];
        occ4_pair(*X, e.k - 1, e.l, ck, cl);
        bytes += lookup_bytes(*X, e.k - 1, e.l);
        const uint64_t occ = e.l - e.k + 1;
        bool allow_diff = true, allow_M = true;
        if (i > 0) {
            const int ii = i - (len - o.seed_len);
            if (W[(size_t)i - 1].bid > m - 1)
                allow_diff = false;
            else if (W[(size_t)i - 1].bid == m - 1 && W[(size_t)i].bid == m - 1 && W[(size_t)i - 1].w == W[(size_t)i].w)
                allow_M = false;
            if (seeding && ii > 0) {
                if (SW[(size_t)ii - 1].bid > m_seed - 1)
                    allow_diff = false;
                else if (SW[(size_t)ii - 1].bid == m_seed - 1 && SW[(size_t)ii].bid == m_seed - 1 &&
                         SW[(size_t)ii - 1].w == SW[(size_t)ii].w)
                    allow_M = false;
            }
        }
        const int gaps = (o.mode & MODE_LOGGAP) ? int_log2((uint32_t)(e.n_gape + e.n_gapo)) / 2 + 1 : e.n_gapo + e.n_gape;
        if (allow_diff && i >= o.indel_end_skip + gaps && len - i >= o.indel_end_skip + gaps) {
            if (e.state == ST_M) {
                if (e.n_gapo < o.max_gapo) {
                    PUSH_OR_FINISH(i, e.k, e.l, e.n_mm, e.n_gapo + 1, e.n_gape, e.n_ins + 1, e.n_del, ST_I, 1);
                    for (int c = 0; c < 4; ++c) {
                        const uint64_t k = X->L2[c] + ck[c] + 1, l = X->L2[c] + cl[c];
                        if (k <= l) PUSH_OR_FINISH(i + 1, k, l, e.n_mm, e.n_gapo + 1, e.n_gape, e.n_ins, e.n_del + 1, ST_D, 1);
                    }
                }
            } else if (e.state == ST_I) {
                if (e.n_gape < o.max_gape)
                    PUSH_OR_FINISH(i, e.k, e.l, e.n_mm, e.n_gapo, e.n_gape + 1, e.n_ins + 1, e.n_del, ST_I, 1);
            } else if (e.state == ST_D) {
                if (e.n_gape < o.max_gape && (e.n_gape + e.n_gapo < max_diff || occ < (uint64_t)(int64_t)o.max_del_occ)) {
                    for (int c = 0; c < 4; ++c) {
                        const uint64_t k = X->L2[c] + ck[c] + 1, l = X->L2[c] + cl[c];
                        if (k <= l) PUSH_OR_FINISH(i + 1, k, l, e.n_mm, e.n_gapo, e.n_gape + 1, e.n_ins, e.n_del + 1, ST_D, 1);
                    }
                }
            }
        }
        const uint8_t qi = seq[i];
        if (allow_diff && allow_M) {
            for (int j = 1; j <= 4; ++j) {
                const int c = (qi + j) & 3;
                const int is_mm = (j != 4 || qi > 3);
                const uint64_t k = X->L2[c] + ck[c] + 1, l = X->L2[c] + cl[c];
                if (k <= l) PUSH_OR_FINISH(i, k, l, e.n_mm + is_mm, e.n_gapo, e.n_gape, e.n_ins, e.n_del, ST_M, is_mm);
            }
        } else if (qi < 4) {
            const int c = qi & 3;
            const uint64_t k = X->L2[c] + ck[c] + 1, l = X->L2[c] + cl[c];
            if (k <= l) PUSH_OR_FINISH(i, k, l, e.n_mm, e.n_gapo, e.n_gape, e.n_ins, e.n_del, ST_M, 0);
        }
    }
#undef PUSH_OR_FINISH

    static int int_log2(uint32_t v) { return v ? 31 - __builtin_clz(v) : 0; }
};

struct ReadOut {
    int32_t rid;
    int32_t n;
    uint64_t off;  // into the thread's record array
};

}  // namespace

extern "C" int galn_cpu_search(const uint32_t *bwt, const galn_bwt_meta_t *meta, const galn_hip_reads *reads,
                               const int32_t *list, int32_t n_list, galn_hip_out *out, int32_t n_threads,
                               int32_t interleave)
{
    if (!bwt || !meta || !reads || !out) return GALN_HIP_E_ARG;
    const int n = list ? n_list : reads->n_reads;
    if (n < 0 || (n > 0 && (!out->n_aln || !out->rec_off))) return GALN_HIP_E_ARG;
    if (reads->n_reads > 0 && (!reads->query || !reads->qoff || !reads->qlen || !reads->max_diff || !reads->batch))
        return GALN_HIP_E_ARG;
    if (reads->n_batches < 1 || !reads->bopts || !reads->n_stacks) return GALN_HIP_E_ARG;
    for (int b = 0; b < reads->n_batches; ++b)
        if (reads->bopts[b].seed_len < 0) return GALN_HIP_E_ARG;
    for (int t = 0; t < n; ++t) {
        const int i = list ? list[t] : t;
        if (i < 0 || i >= reads->n_reads || reads->batch[i] < 0 || reads->batch[i] >= reads->n_batches || reads->qlen[i] < 0)
            return GALN_HIP_E_ARG;
    }
    Idx X;
    X.bwt = bwt;
    X.primary = meta->primary;
    X.seq_len = meta->seq_len;
    for (int c = 0; c < 5; ++c) X.L2[c] = meta->L2[c];
    if (n_threads <= 0) n_threads = (int)std::max(1u, std::thread::hardware_concurrency());
    n_threads = std::max(1, std::min(n_threads, std::max(1, n / 64 + 1)));
    if (interleave <= 0) {
        // reads advanced in turn hide the index's load latency: little of it while the .bwt body
        // stays in the caches, where fewer reads in flight keep each read's state closer (200k
        // aDNA reads, 8 threads: 2 best at a 5 MB body, 4 at 20 MB, 8 at 40 MB)
        const uint64_t body = (meta->seq_len + 127) / 128 * 64;
        interleave = body < (8ull << 20) ? 2 : body < (32ull << 20) ? 4 : 8;
    }
    interleave = std::max(1, std::min(interleave, 64));

    std::atomic<int> next{0};
    constexpr int CLAIM = 64;
    std::vector<std::vector<galn_aln_rec_t>> trecs((size_t)n_threads);
    std::vector<std::vector<ReadOut>> touts((size_t)n_threads);
    auto worker = [&](int tid) {
        std::vector<Search> slot((size_t)interleave);
        for (Search &S : slot) S.want_trace = out->trace != nullptr;
        std::vector<int> rid((size_t)interleave, -1);
        int lo = 0, hi = 0;  // claimed positions not yet started
        auto start = [&](int s) -> bool {
            for (;;) {
                if (lo == hi) {
                    lo = next.fetch_add(CLAIM);
                    if (lo >= n) {
                        lo = hi = n;
                        return false;
                    }
                    hi = std::min(n, lo + CLAIM);
                }
                const int pos = lo++;
                const int i = list ? list[pos] : pos;
                const int b = reads->batch[i];
                rid[(size_t)s] = pos;
                slot[(size_t)s].begin(X, reads->query + reads->qoff[i], reads->qlen[i], reads->bopts[b], reads->n_stacks[b],
                                      reads->max_diff[i]);
                return true;
            }
        };
        auto done = [&](int s) {
            Search &S = slot[(size_t)s];
            const int pos = rid[(size_t)s];
            ReadOut ro;
            ro.rid = pos;
            ro.n = S.result;
            ro.off = trecs[(size_t)tid].size();
            if (S.result > 0) trecs[(size_t)tid].insert(trecs[(size_t)tid].end(), S.recs.begin(), S.recs.end());
            touts[(size_t)tid].push_back(ro);
            out->n_aln[pos] = S.result;
            if (out->n_pops) out->n_pops[pos] = S.pops;
            if (out->trace) out->trace[pos] = S.trace;
            if (out->occ_bytes) out->occ_bytes[pos] = S.bytes;
            rid[(size_t)s] = -1;
        };
        int active = 0;
        for (int s = 0; s < interleave; ++s)
            if (start(s)) ++active;
        while (active > 0) {
            for (int s = 0; s < interleave; ++s) {
                if (rid[(size_t)s] < 0) continue;
                Search &S = slot[(size_t)s];
                for (;;) {
                    if (S.run() == Search::LOOKUP) {
                        prefetch_lookup(X, S.qa, S.qb);
                        break;
                    }
                    done(s);
                    if (!start(s)) {
                        --active;
                        break;
                    }
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < n_threads; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto &t : th) t.join();

    // records in read (list) order
    std::vector<std::pair<int, const ReadOut *>> where((size_t)n, {-1, nullptr});
    for (int t = 0; t < n_threads; ++t)
        for (const ReadOut &ro : touts[(size_t)t]) where[(size_t)ro.rid] = {t, &ro};
    uint64_t total = 0;
    for (int p = 0; p < n; ++p) {
        out->rec_off[p] = total;
        const ReadOut *ro = where[(size_t)p].second;
        if (ro && ro->n > 0) total += (uint64_t)ro->n;
    }
    out->n_rec = total;
    if (total > out->rec_cap || (total > 0 && !out->rec)) return GALN_HIP_E_RECCAP;
    for (int p = 0; p < n; ++p) {
        const ReadOut *ro = where[(size_t)p].second;
        if (ro && ro->n > 0)
            memcpy(out->rec + out->rec_off[p], trecs[(size_t)where[(size_t)p].first].data() + ro->off,
                   (size_t)ro->n * sizeof(galn_aln_rec_t));
    }
    return 0;
}
