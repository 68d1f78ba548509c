/* galn_oracle.c -- CPU oracle for BWA-aln's backtracking search.
 *
 * TEST INFRASTRUCTURE ONLY (see galn_oracle.h): the checker the HIP search is
 * compared with, never part of the product. Compile with -ffp-contract=off so
 * orc_cal_maxdiff rounds as the reference does (reference CMakeLists.txt:22-23).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   orc_occ / orc_occ4 / orc_2occ4  brute-force counts over the decoded BWT of
 *                                   small indexes, at every position
 *   orc_match_exact                 SA-interval size == occurrences of the
 *                                   pattern in the text, by substring scan
 *   orc_cal_width                   its definition (greedy restarts), brute force
 *   orc_cal_maxdiff                 the closed-form Poisson tail and a table
 *   orc_search_read / _batch        golden .sai, per-read pops and trace hashes
 *                                   written by the reference's own C search
 *                                   (oracle/_ref, built from src/aln_host.c and
 *                                   src/bwtgap_hd.h), plus the max_diff = 0 ==
 *                                   exact-match and -N == brute-force k-difference
 *                                   enumeration reductions on tiny texts
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "galn_oracle.h"

#define NEG1 ((uint64_t)-1)

static void *xrealloc(void *p, size_t n)
{
    void *q = realloc(p, n ? n : 1);
    if (!q) { fprintf(stderr, "[oracle] out of memory (%zu bytes)\n", n); abort(); }
    return q;
}

/* ------------------------------------------------------------------------ */
/* Occurrence counting over the BWA .bwt body.
 * Layout (reference src/bwtsearch.h:33-34, src/index/idxcommon.c:144-184): the
 * symbols of the $-less BWT in 128-symbol blocks of 16 u32 words: 4 x u64
 * exclusive per-base counts, then 8 words of 16 symbols, first symbol in
 * bits 31-30. Position p of the $-inclusive BWT lives at p - (p >= primary). */

/* number of symbols equal to c among the first `nsym` (1..16) symbols of word x */
static inline uint32_t word_count(uint32_t x, int c, int nsym)
{
    uint32_t hi = (x >> 1) & 0x55555555u, lo = x & 0x55555555u, sel;
    uint32_t valid = 0x55555555u & ~((nsym >= 16) ? 0u : ((1u << (32 - 2 * nsym)) - 1u));
    switch (c) {
    case 0: sel = ~hi & ~lo; break;
    case 1: sel = ~hi & lo; break;
    case 2: sel = hi & ~lo; break;
    default: sel = hi & lo; break;
    }
    return (uint32_t)__builtin_popcount(sel & valid);
}

/* Occ of all four bases in B[0..p'] for a $-less position p' (block count + in-block scan) */
static inline void block_occ4(const uint32_t *bwt, uint64_t pp, uint64_t cnt[4])
{
    const uint32_t *blk = bwt + ((pp >> 7) << 4);
    const uint64_t *base = (const uint64_t *)blk;
    const uint32_t *w = blk + 8;
    int off = (int)(pp & 127), last = off >> 4, j, c;
    for (c = 0; c < 4; ++c) {
        uint64_t n = base[c];
        for (j = 0; j < last; ++j) n += word_count(w[j], c, 16);
        n += word_count(w[last], c, (off & 15) + 1);
        cnt[c] = n;
    }
}

/* bwt_occ: src/bwtsearch.h:46-70 (upstream bwt.c bwt_occ) */
uint64_t orc_occ(const uint32_t *bwt, const orc_meta_t *m, uint64_t k, int c)
{
    uint64_t cnt[4];
    if (k == m->seq_len) return m->L2[c + 1] - m->L2[c];
    if (k == NEG1) return 0;
    block_occ4(bwt, k - (k >= m->primary), cnt);
    return cnt[c];
}

/* bwt_occ4: src/bwtgap_hd.h:135-155 */
void orc_occ4(const uint32_t *bwt, const orc_meta_t *m, uint64_t k, uint64_t cnt[4])
{
    if (k == NEG1) { cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0; return; }
    block_occ4(bwt, k - (k >= m->primary), cnt);
}

/* bwt_2occ4: src/bwtgap_hd.h:158-192. The shared-block shortcut there gives the
 * same integers as two independent counts, so they are computed independently. */
void orc_2occ4(const uint32_t *bwt, const orc_meta_t *m, uint64_t k, uint64_t l,
               uint64_t cntk[4], uint64_t cntl[4])
{
    orc_occ4(bwt, m, k, cntk);
    orc_occ4(bwt, m, l, cntl);
}

/* algorithmic bytes of one occurrence query at positions (a, b): 64 per distinct
 * 128-symbol block; position -1 needs none */
static inline uint64_t query_bytes(const orc_meta_t *m, uint64_t a, uint64_t b)
{
    uint64_t ba, bb;
    if (a == NEG1 && b == NEG1) return 0;
    if (a == NEG1) return 64;
    if (b == NEG1) return 64;
    ba = (a - (a >= m->primary)) >> 7;
    bb = (b - (b >= m->primary)) >> 7;
    return ba == bb ? 64 : 128;
}

/* the symbol at position x of the $-less BWT (upstream bwt.h bwt_B0) */
static inline int bwt_sym(const uint32_t *bwt, uint64_t x)
{
    const uint32_t w = bwt[((x >> 7) << 4) + 8 + ((x & 127) >> 4)];
    return (int)((w >> (30 - 2 * (x & 15))) & 3u);
}

/* bwt_invPsi: the row one text position to the left (upstream bwt.c:53-59) */
uint64_t orc_inv_psi(const uint32_t *bwt, const orc_meta_t *m, uint64_t k)
{
    uint64_t x;
    int c;
    if (k == m->primary) return 0;
    x = k - (k > m->primary);
    c = bwt_sym(bwt, x);
    return m->L2[c] + orc_occ(bwt, m, k, c);
}

/* bwt_sa: the text position of row k (upstream bwt.c:88-96), with the samples of rows intv, 2 intv,
 * ... in samples[0], samples[1], ... and row 0 (the $ row) counted as -1, as bwt_restore_sa sets it */
uint64_t orc_sa(const uint32_t *bwt, const orc_meta_t *m, const uint64_t *samples, uint64_t intv, uint64_t k)
{
    uint64_t sa = 0, mask = intv - 1;
    while (k & mask) {
        ++sa;
        k = orc_inv_psi(bwt, m, k);
    }
    return sa + (k == 0 ? NEG1 : samples[k / intv - 1]);
}

/* bwa_sa2pos (upstream bwase.c:112-123, bntseq.h bns_depos): the forward-strand position of the
 * first base of a hit of ref_len bases at row k; *strand = 1 forward, 0 reverse; -1 across the
 * forward/reverse boundary */
uint64_t orc_sa2pos(const uint32_t *bwt, const orc_meta_t *m, const uint64_t *samples, uint64_t intv, uint64_t k,
                    int ref_len, int *strand)
{
    const uint64_t l_pac = m->seq_len >> 1;
    uint64_t pos = orc_sa(bwt, m, samples, intv, k);
    int is_rev;
    *strand = 0;
    if (pos < l_pac && l_pac < pos + (uint64_t)(int64_t)ref_len) return NEG1;
    is_rev = pos >= l_pac;
    if (is_rev) pos = (l_pac << 1) - 1 - pos;
    *strand = !is_rev;
    if (is_rev) pos = pos + 1 < (uint64_t)(int64_t)ref_len ? 0 : pos - (uint64_t)(int64_t)ref_len + 1;
    return pos;
}

/* bwt_2occ for one base: src/bwtsearch.h:73-108 */
static inline void occ2(const uint32_t *bwt, const orc_meta_t *m, uint64_t k, uint64_t l, int c,
                        uint64_t *ok, uint64_t *ol)
{
    *ok = orc_occ(bwt, m, k, c);
    *ol = orc_occ(bwt, m, l, c);
}

/* bwt_match_exact: src/bwtsearch.h:114-133 */
uint64_t orc_match_exact(const uint32_t *bwt, const orc_meta_t *m, int len, const uint8_t *str,
                         uint64_t *sa_begin, uint64_t *sa_end)
{
    uint64_t k = 0, l = m->seq_len, ok, ol;
    int i;
    for (i = len - 1; i >= 0; --i) {
        int c = str[i];
        if (c > 3) return 0;
        occ2(bwt, m, k - 1, l, c, &ok, &ol);
        k = m->L2[c] + ok + 1;
        l = m->L2[c] + ol;
        if (k > l) break;
    }
    if (k > l) return 0;
    if (sa_begin) *sa_begin = k;
    if (sa_end) *sa_end = l;
    return l - k + 1;
}

typedef struct { uint64_t bytes, calls; } occ_ctr_t;

/* bwt_match_exact_alt: src/bwtgap_hd.h:195-211. Returns the int the reference
 * returns, (int)(l - k + 1), which the caller tests for non-zero. */
static int match_exact_alt(const uint32_t *bwt, const orc_meta_t *m, int len, const uint8_t *str,
                           uint64_t *k0, uint64_t *l0, occ_ctr_t *ctr)
{
    uint64_t k = *k0, l = *l0, ok, ol;
    int i;
    for (i = len - 1; i >= 0; --i) {
        int c = str[i];
        if (c > 3) return 0;
        ctr->bytes += query_bytes(m, k - 1, l); ctr->calls++;
        occ2(bwt, m, k - 1, l, c, &ok, &ol);
        k = m->L2[c] + ok + 1;
        l = m->L2[c] + ol;
        if (k > l) return 0;
    }
    *k0 = k; *l0 = l;
    return (int)(uint32_t)(l - k + 1);
}

/* bwt_cal_width with the complement undone on the fly: src/bwtgap_hd.h:216-242 */
static int cal_width_ctr(const uint32_t *bwt, const orc_meta_t *m, int len, const uint8_t *str,
                         int comp, orc_width_t *width, occ_ctr_t *ctr)
{
    uint64_t k = 0, l = m->seq_len, ok, ol;
    int i, bid = 0;
    for (i = 0; i < len; ++i) {
        int c = str[i];
        if (comp) c = c > 3 ? 4 : 3 - c;
        if (c < 4) {
            if (ctr) { ctr->bytes += query_bytes(m, k - 1, l); ctr->calls++; }
            occ2(bwt, m, k - 1, l, c, &ok, &ol);
            k = m->L2[c] + ok + 1;
            l = m->L2[c] + ol;
        }
        if (k > l || c > 3) { k = 0; l = m->seq_len; ++bid; }
        width[i].w = l - k + 1;
        width[i].bid = bid;
        width[i].pad_ = 0;
    }
    width[len].w = 0;
    width[len].bid = ++bid;
    width[len].pad_ = 0;
    return bid;
}

int orc_cal_width(const uint32_t *bwt, const orc_meta_t *m, int len, const uint8_t *str,
                  int comp, orc_width_t *width)
{
    return cal_width_ctr(bwt, m, len, str, comp, width, NULL);
}

/* ------------------------------------------------------------------------ */
/* Options. */

/* bwa_cal_maxdiff: src/aln_host.c:9-21. x *= k wraps as a 32-bit int. */
int orc_cal_maxdiff(int l, double err, double thres)
{
    double elambda = exp(-l * err);
    double sum, y = 1.0;
    int k, x = 1;
    for (k = 1, sum = elambda; k < 1000; ++k) {
        y *= l * err;
        x = (int)((unsigned)x * (unsigned)k);
        sum += elambda * y / x;
        if (1.0 - sum < thres) return k;
    }
    return 2;
}

/* aln_batch_opt: src/aln_host.c:23-38 */
int orc_batch_opt(int32_t s_mm, int32_t s_gapo, int32_t s_gape, int32_t mode,
                  int32_t indel_end_skip, int32_t max_del_occ, int32_t max_entries,
                  float fnr, int32_t max_diff, int32_t max_gapo, int32_t max_gape,
                  int32_t max_seed_diff, int32_t seed_len, int32_t max_top2,
                  int32_t max_len, orc_opt_t *b)
{
    b->s_mm = s_mm; b->s_gapo = s_gapo; b->s_gape = s_gape; b->mode = mode;
    b->indel_end_skip = indel_end_skip; b->max_del_occ = max_del_occ; b->max_entries = max_entries;
    b->max_diff = max_diff; b->max_gapo = max_gapo; b->max_gape = max_gape;
    b->max_seed_diff = max_seed_diff; b->seed_len = seed_len; b->max_top2 = max_top2;
    if (fnr > 0.0) b->max_diff = orc_cal_maxdiff(max_len, 0.02, fnr);
    if (b->max_diff < b->max_gapo) b->max_gapo = b->max_diff;
    return (b->max_diff + 1) * b->s_mm + (b->max_gapo + 1) * b->s_gapo + (b->max_gape + 1) * b->s_gape;
}

/* aln_read_maxdiff: src/aln_host.c:40-43 */
int orc_read_maxdiff(float fnr, const orc_opt_t *b, int len)
{
    return (fnr > 0.0) ? orc_cal_maxdiff(len, 0.02, fnr) : b->max_diff;
}

/* ------------------------------------------------------------------------ */
/* The priority stack: one growable LIFO array per score bucket (the shape of
 * upstream BWA's gap_stack_t); pop takes the top of the lowest non-empty
 * bucket -- the order of src/bwtgap_hd.h:264-306. */

typedef struct {
    uint64_t k, l;
    uint32_t info;          /* score<<21 | i */
    int32_t  last_diff_pos;
    uint8_t  n_mm, n_gapo, n_gape, state;
    uint16_t n_ins, n_del;
} ent_t;

typedef struct { ent_t *a; int n, m; } bucket_t;

struct orc_ws {
    bucket_t *b; int nb;
    orc_width_t *w, *sw; int wcap, swcap;
    orc_aln_t *aln; int acap;
};

orc_ws_t *orc_ws_new(void)
{
    orc_ws_t *ws = (orc_ws_t *)calloc(1, sizeof(orc_ws_t));
    if (!ws) abort();
    return ws;
}

void orc_ws_free(orc_ws_t *ws)
{
    int i;
    if (!ws) return;
    for (i = 0; i < ws->nb; ++i) free(ws->b[i].a);
    free(ws->b); free(ws->w); free(ws->sw); free(ws->aln);
    free(ws);
}

typedef struct {
    bucket_t *b;
    int n_stacks, best, n_entries, peak;
    uint64_t pops, trace;
} stk_t;

#define TR_INIT 1469598103934665603ULL
#define TR_MIX(h, v) (((h) ^ (uint64_t)(v)) * 1099511628211ULL)

/* trace hash of one pop: src/bwtgap_hd.h:83-100 (the reference's pop-order check) */
static inline uint64_t trace_pop(uint64_t h, const ent_t *e, int n_entries)
{
    h = TR_MIX(h, e->info);
    h = TR_MIX(h, e->k);
    h = TR_MIX(h, e->l);
    h = TR_MIX(h, (uint32_t)e->n_mm | (uint32_t)e->n_gapo << 8 | (uint32_t)e->n_gape << 16 | (uint32_t)e->state << 24);
    h = TR_MIX(h, (uint64_t)e->n_ins | (uint64_t)e->n_del << 16);
    h = TR_MIX(h, (uint64_t)(uint32_t)e->last_diff_pos);
    h = TR_MIX(h, (uint64_t)(uint32_t)n_entries);
    return h;
}

static inline int score_of(int mm, int go, int ge, const orc_opt_t *p)
{
    return mm * p->s_mm + go * p->s_gapo + ge * p->s_gape;
}

/* gap_push: src/bwtgap_hd.h:264-286. Returns 0 or ORC_ERR_SCORE. */
static int push(stk_t *s, int i, uint64_t k, uint64_t l, int n_mm, int n_gapo, int n_gape,
                int n_ins, int n_del, int state, int is_diff, const orc_opt_t *opt)
{
    int score = score_of(n_mm, n_gapo, n_gape, opt);
    bucket_t *q;
    ent_t *p;
    if (score < 0 || score >= s->n_stacks) return ORC_ERR_SCORE;
    q = &s->b[score];
    if (q->n == q->m) {
        q->m = q->m ? q->m * 2 : 16;
        q->a = (ent_t *)xrealloc(q->a, (size_t)q->m * sizeof(ent_t));
    }
    p = &q->a[q->n++];
    p->info = (uint32_t)score << 21 | (uint32_t)i;
    p->k = k; p->l = l;
    p->n_mm = (uint8_t)n_mm; p->n_gapo = (uint8_t)n_gapo; p->n_gape = (uint8_t)n_gape;
    p->n_ins = (uint16_t)n_ins; p->n_del = (uint16_t)n_del;
    p->state = (uint8_t)state;
    p->last_diff_pos = is_diff ? i : 0;
    ++s->n_entries;
    if (s->n_entries > s->peak) s->peak = s->n_entries;
    if (s->best > score) s->best = score;
    return 0;
}

/* gap_pop: src/bwtgap_hd.h:289-306 */
static void pop(stk_t *s, ent_t *e)
{
    bucket_t *q = &s->b[s->best];
    *e = q->a[--q->n];
    --s->n_entries;
    ++s->pops;
    s->trace = trace_pop(s->trace, e, s->n_entries);
    if (q->n == 0 && s->n_entries) {
        int i;
        for (i = s->best + 1; i < s->n_stacks; ++i)
            if (s->b[i].n) break;
        s->best = i;
    } else if (s->n_entries == 0) s->best = s->n_stacks;
}

/* gap_shadow: src/bwtgap_hd.h:309-320 (x is the int the reference passes) */
static void shadow(int x, uint64_t max, int last_diff_pos, orc_width_t *w)
{
    int i, j;
    uint64_t ux = (uint64_t)(int64_t)x;
    for (i = j = 0; i < last_diff_pos; ++i) {
        if (w[i].w > ux) w[i].w -= ux;
        else if (w[i].w == ux) {
            w[i].bid = 1;
            w[i].w = max - (uint64_t)(++j);
        }
    }
}

/* int_log2: src/bwtgap_hd.h:322-331 (floor(log2(v)), 0 for v == 0) */
static int int_log2(uint32_t v)
{
    int c = 0;
    while (v >>= 1) ++c;
    return c;
}

static uint64_t pack_rec(int n_mm, int n_gapo, int n_gape, int score, int n_ins, int n_del)
{
    /* bwt_aln1_t bit-fields, LSB first: src/bwtgap_hd.h:108-117, src/sai_format.h:36-45 */
    return ((uint64_t)(n_mm & 0xff)) | ((uint64_t)(n_gapo & 0xff) << 8) | ((uint64_t)(n_gape & 0xff) << 16)
         | ((uint64_t)(score & 0xfffff) << 24) | ((uint64_t)(n_ins & 0x3ff) << 44)
         | ((uint64_t)(n_del & 0x3ff) << 54);
}

#define PUSH(...) do { int r_ = push(&st, __VA_ARGS__, opt); if (r_) { ret = r_; goto done; } } while (0)

/* bwt_match_gap (src/bwtgap_hd.h:338-473) driven as aln_search_host drives it
 * (src/aln_host.c:141-201), without any capacity limit, so the result is the
 * reference's final (re-run-until-it-fits) result. */
int orc_search_read(const uint32_t *bwt, const orc_meta_t *bm, const orc_opt_t *bopt,
                    int n_stacks, int read_max_diff, int glob_seed_len,
                    int len, const uint8_t *seq, orc_ws_t *ws,
                    const orc_aln_t **aln_out, orc_stats_t *stats)
{
    orc_opt_t lo = *bopt, *opt = &lo;
    orc_width_t *width, *seed_width = NULL;
    occ_ctr_t ctr = {0, 0};
    stk_t st;
    int best_score, best_diff, max_diff, best_cnt = 0, n_aln = 0, j, nN, ret = 0;

    lo.max_diff = read_max_diff;
    lo.seed_len = glob_seed_len < len ? glob_seed_len : 0x7fffffff;
    if (ws->wcap < len + 1) { ws->wcap = len + 1; ws->w = (orc_width_t *)xrealloc(ws->w, (size_t)ws->wcap * sizeof(orc_width_t)); }
    if (len > glob_seed_len && ws->swcap < glob_seed_len + 1) {
        ws->swcap = glob_seed_len + 1;
        ws->sw = (orc_width_t *)xrealloc(ws->sw, (size_t)ws->swcap * sizeof(orc_width_t));
    }
    if (ws->nb < n_stacks) {
        ws->b = (bucket_t *)xrealloc(ws->b, (size_t)n_stacks * sizeof(bucket_t));
        memset(ws->b + ws->nb, 0, (size_t)(n_stacks - ws->nb) * sizeof(bucket_t));
        ws->nb = n_stacks;
    }
    width = ws->w;
    cal_width_ctr(bwt, bm, len, seq, 1, width, &ctr);
    if (len > glob_seed_len) {
        cal_width_ctr(bwt, bm, glob_seed_len, seq + (len - glob_seed_len), 1, ws->sw, &ctr);
        seed_width = ws->sw;
    }

    st.b = ws->b; st.n_stacks = n_stacks; st.best = n_stacks; st.n_entries = 0; st.peak = 0;
    st.pops = 0; st.trace = TR_INIT;
    for (j = 0; j < n_stacks; ++j) st.b[j].n = 0;

    best_score = score_of(opt->max_diff + 1, opt->max_gapo + 1, opt->max_gape + 1, opt);
    best_diff = opt->max_diff + 1; max_diff = opt->max_diff;
    (void)best_diff;

    for (j = nN = 0; j < len; ++j) if (seq[j] > 3) ++nN;
    if (nN > max_diff) { ret = 0; goto done; }

    PUSH(len, 0, bm->seq_len, 0, 0, 0, 0, 0, 0, 0);

    while (st.n_entries) {
        ent_t e;
        int i, m, m_seed = 0, hit_found, allow_diff, allow_M, tmp;
        uint64_t k, l, cnt_k[4], cnt_l[4], occ;

        if (st.n_entries > opt->max_entries) break;
        pop(&st, &e);
        k = e.k; l = e.l;
        i = (int)(e.info & 0xffff);
        if (!(opt->mode & ORC_MODE_NONSTOP) && e.info >> 21 > (uint32_t)(best_score + opt->s_mm)) break;

        m = max_diff - (e.n_mm + e.n_gapo);
        if (opt->mode & ORC_MODE_GAPE) m -= e.n_gape;
        if (m < 0) continue;
        if (seed_width) {
            m_seed = opt->max_seed_diff - (e.n_mm + e.n_gapo);
            if (opt->mode & ORC_MODE_GAPE) m_seed -= e.n_gape;
        }
        if (i > 0 && m < width[i - 1].bid) continue;

        hit_found = 0;
        if (i == 0) hit_found = 1;
        else if (m == 0 && (e.state == 0 || (opt->mode & ORC_MODE_GAPE) || e.n_gape == opt->max_gape)) {
            if (match_exact_alt(bwt, bm, i, seq, &k, &l, &ctr)) hit_found = 1;
            else continue;
        }

        if (hit_found) {
            int score = score_of(e.n_mm, e.n_gapo, e.n_gape, opt);
            int do_add = 1;
            if (n_aln == 0) {
                best_score = score;
                best_diff = e.n_mm + e.n_gapo;
                if (opt->mode & ORC_MODE_GAPE) best_diff += e.n_gape;
                if (!(opt->mode & ORC_MODE_NONSTOP))
                    max_diff = (best_diff + 1 > opt->max_diff) ? opt->max_diff : best_diff + 1;
            }
            /* int best_cnt += uint64: wraps modulo 2^32 as on every target the reference builds for */
            if (score == best_score) best_cnt = (int)(uint32_t)((uint64_t)(int64_t)best_cnt + (l - k + 1));
            else if (best_cnt > opt->max_top2) break;
            if (e.n_gapo) {
                for (j = 0; j != n_aln; ++j)
                    if (ws->aln[j].k == k && ws->aln[j].l == l) break;
                if (j < n_aln) do_add = 0;
            }
            if (do_add) {
                shadow((int)(uint32_t)(l - k + 1), bm->seq_len, e.last_diff_pos, width);
                if (n_aln == ws->acap) {
                    ws->acap = ws->acap ? ws->acap * 2 : 64;
                    ws->aln = (orc_aln_t *)xrealloc(ws->aln, (size_t)ws->acap * sizeof(orc_aln_t));
                }
                ws->aln[n_aln].packed = pack_rec(e.n_mm, e.n_gapo, e.n_gape, score, e.n_ins, e.n_del);
                ws->aln[n_aln].k = k; ws->aln[n_aln].l = l;
                ++n_aln;
            }
            continue;
        }

        --i;
        ctr.bytes += query_bytes(bm, k - 1, l); ctr.calls++;
        orc_2occ4(bwt, bm, k - 1, l, cnt_k, cnt_l);
        occ = l - k + 1;
        allow_diff = allow_M = 1;
        if (i > 0) {
            if (width[i - 1].bid > m - 1) allow_diff = 0;
            else if (width[i - 1].bid == m - 1 && width[i].bid == m - 1 && width[i - 1].w == width[i].w) allow_M = 0;
            if (seed_width) {
                int ii = i - (len - opt->seed_len);
                if (ii > 0) {
                    if (seed_width[ii - 1].bid > m_seed - 1) allow_diff = 0;
                    else if (seed_width[ii - 1].bid == m_seed - 1 && seed_width[ii].bid == m_seed - 1
                             && seed_width[ii - 1].w == seed_width[ii].w) allow_M = 0;
                }
            }
        }
        tmp = (opt->mode & ORC_MODE_LOGGAP) ? int_log2((uint32_t)(e.n_gape + e.n_gapo)) / 2 + 1 : e.n_gapo + e.n_gape;
        if (allow_diff && i >= opt->indel_end_skip + tmp && len - i >= opt->indel_end_skip + tmp) {
            if (e.state == 0) {
                if (e.n_gapo < opt->max_gapo) {
                    PUSH(i, k, l, e.n_mm, e.n_gapo + 1, e.n_gape, e.n_ins + 1, e.n_del, 1, 1);
                    for (j = 0; j != 4; ++j) {
                        uint64_t kk = bm->L2[j] + cnt_k[j] + 1, ll = bm->L2[j] + cnt_l[j];
                        if (kk <= ll) PUSH(i + 1, kk, ll, e.n_mm, e.n_gapo + 1, e.n_gape, e.n_ins, e.n_del + 1, 2, 1);
                    }
                }
            } else if (e.state == 1) {
                if (e.n_gape < opt->max_gape)
                    PUSH(i, k, l, e.n_mm, e.n_gapo, e.n_gape + 1, e.n_ins + 1, e.n_del, 1, 1);
            } else if (e.state == 2) {
                if (e.n_gape < opt->max_gape) {
                    if (e.n_gape + e.n_gapo < max_diff || occ < (uint64_t)(int64_t)opt->max_del_occ) {
                        for (j = 0; j != 4; ++j) {
                            uint64_t kk = bm->L2[j] + cnt_k[j] + 1, ll = bm->L2[j] + cnt_l[j];
                            if (kk <= ll) PUSH(i + 1, kk, ll, e.n_mm, e.n_gapo, e.n_gape + 1, e.n_ins, e.n_del + 1, 2, 1);
                        }
                    }
                }
            }
        }
        if (allow_diff && allow_M) {
            for (j = 1; j <= 4; ++j) {
                int c = (seq[i] + j) & 3;
                int is_mm = (j != 4 || seq[i] > 3);
                uint64_t kk = bm->L2[c] + cnt_k[c] + 1, ll = bm->L2[c] + cnt_l[c];
                if (kk <= ll) PUSH(i, kk, ll, e.n_mm + is_mm, e.n_gapo, e.n_gape, e.n_ins, e.n_del, 0, is_mm);
            }
        } else if (seq[i] < 4) {
            int c = seq[i] & 3;
            uint64_t kk = bm->L2[c] + cnt_k[c] + 1, ll = bm->L2[c] + cnt_l[c];
            if (kk <= ll) PUSH(i, kk, ll, e.n_mm, e.n_gapo, e.n_gape, e.n_ins, e.n_del, 0, 0);
        }
    }
    ret = n_aln;
done:
    if (stats) {
        stats->pops = st.pops;
        stats->trace = st.trace;
        stats->occ_bytes = ctr.bytes;
        stats->occ_calls = ctr.calls;
        stats->peak = st.peak;
    }
    if (aln_out) *aln_out = ws->aln;
    return ret;
}
#undef PUSH

/* ------------------------------------------------------------------------ */
/* Batches on host threads. */

struct orc_result {
    int n;
    int32_t *n_aln;
    uint64_t *rec_off;
    orc_aln_t *rec;
    uint64_t nrec;
    uint64_t *pops, *trace, *occ_bytes, *occ_calls;
    int32_t *peak;
};

typedef struct { orc_aln_t *a; int n; } rbuf_t;

#ifdef _OPENMP
#include <omp.h>
#endif

orc_result_t *orc_search_batch(const uint32_t *bwt, const orc_meta_t *m,
                               int n, const uint8_t *query, const uint64_t *qoff,
                               const int32_t *qlen, const int32_t *max_diff, const int32_t *bid,
                               const orc_opt_t *bopts, const int32_t *n_stacks, int glob_seed_len,
                               int n_threads)
{
    orc_result_t *r = (orc_result_t *)calloc(1, sizeof(orc_result_t));
    rbuf_t *per = (rbuf_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(rbuf_t));
    int i;
    uint64_t tot = 0;
    r->n = n;
    r->n_aln = (int32_t *)xrealloc(NULL, (size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    r->rec_off = (uint64_t *)xrealloc(NULL, (size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    r->pops = (uint64_t *)xrealloc(NULL, (size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    r->trace = (uint64_t *)xrealloc(NULL, (size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    r->occ_bytes = (uint64_t *)xrealloc(NULL, (size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    r->occ_calls = (uint64_t *)xrealloc(NULL, (size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    r->peak = (int32_t *)xrealloc(NULL, (size_t)(n > 0 ? n : 1) * sizeof(int32_t));
#ifdef _OPENMP
    if (n_threads < 1) n_threads = omp_get_max_threads();
    #pragma omp parallel num_threads(n_threads)
#endif
    {
        orc_ws_t *ws = orc_ws_new();
#ifdef _OPENMP
        #pragma omp for schedule(dynamic, 16)
#endif
        for (i = 0; i < n; ++i) {
            const orc_aln_t *a;
            orc_stats_t s;
            int b = bid[i];
            int na = orc_search_read(bwt, m, &bopts[b], n_stacks[b], max_diff[i], glob_seed_len,
                                     qlen[i], query + qoff[i], ws, &a, &s);
            r->n_aln[i] = na;
            r->pops[i] = s.pops; r->trace[i] = s.trace;
            r->occ_bytes[i] = s.occ_bytes; r->occ_calls[i] = s.occ_calls; r->peak[i] = s.peak;
            if (na > 0) {
                per[i].a = (orc_aln_t *)xrealloc(NULL, (size_t)na * sizeof(orc_aln_t));
                memcpy(per[i].a, a, (size_t)na * sizeof(orc_aln_t));
                per[i].n = na;
            }
        }
        orc_ws_free(ws);
    }
    (void)n_threads;
    for (i = 0; i < n; ++i) { r->rec_off[i] = tot; tot += (uint64_t)per[i].n; }
    r->nrec = tot;
    r->rec = (orc_aln_t *)xrealloc(NULL, (size_t)(tot ? tot : 1) * sizeof(orc_aln_t));
    for (i = 0; i < n; ++i) {
        if (per[i].n) memcpy(r->rec + r->rec_off[i], per[i].a, (size_t)per[i].n * sizeof(orc_aln_t));
        free(per[i].a);
    }
    free(per);
    return r;
}

uint64_t orc_result_nrec(const orc_result_t *r) { return r->nrec; }

void orc_result_copy(const orc_result_t *r, int32_t *n_aln, uint64_t *rec_off, orc_aln_t *rec,
                     uint64_t *pops, uint64_t *trace, uint64_t *occ_bytes, uint64_t *occ_calls,
                     int32_t *peak)
{
    size_t n = (size_t)r->n;
    if (n_aln) memcpy(n_aln, r->n_aln, n * sizeof(int32_t));
    if (rec_off) memcpy(rec_off, r->rec_off, n * sizeof(uint64_t));
    if (rec && r->nrec) memcpy(rec, r->rec, (size_t)r->nrec * sizeof(orc_aln_t));
    if (pops) memcpy(pops, r->pops, n * sizeof(uint64_t));
    if (trace) memcpy(trace, r->trace, n * sizeof(uint64_t));
    if (occ_bytes) memcpy(occ_bytes, r->occ_bytes, n * sizeof(uint64_t));
    if (occ_calls) memcpy(occ_calls, r->occ_calls, n * sizeof(uint64_t));
    if (peak) memcpy(peak, r->peak, n * sizeof(int32_t));
}

void orc_result_free(orc_result_t *r)
{
    if (!r) return;
    free(r->n_aln); free(r->rec_off); free(r->rec); free(r->pops); free(r->trace);
    free(r->occ_bytes); free(r->occ_calls); free(r->peak);
    free(r);
}
