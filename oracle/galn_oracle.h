/* galn_oracle.h -- CPU oracle for BWA-aln's FM-index backtracking search.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the search that
 * G-ALN 0.2.7 runs on its CPU path (reference src/bwtgap_hd.h,
 * src/bwtsearch.h, src/aln_host.c). It exists to check the HIP search
 * against; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it. The product library never links or calls it.
 *
 * Every function cites the reference lines it follows. The data structures
 * are this file's own (per-bucket growable arrays for the priority stack,
 * per-symbol-pair popcounts for Occ); only the observable behaviour --
 * records, their order, pop count, trace hash -- is meant to be the same.
 */
#ifndef GALN_ORACLE_H
#define GALN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* index descriptor: same fields and order as bwtgpu_meta_t (src/bwtsearch.h:26-31) */
typedef struct {
    uint64_t primary;
    uint64_t L2[5];
    uint64_t seq_len;
    uint64_t bwt_size;
} orc_meta_t;

/* search options: same fields and order as bwtgap_opt_t (src/bwtgap_hd.h:52-59) */
typedef struct {
    int32_t s_mm, s_gapo, s_gape;
    int32_t mode;
    int32_t indel_end_skip, max_del_occ, max_entries;
    int32_t max_diff, max_gapo, max_gape;
    int32_t max_seed_diff, seed_len;
    int32_t max_top2;
} orc_opt_t;

/* bwt_width_t (src/bwtgap_hd.h:46-49) */
typedef struct {
    uint64_t w;
    int32_t bid;
    int32_t pad_;
} orc_width_t;

/* one .sai record, 24 bytes (src/sai_format.h:30-33) */
typedef struct {
    uint64_t packed;
    uint64_t k, l;
} orc_aln_t;

#define ORC_MODE_GAPE     0x01
#define ORC_MODE_COMPREAD 0x02
#define ORC_MODE_LOGGAP   0x04
#define ORC_MODE_NONSTOP  0x10

#define ORC_ERR_SCORE (-3)

/* ---- occurrence primitives ---- */
uint64_t orc_occ(const uint32_t *bwt, const orc_meta_t *m, uint64_t k, int c);
void     orc_occ4(const uint32_t *bwt, const orc_meta_t *m, uint64_t k, uint64_t cnt[4]);
void     orc_2occ4(const uint32_t *bwt, const orc_meta_t *m, uint64_t k, uint64_t l,
                   uint64_t cntk[4], uint64_t cntl[4]);
uint64_t orc_match_exact(const uint32_t *bwt, const orc_meta_t *m, int len, const uint8_t *str,
                         uint64_t *sa_begin, uint64_t *sa_end);
int      orc_cal_width(const uint32_t *bwt, const orc_meta_t *m, int len, const uint8_t *str,
                       int comp, orc_width_t *width);

/* bwt_invPsi and bwt_sa (upstream bwt.c:53-59, :88-96): LF walk to a sampled row */
uint64_t orc_inv_psi(const uint32_t *bwt, const orc_meta_t *m, uint64_t k);
uint64_t orc_sa(const uint32_t *bwt, const orc_meta_t *m, const uint64_t *samples, uint64_t intv, uint64_t k);
/* bwa_sa2pos (upstream bwase.c:112-123) */
uint64_t orc_sa2pos(const uint32_t *bwt, const orc_meta_t *m, const uint64_t *samples, uint64_t intv, uint64_t k,
                    int ref_len, int *strand);

/* ---- options ---- */
int orc_cal_maxdiff(int l, double err, double thres);
/* per-batch options from the .sai header fields; returns n_stacks */
int orc_batch_opt(int32_t s_mm, int32_t s_gapo, int32_t s_gape, int32_t mode,
                  int32_t indel_end_skip, int32_t max_del_occ, int32_t max_entries,
                  float fnr, int32_t max_diff, int32_t max_gapo, int32_t max_gape,
                  int32_t max_seed_diff, int32_t seed_len, int32_t max_top2,
                  int32_t max_len, orc_opt_t *out);
int orc_read_maxdiff(float fnr, const orc_opt_t *b, int len);

/* ---- one read, uncapped ---- */
typedef struct orc_ws orc_ws_t;
orc_ws_t *orc_ws_new(void);
void      orc_ws_free(orc_ws_t *ws);

typedef struct {
    uint64_t pops;       /* popped entries */
    uint64_t trace;      /* FNV-1a over every popped entry */
    uint64_t occ_bytes;  /* algorithmic bytes: 64 per distinct Occ block per occurrence query */
    uint64_t occ_calls;  /* occurrence queries (cal_width steps + expansions + exact-tail steps) */
    int32_t  peak;       /* largest number of entries held at once */
} orc_stats_t;

/* Searches one read (query = complement(nt4(reverse(read))), 0..3, >3 = N).
 * bopt: batch options; max_diff / seed_len are applied per read as the
 * reference's aln_search_host does (src/aln_host.c:141-153).
 * Returns n_aln (records left in *aln, valid until the next call on ws) or
 * ORC_ERR_SCORE. */
int orc_search_read(const uint32_t *bwt, const orc_meta_t *m, const orc_opt_t *bopt,
                    int n_stacks, int read_max_diff, int glob_seed_len,
                    int len, const uint8_t *query, orc_ws_t *ws,
                    const orc_aln_t **aln, orc_stats_t *stats);

/* ---- a batch of reads (SoA, as the HIP library takes them), on host threads ---- */
typedef struct orc_result orc_result_t;
orc_result_t *orc_search_batch(const uint32_t *bwt, const orc_meta_t *m,
                               int n, const uint8_t *query, const uint64_t *qoff,
                               const int32_t *qlen, const int32_t *max_diff, const int32_t *bid,
                               const orc_opt_t *bopts, const int32_t *n_stacks, int glob_seed_len,
                               int n_threads);
uint64_t orc_result_nrec(const orc_result_t *r);
/* copies out: n_aln[n], rec_off[n] (record index of each read's first record),
 * rec[nrec], pops/trace/occ_bytes/occ_calls[n] (any may be NULL) */
void orc_result_copy(const orc_result_t *r, int32_t *n_aln, uint64_t *rec_off, orc_aln_t *rec,
                     uint64_t *pops, uint64_t *trace, uint64_t *occ_bytes, uint64_t *occ_calls,
                     int32_t *peak);
void orc_result_free(orc_result_t *r);

#ifdef __cplusplus
}
#endif

#endif
