/* galn_hip.h -- C ABI of the MI355X (gfx950) BWA-aln backtracking search.
 *
 * The library searches short reads against a BWA .bwt index resident in HBM
 * and returns, per read, the SA intervals (.sai records) that BWA aln's
 * bwt_match_gap finds, byte-identical to the reference's CPU path.
 *
 * Plain pointers and sizes only. Host buffers are caller-owned; device
 * buffers (index, reads, arenas, queues) are owned by the library handles.
 * Every call returns 0 or a negative GALN_HIP_E_* code. Calls on one index
 * handle are serialised by the caller; the serving API is thread-safe.
 *
 * Interfaces this replaces in the reference (bilalbioinfo/galn 0.2.7):
 *   galn_hip_index_open/close   index upload of run_device, src/galn.cu:804-825, :1190-1192
 *   galn_hip_search_plain       aln_search_host (src/aln_host.h:46-48) per read, and the
 *                               light tier launch light_kernel, src/galn.cu:835-905
 *   galn_hip_search_batched     phase W/G persistent warp engine gapw_kernel with its
 *                               device work list, src/gapw.cuh:716-795, src/galn.cu:907-1180
 *   galn_hip_readset_*          the device read arrays of run_device, src/galn.cu:780-825
 *   galn_hip_server_*           no counterpart (the reference is a one-shot CLI); serving loop
 *                               with the index resident in HBM and continuous batching
 */
#ifndef GALN_HIP_H
#define GALN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GALN_HIP_ABI_VERSION 5

/* ---- library return codes ---- */
#define GALN_HIP_OK             0
#define GALN_HIP_E_ARG        (-1)   /* bad argument */
#define GALN_HIP_E_DEVICE     (-2)   /* no such device / HIP runtime error */
#define GALN_HIP_E_NOMEM      (-3)   /* device memory exhausted */
#define GALN_HIP_E_UNSUPPORTED (-4)  /* a read or option outside the kernel's limits */
#define GALN_HIP_E_RECCAP     (-5)   /* caller's record array too small: out->n_rec says how many */
#define GALN_HIP_E_BUSY       (-6)   /* serving: admission limit reached, or result not ready;
                                        index_close: read sets or servers still open */

/* ---- per-read result codes (out->n_aln[i] < 0), as src/bwtgap_hd.h:41-43 ---- */
#define GALN_READ_ERR_SCORE   (-3)   /* a push with score outside [0, n_stacks): final */
#define GALN_READ_ERR_UNSUP   (-5)   /* read outside the kernel's limits (see galn_hip_limits) */

/* search flags */
#define GALN_HIP_TRACE   0x1u        /* compute the per-read trace hash (pop-order check) */
#define GALN_HIP_NOFUSE  0x2u        /* never keep the exact child in registers (testing) */
#define GALN_HIP_STATS   0x4u        /* count the wave-per-read engine's rounds and clocks */
#define GALN_HIP_GRAPH   0x8u        /* batched form: run the launches as a captured hipGraph (replayed
                                        while the read set, tuning and buffers are unchanged) */

/* Index descriptor: field order of bwtgpu_meta_t, src/bwtsearch.h:26-31. */
typedef struct {
    uint64_t primary;     /* row of $ */
    uint64_t L2[5];       /* C(): L2[0] = 0, L2[4] = seq_len */
    uint64_t seq_len;     /* BWT length without $ */
    uint64_t bwt_size;    /* u32 words of the .bwt body */
} galn_bwt_meta_t;

/* Search options of one bwa read batch: field order of bwtgap_opt_t, src/bwtgap_hd.h:52-59.
 * max_diff is the batch value; the per-read value comes in galn_hip_reads.max_diff. */
typedef struct {
    int32_t s_mm, s_gapo, s_gape;
    int32_t mode;
    int32_t indel_end_skip, max_del_occ, max_entries;
    int32_t max_diff, max_gapo, max_gape;
    int32_t max_seed_diff, seed_len;
    int32_t max_top2;
} galn_search_opt_t;

/* One .sai record (bwt_aln1_t, 24 bytes): packed = n_mm:8 n_gapo:8 n_gape:8 score:20
 * n_ins:10 n_del:10 (LSB first), then the SA interval [k, l]. src/sai_format.h:30-45. */
typedef struct {
    uint64_t packed;
    uint64_t k, l;
} galn_aln_rec_t;

/* Reads as structure-of-arrays. query: concatenated search queries, each
 * complement(nt4(reverse(read))) with 0..3 for A,C,G,T and 4 for anything else
 * (src/host_io.c:178-182). */
typedef struct {
    int32_t n_reads;
    const uint8_t *query;
    const uint64_t *qoff;      /* [n_reads] offset of each query in `query` */
    const int32_t *qlen;       /* [n_reads] >= 1 */
    const int32_t *max_diff;   /* [n_reads] per-read max_diff (aln_read_maxdiff) */
    const int32_t *batch;      /* [n_reads] index into bopts / n_stacks */
    int32_t n_batches;
    const galn_search_opt_t *bopts;   /* [n_batches] (aln_batch_opt; seed_len is the -l value,
                                         and a read is seeded when seed_len < its length) */
    const int32_t *n_stacks;          /* [n_batches] */
} galn_hip_reads;

/* Results, caller-owned. Records of read i are rec[rec_off[i] .. rec_off[i] + n_aln[i]). */
typedef struct {
    int32_t *n_aln;            /* [n_reads] record count, or a GALN_READ_ERR_* code */
    uint64_t *rec_off;         /* [n_reads] */
    galn_aln_rec_t *rec;       /* [rec_cap] */
    uint64_t rec_cap;
    uint64_t n_rec;            /* out: records written (or needed, with GALN_HIP_E_RECCAP) */
    uint64_t *n_pops;          /* optional [n_reads]: popped entries */
    uint64_t *trace;           /* optional [n_reads]: trace hash (needs GALN_HIP_TRACE) */
    uint64_t *occ_bytes;       /* optional [n_reads]: algorithmic index bytes, 64 per distinct
                                  128-symbol block per occurrence query the search performs */
} galn_hip_out;

/* Kernel limits; reads beyond them get GALN_READ_ERR_UNSUP. */
typedef struct {
    int32_t max_read_len;      /* 1023 */
    int32_t max_n_stacks;      /* 256 */
    int32_t max_diff;          /* 126 (n_mm may reach max_diff + 1 in 7 bits) */
    int32_t max_gapo_plus_gape;/* 255 */
    uint64_t max_seq_len;      /* 2^38 - 1 */
} galn_hip_limits_t;

/* Tuning of the batched form (0 = library default / autotuned). */
typedef struct {
    int32_t block_threads;     /* threads per workgroup (64 or 128; default 64) */
    int32_t waves_per_cu;      /* persistent grid size = waves_per_cu * CUs */
    int32_t chunks_per_lane;   /* stack chunks (16 entries each) per lane in the first tier */
    int32_t rec_per_lane;      /* record slots per lane in the first tier */
    int32_t sort_by_cost;      /* 1: claim reads longest-first */
    int32_t heavy_pops;        /* pops after which a read moves to the wave-per-read engine
                                  (0: default 512, -1: never) */
    int32_t heavy_waves_per_simd; /* wave-per-read engine occupancy: 4 (0: default) or 5 (fewer
                                  registers, more of them spilled; more loads in flight, which pays
                                  for indexes in HBM) */
} galn_hip_tune_t;

/* Serving loop configuration (0 = default). */
typedef struct {
    int32_t max_launch_reads;  /* reads searched together in one launch; 0 (default): adaptive,
                                  the reads the server finished in its last ~0.5 s of busy time,
                                  from 262144 to 1048576 (larger launches serve more reads/s where
                                  no single search is long, smaller ones where some take seconds) */
    int32_t quantum_reads;     /* reads a request gets per launch before the next request's
                                  reads are taken (default 16384) */
    uint64_t max_queued_bytes; /* admission: device bytes of the requests not yet collected
                                  (query bytes + 176 per read: inputs, outputs and two
                                  record slots); a submit beyond it returns
                                  GALN_HIP_E_BUSY (default 4 GiB) */
    uint32_t flags;            /* search flags of every launch */
    int32_t form;              /* 1 = batched (default), 0 = plain */
    int32_t slots;             /* launches on the device at once, each with its own worker thread,
                                  stream and stack pool; slot 0 uses the index's (default 3, at
                                  most 16; HIP spreads streams over GPU_MAX_HW_QUEUES hardware
                                  queues, 4 by default, and slots whose streams share one run
                                  one after the other); before its first launch a slot reserves
                                  the search buffers of a full launch (default launch sizes: a
                                  16 GiB stack pool, ~2.3 GB of hand-off slots, ~1.2 GB of
                                  wave-engine scratch) */
    int32_t heavy_pops;        /* batched form: pops after which a read moves to the wave-per-read
                                  engine, as galn_hip_tune_t.heavy_pops (0: the library default) */
} galn_hip_server_cfg_t;

typedef struct galn_hip_index galn_hip_index;
typedef struct galn_hip_readset galn_hip_readset;
typedef struct galn_hip_server galn_hip_server;

int galn_hip_abi_version(void);
void galn_hip_get_limits(galn_hip_limits_t *lim);
/* Last error text of this thread (static storage). */
const char *galn_hip_last_error(void);

/* Upload a .bwt body (bwt_size words, as load_bwt returns it) to `device`; the device copy keeps
   each block's symbols as two bit planes (the host buffer is not changed). */
int galn_hip_index_open(const uint32_t *bwt, const galn_bwt_meta_t *meta, int device,
                        galn_hip_index **out);
/* Returns GALN_HIP_E_BUSY (and keeps the index) while read sets or servers of it are open. */
int galn_hip_index_close(galn_hip_index *idx);
int galn_hip_index_device(const galn_hip_index *idx);

/* Plain form: reads uploaded, one launch per 0x40000 reads (thread per read), re-runs of
 * reads that outgrow the first tier in larger tiers, results copied to `out`. */
int galn_hip_search_plain(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_out *out,
                          uint32_t flags);

/* Batched form: one persistent launch over all reads with a device work queue
 * (lanes claim the next read as soon as theirs is done). Same results. */
int galn_hip_search_batched(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_out *out,
                            uint32_t flags, const galn_hip_tune_t *tune);

/* Search that allocates nothing. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same records, pop counts and trace hashes as galn_hip_search_plain, from
 * the thread-per-read engine with every read claimed from a device counter.
 *   reads:  query, qoff, qlen, max_diff, batch in DEVICE memory; bopts and n_stacks in host memory.
 *   out:    n_aln, rec_off, rec, and the optional n_pops, trace, occ_bytes in DEVICE memory.
 *           Read i's records are rec[rec_off[i] .. rec_off[i] + n_aln[i]) (each read's records are
 *           contiguous; reads follow each other in no set order; rec_off is 0 for reads without
 *           records). out->n_rec (host) = records written; with GALN_HIP_E_RECCAP, at least that
 *           many are needed.
 * Everything else lives in one static device array of 256 MiB per GPU, held from the library's
 * first use on that GPU until the process ends: at most 1,048,576 reads and 4096 batches per call
 * (beyond: GALN_HIP_E_ARG); reads whose stacks outgrow the array's pool are re-run with fewer
 * lanes sharing it. Calls on one GPU share its array and take turns (a lock per device); the
 * array keeps the last call's contents. Synchronous: returns after the outputs are written. */
int galn_hip_search_device_noalloc(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_out *out,
                                   uint32_t flags);

/* Device-resident reads: upload once, search many times (benchmarks, serving). */
int galn_hip_readset_upload(galn_hip_index *idx, const galn_hip_reads *reads, galn_hip_readset **out);
int galn_hip_readset_free(galn_hip_readset *rs);
/* Replace the reads of a read set, reusing its device arrays when they are large enough
 * (the serving loop's launches). */
int galn_hip_readset_reload(galn_hip_readset *rs, const galn_hip_reads *reads);
/* form: 0 = plain, 1 = batched. Results stay on the device until galn_hip_readset_results. */
int galn_hip_readset_search(galn_hip_index *idx, galn_hip_readset *rs, int form, uint32_t flags,
                            const galn_hip_tune_t *tune);
int galn_hip_readset_results(galn_hip_readset *rs, galn_hip_out *out);
/* Autotuner of the batched form: searches the read set once per candidate -- the pops after
 * which a read moves to the wave-per-read engine (256 .. 4096, when the options allow that
 * engine), then 64- or 128-thread workgroups for the best -- and keeps the fastest (host clock
 * around each search). Later batched searches of this read set that pass no tuning use it.
 * The chosen tuning is copied to *out (may be NULL). Results are those of the last trial. */
int galn_hip_readset_autotune(galn_hip_index *idx, galn_hip_readset *rs, uint32_t flags, galn_hip_tune_t *out);
/* Time of the last search's kernels in ms, measured with HIP events on the
 * search stream (first tier / all tiers). */
int galn_hip_readset_kernel_ms(const galn_hip_readset *rs, double *first_tier_ms, double *total_ms);
/* Per kernel launch of the last search: tier (0 = the first thread-per-read launch, 1.. its
 * re-runs; 10 = the wave-per-read launch, 11.. its re-runs), duration in ms (HIP events on the
 * search stream) and the algorithmic index bytes of the reads it finished. Returns the number
 * of launches (fills at most `cap`). */
int galn_hip_readset_launches(const galn_hip_readset *rs, int cap, int32_t *tier, double *ms, uint64_t *algo_bytes);
/* Counters of the wave-per-read engine in the last search made with GALN_HIP_STATS: rounds,
 * chains started, chains committed, steps committed, longest chain per round (summed), then
 * clock ticks (100 MHz, summed over waves) in speculation, in commits, and in the commits'
 * parts: plan, piece allocation, placement, state update; granules taken; ticks in read
 * set-up (claim, widths), in round starts, between a commit and the next round (hits); reads. */
int galn_hip_readset_engine_stats(galn_hip_readset *rs, uint64_t out[16]);

/* Serving: the index stays resident, requests (each a galn_hip_reads with its own options)
 * are queued by submit and searched by a worker thread in launches of up to
 * max_launch_reads reads; each launch takes up to quantum_reads reads of every waiting
 * request in arrival order, so a small request is not held behind a large one, and reads
 * of different requests share a launch (continuous batching). A read's records are the
 * same whether it is searched alone or in a mix. While a server is open, the index handle
 * belongs to it: make no other call on that index until galn_hip_server_close.
 * submit copies the request (the caller's arrays may be reused on return) and returns
 * its id; wait blocks up to timeout_ms (0 = poll, < 0 = forever) and returns 1 when the
 * request is finished, 0 if not yet; result copies the records out like
 * galn_hip_readset_results, gives the latency (submit to finish, ms) and releases the
 * request (unless it returns GALN_HIP_E_RECCAP). A request whose search failed reports
 * the error from wait and result. Thread-safe. */
int galn_hip_server_open(galn_hip_index *idx, const galn_hip_server_cfg_t *cfg, galn_hip_server **out);
int galn_hip_server_submit(galn_hip_server *srv, const galn_hip_reads *reads, uint64_t *id);
int galn_hip_server_wait(galn_hip_server *srv, uint64_t id, int timeout_ms);
int galn_hip_server_result(galn_hip_server *srv, uint64_t id, galn_hip_out *out, double *latency_ms);
/* Counters since open: launches, reads searched, requests finished, summed launch ms (host
 * clock around each launch), most launches in flight at once, slots, the current launch size
 * (reads), 0. */
int galn_hip_server_stats(galn_hip_server *srv, double out[8]);
/* Finishes the requests already submitted, then stops the worker. */
int galn_hip_server_close(galn_hip_server *srv);

/* SA samples for the position step: samples[j] is the text position of the suffix at BWT row
 * (j + 1) * intv, j < (seq_len + intv) / intv - 1 -- the body of a .sa file after its 56-byte
 * header (bwt_dump_sa/bwt_restore_sa, src/index/idxcommon.c:137-153), which galn_hip_build_bwt
 * also returns. Uploaded once; kept with the index. */
int galn_hip_index_set_sa(galn_hip_index *idx, const uint64_t *samples, uint64_t n_samples, uint32_t intv);

/* Text positions of BWT rows on the GPU (replaces bwt_sa / bwa_sa2pos of samse, upstream
 * src/samse/bwa_upstream/bwt.c:53-96 and bwase.c:112-123, and the LF-walk kernel of
 * src/index/gpu_index.cu:93-130): from each row (<= seq_len, e.g. a .sai record's k) LF steps to the
 * nearest sampled row. pos[i] = bwt_sa(rows[i]) on the forward-reverse text (row 0 gives all ones,
 * bwa's -1). With ref_len (and strand) given: bwa_sa2pos instead -- the forward-strand position of
 * the hit's first base for a hit of ref_len[i] reference bases, strand[i] = 1 forward / 0 reverse,
 * and all ones for a hit across the forward/reverse boundary. Host arrays. */
int galn_hip_sa_positions(galn_hip_index *idx, const uint64_t *rows, uint64_t n, uint64_t *pos, const int32_t *ref_len,
                          int32_t *strand);

/* u32 words of a .bwt body for a BWT of n symbols: 8 * (ceil(n/128) + 1) + ceil(n/16)
 * (galn_bwt_words, src/index/idxcommon.h:40-44). */
uint64_t galn_hip_bwt_words(uint64_t n);

/* The .bwt of T$ with T = F . revcomp(F), built on `device` (replaces the index builders of
 * `galn index`: src/index/cpu_ref.c:56-106, src/index/gpu_sa.h:12-20 galn_gpu_bwt,
 * src/index/gpu_index.cu:211-408).
 *   fwd:  the forward reference, `len` 2-bit codes (0..3; ambiguous bases already replaced by
 *         lrand48() & 3 as bns_fasta2bntseq does). The BWT has n = 2 * len symbols; any n that
 *         fits the device (n/4 bytes of text, n/4 of BWT, n/2 of output, ~100 bytes per suffix of
 *         one chunk of at most 2^31 suffixes) is supported, 2^32 and beyond included.
 *   bwt:  caller-owned, galn_hip_bwt_words(n) words: the body of the .bwt file (what follows its
 *         40-byte header: primary, L2[1..4]). meta gets primary, L2, seq_len = n, bwt_size.
 *   sa_samples: optional, (n + 32) / 32 - 1 words: the suffix at rows 32, 64, ... (the .sa samples
 *         of bwt_cal_sa(bwt, 32)).
 *   stats: optional [4]: chunks, most refinement rounds of a chunk, largest chunk, 0.
 * Byte-identical to the reference builder's .bwt. */
int galn_hip_build_bwt(const uint8_t *fwd, uint64_t len, int device, galn_bwt_meta_t *meta, uint32_t *bwt,
                       uint64_t *sa_samples, double *stats);

/* Random-access ceiling of `device`: pointer chase over a `ws_bytes` buffer of
 * `line_bytes` (64 or 128) lines in one pseudo-random cycle, waves_per_cu * CUs
 * independent chains of `iters` dependent whole-line loads each. Returns the
 * saturated line gathers per second. No counterpart in the reference (measurement). */
int galn_hip_gather_ceiling(int device, uint64_t ws_bytes, int line_bytes, int waves_per_cu, int iters,
                            double *lines_per_s, double *ms);

#ifdef __cplusplus
}
#endif

#endif
