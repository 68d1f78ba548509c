/* galn_host.h -- host side of `galn aln`: options, index and read input, .sai output.
 *
 * Everything here runs on the CPU; the search runs on the GPU through galn_hip.h, and on the
 * CPU engine (galn_cpu.h) for --cpu-only and for the reads outside the GPU kernels' limits.
 * Interfaces this replaces in the reference (bilalbioinfo/galn 0.2.7):
 *   galn_gap_opt_t / galn_gap_opt_default   sai_gap_opt_t, opt_default   src/sai_format.h:17-27, src/opt_build.h:23-36
 *   galn_parse_aln_args                     parse_args                   src/galn.cu:185-250
 *   galn_cal_maxdiff                        aln_cal_maxdiff              src/aln_host.c:9-21
 *   galn_batch_opt / galn_read_maxdiff      aln_batch_opt / aln_read_maxdiff  src/aln_host.c:23-43
 *   galn_load_bwt                           load_bwt                     src/host_io.c:29-76
 *   galn_read_fastq                         read_fastq                   src/host_io.c:151-205
 *   galn_sai_write                          sai_write_header / sai_write_read  src/sai_writer.h:9-21
 */
#ifndef GALN_HOST_H
#define GALN_HOST_H

#include <stdint.h>
#include <stdio.h>
#include "galn_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GALN_VERSION "0.2.7"
#define GALN_HOST_ALN_BATCH 0x40000   /* bwa aln reads per batch */

/* gap_opt_t: the 64-byte block written verbatim after "SAI\1". Field order of upstream bwtaln.h. */
typedef struct {
    int32_t s_mm, s_gapo, s_gape;
    int32_t mode;
    int32_t indel_end_skip, max_del_occ, max_entries;
    float fnr;
    int32_t max_diff, max_gapo, max_gape;
    int32_t max_seed_diff, seed_len;
    int32_t n_threads;
    int32_t max_top2;
    int32_t trim_qual;
} galn_gap_opt_t;

#define GALN_MODE_GAPE     0x01
#define GALN_MODE_COMPREAD 0x02
#define GALN_MODE_LOGGAP   0x04
#define GALN_MODE_NONSTOP  0x10

/* search forms */
#define GALN_FORM_PLAIN   0
#define GALN_FORM_BATCHED 1

typedef struct {
    galn_gap_opt_t opt;
    const char *prefix;     /* index prefix (<prefix>.bwt) */
    const char *reads;      /* FASTQ/FASTA, plain or gzip */
    const char *out;        /* -f FILE, or NULL for stdout */
    int device;             /* --device N */
    int form;               /* --form plain|batched */
    int verbose;            /* -v */
    int trace;              /* --trace: also compute per-read trace hashes (printed with -v) */
    int cpu_only;           /* --cpu-only: every read on the CPU engine, no GPU */
    int verify;             /* --verify: re-run the GPU's reads on the CPU engine and compare
                               records, pops and traces (exit 6 on a difference) */
    int cpu_threads;        /* --cpu-workers N: CPU engine threads [all hardware threads] */
    int cpu_interleave;     /* --cpu-interleave N: reads a CPU thread advances in turn [0: 2, 4 or 8
                               by the index's size] */
    int block_batches;      /* --block-batches N: the GPU path streams the input in blocks of N bwa
                               batches (0x40000 reads each), searched while the next are read and
                               the last written [4]; 0: read the whole input first */
} galn_aln_args_t;

void galn_gap_opt_default(galn_gap_opt_t *o);

/* `galn aln` arguments (argv without "galn aln"). Options may come before or after the two
 * positional arguments. Returns 0, or -1 with a message on stderr (usage error, exit code 1). */
int galn_parse_aln_args(int argc, char **argv, galn_aln_args_t *a);

int galn_cal_maxdiff(int l, double err, double thres);
/* per-batch search options from the header options and the batch's longest read; returns n_stacks */
int galn_batch_opt(const galn_gap_opt_t *o, int max_len, galn_search_opt_t *b);
int galn_read_maxdiff(const galn_gap_opt_t *o, const galn_search_opt_t *b, int len);

/* Options for a whole read set: reads carry their bwa batch index (batch[i], non-decreasing);
 * fills bopts[n_batches], n_stacks[n_batches] and max_diff[n]. Returns n_batches (>= 1). */
int galn_prepare_batches(const galn_gap_opt_t *o, int n, const int32_t *qlen, const int32_t *batch,
                         int max_batches, galn_search_opt_t *bopts, int32_t *n_stacks, int32_t *max_diff);

/* .bwt: returns the u32 body (free with galn_free_bwt) and fills the descriptor; NULL on error. */
uint32_t *galn_load_bwt(const char *fn, galn_bwt_meta_t *m);
void galn_free_bwt(uint32_t *bwt);

/* Reads with bwa_read_seq's rules: empty records skipped, a malformed record (kseq -2) ends the
 * current batch, a batch that would be empty ends the input, 0x40000 reads per batch.
 * Queries are complement(nt4(reverse(read))). */
typedef struct galn_readbuf galn_readbuf_t;
galn_readbuf_t *galn_read_fastq(const char *fn);   /* NULL on open / decompression error */
int32_t galn_readbuf_n(const galn_readbuf_t *r);
int32_t galn_readbuf_n_batches(const galn_readbuf_t *r);
const uint8_t *galn_readbuf_query(const galn_readbuf_t *r);
const uint64_t *galn_readbuf_qoff(const galn_readbuf_t *r);
const int32_t *galn_readbuf_qlen(const galn_readbuf_t *r);
const int32_t *galn_readbuf_batch(const galn_readbuf_t *r);
const char *galn_readbuf_name(const galn_readbuf_t *r, int32_t i);
void galn_readbuf_free(galn_readbuf_t *r);

/* The .sai stream: header then, per read in input order, int32 n_aln and its records; a read
 * with n_aln < 0 is written with 0 records. Returns 0, or 5 on a write error. */
int galn_sai_write(FILE *fp, const galn_gap_opt_t *o, int32_t n, const int32_t *n_aln, const uint64_t *rec_off,
                   const galn_aln_rec_t *rec);

/* `galn aln` itself; returns the process exit code (src/galn.cu exit codes). */
int galn_aln_main(int argc, char **argv);

#ifdef __cplusplus
}
#endif

#endif
