/* ref_harness.c -- drives the reference's own C search (aln_host.c, aln_il.c,
 * host_io.c from G-ALN 0.2.7) as `galn aln --cpu-only` does, to write golden
 * .sai files and per-read statistics, and to time the reference CPU path.
 *
 * TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile into oracle/_ref/ from the
 * reference sources where they lie; never part of the product.
 *
 * The driver steps restate src/galn.cu: options (parse_args :185-250), batch
 * options (:605-621), one search per read (aln_search_host, or the interleaved
 * engine as the CPU workers run it :686-758), .sai in input order (:1250-1263).
 *
 * usage: galn_ref [aln options] [--threads N] [--il N] [--stats out.tsv]
 *                 [--limit N] [-f out.sai] <prefix> <reads.fq>
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <time.h>
#include <omp.h>
#include "host_io.h"
#include "bwtgap_hd.h"
#include "aln_host.h"
#include "aln_il.h"
#include "sai_writer.h"
#include "opt_build.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

int main(int argc, char **argv)
{
    sai_gap_opt_t o = opt_default();
    const char *prefix = NULL, *fq = NULL, *outfn = NULL, *statsfn = NULL;
    int npos = 0, opte = -1, threads = 0, il = 0;
    long limit = -1;
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i], *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (a[0] != '-' || !a[1]) {
            if (npos == 0) prefix = a; else if (npos == 1) fq = a;
            else { fprintf(stderr, "unexpected argument %s\n", a); return 1; }
            ++npos; continue;
        }
        if (!strcmp(a, "-L")) { o.mode |= OPT_MODE_LOGGAP; continue; }
        if (!strcmp(a, "-N")) { o.mode |= OPT_MODE_NONSTOP; o.max_top2 = 0x7fffffff; continue; }
        if (!strcmp(a, "--cpu-only")) continue;
        if (!v) { fprintf(stderr, "option %s needs a value\n", a); return 1; }
        if (!strcmp(a, "-n")) {
            if (strstr(v, ".")) o.fnr = (float)atof(v), o.max_diff = -1;
            else o.max_diff = atoi(v), o.fnr = -1.0f;
        }
        else if (!strcmp(a, "-o")) o.max_gapo = atoi(v);
        else if (!strcmp(a, "-e")) opte = atoi(v);
        else if (!strcmp(a, "-M")) o.s_mm = atoi(v);
        else if (!strcmp(a, "-O")) o.s_gapo = atoi(v);
        else if (!strcmp(a, "-E")) o.s_gape = atoi(v);
        else if (!strcmp(a, "-d")) o.max_del_occ = atoi(v);
        else if (!strcmp(a, "-i")) o.indel_end_skip = atoi(v);
        else if (!strcmp(a, "-l")) o.seed_len = atoi(v);
        else if (!strcmp(a, "-k")) o.max_seed_diff = atoi(v);
        else if (!strcmp(a, "-m")) o.max_entries = atoi(v);
        else if (!strcmp(a, "-dt") || !strcmp(a, "-t")) o.n_threads = atoi(v);
        else if (!strcmp(a, "-R")) o.max_top2 = atoi(v);
        else if (!strcmp(a, "-f")) outfn = v;
        else if (!strcmp(a, "--threads")) threads = atoi(v);
        else if (!strcmp(a, "--il")) il = atoi(v);
        else if (!strcmp(a, "--stats")) statsfn = v;
        else if (!strcmp(a, "--limit")) limit = atol(v);
        else { fprintf(stderr, "unsupported option %s\n", a); return 1; }
        ++i;
    }
    if (opte > 0) { o.max_gape = opte; o.mode &= ~OPT_MODE_GAPE; }
    if (o.n_threads < 1) o.n_threads = 1;
    if (npos != 2) { fprintf(stderr, "usage: galn_ref [options] <prefix> <reads.fq>\n"); return 1; }
    if (threads > 0) omp_set_num_threads(threads);

    char bwtfn[4096];
    snprintf(bwtfn, sizeof(bwtfn), "%s.bwt", prefix);
    bwtgpu_meta_t meta;
    uint32_t *bwt = load_bwt(bwtfn, &meta);
    if (!bwt) return 2;
    int nreads = 0;
    read_t *reads = read_fastq(fq, &nreads);
    if (!reads) return 3;
    if (limit >= 0 && limit < nreads) nreads = (int)limit;
    uint32_t cnt_table[256];
    bwtgap_gen_cnt_table(cnt_table);

    int *bstart = NULL;
    int nbatch = read_batch_ranges(reads, nreads, &bstart);
    bwtgap_opt_t *bopts = (bwtgap_opt_t *)calloc(nbatch > 0 ? nbatch : 1, sizeof(bwtgap_opt_t));
    int *bstacks = (int *)calloc(nbatch > 0 ? nbatch : 1, sizeof(int));
    int *rbatch = (int *)calloc(nreads > 0 ? nreads : 1, sizeof(int));
    int max_stacks = 1, max_len = 1;
    for (int b = 0; b < nbatch; ++b) {
        int base = bstart[b], n = bstart[b + 1] - base;
        bstacks[b] = aln_batch_opt(&o, reads + base, n, &bopts[b]);
        if (bstacks[b] > max_stacks) max_stacks = bstacks[b];
        for (int t = 0; t < n; ++t) {
            rbatch[base + t] = b;
            if (reads[base + t].len > max_len) max_len = reads[base + t].len;
        }
    }

    int *res_n = (int *)malloc(sizeof(int) * (nreads > 0 ? nreads : 1));
    bwtgap_aln_t **res_a = (bwtgap_aln_t **)calloc(nreads > 0 ? nreads : 1, sizeof(void *));
    unsigned long long *res_pops = (unsigned long long *)calloc(nreads > 0 ? nreads : 1, 8);
    unsigned long long *res_trace = (unsigned long long *)calloc(nreads > 0 ? nreads : 1, 8);
    int *res_peak = (int *)calloc(nreads > 0 ? nreads : 1, sizeof(int));
    double t0 = now_s();
    if (il <= 0) {
        #pragma omp parallel
        {
            aln_ws_t ws;
            aln_ws_init(&ws);
            #pragma omp for schedule(dynamic, 16)
            for (int r = 0; r < nreads; ++r) {
                int b = rbatch[r];
                int n = aln_search_host(bwt, &meta, cnt_table, &o, &bopts[b], bstacks[b], &reads[r], &ws);
                res_n[r] = n;
                res_pops[r] = ws.last_pops; res_trace[r] = ws.last_trace; res_peak[r] = ws.last_peak;
                if (n > 0) {
                    res_a[r] = (bwtgap_aln_t *)malloc(sizeof(bwtgap_aln_t) * n);
                    memcpy(res_a[r], ws.aln, sizeof(bwtgap_aln_t) * n);
                }
            }
            aln_ws_free(&ws);
        }
    } else {
        /* interleaved engine, as the reference's CPU workers run it */
        int next = 0;
        #pragma omp parallel
        {
            aln_il_t *ctx = (aln_il_t *)malloc(sizeof(aln_il_t));
            aln_il_init(ctx, bwt, &meta, cnt_table, il, o.max_entries + 16, max_stacks, max_len, o.seed_len);
            int active = 0, exhausted = 0;
            for (;;) {
                for (int s = 0; s < ctx->n && !exhausted; ++s) {
                    if (ctx->slot[s].status != ALN_IL_IDLE) continue;
                    int r;
                    #pragma omp atomic capture
                    r = next++;
                    if (r >= nreads) { exhausted = 1; break; }
                    int b = rbatch[r];
                    aln_il_start(ctx, s, &o, &bopts[b], bstacks[b], &reads[r], r);
                    ++active;
                }
                if (!active) break;
                aln_il_round(ctx);
                for (int s = 0; s < ctx->n; ++s) {
                    aln_il_slot_t *sl = &ctx->slot[s];
                    if (sl->status != ALN_IL_DONE) continue;
                    int r = sl->tag, n = sl->ret;
                    res_n[r] = n;
                    res_pops[r] = sl->st.n_pops; res_trace[r] = sl->st.trace; res_peak[r] = sl->ws.last_peak;
                    if (n > 0) {
                        res_a[r] = (bwtgap_aln_t *)malloc(sizeof(bwtgap_aln_t) * n);
                        memcpy(res_a[r], sl->ws.aln, sizeof(bwtgap_aln_t) * n);
                    }
                    sl->status = ALN_IL_IDLE;
                    --active;
                }
            }
            aln_il_free(ctx);
            free(ctx);
        }
    }
    double t1 = now_s();
    fprintf(stderr, "[galn_ref] %d reads searched in %.3f s on %d threads (%s)\n", nreads, t1 - t0,
            omp_get_max_threads(), il > 0 ? "interleaved" : "plain");

    FILE *out = outfn ? fopen(outfn, "wb") : stdout;
    if (!out) return 5;
    sai_write_header(out, &o);
    int err = 0;
    for (int r = 0; r < nreads; ++r) {
        int n = res_n[r];
        if (n < 0) { err = 4; n = 0; }
        sai_write_read(out, n, (const sai_aln1_t *)res_a[r]);
    }
    if (outfn) fclose(out); else fflush(out);
    if (statsfn) {
        FILE *fp = fopen(statsfn, "w");
        fprintf(fp, "read\tn_aln\tpops\ttrace\tpeak\n");
        for (int r = 0; r < nreads; ++r)
            fprintf(fp, "%d\t%d\t%llu\t%016llx\t%d\n", r, res_n[r], res_pops[r], res_trace[r], res_peak[r]);
        fclose(fp);
    }
    return err;
}
