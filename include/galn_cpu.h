/* galn_cpu.h -- C ABI of the CPU search engine (libgaln_host.so).
 *
 * BWA aln's bwt_match_gap on the host, with the same records (in the order found), pop counts
 * and trace hashes as the reference's CPU search. `galn aln --cpu-only` runs every read here;
 * the GPU path runs here the reads its kernels do not take (galn_hip_get_limits).
 *
 * Interfaces this replaces in the reference (bilalbioinfo/galn 0.2.7):
 *   galn_cpu_search   aln_search_host (src/aln_host.h:46-48, src/aln_host.c:87-147) per read, driven
 *                     by the interleaved worker engine aln_il_* (src/aln_il.c:15-302) on the CPU
 *                     worker threads of run (src/galn.cu:1034-1053, :1151-1180)
 */
#ifndef GALN_CPU_H
#define GALN_CPU_H

#include <stdint.h>
#include "galn_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Searches reads (all of them, or the n_list read ids in `list`) against the .bwt body `bwt`
 * (host memory, the layout galn_load_bwt returns). Results go to `out` indexed by position:
 * read i, or list[i] at position i. Records are in position order; per-read codes as the GPU's
 * (GALN_READ_ERR_SCORE for a push outside [0, n_stacks)); n_pops, trace and occ_bytes are
 * filled when given (the trace hash is computed only then). n_threads <= 0: every hardware
 * thread; interleave: reads one thread advances in turn (<= 0: by the .bwt body's size, 2 under
 * 8 MiB, 4 under 32 MiB, else 8). Returns 0, GALN_HIP_E_ARG, or GALN_HIP_E_RECCAP with
 * out->n_rec = records needed. */
int galn_cpu_search(const uint32_t *bwt, const galn_bwt_meta_t *meta, const galn_hip_reads *reads,
                    const int32_t *list, int32_t n_list, galn_hip_out *out, int32_t n_threads, int32_t interleave);

#ifdef __cplusplus
}
#endif

#endif
