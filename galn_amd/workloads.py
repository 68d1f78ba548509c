"""The aln workloads, built from seeds at any scale (tests use scaled copies).

A workload is a reference (ASCII bases), its .bwt (written under ``workdir``),
reads (padded ASCII matrix + lengths), the aln option list and the derived
search inputs: queries, offsets, bwa batch ids, per-batch options, n_stacks and
per-read max_diff (computed by the product's host library).
"""
from __future__ import annotations

import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import index as ix
from . import native as nv
from . import synth

ADNA_OPTS = ["-l", "1024", "-n", "0.01", "-o", "2"]
DEFAULT_OPTS: list[str] = []


@dataclass
class Workload:
    name: str
    prefix: str
    meta: dict
    words: np.ndarray
    reads: np.ndarray
    lens: np.ndarray
    opts: list[str]
    query: np.ndarray = field(default=None)
    qoff: np.ndarray = field(default=None)
    batch: np.ndarray = field(default=None)
    gap_opt: object = field(default=None)
    bopts: object = field(default=None)
    nstacks: np.ndarray = field(default=None)
    maxdiff: np.ndarray = field(default=None)
    timings: dict = field(default_factory=dict)
    ref: np.ndarray = field(default=None)

    def prepare(self) -> "Workload":
        t = time.time()
        self.query, self.qoff = synth.encode_queries(self.reads, self.lens)
        n = int(self.lens.size)
        self.batch = (np.arange(n, dtype=np.int64) // 0x40000).astype(np.int32)
        self.gap_opt = nv.gap_opt_from_args(self.opts)
        self.bopts, self.nstacks, self.maxdiff = nv.prepare_batches(self.gap_opt, self.lens.astype(np.int32), self.batch)
        self.timings["prepare_s"] = time.time() - t
        return self

    @property
    def n(self) -> int:
        return int(self.lens.size)

    @property
    def seed_len(self) -> int:
        return int(self.gap_opt.seed_len)

    def readset(self, index: "nv.Index") -> "nv.ReadSet":
        return nv.ReadSet(index, self.query, self.qoff, self.lens.astype(np.int32), self.maxdiff, self.batch,
                          self.bopts, self.nstacks)

    def fastq(self, path: str, start: int = 0, count: int | None = None) -> str:
        end = self.n if count is None else min(self.n, start + count)
        synth.write_fastq(path, self.reads[start:end], self.lens[start:end])
        return path


def _index(ref: np.ndarray, prefix: str, device: str) -> tuple[dict, np.ndarray]:
    """Build <prefix>.bwt, or reuse one an earlier call finished (the name carries the
    generator's size and seed; <prefix>.done is written after a complete build). On a GPU the
    HIP builder's words are returned as built (the file is written for the CLI and the CPU
    baseline)."""
    done = prefix + ".done"
    if os.path.exists(done) and os.path.exists(prefix + ".bwt"):
        return ix.read_bwt(prefix + ".bwt")
    if str(device).startswith("cuda"):
        m, words, _, _ = nv.build_bwt(ix.encode_forward(ref), 0)
        L2 = [int(m.L2[i]) for i in range(5)]
        ix.write_bwt(prefix + ".bwt", int(m.primary), L2, words)
        open(done, "w").close()
        return {"primary": int(m.primary), "L2": L2, "seq_len": L2[4], "bwt_size": int(words.size)}, words
    ix.build_index(ref, prefix, device=device)
    open(done, "w").close()
    return ix.read_bwt(prefix + ".bwt")


def plain_small(workdir: str, ref_bp: int = 5_000_000, n_reads: int = 100_000, read_len: int = 100,
                seeds: tuple[int, int] = (101, 102), device: str = "cpu") -> Workload:
    """i.i.d. reference, fixed-length reads, default aln options."""
    os.makedirs(workdir, exist_ok=True)
    t = time.time()
    ref = synth.plain_reference(seeds[0], ref_bp)
    prefix = os.path.join(workdir, f"plain_{ref_bp}_{seeds[0]}")
    meta, words = _index(ref, prefix, device)
    t_idx = time.time() - t
    reads, lens = synth.plain_reads(seeds[1], ref, n_reads, length=read_len)
    w = Workload("plain_small", prefix, meta, words, reads, lens, list(DEFAULT_OPTS))
    w.timings.update(index_s=t_idx, reads_s=time.time() - t - t_idx)
    return w.prepare()


def adna_mixed(workdir: str, ref_bp: int = 100_000_000, n_reads: int = 10_000_000,
               seeds: tuple[int, int] = (201, 202), device: str = "cpu", workers: int = 8) -> Workload:
    """Duplicated/N-run 41% GC reference, aDNA-like mixed-length reads, aln -l 1024 -n 0.01 -o 2."""
    os.makedirs(workdir, exist_ok=True)
    t = time.time()
    ref = synth.adna_reference(seeds[0], ref_bp)
    prefix = os.path.join(workdir, f"adna_{ref_bp}_{seeds[0]}")
    meta, words = _index(ref, prefix, device)
    t_idx = time.time() - t
    reads, lens = synth.adna_reads(seeds[1], ref, n_reads, workers=workers)
    w = Workload("adna_mixed", prefix, meta, words, reads, lens, list(ADNA_OPTS))
    w.timings.update(index_s=t_idx, reads_s=time.time() - t - t_idx)
    return w.prepare()


def human_scale(workdir: str, ref_bp: int = 3_100_000_000, n_reads: int = 50_000_000,
                seeds: tuple[int, int] = (301, 302), device: str = "cuda", workers: int = 8,
                lo: int = 30, hi: int = 150) -> Workload:
    """Human-scale reference (24 contigs, 41% GC, 45% interspersed repeats from 50 families at
    2-20% divergence, tandem repeats of period 2-200); reads of uniform length lo..hi with
    deamination and substitutions; aln -l 1024 -n 0.01 -o 2."""
    os.makedirs(workdir, exist_ok=True)
    t = time.time()
    ref = np.concatenate([c for _, c in synth.human_like_reference(seeds[0], ref_bp)])
    t_ref = time.time() - t
    prefix = os.path.join(workdir, f"human_{ref_bp}_{seeds[0]}")
    meta, words = _index(ref, prefix, device)
    t_idx = time.time() - t - t_ref
    reads, lens = synth.uniform_length_reads(seeds[1], ref, n_reads, lo, hi, workers=workers)
    w = Workload("human_scale", prefix, meta, words, reads, lens, list(ADNA_OPTS))
    w.timings.update(ref_s=t_ref, index_s=t_idx, reads_s=time.time() - t - t_ref - t_idx)
    w.ref = ref
    return w.prepare()


DEEP_OPTS = ["-n", "0.04", "-o", "2", "-e", "6", "-k", "2"]


def deep_copies(workdir: str, base_bp: int = 2_500_000_000, copies: int = 8, n_pairs: int = 20_000_000,
                seeds: tuple[int, int] = (501, 502), device: str = "cuda", workers: int = 8, progress=None) -> Workload:
    """Deep repeats: a random base genome and `copies` copies with 1-3% substitutions and 0.2%
    indels (20 Gbp at the defaults, n = 4e10 > 2^32); 2 x 150 bp pairs (insert N(350, 50), 2%
    errors), aln -n 0.04 -o 2 -e 6 -k 2. The two mates' files are one read set here, mate 1 then
    mate 2: every read is 150 bp, so each batch's options equal those of two separate runs."""
    os.makedirs(workdir, exist_ok=True)
    t = time.time()
    def note(i):
        if progress:
            progress(f"copy {i + 1} of {copies} made ({time.time() - t:.0f} s)")

    ref = np.concatenate([c for _, c in synth.diverged_copies_reference(seeds[0], base_bp, copies, note)])
    t_ref = time.time() - t
    prefix = os.path.join(workdir, f"deep_{base_bp}x{copies}_{seeds[0]}")
    meta, words = _index(ref, prefix, device)
    t_idx = time.time() - t - t_ref
    m1, l1, m2, l2 = synth.paired_reads_chunked(seeds[1], ref, n_pairs, workers=workers)
    reads = np.concatenate([m1, m2])
    del m1, m2
    lens = np.concatenate([l1, l2])
    w = Workload("deep_copies", prefix, meta, words, reads, lens, list(DEEP_OPTS))
    w.timings.update(ref_s=t_ref, index_s=t_idx, reads_s=time.time() - t - t_ref - t_idx)
    w.ref = ref
    return w.prepare()
