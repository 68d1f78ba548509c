"""Seeded synthetic references and reads for the aln workloads.

Workloads (sizes are parameters, so tests run scaled copies of the same models):

* ``plain_small``   i.i.d. uniform ACGT reference, one contig; fixed-length reads
  from both strands with substitutions and short indels (default aln options).
* ``adna_mixed``    41% GC reference with duplicated segments (300 bp - 5 kb at
  0-5% divergence) and N runs; log-normal read lengths (median 45, clipped to
  30-150), 5' C->T / 3' G->A deamination, random substitutions, and a share of
  reads from an unrelated random sequence (``aln -l 1024 -n 0.01 -o 2``).

Everything is a pure function of the seeds. Reads are returned as a padded
uint8 ASCII matrix plus lengths; :func:`encode_queries` turns them into the
search's query encoding and :func:`write_fastq` into FASTQ.
"""
from __future__ import annotations

import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)
_COMP = np.zeros(256, dtype=np.uint8)
for _a, _b in zip(b"ACGTN", b"TGCAN"):
    _COMP[_a] = _b
_NT4 = np.full(256, 4, dtype=np.uint8)
for _i, _c in enumerate(b"ACGT"):
    _NT4[_c] = _i
    _NT4[_c + 32] = _i


def random_seq(rng: np.random.Generator, n: int, gc: float = 0.5, step: int = 1 << 26) -> np.ndarray:
    """n bases with the given GC share. Drawn in pieces of `step` (choice with p takes one
    uniform double per base, in order, so the pieces give the same bases as one call)."""
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    out = np.empty(n, dtype=np.uint8)
    for s in range(0, n, step):
        out[s:s + step] = ACGT[rng.choice(4, size=min(step, n - s), p=p).astype(np.uint8)]
    return out


def mutate_subs(rng: np.random.Generator, s: np.ndarray, rate: float) -> np.ndarray:
    """Substitute each base with probability ``rate`` by a uniformly chosen other base."""
    s = s.copy()
    hit = np.flatnonzero(rng.random(s.size) < rate)
    if hit.size:
        cur = _NT4[s[hit]]
        cur = np.where(cur > 3, 0, cur)
        s[hit] = ACGT[(cur + rng.integers(1, 4, size=hit.size)) & 3]
    return s


def plain_reference(seed: int, length: int) -> np.ndarray:
    """i.i.d. uniform ACGT reference (one contig)."""
    return random_seq(np.random.default_rng(seed), length, 0.5)


def adna_reference(seed: int, length: int, gc: float = 0.41, dup_frac: float = 0.10,
                   n_frac: float = 0.01) -> np.ndarray:
    """41% GC reference with duplicated diverged segments and N runs."""
    rng = np.random.default_rng(seed)
    s = random_seq(rng, length, gc)
    target = int(length * dup_frac)
    done = 0
    while done < target:
        seg = int(rng.integers(300, 5001))
        if seg * 2 >= length:
            break
        src = int(rng.integers(0, length - seg))
        dst = int(rng.integers(0, length - seg))
        div = float(rng.uniform(0.0, 0.05))
        s[dst:dst + seg] = mutate_subs(rng, s[src:src + seg], div)
        done += seg
    target = int(length * n_frac)
    done = 0
    while done < target:
        run = int(rng.integers(100, 5001))
        run = min(run, target - done) if target - done >= 100 else target - done
        pos = int(rng.integers(0, max(1, length - run)))
        s[pos:pos + run] = ord("N")
        done += run
    return s


def revcomp_rows(m: np.ndarray, lens: np.ndarray) -> np.ndarray:
    """Reverse-complement each row's first ``lens[i]`` bases (padded matrix)."""
    n, w = m.shape
    idx = lens[:, None] - 1 - np.arange(w)[None, :]
    valid = idx >= 0
    out = np.zeros_like(m)
    out[valid] = _COMP[m[np.nonzero(valid)[0], idx[valid]]]
    return out


def _sample_fragments(rng, ref: np.ndarray, lens: np.ndarray, width: int, extra: int):
    n = lens.size
    hi = ref.size - lens - extra
    pos = (rng.random(n) * np.maximum(hi, 1)).astype(np.int64)
    cols = np.arange(width + extra)[None, :]
    frag = ref[np.minimum(pos[:, None] + cols, ref.size - 1)]
    return frag, pos


def plain_reads(seed: int, ref: np.ndarray, n: int, length: int = 100, sub_rate: float = 0.01,
                indel_rate: float = 0.001) -> tuple[np.ndarray, np.ndarray]:
    """Fixed-length reads from both strands, substitutions and 1-3 bp indels."""
    rng = np.random.default_rng(seed)
    extra = 32
    lens = np.full(n, length, dtype=np.int64)
    frag, _ = _sample_fragments(rng, ref, lens, length, extra)
    flen = np.full(n, length + extra, dtype=np.int64)
    strand = rng.random(n) < 0.5
    frag[strand] = revcomp_rows(frag[strand], flen[strand])
    out = frag[:, :length].copy()
    # indels: a per-base event with probability indel_rate, applied along the fragment
    ev = rng.random((n, length)) < indel_rate
    for r in np.flatnonzero(ev.any(axis=1)):
        src = frag[r]
        buf = []
        j = 0
        for p in range(length):
            if len(buf) >= length or j >= src.size:
                break
            if ev[r, p]:
                k = int(rng.integers(1, 4))
                if rng.random() < 0.5:
                    buf.extend(ACGT[rng.integers(0, 4, size=k)].tolist())
                else:
                    j += k
            if j < src.size:
                buf.append(int(src[j]))
                j += 1
        while len(buf) < length and j < src.size:
            buf.append(int(src[j]))
            j += 1
        out[r, :] = np.array(buf[:length], dtype=np.uint8)
    sub = rng.random((n, length)) < sub_rate
    rows, cols = np.nonzero(sub)
    cur = _NT4[out[rows, cols]]
    cur = np.where(cur > 3, 0, cur)
    out[rows, cols] = ACGT[(cur + rng.integers(1, 4, size=rows.size)) & 3]
    return out, lens


def adna_lengths(rng, n: int, median: float = 45.0, sigma: float = 0.45,
                 lo: int = 30, hi: int = 150) -> np.ndarray:
    raw = np.exp(rng.normal(np.log(median), sigma, size=n))
    return np.clip(np.rint(raw), lo, hi).astype(np.int64)


def _adna_chunk(args):
    """One chunk of adna_reads; a pure function of (seed, chunk index)."""
    seed, c, ref, m, lens, random_frac, sub_rate, deam, hi = args
    rng = np.random.default_rng([seed, c])
    L = ref.size
    win = np.lib.stride_tricks.sliding_window_view(ref, hi)
    # window [s, s + hi): forward reads start at s, reverse reads end at s + hi
    start = (rng.random(m) * (L - hi)).astype(np.int64)
    frag = win[start].copy()
    rev = rng.random(m) < 0.5
    if rev.any():
        # revcomp of the window's last len bases = the first len bases of revcomp(window)
        sub = frag[rev]
        frag[rev] = _COMP[sub[:, ::-1]]
    unrel = rng.random(m) < random_frac
    if unrel.any():
        frag[unrel] = random_seq(rng, int(unrel.sum()) * hi, 0.41).reshape(-1, hi)
    # deamination: 5' C->T with probability deam * 0.5^i at position i, mirrored G->A at the 3' end
    # (positions beyond 16 from either end are left alone: probability < 5e-6)
    nd = min(16, hi)
    pdec = (deam * 0.5 ** np.arange(nd)).astype(np.float32)
    u5 = rng.random((m, nd), dtype=np.float32) < pdec[None, :]
    head = frag[:, :nd]
    ct = u5 & (head == ord("C")) & (np.arange(nd)[None, :] < lens[:, None])
    head[ct] = ord("T")
    u3 = rng.random((m, nd), dtype=np.float32) < pdec[None, :]
    pos3 = lens[:, None] - 1 - np.arange(nd)[None, :]  # j-th base from the 3' end
    ok3 = pos3 >= 0
    rows = np.nonzero(u3 & ok3)[0]
    cols = pos3[u3 & ok3]
    isg = frag[rows, cols] == ord("G")
    frag[rows[isg], cols[isg]] = ord("A")
    # substitutions: sparse events over the read bodies
    total = int(lens.sum())
    ne = rng.binomial(total, sub_rate)
    flat = rng.integers(0, total, size=ne)
    offs = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    r = np.searchsorted(offs, flat, side="right") - 1
    cpos = flat - offs[r]
    cur = _NT4[frag[r, cpos]]
    keep_n = cur > 3
    newb = ACGT[(np.where(keep_n, 0, cur) + rng.integers(1, 4, size=ne)) & 3]
    frag[r, cpos] = np.where(keep_n, frag[r, cpos], newb)
    frag[np.arange(hi)[None, :] >= lens[:, None]] = 0
    return frag


def adna_reads(seed: int, ref: np.ndarray, n: int, random_frac: float = 0.05, sub_rate: float = 0.005,
               deam: float = 0.30, lo: int = 30, hi: int = 150, median: float = 45.0,
               chunk: int = 1 << 18, workers: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """aDNA-like reads: log-normal lengths, deamination at both ends, substitutions,
    and a share of reads drawn from an unrelated random sequence. Chunks are
    independent (seeded by (seed, chunk)), so ``workers`` only changes the speed."""
    rng = np.random.default_rng(seed)
    lens_all = adna_lengths(rng, n, median=median, lo=lo, hi=hi)
    out = np.zeros((n, hi), dtype=np.uint8)
    jobs = [(seed, c, ref, min(chunk, n - s), lens_all[s:s + chunk], random_frac, sub_rate, deam, hi)
            for c, s in enumerate(range(0, n, chunk))]
    if workers > 1 and len(jobs) > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(workers) as ex:
            for (c, s), frag in zip(enumerate(range(0, n, chunk)), ex.map(_adna_chunk, jobs)):
                out[s:s + frag.shape[0]] = frag
    else:
        for (c, s), job in zip(enumerate(range(0, n, chunk)), jobs):
            out[s:s + job[3]] = _adna_chunk(job)
    return out, lens_all


def encode_queries(reads: np.ndarray, lens: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Search queries: q[j] = complement(nt4(read[len-1-j])), N -> 4 (reference
    src/host_io.c:178-182). Returns (bytes, u64 offsets). The bytes are the
    reversed concatenation of the reads, so read i's query starts at
    total - (start of read i) - len(i)."""
    n, w = reads.shape
    lens = np.asarray(lens, dtype=np.int64)
    mask = np.arange(w)[None, :] < lens[:, None]
    cat = reads[mask]
    b = _NT4[cat[::-1]]
    q = np.where(b > 3, 4, 3 - b).astype(np.uint8)
    starts = np.zeros(n, dtype=np.int64)
    if n:
        starts[1:] = np.cumsum(lens[:-1])
    qoff = (int(lens.sum()) - starts - lens).astype(np.uint64)
    return q, qoff


def write_fasta(path: str, contigs: list[tuple[str, np.ndarray]], width: int = 60) -> None:
    with open(path, "wb") as fh:
        for name, s in contigs:
            fh.write(b">" + name.encode() + b"\n")
            b = np.ascontiguousarray(s, dtype=np.uint8).tobytes()
            for i in range(0, len(b), width):
                fh.write(b[i:i + width] + b"\n")


def write_fastq(path: str, reads: np.ndarray, lens: np.ndarray, prefix: str = "r") -> None:
    with open(path, "wb") as fh:
        parts = []
        for i in range(reads.shape[0]):
            L = int(lens[i])
            s = reads[i, :L].tobytes()
            parts.append(b"@%s%d\n%s\n+\n%s\n" % (prefix.encode(), i, s, b"I" * L))
            if len(parts) >= 65536:
                fh.write(b"".join(parts))
                parts = []
        fh.write(b"".join(parts))


def human_like_reference(seed: int, length: int, n_contigs: int = 24, gc: float = 0.41,
                         repeat_frac: float = 0.45, n_families: int = 50, tandem_frac: float = 0.03
                         ) -> list[tuple[str, np.ndarray]]:
    """Contigs with interspersed repeats from seed families (2-20% divergence) and
    tandem repeats of period 2-200."""
    rng = np.random.default_rng(seed)
    s = random_seq(rng, length, gc)
    fams = [random_seq(rng, int(rng.integers(300, 6001)), gc) for _ in range(n_families)]
    target, done = int(length * repeat_frac), 0
    while done < target:
        f = fams[int(rng.integers(0, n_families))]
        cut = int(rng.integers(0, max(1, f.size // 3)))
        piece = f[cut:]
        if piece.size >= length:
            break
        dst = int(rng.integers(0, length - piece.size))
        s[dst:dst + piece.size] = mutate_subs(rng, piece, float(rng.uniform(0.02, 0.20)))
        done += piece.size
    target, done = int(length * tandem_frac), 0
    while done < target:
        period = int(rng.integers(2, 201))
        unit = random_seq(rng, period, gc)
        copies = int(rng.integers(3, 40))
        rep = np.tile(unit, copies)
        if rep.size >= length:
            continue
        dst = int(rng.integers(0, length - rep.size))
        s[dst:dst + rep.size] = mutate_subs(rng, rep, 0.01)
        done += rep.size
    # cut points 1 .. length-1 (choice over the count draws the same indexes as over the array)
    cuts = np.sort(1 + rng.choice(length - 1, size=n_contigs - 1, replace=False))
    parts = np.split(s, cuts)
    return [(f"chr{i + 1}", p) for i, p in enumerate(parts)]


def _indel_copy(rng, base: np.ndarray, sub_rate: float, indel_rate: float) -> np.ndarray:
    """A diverged copy: substitutions, then 1-3 bp insertions/deletions at indel_rate per base."""
    s = mutate_subs(rng, base, sub_rate)
    ev = np.flatnonzero(rng.random(s.size) < indel_rate)
    if ev.size == 0:
        return s
    out, last = [], 0
    for p in ev:
        out.append(s[last:p])
        k = int(rng.integers(1, 4))
        if rng.random() < 0.5:
            out.append(ACGT[rng.integers(0, 4, size=k)])
            last = p
        else:
            last = min(s.size, p + k)
    out.append(s[last:])
    return np.concatenate(out)


def diverged_copies_reference(seed: int, base_len: int, copies: int = 8, progress=None) -> list[tuple[str, np.ndarray]]:
    """A random base genome and `copies` copies of it, each with 1-3% substitutions and 0.2% indels.
    `progress(i)` is called after copy i is made."""
    rng = np.random.default_rng(seed)
    base = random_seq(rng, base_len, 0.5)
    out = []
    for i in range(copies):
        out.append((f"copy{i + 1}", _indel_copy(rng, base, float(rng.uniform(0.01, 0.03)), 0.002)))
        if progress:
            progress(i)
    return out


def paired_reads(seed: int, ref: np.ndarray, n_pairs: int, length: int = 150, insert_mean: float = 350.0,
                 insert_sd: float = 50.0, err: float = 0.02) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Paired-end reads: fragment length ~ N(insert_mean, insert_sd); mate 1 from the fragment's
    start, mate 2 the reverse complement of its end; substitutions at rate `err`."""
    rng = np.random.default_rng(seed)
    ins = np.clip(np.rint(rng.normal(insert_mean, insert_sd, n_pairs)), length, 4 * insert_mean).astype(np.int64)
    start = (rng.random(n_pairs) * (ref.size - ins - 1)).astype(np.int64)
    cols = np.arange(length)[None, :]
    m1 = ref[start[:, None] + cols]
    m2 = _COMP[ref[(start + ins - 1)[:, None] - cols]]
    flip = rng.random(n_pairs) < 0.5
    m1[flip], m2[flip] = m2[flip].copy(), m1[flip].copy()
    lens = np.full(n_pairs, length, dtype=np.int64)
    out = []
    for m in (m1, m2):
        mm = m.copy()
        sub = rng.random(mm.shape) < err
        rows, cc = np.nonzero(sub)
        cur = _NT4[mm[rows, cc]]
        cur = np.where(cur > 3, 0, cur)
        mm[rows, cc] = ACGT[(cur + rng.integers(1, 4, size=rows.size)) & 3]
        out.append(mm)
    return out[0], lens, out[1], lens.copy()


def uniform_length_reads(seed: int, ref: np.ndarray, n: int, lo: int = 30, hi: int = 150, sub_rate: float = 0.005,
                         deam: float = 0.30, random_frac: float = 0.0, workers: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Reads of uniform length in [lo, hi], both strands, deamination and substitutions. Chunks
    are seeded by (seed, chunk), so ``workers`` only changes the speed."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, size=n).astype(np.int64)
    out = np.zeros((n, hi), dtype=np.uint8)
    chunk = 1 << 18
    jobs = [(seed, c, ref, min(chunk, n - s), lens[s:s + chunk], random_frac, sub_rate, deam, hi)
            for c, s in enumerate(range(0, n, chunk))]
    starts = list(range(0, n, chunk))
    if workers > 1 and len(jobs) > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(workers) as ex:
            for s, frag in zip(starts, ex.map(_adna_chunk, jobs)):
                out[s:s + frag.shape[0]] = frag
    else:
        for s, job in zip(starts, jobs):
            out[s:s + job[3]] = _adna_chunk(job)
    return out, lens


def _paired_chunk(args):
    seed, c, ref, m, length, insert_mean, insert_sd, err = args
    return paired_reads((seed, c), ref, m, length, insert_mean, insert_sd, err)


def paired_reads_chunked(seed: int, ref: np.ndarray, n_pairs: int, length: int = 150, insert_mean: float = 350.0,
                         insert_sd: float = 50.0, err: float = 0.02, chunk: int = 1 << 18,
                         workers: int = 1) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """paired_reads for many pairs: chunks seeded by (seed, chunk), generated on `workers` threads."""
    starts = list(range(0, n_pairs, chunk))
    jobs = [(seed, c, ref, min(chunk, n_pairs - s), length, insert_mean, insert_sd, err) for c, s in enumerate(starts)]
    m1 = np.empty((n_pairs, length), np.uint8)
    m2 = np.empty((n_pairs, length), np.uint8)
    if workers > 1 and len(jobs) > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(workers) as ex:
            parts = ex.map(_paired_chunk, jobs)
            for s, (a, _, b, _) in zip(starts, parts):
                m1[s:s + a.shape[0]] = a
                m2[s:s + b.shape[0]] = b
    else:
        for s, job in zip(starts, jobs):
            a, _, b, _ = _paired_chunk(job)
            m1[s:s + a.shape[0]] = a
            m2[s:s + b.shape[0]] = b
    lens = np.full(n_pairs, length, dtype=np.int64)
    return m1, lens, m2, lens.copy()
