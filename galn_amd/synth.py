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


def random_seq(rng: np.random.Generator, n: int, gc: float = 0.5) -> np.ndarray:
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return ACGT[rng.choice(4, size=n, p=p).astype(np.uint8)]


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


def adna_reads(seed: int, ref: np.ndarray, n: int, random_frac: float = 0.05, sub_rate: float = 0.005,
               deam: float = 0.30, lo: int = 30, hi: int = 150, median: float = 45.0,
               chunk: int = 1 << 20) -> tuple[np.ndarray, np.ndarray]:
    """aDNA-like reads: log-normal lengths, deamination at both ends, substitutions,
    and a share of reads drawn from an unrelated random sequence."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, hi), dtype=np.uint8)
    lens_all = adna_lengths(rng, n, median=median, lo=lo, hi=hi)
    decay = deam * 0.5 ** np.arange(hi)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        m = e - s
        lens = lens_all[s:e]
        frag, _ = _sample_fragments(rng, ref, lens, hi, 0)
        strand = rng.random(m) < 0.5
        frag[strand] = revcomp_rows(frag[strand], lens[strand])
        unrel = rng.random(m) < random_frac
        if unrel.any():
            frag[unrel] = random_seq(rng, int(unrel.sum()) * hi, 0.41).reshape(-1, hi)
        cols = np.arange(hi)[None, :]
        inread = cols < lens[:, None]
        # 5' C->T with probability deam * 0.5^i at position i
        ct = (frag == ord("C")) & (rng.random((m, hi)) < decay[None, :]) & inread
        frag[ct] = ord("T")
        # 3' G->A mirrored: position j from the 3' end
        from3 = lens[:, None] - 1 - cols
        p3 = np.where(inread, deam * 0.5 ** np.clip(from3, 0, 60), 0.0)
        ga = (frag == ord("G")) & (rng.random((m, hi)) < p3)
        frag[ga] = ord("A")
        sub = (rng.random((m, hi)) < sub_rate) & inread
        rows, cc = np.nonzero(sub)
        cur = _NT4[frag[rows, cc]]
        keep_n = cur > 3
        cur = np.where(keep_n, 0, cur)
        newb = ACGT[(cur + rng.integers(1, 4, size=rows.size)) & 3]
        frag[rows, cc] = np.where(keep_n, frag[rows, cc], newb)
        frag[~inread] = 0
        out[s:e] = frag
    return out, lens_all


def encode_queries(reads: np.ndarray, lens: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Search queries: q[j] = complement(nt4(read[len-1-j])), N -> 4 (reference
    src/host_io.c:178-182). Returns (concatenated bytes, u64 offsets)."""
    n, w = reads.shape
    lens = np.asarray(lens, dtype=np.int64)
    idx = lens[:, None] - 1 - np.arange(w)[None, :]
    valid = idx >= 0
    b = _NT4[reads[np.nonzero(valid)[0], idx[valid]]]
    q = np.where(b > 3, 4, 3 - b).astype(np.uint8)
    off = np.zeros(n, dtype=np.uint64)
    if n:
        off[1:] = np.cumsum(lens[:-1]).astype(np.uint64)
    return q, off


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
