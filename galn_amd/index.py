"""BWA-format ``.bwt`` index builder (tooling for the search path).

Builds the file that ``galn aln`` reads: the BWT of T$ with
T = F . revcomp(F) (F = all contigs concatenated, ambiguous bases replaced by
``lrand48() & 3`` after ``srand48(11)``), ``$`` removed, packed into BWA's
64-byte blocks. Same on-disk result as the reference's CPU builder
(src/index/cpu_ref.c:56-106, src/index/idxcommon.c:128-184, upstream
bntseq.c bns_fasta2bntseq): any correct suffix array gives the same BWT.

On a GPU the HIP builder of libgaln_hip.so (galn_hip_build_bwt: bucketed
radix sorts of 27-symbol keys with refinement, any length that fits HBM) runs.
On the CPU the suffix array is built by prefix doubling over torch tensors
(sort key = (rank[i], rank[i+h]) until all ranks are distinct), for the small
texts of the CPU tests.
"""
from __future__ import annotations

import struct

import numpy as np
import torch

# BWA nst_nt4_table: A/C/G/T (either case) -> 0..3, everything else -> 4
NT4 = np.full(256, 4, dtype=np.uint8)
for _i, _c in enumerate("ACGT"):
    NT4[ord(_c)] = _i
    NT4[ord(_c.lower())] = _i


def lrand48_low2(count: int, seed: int = 11) -> np.ndarray:
    """``lrand48() & 3`` for ``count`` draws after ``srand48(seed)``.

    POSIX drand48 family: X' = (0x5DEECE66D * X + 0xB) mod 2^48, X0 = seed<<16 | 0x330E,
    lrand48 returns X' >> 17. Used for ambiguous bases as upstream bntseq.c:266,296.
    """
    a, c, mask = 0x5DEECE66D, 0xB, (1 << 48) - 1
    x = ((seed << 16) | 0x330E) & mask
    out = np.empty(count, dtype=np.uint8)
    for i in range(count):
        x = (a * x + c) & mask
        out[i] = (x >> 17) & 3
    return out


def encode_forward(seq: bytes | np.ndarray) -> np.ndarray:
    """ASCII reference bases (all contigs concatenated) -> 2-bit codes as bwa packs them."""
    s = np.frombuffer(seq, dtype=np.uint8) if isinstance(seq, (bytes, bytearray)) else np.asarray(seq, dtype=np.uint8)
    f = NT4[s]
    amb = np.flatnonzero(f >= 4)
    if amb.size:
        f = f.copy()
        f[amb] = lrand48_low2(int(amb.size))
    return f


def read_fasta(path: str) -> tuple[list[str], bytes]:
    """Names and the concatenated sequence of a FASTA file (plain text)."""
    names, parts = [], []
    with open(path, "rb") as fh:
        for line in fh:
            if line.startswith(b">"):
                names.append(line[1:].split()[0].decode())
            else:
                parts.append(line.strip())
    return names, b"".join(parts)


def suffix_array(text: torch.Tensor) -> torch.Tensor:
    """Suffix array of ``text`` (small integer symbols), end of string smallest.

    Prefix doubling: ranks of 2h-prefixes from sorted (rank[i], rank[i+h]) pairs.
    Keys are packed into one int64, so ``len(text)**2`` must stay below 2**62.
    """
    n = int(text.numel())
    if n == 0:
        return torch.empty(0, dtype=torch.int64, device=text.device)
    if n >= (1 << 31):
        raise ValueError("suffix_array: text of 2^31 symbols or more needs the two-key sort")
    rank = text.to(torch.int64)
    mult = max(n, int(rank.max()) + 1) + 1  # second keys lie in [0, mult)
    h = 1
    while True:
        if h > 2 * n:
            raise RuntimeError("suffix_array: ranks failed to separate")
        second = torch.zeros_like(rank)
        if h < n:
            second[: n - h] = rank[h:] + 1  # 0 = past the end (smallest)
        key = rank * mult + second
        sa = torch.argsort(key)
        skey = key[sa]
        del key, second
        new = torch.ones(n, dtype=torch.int64, device=text.device)
        new[1:] = (skey[1:] != skey[:-1]).to(torch.int64)
        del skey
        new = torch.cumsum(new, 0) - 1
        rank = torch.empty_like(new)
        rank[sa] = new
        top = int(new[-1])
        del new
        if top == n - 1:
            return sa
        h *= 2


def bwt_from_forward(f2: np.ndarray, device: str | torch.device = "cpu") -> tuple[int, np.ndarray]:
    """(primary, L) for forward 2-bit codes: L is the BWT of T$ with $ removed."""
    f = torch.as_tensor(np.ascontiguousarray(f2), dtype=torch.uint8, device=device)
    text = torch.cat([f, (3 - f).flip(0)])  # T = F . revcomp(F)
    n = int(text.numel())
    sa = suffix_array(text)
    # row 0 of the (n+1)-row matrix is the $ suffix; row r > 0 is SA[r-1]
    j0 = int(torch.nonzero(sa == 0)[0, 0])
    primary = j0 + 1
    prev = sa - 1
    prev[j0] = 0  # placeholder; this row is the $ and is dropped below
    lrows = text[prev]  # chars of rows 1..n
    lfull = torch.cat([text[n - 1 : n], lrows])  # rows 0..n
    keep = torch.ones(n + 1, dtype=torch.bool, device=text.device)
    keep[primary] = False
    L = lfull[keep].to("cpu").numpy().astype(np.uint8)
    return primary, L


def pack_bwt(L: np.ndarray) -> tuple[list[int], np.ndarray]:
    """BWA .bwt body from the $-less BWT: L2[0..4] and the u32 words
    (src/index/idxcommon.c:144-184: 128-symbol blocks of [4 x u64 counts][8 x u32]
    with the first symbol in bits 31-30, then the 4 x u64 totals)."""
    n = int(L.size)
    nb = (n + 127) // 128
    padded = np.zeros(nb * 128, dtype=np.uint32)
    padded[:n] = L
    sym = padded.reshape(nb * 8, 16)
    shifts = (30 - 2 * np.arange(16, dtype=np.uint32)).astype(np.uint32)
    words = np.bitwise_or.reduce(sym << shifts, axis=1).astype(np.uint32).reshape(nb, 8)
    onehot = np.zeros((nb, 4), dtype=np.uint64)
    blk = np.arange(n) // 128
    for c in range(4):
        onehot[:, c] = np.bincount(blk[L == c], minlength=nb)[:nb]
    totals = onehot.sum(axis=0)
    excl = np.cumsum(onehot, axis=0) - onehot
    body = np.empty((nb, 16), dtype=np.uint32)
    body[:, :8] = excl.view(np.uint32).reshape(nb, 8)
    body[:, 8:] = words
    n_words = 8 * (nb + 1) + (n + 15) // 16
    out = np.empty(n_words, dtype=np.uint32)
    if nb:
        last_syms = (n - 128 * (nb - 1) + 15) // 16
        flat = body.reshape(-1)
        head = 16 * (nb - 1) + 8 + last_syms
        out[:head] = flat[:head]
        out[head:] = totals.astype(np.uint64).view(np.uint32)
    else:
        out[:] = totals.astype(np.uint64).view(np.uint32)
    L2 = [0, int(totals[0])]
    for c in range(1, 4):
        L2.append(L2[-1] + int(totals[c]))
    return L2, out


def write_bwt(path: str, primary: int, L2: list[int], words: np.ndarray) -> None:
    """bwt_dump_bwt format: primary, L2[1..4] (u64), then the words."""
    with open(path, "wb") as fh:
        fh.write(struct.pack("<5Q", primary, *L2[1:5]))
        fh.write(np.ascontiguousarray(words, dtype=np.uint32).tobytes())


def read_bwt(path: str) -> tuple[dict, np.ndarray]:
    """Descriptor (primary, L2, seq_len, bwt_size) and the u32 body of a .bwt file."""
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = struct.unpack("<5Q", raw[:40].tobytes())
    words = raw[40:].view(np.uint32)
    L2 = [0] + list(hdr[1:5])
    meta = {"primary": hdr[0], "L2": L2, "seq_len": L2[4], "bwt_size": int(words.size)}
    return meta, words


def write_sa(path: str, primary: int, L2: list[int], n: int, samples: np.ndarray) -> None:
    """bwt_dump_sa format (src/index/idxcommon.c:137-153): primary, L2[1..4], the interval 32,
    n, then the suffix at rows 32, 64, ... (all u64)."""
    with open(path, "wb") as fh:
        fh.write(struct.pack("<7Q", primary, *L2[1:5], 32, n))
        fh.write(np.ascontiguousarray(samples, dtype=np.uint64).tobytes())


def build_index(seq: bytes | np.ndarray, prefix: str, device: str | torch.device = "cpu", want_sa: bool = False) -> dict:
    """Build ``<prefix>.bwt`` (and with `want_sa` ``<prefix>.sa``) from ASCII reference bases;
    returns the descriptor. On a GPU device the HIP builder (galn_hip_build_bwt) runs; on the
    CPU the prefix doubling above (small texts)."""
    f2 = encode_forward(seq)
    dev = torch.device(device)
    if dev.type == "cuda":
        from . import native as nv
        m, words, sa, _ = nv.build_bwt(f2, dev.index or 0, want_sa=want_sa)
        L2 = [int(m.L2[i]) for i in range(5)]
        primary = int(m.primary)
        write_bwt(prefix + ".bwt", primary, L2, words)
        if want_sa:
            write_sa(prefix + ".sa", primary, L2, int(m.seq_len), sa)
        return {"primary": primary, "L2": L2, "seq_len": L2[4], "bwt_size": int(words.size)}
    if want_sa:
        raise ValueError("build_index: SA samples come from the GPU builder only")
    primary, L = bwt_from_forward(f2, device=device)
    L2, words = pack_bwt(L)
    write_bwt(prefix + ".bwt", primary, L2, words)
    return {"primary": primary, "L2": L2, "seq_len": L2[4], "bwt_size": int(words.size)}


def build_index_from_fasta(fasta: str, prefix: str, device: str | torch.device = "cpu") -> dict:
    _, seq = read_fasta(fasta)
    return build_index(seq, prefix, device=device)


def decode_bwt(meta: dict, words: np.ndarray) -> np.ndarray:
    """The $-less BWT symbols back out of a packed body (for brute-force checks)."""
    n = int(meta["seq_len"])
    nb = (n + 127) // 128
    out = np.empty(n, dtype=np.uint8)
    for b in range(nb):
        s = b * 128
        e = min(s + 128, n)
        w = words[16 * b + 8 : 16 * b + 8 + (e - s + 15) // 16]
        sy = ((w[:, None] >> (30 - 2 * np.arange(16, dtype=np.uint32))[None, :]) & 3).reshape(-1)
        out[s:e] = sy[: e - s]
    return out
