"""The CPU oracle, pinned: against the reference's golden outputs, brute force on tiny
indexes, and the closed form of the Poisson tail."""
from __future__ import annotations

import math

import numpy as np
import pytest

import fixtures as fx
import oracle as orc
from conftest import golden_inputs, golden_runs, load_reads, run_opts
from galn_amd import index as ix


@pytest.mark.parametrize("name,key", golden_runs())
def test_oracle_matches_reference_golden(name, key, golden_dir):
    """.sai md5, per-read pop count and trace hash equal the reference's own C search."""
    w = golden_inputs(name, golden_dir)
    opts, mate = run_opts(w["meta"], key)
    r = load_reads(w["reads"][mate])
    meta_bwt, words = ix.read_bwt(w["bwt"])
    o = orc.parse_aln_opts(opts)
    bopts, nst, maxd = orc.batch_options(o, r["qlen"], r["batch"])
    res = orc.search(words, meta_bwt, r["query"], r["qoff"], r["qlen"], maxd, r["batch"], bopts, nst, o["seed_len"])
    exp = w["expected"]
    assert fx.md5_bytes(orc.sai_bytes(o, res.n_aln, res.rec_off, res.rec)) == w["meta"]["runs"][key]["sai_md5"]
    np.testing.assert_array_equal(res.n_aln, exp[key + "_n_aln"])
    np.testing.assert_array_equal(res.pops, exp[key + "_pops"])
    np.testing.assert_array_equal(res.trace, exp[key + "_trace"])


# ---------------------------------------------------------------------------
# brute force on tiny indexes


def tiny_index(seed: int, n: int):
    rng = np.random.default_rng(seed)
    f = rng.integers(0, 4, size=n).astype(np.uint8)
    seq = np.frombuffer(b"ACGT", np.uint8)[f].tobytes()
    text = np.concatenate([f, 3 - f[::-1]])
    primary, L = ix.bwt_from_forward(f)
    L2, words = ix.pack_bwt(L)
    meta = {"primary": primary, "L2": L2, "seq_len": L2[4], "bwt_size": int(words.size)}
    return seq, text, meta, words, L


def full_bwt(L: np.ndarray, primary: int) -> np.ndarray:
    """$-inclusive BWT as symbols (4 = $)."""
    return np.insert(L.astype(np.int64), primary, 4)


@pytest.mark.parametrize("seed,n", [(1, 300), (2, 1000), (3, 129)])
def test_occ_brute_force(seed, n):
    """Occ(c, k) = #c in B[0..k] over the $-inclusive BWT, at every position, all four bases."""
    _, _, meta, words, L = tiny_index(seed, n)
    B = full_bwt(L, meta["primary"])
    lib = orc.lib()
    m = orc.make_meta(meta)
    import ctypes as C
    cnt = (C.c_uint64 * 4)()
    cntk = (C.c_uint64 * 4)()
    cntl = (C.c_uint64 * 4)()
    wp = words.ctypes.data_as(C.c_void_p)
    prefix = np.zeros((B.size + 1, 4), np.int64)
    for c in range(4):
        prefix[1:, c] = np.cumsum(B == c)
    for k in range(-1, B.size):
        kk = k % (1 << 64)
        want = prefix[k + 1] if k >= 0 else np.zeros(4, np.int64)
        for c in range(4):
            assert lib.orc_occ(wp, C.byref(m), kk, c) == want[c], (k, c)
        lib.orc_occ4(wp, C.byref(m), kk, cnt)
        assert list(cnt) == list(want), k
    # total counts at seq_len equal L2 differences
    for c in range(4):
        assert lib.orc_occ(wp, C.byref(m), meta["seq_len"], c) == meta["L2"][c + 1] - meta["L2"][c]
    rng = np.random.default_rng(seed)
    for _ in range(200):
        k, l = sorted(int(x) for x in rng.integers(-1, B.size, size=2))
        lib.orc_2occ4(wp, C.byref(m), k % (1 << 64), l, cntk, cntl)
        assert list(cntk) == list(prefix[k + 1]) and list(cntl) == list(prefix[l + 1])


def count_occurrences(text: np.ndarray, pat: np.ndarray) -> int:
    if pat.size == 0:
        return text.size + 1
    w = np.lib.stride_tricks.sliding_window_view(text, pat.size)
    return int((w == pat[None, :]).all(axis=1).sum())


def test_match_exact_counts_occurrences():
    """The SA interval size of a pattern = its occurrences in T = F.revcomp(F)."""
    _, text, meta, words, _ = tiny_index(4, 2000)
    rng = np.random.default_rng(5)
    import ctypes as C
    lib = orc.lib()
    m = orc.make_meta(meta)
    wp = words.ctypes.data_as(C.c_void_p)
    for t in range(300):
        L = int(rng.integers(1, 12))
        if t % 2:
            p = int(rng.integers(0, text.size - L))
            pat = text[p:p + L].copy()
        else:
            pat = rng.integers(0, 4, size=L).astype(np.uint8)
        a = np.ascontiguousarray(pat, dtype=np.uint8)
        got = lib.orc_match_exact(wp, C.byref(m), L, a.ctypes.data_as(C.c_void_p), None, None)
        assert got == count_occurrences(text, pat)


def test_cal_width_definition():
    """width[i]: greedy left-to-right exact matching of the reversed read, restarting (bid + 1)
    when the substring since the last restart stops occurring in T; w = its occurrence count."""
    _, text, meta, words, _ = tiny_index(6, 1500)
    rng = np.random.default_rng(7)
    import ctypes as C
    lib = orc.lib()
    m = orc.make_meta(meta)
    wp = words.ctypes.data_as(C.c_void_p)
    for _ in range(60):
        L = int(rng.integers(5, 40))
        p = int(rng.integers(0, text.size - L))
        read = text[p:p + L].copy()
        for _j in range(int(rng.integers(0, 4))):
            read[int(rng.integers(0, L))] = int(rng.integers(0, 5))
        # the stored query is complement(reverse(read)); cal_width undoes the complement
        query = np.where(read[::-1] > 3, 4, 3 - read[::-1]).astype(np.uint8)
        width = (orc.Width * (L + 1))()
        lib.orc_cal_width(wp, C.byref(m), L, query.ctypes.data_as(C.c_void_p), 1, width)
        rev = read[::-1]
        bid, start = 0, 0
        for i in range(L):
            c = rev[i]
            ok = c < 4 and count_occurrences(text, rev[start:i + 1][::-1]) > 0
            if not ok:
                bid += 1
                start = i + 1
                w_exp = meta["seq_len"] + 1
            else:
                w_exp = count_occurrences(text, rev[start:i + 1][::-1])
            assert width[i].bid == bid and width[i].w == w_exp, (i, width[i].bid, bid, width[i].w, w_exp)
        assert width[L].w == 0 and width[L].bid == bid + 1


def poisson_maxdiff(L: int, err: float, thres: float) -> int:
    """Smallest k with P(X > k) < thres for X ~ Poisson(L * err) (k < 1000), else 2."""
    lam = L * err
    s = 0.0
    for k in range(0, 1000):
        s += math.exp(-lam) * lam ** k / math.factorial(k) if k < 170 else 0.0
        if k >= 1 and 1.0 - s < thres:
            return k
    return 2


def wrapped_maxdiff(L: int, err: float, thres: float) -> int:
    """The same tail sum with k! kept in a wrapping 32-bit int, as upstream computes it."""
    elam = math.exp(-L * err)
    s, y, x = elam, 1.0, 1
    for k in range(1, 1000):
        y *= L * err
        x = (x * k) & 0xFFFFFFFF
        xs = x - (1 << 32) if x >= (1 << 31) else x
        s += elam * y / xs
        if 1.0 - s < thres:
            return k
    return 2


@pytest.mark.parametrize("fnr", [0.04, 0.01, 0.02, 0.001])
def test_cal_maxdiff_closed_form(fnr):
    """Equal to the Poisson tail wherever k! fits a 32-bit int (k <= 12); beyond that
    upstream's wrapped factorial decides, and the oracle follows it."""
    thr = float(np.float32(fnr))
    for L in range(1, 1025):
        got = orc.cal_maxdiff(L, 0.02, thr)
        exact = poisson_maxdiff(L, 0.02, thr)
        if exact <= 12:
            assert got == exact, L
        assert got == wrapped_maxdiff(L, 0.02, thr), L


def test_cal_maxdiff_table():
    # read lengths of the workloads and their max_diff (-n 0.04 / -n 0.01)
    t04 = {30: 2, 45: 3, 100: 5, 150: 6, 250: 9}
    t01 = {30: 3, 45: 4, 100: 6, 150: 8, 250: 11}
    for L, v in t04.items():
        assert orc.cal_maxdiff(L, 0.02, float(np.float32(0.04))) == v
    for L, v in t01.items():
        assert orc.cal_maxdiff(L, 0.02, float(np.float32(0.01))) == v


def suffix_interval(sa_text: np.ndarray, sa: np.ndarray, pat: np.ndarray) -> tuple[int, int] | None:
    """SA rows (1-based, row 0 = $) of the suffixes starting with pat, by binary search."""
    n = sa_text.size

    def cmp(i):  # suffix i vs pat: -1, 0 (prefix), 1
        s = sa_text[i:i + pat.size]
        for a, b in zip(s, pat):
            if a != b:
                return -1 if a < b else 1
        return 0 if s.size == pat.size else -1

    lo, hi = 0, n
    while lo < hi:
        mid = (lo + hi) // 2
        if cmp(int(sa[mid])) < 0:
            lo = mid + 1
        else:
            hi = mid
    first = lo
    lo, hi = first, n
    while lo < hi:
        mid = (lo + hi) // 2
        if cmp(int(sa[mid])) <= 0:
            lo = mid + 1
        else:
            hi = mid
    if lo == first:
        return None
    return first + 1, lo


def test_max_diff_zero_is_exact_match():
    """With -n 0 the search reduces to exact matching: one record, the SA interval of the
    query (found here by sorting the suffixes of the text directly), or none."""
    import torch
    _, text, meta, words, _ = tiny_index(8, 3000)
    sa = ix.suffix_array(torch.as_tensor(text)).numpy()
    rng = np.random.default_rng(9)
    reads = []
    for t in range(200):
        L = int(rng.integers(8, 30))
        p = int(rng.integers(0, text.size - L))
        r = text[p:p + L].copy()
        if t % 3 == 0:
            r[int(rng.integers(0, L))] ^= 1
        reads.append(r)
    lens = np.array([r.size for r in reads], np.int32)
    query = np.concatenate([np.where(r[::-1] > 3, 4, 3 - r[::-1]) for r in reads]).astype(np.uint8)
    qoff = np.zeros(lens.size, np.uint64)
    qoff[1:] = np.cumsum(lens[:-1])
    o = orc.parse_aln_opts(["-n", "0"])
    bid = orc.batch_ids(lens.size)
    bopts, nst, maxd = orc.batch_options(o, lens, bid)
    res = orc.search(words, meta, query, qoff, lens, maxd, bid, bopts, nst, o["seed_len"])
    for i, r in enumerate(reads):
        # backward search over the query matches the query itself: revcomp(read)
        iv = suffix_interval(text, sa, (3 - r[::-1]).astype(np.uint8))
        if iv is None:
            assert res.n_aln[i] == 0
        else:
            assert res.n_aln[i] == 1
            rec = res.records(i)[0]
            assert (int(rec[1]), int(rec[2])) == iv
            assert int(rec[0]) == 0  # no mismatch, no gap, score 0


def _sa_index(f2):
    """(meta, words, full suffix array, samples of rows 32, 64, ...) of T = F . revcomp(F)."""
    import torch
    f = torch.as_tensor(f2, dtype=torch.uint8)
    text = torch.cat([f, (3 - f).flip(0)])
    n = int(text.numel())
    sa = ix.suffix_array(text).numpy().astype(np.uint64)
    primary, L = ix.bwt_from_forward(f2)
    L2, words = ix.pack_bwt(L)
    meta = {"primary": primary, "L2": L2, "seq_len": n, "bwt_size": int(words.size)}
    return meta, words, sa, sa[32 * np.arange(1, (n + 32) // 32) - 1]


@pytest.mark.parametrize("kind", ["random", "repeats"])
def test_oracle_sa_equals_sorted_suffixes(kind):
    """bwt_sa of every row (upstream bwt.c:88-96) is the suffix sorted to that row; row 0 is bwa's -1."""
    rng = np.random.default_rng(11)
    f2 = rng.integers(0, 4, 4001).astype(np.uint8)
    if kind == "repeats":
        f2[2000:3000] = f2[100:1100]
        f2[3500:3800] = np.tile(np.array([0, 1], np.uint8), 150)
    meta, words, sa, samples = _sa_index(f2)
    pos = orc.sa_positions(words, meta, samples, np.arange(meta["seq_len"] + 1))
    np.testing.assert_array_equal(pos, np.concatenate([[np.uint64(2**64 - 1)], sa]))


def test_oracle_sa2pos_exact_hits_land_on_the_read(golden_dir):
    """bwa_sa2pos (upstream bwase.c:112-123): for every exact hit (no mismatch or gap), the forward
    reference at the position is the read itself on strand 0 and its reverse complement on strand 1."""
    w = golden_inputs("plain_small", golden_dir)
    contigs = fx.reference_contigs(fx.SPECS["plain_small"])
    ref = np.concatenate([c for _, c in contigs])
    f2 = ix.encode_forward(ref)
    meta, words, _, samples = _sa_index(f2)
    r = load_reads(w["reads"][0])
    o = orc.parse_aln_opts([])
    res = orc.run_workload(words, meta, r["query"], r["qoff"], r["qlen"], o)
    comp = np.array([3, 2, 1, 0, 4], np.uint8)
    checked = 0
    for i in range(300):
        L = int(r["qlen"][i])
        q = r["query"][int(r["qoff"][i]):int(r["qoff"][i]) + L]
        read = comp[q[::-1]]  # the query is the read's reverse complement
        for rec in res.records(i):
            if int(rec[0]) & 0xFFFFFF:  # mismatches or gaps
                continue
            for k in range(int(rec[1]), min(int(rec[2]), int(rec[1]) + 3) + 1):
                (p,), (s,) = orc.sa2pos(words, meta, samples, [k], L)
                seg = f2[int(p):int(p) + L]
                assert np.array_equal(seg, read if s == 0 else comp[read[::-1]]), (i, k, int(p), int(s))
                checked += 1
    assert checked > 100
