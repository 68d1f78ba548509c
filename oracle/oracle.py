"""Python side of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (the C restatement in galn_oracle.c) and restates
the host-side steps of the reference's ``galn aln`` that feed the search:
option parsing into the 64-byte ``gap_opt_t`` header (reference
src/galn.cu:185-250, src/opt_build.h:23-36), bwa's read batches
(src/host_io.c:188-224), per-batch options (src/galn.cu:605-621) and the .sai
byte stream (src/sai_writer.h:9-21). Only tests/, ``__graft_entry__.smoke()``
and bench.py's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "galn_ref")
REF_INDEX_BIN = os.path.join(HERE, "_ref", "galn_index_ref")

HOST_ALN_BATCH = 0x40000
MODE_GAPE, MODE_COMPREAD, MODE_LOGGAP, MODE_NONSTOP = 0x01, 0x02, 0x04, 0x10


class Meta(C.Structure):
    _fields_ = [("primary", C.c_uint64), ("L2", C.c_uint64 * 5), ("seq_len", C.c_uint64),
                ("bwt_size", C.c_uint64)]


class Opt(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "s_mm", "s_gapo", "s_gape", "mode", "indel_end_skip", "max_del_occ", "max_entries",
        "max_diff", "max_gapo", "max_gape", "max_seed_diff", "seed_len", "max_top2")]


class Width(C.Structure):
    _fields_ = [("w", C.c_uint64), ("bid", C.c_int32), ("pad_", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_occ.restype = C.c_uint64
        L.orc_occ.argtypes = [P, C.POINTER(Meta), C.c_uint64, C.c_int]
        L.orc_occ4.argtypes = [P, C.POINTER(Meta), C.c_uint64, C.POINTER(C.c_uint64)]
        L.orc_2occ4.argtypes = [P, C.POINTER(Meta), C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64),
                                C.POINTER(C.c_uint64)]
        L.orc_match_exact.restype = C.c_uint64
        L.orc_match_exact.argtypes = [P, C.POINTER(Meta), C.c_int, P, C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]
        L.orc_cal_width.restype = C.c_int
        L.orc_cal_width.argtypes = [P, C.POINTER(Meta), C.c_int, P, C.c_int, P]
        L.orc_cal_maxdiff.restype = C.c_int
        L.orc_cal_maxdiff.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_search_batch.restype = P
        L.orc_search_batch.argtypes = [P, C.POINTER(Meta), C.c_int, P, P, P, P, P, P, P, C.c_int, C.c_int]
        L.orc_result_nrec.restype = C.c_uint64
        L.orc_result_nrec.argtypes = [P]
        L.orc_result_copy.argtypes = [P] * 9  # handle + 8 outputs
        L.orc_result_free.argtypes = [P]
        L.orc_inv_psi.restype = C.c_uint64
        L.orc_inv_psi.argtypes = [P, C.POINTER(Meta), C.c_uint64]
        L.orc_sa.restype = C.c_uint64
        L.orc_sa.argtypes = [P, C.POINTER(Meta), P, C.c_uint64, C.c_uint64]
        L.orc_sa2pos.restype = C.c_uint64
        L.orc_sa2pos.argtypes = [P, C.POINTER(Meta), P, C.c_uint64, C.c_uint64, C.c_int, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def make_meta(meta: dict) -> Meta:
    m = Meta()
    m.primary = meta["primary"]
    for i in range(5):
        m.L2[i] = meta["L2"][i]
    m.seq_len = meta["seq_len"]
    m.bwt_size = meta["bwt_size"]
    return m


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def sa_positions(words: np.ndarray, meta: dict, samples: np.ndarray, rows, intv: int = 32) -> np.ndarray:
    """bwt_sa for each row (upstream bwt.c:88-96): text position of the suffix at that row, from the
    SA samples of rows intv, 2 intv, ... (row 0, the $ row, gives all ones as bwa's -1)."""
    L = lib()
    m = make_meta(meta)
    w = np.ascontiguousarray(words, dtype=np.uint32)
    sm = np.ascontiguousarray(samples, dtype=np.uint64)
    return np.array([L.orc_sa(_ptr(w), C.byref(m), _ptr(sm), intv, int(k)) for k in rows], dtype=np.uint64)


def sa2pos(words: np.ndarray, meta: dict, samples: np.ndarray, rows, ref_len, intv: int = 32):
    """bwa_sa2pos for each row (upstream bwase.c:112-123): (forward-strand position of the hit's
    first base, strand 1 forward / 0 reverse); all ones across the forward/reverse boundary."""
    L = lib()
    m = make_meta(meta)
    w = np.ascontiguousarray(words, dtype=np.uint32)
    sm = np.ascontiguousarray(samples, dtype=np.uint64)
    rl = np.broadcast_to(np.asarray(ref_len), np.shape(rows))
    pos, st = [], []
    s = C.c_int()
    for k, r in zip(rows, rl):
        pos.append(L.orc_sa2pos(_ptr(w), C.byref(m), _ptr(sm), intv, int(k), int(r), C.byref(s)))
        st.append(s.value)
    return np.array(pos, dtype=np.uint64), np.array(st, dtype=np.int32)


# --------------------------------------------------------------------------
# options (reference src/opt_build.h:23-36, src/galn.cu:185-250)

GAP_OPT_FIELDS = ("s_mm", "s_gapo", "s_gape", "mode", "indel_end_skip", "max_del_occ", "max_entries",
                  "fnr", "max_diff", "max_gapo", "max_gape", "max_seed_diff", "seed_len", "n_threads",
                  "max_top2", "trim_qual")


def default_gap_opt() -> dict:
    return dict(s_mm=3, s_gapo=11, s_gape=4, mode=MODE_GAPE | MODE_COMPREAD, indel_end_skip=5,
                max_del_occ=10, max_entries=2000000, fnr=float(np.float32(0.04)), max_diff=-1,
                max_gapo=1, max_gape=6, max_seed_diff=2, seed_len=32, n_threads=1, max_top2=30,
                trim_qual=0)


def parse_aln_opts(args: list[str]) -> dict:
    """The `galn aln` option list (without positionals) -> gap_opt_t fields."""
    o = default_gap_opt()
    opte = -1
    i = 0
    while i < len(args):
        a = args[i]
        if a == "-L":
            o["mode"] |= MODE_LOGGAP
        elif a == "-N":
            o["mode"] |= MODE_NONSTOP
            o["max_top2"] = 0x7FFFFFFF
        else:
            v = args[i + 1]
            i += 1
            if a == "-n":
                if "." in v:
                    o["fnr"], o["max_diff"] = float(np.float32(float(v))), -1
                else:
                    o["max_diff"], o["fnr"] = int(v), -1.0
            elif a == "-e":
                opte = int(v)
            elif a == "-q":
                if int(v) >= 1:
                    raise ValueError("-q (quality trimming) is not supported")
            else:
                key = {"-o": "max_gapo", "-M": "s_mm", "-O": "s_gapo", "-E": "s_gape", "-d": "max_del_occ",
                       "-i": "indel_end_skip", "-l": "seed_len", "-k": "max_seed_diff", "-m": "max_entries",
                       "-dt": "n_threads", "-t": "n_threads", "-R": "max_top2"}[a]
                o[key] = int(v)
        i += 1
    if opte > 0:
        o["max_gape"] = opte
        o["mode"] &= ~MODE_GAPE
    if o["n_threads"] < 1:
        o["n_threads"] = 1
    return o


def sai_header(o: dict) -> bytes:
    vals = [o[f] for f in GAP_OPT_FIELDS]
    return b"SAI\x01" + struct.pack("<7if8i", *vals)


def cal_maxdiff(length: int, err: float, thres: float) -> int:
    return lib().orc_cal_maxdiff(int(length), float(err), float(thres))


def batch_ids(n: int) -> np.ndarray:
    """bwa's 0x40000-read batches for a clean input (no empty or malformed records)."""
    return (np.arange(n, dtype=np.int64) // HOST_ALN_BATCH).astype(np.int32)


def batch_options(o: dict, lens: np.ndarray, bid: np.ndarray):
    """Per-batch search options (Opt array), n_stacks, and per-read max_diff
    (reference src/aln_host.c:23-43)."""
    nb = int(bid.max()) + 1 if bid.size else 1
    bopts = (Opt * nb)()
    nst = np.zeros(nb, dtype=np.int32)
    maxd = np.zeros(lens.size, dtype=np.int32)
    fnr = float(np.float32(o["fnr"]))
    for b in range(nb):
        sel = bid == b
        ml = int(lens[sel].max()) if sel.any() else 0
        bo = bopts[b]
        for f in ("s_mm", "s_gapo", "s_gape", "mode", "indel_end_skip", "max_del_occ", "max_entries",
                  "max_diff", "max_gapo", "max_gape", "max_seed_diff", "seed_len", "max_top2"):
            setattr(bo, f, int(o[f]))
        if fnr > 0.0:
            bo.max_diff = cal_maxdiff(ml, 0.02, fnr)
        if bo.max_diff < bo.max_gapo:
            bo.max_gapo = bo.max_diff
        nst[b] = (bo.max_diff + 1) * bo.s_mm + (bo.max_gapo + 1) * bo.s_gapo + (bo.max_gape + 1) * bo.s_gape
        if fnr > 0.0:
            ul = np.unique(lens[sel])
            table = {int(x): cal_maxdiff(int(x), 0.02, fnr) for x in ul}
            maxd[sel] = np.array([table[int(x)] for x in lens[sel]], dtype=np.int32)
        else:
            maxd[sel] = bo.max_diff
    return bopts, nst, maxd


# --------------------------------------------------------------------------
# search


class OracleResult:
    def __init__(self, n_aln, rec_off, rec, pops, trace, occ_bytes, occ_calls, peak):
        self.n_aln, self.rec_off, self.rec = n_aln, rec_off, rec
        self.pops, self.trace, self.occ_bytes, self.occ_calls, self.peak = pops, trace, occ_bytes, occ_calls, peak

    def records(self, i: int) -> np.ndarray:
        n = max(int(self.n_aln[i]), 0)
        o = int(self.rec_off[i])
        return self.rec[o:o + n]


def search(words: np.ndarray, meta: dict, query: np.ndarray, qoff: np.ndarray, qlen: np.ndarray,
           maxdiff: np.ndarray, bid: np.ndarray, bopts, nstacks: np.ndarray, seed_len: int,
           threads: int = 0) -> OracleResult:
    L = lib()
    m = make_meta(meta)
    n = int(qlen.size)
    q = np.ascontiguousarray(query, dtype=np.uint8)
    qo = np.ascontiguousarray(qoff, dtype=np.uint64)
    ql = np.ascontiguousarray(qlen, dtype=np.int32)
    md = np.ascontiguousarray(maxdiff, dtype=np.int32)
    bi = np.ascontiguousarray(bid, dtype=np.int32)
    ns = np.ascontiguousarray(nstacks, dtype=np.int32)
    w = np.ascontiguousarray(words, dtype=np.uint32)
    h = L.orc_search_batch(_ptr(w), C.byref(m), n, _ptr(q), _ptr(qo), _ptr(ql), _ptr(md), _ptr(bi),
                           C.cast(bopts, C.c_void_p), _ptr(ns), int(seed_len), int(threads))
    nrec = int(L.orc_result_nrec(h))
    n_aln = np.zeros(n, np.int32)
    rec_off = np.zeros(n, np.uint64)
    rec = np.zeros((max(nrec, 1), 3), np.uint64)
    pops = np.zeros(n, np.uint64)
    trace = np.zeros(n, np.uint64)
    ob = np.zeros(n, np.uint64)
    oc = np.zeros(n, np.uint64)
    peak = np.zeros(n, np.int32)
    L.orc_result_copy(h, _ptr(n_aln), _ptr(rec_off), _ptr(rec), _ptr(pops), _ptr(trace), _ptr(ob), _ptr(oc),
                      _ptr(peak))
    L.orc_result_free(h)
    return OracleResult(n_aln, rec_off, rec[:nrec], pops, trace, ob, oc, peak)


def sai_bytes(o: dict, n_aln: np.ndarray, rec_off: np.ndarray, rec: np.ndarray) -> bytes:
    """The .sai stream for per-read results in input order; an errored read gets
    0 records (reference src/galn.cu:1250-1263)."""
    parts = [sai_header(o)]
    for i in range(n_aln.size):
        n = int(n_aln[i])
        if n <= 0:
            parts.append(struct.pack("<i", 0))
            continue
        o0 = int(rec_off[i])
        parts.append(struct.pack("<i", n))
        parts.append(np.ascontiguousarray(rec[o0:o0 + n], dtype=np.uint64).tobytes())
    return b"".join(parts)


def run_workload(words, meta, query, qoff, qlen, opts: dict, threads: int = 0):
    """Full oracle run of a clean read set: batches, options, search."""
    bid = batch_ids(qlen.size)
    bopts, nst, maxd = batch_options(opts, qlen, bid)
    res = search(words, meta, query, qoff, qlen, maxd, bid, bopts, nst, opts["seed_len"], threads)
    return res
