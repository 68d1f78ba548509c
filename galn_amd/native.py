"""ctypes mirror of the two C ABIs (include/galn_hip.h, include/galn_host.h).

The search runs only through ``libgaln_hip.so``; this module raises if the
library is missing. Structures mirror the headers field for field.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libgaln_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "libgaln_host.so")
GALN_BIN = os.path.join(PKG, "bin", "galn")

GALN_HIP_TRACE = 0x1
GALN_HIP_NOFUSE = 0x2
GALN_HIP_STATS = 0x4
GALN_HIP_GRAPH = 0x8
FORM_PLAIN, FORM_BATCHED = 0, 1
E_RECCAP = -5

HIP_SYMBOLS = (
    "galn_hip_abi_version", "galn_hip_get_limits", "galn_hip_last_error", "galn_hip_index_open",
    "galn_hip_index_close", "galn_hip_index_device", "galn_hip_search_plain", "galn_hip_search_batched",
    "galn_hip_readset_upload", "galn_hip_readset_free", "galn_hip_readset_search", "galn_hip_readset_results",
    "galn_hip_readset_reload", "galn_hip_readset_autotune", "galn_hip_readset_kernel_ms", "galn_hip_readset_launches",
    "galn_hip_readset_engine_stats", "galn_hip_server_open", "galn_hip_server_submit", "galn_hip_server_wait",
    "galn_hip_server_result", "galn_hip_server_stats", "galn_hip_server_close", "galn_hip_gather_ceiling",
    "galn_hip_search_device_noalloc", "galn_hip_bwt_words", "galn_hip_build_bwt", "galn_hip_index_set_sa",
    "galn_hip_sa_positions",
)
HOST_SYMBOLS = (
    "galn_gap_opt_default", "galn_parse_aln_args", "galn_cal_maxdiff", "galn_batch_opt", "galn_read_maxdiff",
    "galn_prepare_batches", "galn_load_bwt", "galn_free_bwt", "galn_read_fastq", "galn_readbuf_n",
    "galn_readbuf_n_batches", "galn_readbuf_query", "galn_readbuf_qoff", "galn_readbuf_qlen", "galn_readbuf_batch",
    "galn_readbuf_name", "galn_readbuf_free", "galn_sai_write", "galn_aln_main", "galn_cpu_search",
)


class BwtMeta(C.Structure):
    _fields_ = [("primary", C.c_uint64), ("L2", C.c_uint64 * 5), ("seq_len", C.c_uint64), ("bwt_size", C.c_uint64)]


class SearchOpt(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "s_mm", "s_gapo", "s_gape", "mode", "indel_end_skip", "max_del_occ", "max_entries", "max_diff",
        "max_gapo", "max_gape", "max_seed_diff", "seed_len", "max_top2")]


class AlnRec(C.Structure):
    _fields_ = [("packed", C.c_uint64), ("k", C.c_uint64), ("l", C.c_uint64)]


class Reads(C.Structure):
    _fields_ = [("n_reads", C.c_int32), ("query", C.c_void_p), ("qoff", C.c_void_p), ("qlen", C.c_void_p),
                ("max_diff", C.c_void_p), ("batch", C.c_void_p), ("n_batches", C.c_int32), ("bopts", C.c_void_p),
                ("n_stacks", C.c_void_p)]


class Out(C.Structure):
    _fields_ = [("n_aln", C.c_void_p), ("rec_off", C.c_void_p), ("rec", C.c_void_p), ("rec_cap", C.c_uint64),
                ("n_rec", C.c_uint64), ("n_pops", C.c_void_p), ("trace", C.c_void_p), ("occ_bytes", C.c_void_p)]


class Limits(C.Structure):
    _fields_ = [("max_read_len", C.c_int32), ("max_n_stacks", C.c_int32), ("max_diff", C.c_int32),
                ("max_gapo_plus_gape", C.c_int32), ("max_seq_len", C.c_uint64)]


class Tune(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("block_threads", "waves_per_cu", "chunks_per_lane", "rec_per_lane",
                                         "sort_by_cost", "heavy_pops", "heavy_waves_per_simd")]


class ServerCfg(C.Structure):
    _fields_ = [("max_launch_reads", C.c_int32), ("quantum_reads", C.c_int32), ("max_queued_bytes", C.c_uint64),
                ("flags", C.c_uint32), ("form", C.c_int32), ("slots", C.c_int32), ("heavy_pops", C.c_int32)]


class GapOpt(C.Structure):
    _fields_ = [("s_mm", C.c_int32), ("s_gapo", C.c_int32), ("s_gape", C.c_int32), ("mode", C.c_int32),
                ("indel_end_skip", C.c_int32), ("max_del_occ", C.c_int32), ("max_entries", C.c_int32),
                ("fnr", C.c_float), ("max_diff", C.c_int32), ("max_gapo", C.c_int32), ("max_gape", C.c_int32),
                ("max_seed_diff", C.c_int32), ("seed_len", C.c_int32), ("n_threads", C.c_int32),
                ("max_top2", C.c_int32), ("trim_qual", C.c_int32)]


class AlnArgs(C.Structure):
    _fields_ = [("opt", GapOpt), ("prefix", C.c_char_p), ("reads", C.c_char_p), ("out", C.c_char_p),
                ("device", C.c_int), ("form", C.c_int), ("verbose", C.c_int), ("trace", C.c_int),
                ("cpu_only", C.c_int), ("verify", C.c_int), ("cpu_threads", C.c_int), ("cpu_interleave", C.c_int),
                ("block_batches", C.c_int)]


class GalnError(RuntimeError):
    pass


_hip = None
_host = None


def hip_lib():
    """libgaln_hip.so; raises if it is not built (there is no fallback search)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise GalnError(f"{HIP_LIB} is missing: build it with `python -m galn_amd.build`")
        L = C.CDLL(HIP_LIB)
        P = C.c_void_p
        L.galn_hip_last_error.restype = C.c_char_p
        L.galn_hip_index_open.argtypes = [P, C.POINTER(BwtMeta), C.c_int, C.POINTER(P)]
        L.galn_hip_index_close.argtypes = [P]
        L.galn_hip_readset_upload.argtypes = [P, C.POINTER(Reads), C.POINTER(P)]
        L.galn_hip_readset_free.argtypes = [P]
        L.galn_hip_readset_reload.argtypes = [P, C.POINTER(Reads)]
        L.galn_hip_readset_autotune.argtypes = [P, P, C.c_uint32, C.POINTER(Tune)]
        L.galn_hip_server_open.argtypes = [P, C.POINTER(ServerCfg), C.POINTER(P)]
        L.galn_hip_server_submit.argtypes = [P, C.POINTER(Reads), C.POINTER(C.c_uint64)]
        L.galn_hip_server_wait.argtypes = [P, C.c_uint64, C.c_int]
        L.galn_hip_server_result.argtypes = [P, C.c_uint64, C.POINTER(Out), C.POINTER(C.c_double)]
        L.galn_hip_server_stats.argtypes = [P, P]
        L.galn_hip_server_close.argtypes = [P]
        L.galn_hip_readset_search.argtypes = [P, P, C.c_int, C.c_uint32, C.POINTER(Tune)]
        L.galn_hip_readset_results.argtypes = [P, C.POINTER(Out)]
        L.galn_hip_readset_kernel_ms.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.galn_hip_readset_launches.argtypes = [P, C.c_int, P, P, P]
        L.galn_hip_readset_engine_stats.argtypes = [P, P]
        L.galn_hip_search_plain.argtypes = [P, C.POINTER(Reads), C.POINTER(Out), C.c_uint32]
        L.galn_hip_search_batched.argtypes = [P, C.POINTER(Reads), C.POINTER(Out), C.c_uint32, C.POINTER(Tune)]
        L.galn_hip_search_device_noalloc.argtypes = [P, C.POINTER(Reads), C.POINTER(Out), C.c_uint32]
        L.galn_hip_bwt_words.restype = C.c_uint64
        L.galn_hip_bwt_words.argtypes = [C.c_uint64]
        L.galn_hip_build_bwt.argtypes = [P, C.c_uint64, C.c_int, C.POINTER(BwtMeta), P, P, P]
        L.galn_hip_index_set_sa.argtypes = [P, P, C.c_uint64, C.c_uint32]
        L.galn_hip_sa_positions.argtypes = [P, P, C.c_uint64, P, P, P]
        L.galn_hip_get_limits.argtypes = [C.POINTER(Limits)]
        L.galn_hip_gather_ceiling.argtypes = [C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int,
                                              C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _hip = L
    return _hip


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise GalnError(f"{HOST_LIB} is missing: build it with `python -m galn_amd.build`")
        L = C.CDLL(HOST_LIB)
        P = C.c_void_p
        L.galn_cal_maxdiff.restype = C.c_int
        L.galn_cal_maxdiff.argtypes = [C.c_int, C.c_double, C.c_double]
        L.galn_gap_opt_default.argtypes = [C.POINTER(GapOpt)]
        L.galn_parse_aln_args.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(AlnArgs)]
        L.galn_batch_opt.argtypes = [C.POINTER(GapOpt), C.c_int, C.POINTER(SearchOpt)]
        L.galn_prepare_batches.argtypes = [C.POINTER(GapOpt), C.c_int, P, P, C.c_int, P, P, P]
        L.galn_load_bwt.restype = P
        L.galn_load_bwt.argtypes = [C.c_char_p, C.POINTER(BwtMeta)]
        L.galn_free_bwt.argtypes = [P]
        L.galn_read_fastq.restype = P
        L.galn_read_fastq.argtypes = [C.c_char_p]
        for f in ("galn_readbuf_n", "galn_readbuf_n_batches"):
            getattr(L, f).restype = C.c_int32
            getattr(L, f).argtypes = [P]
        for f in ("galn_readbuf_query", "galn_readbuf_qoff", "galn_readbuf_qlen", "galn_readbuf_batch"):
            getattr(L, f).restype = P
            getattr(L, f).argtypes = [P]
        L.galn_readbuf_name.restype = C.c_char_p
        L.galn_readbuf_name.argtypes = [P, C.c_int32]
        L.galn_readbuf_free.argtypes = [P]
        L.galn_aln_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        L.galn_cpu_search.argtypes = [P, C.POINTER(BwtMeta), C.POINTER(Reads), P, C.c_int32, C.POINTER(Out),
                                      C.c_int32, C.c_int32]
        _host = L
    return _host


def last_error() -> str:
    return hip_lib().galn_hip_last_error().decode(errors="replace")


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def gather_ceiling(ws_bytes: int, line_bytes: int = 64, waves_per_cu: int = 16, iters: int = 2000,
                   device: int = 0) -> tuple[float, float]:
    """(line gathers per second, ms) of the pointer-chase microbenchmark."""
    lps, ms = C.c_double(), C.c_double()
    rc = hip_lib().galn_hip_gather_ceiling(device, int(ws_bytes), int(line_bytes), int(waves_per_cu), int(iters),
                                           C.byref(lps), C.byref(ms))
    if rc != 0:
        raise GalnError(f"gather_ceiling failed ({rc})")
    return lps.value, ms.value


def limits() -> dict:
    lim = Limits()
    hip_lib().galn_hip_get_limits(C.byref(lim))
    return {f: getattr(lim, f) for f, _ in Limits._fields_}


def build_bwt(fwd2: np.ndarray, device: int = 0, want_sa: bool = False) -> tuple[BwtMeta, np.ndarray, np.ndarray | None, dict]:
    """galn_hip_build_bwt: the .bwt body of T$ (T = F . revcomp(F)) for forward 2-bit codes,
    built on `device`; with `want_sa` also the SA samples of rows 32, 64, ... Returns (meta,
    body words, samples or None, {chunks, rounds, largest_chunk})."""
    f = np.ascontiguousarray(fwd2, dtype=np.uint8)
    n = 2 * int(f.size)
    L = hip_lib()
    words = np.empty(int(L.galn_hip_bwt_words(n)), np.uint32)
    sa = np.empty(max((n + 32) // 32 - 1, 1), np.uint64) if want_sa else None
    st = np.zeros(4, np.float64)
    m = BwtMeta()
    rc = L.galn_hip_build_bwt(_p(f), int(f.size), int(device), C.byref(m), _p(words), _p(sa) if sa is not None else None,
                              _p(st))
    if rc != 0:
        raise GalnError(f"build_bwt failed ({rc}): {last_error()}")
    info = {"chunks": int(st[0]), "rounds": int(st[1]), "largest_chunk": int(st[2])}
    return m, words, (sa[: (n + 32) // 32 - 1] if sa is not None else None), info


# --------------------------------------------------------------------------
# host side


def parse_aln_args(args: list[str]) -> tuple[int, AlnArgs]:
    """galn_parse_aln_args over a Python list (without "galn aln")."""
    arr = (C.c_char_p * max(1, len(args)))(*[a.encode() for a in args])
    out = AlnArgs()
    rc = host_lib().galn_parse_aln_args(len(args), arr, C.byref(out))
    out._keep = arr  # the parsed strings point into arr
    return rc, out


def gap_opt_from_args(opt_args: list[str]) -> GapOpt:
    rc, a = parse_aln_args(list(opt_args) + ["prefix", "reads"])
    if rc != 0:
        raise GalnError(f"bad aln options: {opt_args}")
    return a.opt


def gap_opt_bytes(o: GapOpt) -> bytes:
    return bytes(memoryview(o))


def prepare_batches(o: GapOpt, qlen: np.ndarray, batch: np.ndarray):
    """Per-batch SearchOpt array, n_stacks and per-read max_diff (galn_prepare_batches)."""
    n = int(qlen.size)
    ql = np.ascontiguousarray(qlen, dtype=np.int32)
    bi = np.ascontiguousarray(batch, dtype=np.int32)
    nb = int(bi[-1]) + 1 if n else 1
    bopts = (SearchOpt * nb)()
    nst = np.zeros(nb, np.int32)
    md = np.zeros(max(n, 1), np.int32)
    got = host_lib().galn_prepare_batches(C.byref(o), n, _p(ql), _p(bi), nb, C.cast(bopts, C.c_void_p), _p(nst), _p(md))
    if got != nb:
        raise GalnError(f"prepare_batches: batch ids not 0..{nb - 1} in order")
    return bopts, nst, md[:n]


def load_bwt(path: str) -> tuple[BwtMeta, np.ndarray]:
    m = BwtMeta()
    L = host_lib()
    p = L.galn_load_bwt(path.encode(), C.byref(m))
    if not p:
        raise GalnError(f"cannot load {path}")
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(int(m.bwt_size),)).copy()
    L.galn_free_bwt(p)
    return m, arr


def meta_from_dict(d: dict) -> BwtMeta:
    m = BwtMeta()
    m.primary = int(d["primary"])
    for i in range(5):
        m.L2[i] = int(d["L2"][i])
    m.seq_len = int(d["seq_len"])
    m.bwt_size = int(d["bwt_size"])
    return m


def read_fastq(path: str) -> dict:
    """galn_read_fastq: queries, offsets, lengths, batch ids and names."""
    L = host_lib()
    h = L.galn_read_fastq(path.encode())
    if not h:
        raise GalnError(f"cannot read {path}")
    try:
        n = L.galn_readbuf_n(h)
        nb = L.galn_readbuf_n_batches(h)

        def arr(fn, ctype, count):
            if count == 0:
                return np.zeros(0, dtype=np.ctypeslib.as_ctypes_type(ctype) and ctype)
            return np.ctypeslib.as_array(C.cast(getattr(L, fn)(h), C.POINTER(np.ctypeslib.as_ctypes_type(ctype))),
                                         shape=(count,)).copy()

        qlen = arr("galn_readbuf_qlen", np.int32, n)
        qoff = arr("galn_readbuf_qoff", np.uint64, n)
        total = int(qoff[-1] + qlen[-1]) if n else 0
        query = arr("galn_readbuf_query", np.uint8, total)
        batch = arr("galn_readbuf_batch", np.int32, n)
        names = [L.galn_readbuf_name(h, i).decode() for i in range(n)] if n <= 1_000_000 else None
    finally:
        L.galn_readbuf_free(h)
    return {"n": n, "n_batches": nb, "query": query, "qoff": qoff, "qlen": qlen, "batch": batch, "names": names}


# --------------------------------------------------------------------------
# search


class Index:
    """A .bwt body resident in one device's HBM. Closing it first closes the read sets and
    servers opened on it (the library refuses to close an index that still has them)."""

    def __init__(self, words: np.ndarray, meta: BwtMeta | dict, device: int = 0):
        self._children = weakref.WeakSet()
        self.meta = meta if isinstance(meta, BwtMeta) else meta_from_dict(meta)
        self._words = np.ascontiguousarray(words, dtype=np.uint32)
        h = C.c_void_p()
        rc = hip_lib().galn_hip_index_open(_p(self._words), C.byref(self.meta), int(device), C.byref(h))
        if rc != 0:
            raise GalnError(f"index_open failed ({rc}): {last_error()}")
        self.h = h
        self._words = None

    def set_sa(self, samples: np.ndarray, intv: int = 32) -> None:
        """galn_hip_index_set_sa: the SA samples of rows intv, 2 intv, ... (a .sa file's body)."""
        sm = np.ascontiguousarray(samples, dtype=np.uint64)
        rc = hip_lib().galn_hip_index_set_sa(self.h, _p(sm), int(sm.size), int(intv))
        if rc != 0:
            raise GalnError(f"index_set_sa failed ({rc}): {last_error()}")

    def sa_positions(self, rows, ref_len=None):
        """galn_hip_sa_positions: bwt_sa of each row, or with ref_len (per row) bwa_sa2pos's
        (position, strand)."""
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        pos = np.empty(max(r.size, 1), np.uint64)
        if ref_len is None:
            rc = hip_lib().galn_hip_sa_positions(self.h, _p(r), int(r.size), _p(pos), None, None)
        else:
            rl = np.ascontiguousarray(np.broadcast_to(ref_len, r.shape), dtype=np.int32)
            st = np.empty(max(r.size, 1), np.int32)
            rc = hip_lib().galn_hip_sa_positions(self.h, _p(r), int(r.size), _p(pos), _p(rl), _p(st))
        if rc != 0:
            raise GalnError(f"sa_positions failed ({rc}): {last_error()}")
        return pos[: r.size] if ref_len is None else (pos[: r.size], st[: r.size])

    def close(self):
        if self.h:
            for c in list(self._children):
                c.close()
            rc = hip_lib().galn_hip_index_close(self.h)
            if rc != 0:
                raise GalnError(f"index_close failed ({rc}): {last_error()}")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Results:
    def __init__(self, n_aln, rec_off, rec, pops, trace, occ_bytes):
        self.n_aln, self.rec_off, self.rec = n_aln, rec_off, rec
        self.pops, self.trace, self.occ_bytes = pops, trace, occ_bytes

    def records(self, i: int) -> np.ndarray:
        n = max(int(self.n_aln[i]), 0)
        o = int(self.rec_off[i])
        return self.rec[o:o + n]


def make_reads(query, qoff, qlen, max_diff, batch, bopts, n_stacks) -> tuple[Reads, list]:
    """A Reads structure over contiguous copies of the arrays (returned with it: keep them
    alive while the structure is used)."""
    keep = [np.ascontiguousarray(query, dtype=np.uint8), np.ascontiguousarray(qoff, dtype=np.uint64),
            np.ascontiguousarray(qlen, dtype=np.int32), np.ascontiguousarray(max_diff, dtype=np.int32),
            np.ascontiguousarray(batch, dtype=np.int32), np.ascontiguousarray(n_stacks, dtype=np.int32), bopts]
    q, qo, ql, md, bi, ns, _ = keep
    R = Reads()
    R.n_reads = int(ql.size)
    R.query, R.qoff, R.qlen, R.max_diff, R.batch = _p(q), _p(qo), _p(ql), _p(md), _p(bi)
    R.n_batches = int(ns.size)
    R.bopts = C.cast(bopts, C.c_void_p)
    R.n_stacks = _p(ns)
    return R, keep


class ReadSet:
    """Reads resident on the device (galn_hip_readset_*)."""

    def __init__(self, index: Index, query, qoff, qlen, max_diff, batch, bopts, n_stacks):
        self.index = index
        R, keep = make_reads(query, qoff, qlen, max_diff, batch, bopts, n_stacks)
        self.n = int(R.n_reads)
        h = C.c_void_p()
        rc = hip_lib().galn_hip_readset_upload(index.h, C.byref(R), C.byref(h))
        if rc != 0:
            raise GalnError(f"readset_upload failed ({rc}): {last_error()}")
        self.h = h
        index._children.add(self)

    def reload(self, query, qoff, qlen, max_diff, batch, bopts, n_stacks) -> None:
        """Replace the reads, reusing the device arrays when they are large enough."""
        R, keep = make_reads(query, qoff, qlen, max_diff, batch, bopts, n_stacks)
        rc = hip_lib().galn_hip_readset_reload(self.h, C.byref(R))
        if rc != 0:
            raise GalnError(f"readset_reload failed ({rc}): {last_error()}")
        self.n = int(R.n_reads)

    def search(self, form: int = FORM_BATCHED, flags: int = 0, tune: Tune | None = None) -> None:
        rc = hip_lib().galn_hip_readset_search(self.index.h, self.h, int(form), int(flags),
                                               C.byref(tune) if tune is not None else None)
        if rc != 0:
            raise GalnError(f"readset_search failed ({rc}): {last_error()}")

    def autotune(self, flags: int = 0) -> dict:
        """galn_hip_readset_autotune: the chosen tuning (later searches without one use it)."""
        t = Tune()
        rc = hip_lib().galn_hip_readset_autotune(self.index.h, self.h, int(flags), C.byref(t))
        if rc != 0:
            raise GalnError(f"readset_autotune failed ({rc}): {last_error()}")
        return {f: getattr(t, f) for f, _ in Tune._fields_}

    def launches(self) -> list[dict]:
        """[{tier, ms, algo_bytes}] for every kernel launch of the last search."""
        L = hip_lib()
        n = L.galn_hip_readset_launches(self.h, 0, None, None, None)
        tier = np.zeros(max(n, 1), np.int32)
        ms = np.zeros(max(n, 1), np.float64)
        ab = np.zeros(max(n, 1), np.uint64)
        L.galn_hip_readset_launches(self.h, n, _p(tier), _p(ms), _p(ab))
        return [{"tier": int(tier[i]), "ms": float(ms[i]), "algo_bytes": int(ab[i])} for i in range(n)]

    def engine_stats(self) -> dict:
        """Counters of the wave-per-read engine in the last search."""
        v = np.zeros(16, np.uint64)
        hip_lib().galn_hip_readset_engine_stats(self.h, _p(v))
        keys = ("rounds", "chains", "chains_committed", "steps_committed", "critical_steps", "spec_ticks",
                "commit_ticks", "plan_ticks", "alloc_ticks", "place_ticks", "update_ticks", "granules", "setup_ticks",
                "start_ticks", "after_ticks", "reads")
        return {k: int(v[i]) for i, k in enumerate(keys)}

    def kernel_ms(self) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        hip_lib().galn_hip_readset_kernel_ms(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def result_buffers(self, want_pops=True, want_trace=False, want_bytes=False, rec_cap: int | None = None,
                       alloc=np.empty) -> dict:
        """Arrays for `results(out=...)`, to be reused across searches of this read set;
        `alloc(shape, dtype)` makes each one (e.g. in pinned host memory)."""
        n1 = max(self.n, 1)
        out = {"n_aln": alloc((n1,), np.int32), "rec_off": alloc((n1,), np.uint64),
               "rec": alloc((max(rec_cap if rec_cap is not None else 2 * self.n, 1), 3), np.uint64)}
        for k, want in (("pops", want_pops), ("trace", want_trace), ("bytes", want_bytes)):
            if want:
                out[k] = alloc((n1,), np.uint64)
        return out

    def results(self, want_pops=True, want_trace=False, want_bytes=False, rec_cap: int | None = None,
                out: dict | None = None) -> Results:
        """Records in read order. With `out` (from `result_buffers`) the arrays are filled in
        place, and the returned views share them until the next call."""
        n = self.n
        # every element is written by the library
        def get(k, size, dt):
            if out is None:
                return np.empty(size, dt)
            if k not in out:
                raise GalnError(f"results: no '{k}' array in the output buffers")
            return out[k]

        n_aln = get("n_aln", max(n, 1), np.int32)
        rec_off = get("rec_off", max(n, 1), np.uint64)
        pops = get("pops", max(n, 1), np.uint64) if want_pops else None
        trace = get("trace", max(n, 1), np.uint64) if want_trace else None
        ob = get("bytes", max(n, 1), np.uint64) if want_bytes else None
        for a_, dt in ((n_aln, np.int32), (rec_off, np.uint64), (pops, np.uint64), (trace, np.uint64), (ob, np.uint64)):
            if a_ is not None and (a_.dtype != dt or a_.shape[0] < max(n, 1) or not a_.flags.c_contiguous):
                raise GalnError("results: an output array has the wrong type or size")
        cap = rec_cap if rec_cap is not None else max(2 * n, 1)
        if out is not None:
            cap = out["rec"].shape[0]
        while True:
            if out is not None and cap <= out["rec"].shape[0]:
                rec = out["rec"]
                if rec.dtype != np.uint64 or rec.ndim != 2 or rec.shape[1] != 3 or not rec.flags.c_contiguous:
                    raise GalnError("results: the record array has the wrong type or shape")
            else:
                rec = np.empty((max(cap, 1), 3), np.uint64)
            O = Out()
            O.n_aln, O.rec_off, O.rec, O.rec_cap = _p(n_aln), _p(rec_off), _p(rec), cap
            O.n_pops = _p(pops) if pops is not None else None
            O.trace = _p(trace) if trace is not None else None
            O.occ_bytes = _p(ob) if ob is not None else None
            rc = hip_lib().galn_hip_readset_results(self.h, C.byref(O))
            if rc == E_RECCAP:
                cap = int(O.n_rec)
                if out is not None:
                    out["rec"] = np.empty((max(cap, 1), 3), np.uint64)  # grown; kept for later calls
                continue
            if rc != 0:
                raise GalnError(f"readset_results failed ({rc}): {last_error()}")
            return Results(n_aln[:n], rec_off[:n], rec[:int(O.n_rec)],
                           pops[:n] if pops is not None else None, trace[:n] if trace is not None else None,
                           ob[:n] if ob is not None else None)

    def close(self):
        if self.h:
            hip_lib().galn_hip_readset_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop (galn_hip_server_*): the index stays resident; requests are queued
    and searched in shared launches, up to `slots` launches on the device at once.
    `max_launch_reads` 0 sizes launches from the server's recent throughput (262,144 to
    1,048,576 reads); a positive value fixes it."""

    def __init__(self, index: Index, max_launch_reads: int = 0, quantum_reads: int = 0, max_queued_bytes: int = 0,
                 flags: int = 0, form: int = FORM_BATCHED, slots: int = 0, heavy_pops: int = 0):
        cfg = ServerCfg()
        cfg.max_launch_reads, cfg.quantum_reads = int(max_launch_reads), int(quantum_reads)
        cfg.max_queued_bytes, cfg.flags, cfg.form = int(max_queued_bytes), int(flags), int(form)
        cfg.slots, cfg.heavy_pops = int(slots), int(heavy_pops)
        h = C.c_void_p()
        rc = hip_lib().galn_hip_server_open(index.h, C.byref(cfg), C.byref(h))
        if rc != 0:
            raise GalnError(f"server_open failed ({rc}): {last_error()}")
        self.h = h
        self.index = index
        index._children.add(self)
        self.n_of: dict[int, int] = {}

    def submit(self, query, qoff, qlen, max_diff, batch, bopts, n_stacks) -> int | None:
        """The request id, or None when the admission limit is reached (GALN_HIP_E_BUSY)."""
        R, keep = make_reads(query, qoff, qlen, max_diff, batch, bopts, n_stacks)
        rid = C.c_uint64()
        rc = hip_lib().galn_hip_server_submit(self.h, C.byref(R), C.byref(rid))
        if rc == -6:
            return None
        if rc != 0:
            raise GalnError(f"server_submit failed ({rc}): {last_error()}")
        self.n_of[rid.value] = int(R.n_reads)
        return rid.value

    def wait(self, rid: int, timeout_ms: int = -1) -> bool:
        rc = hip_lib().galn_hip_server_wait(self.h, int(rid), int(timeout_ms))
        if rc < 0:
            raise GalnError(f"request {rid} failed ({rc}): {last_error()}")
        return rc == 1

    def result(self, rid: int, want_trace: bool = False) -> tuple[Results, float]:
        """(results, latency in ms) of a finished request; releases it."""
        n = self.n_of[rid]
        n_aln = np.zeros(max(n, 1), np.int32)
        rec_off = np.zeros(max(n, 1), np.uint64)
        pops = np.zeros(max(n, 1), np.uint64)
        trace = np.zeros(max(n, 1), np.uint64) if want_trace else None
        cap = max(2 * n, 1)
        lat = C.c_double()
        while True:
            rec = np.zeros((cap, 3), np.uint64)
            O = Out()
            O.n_aln, O.rec_off, O.rec, O.rec_cap = _p(n_aln), _p(rec_off), _p(rec), cap
            O.n_pops = _p(pops)
            O.trace = _p(trace) if trace is not None else None
            rc = hip_lib().galn_hip_server_result(self.h, int(rid), C.byref(O), C.byref(lat))
            if rc == E_RECCAP:
                cap = max(int(O.n_rec), 1)
                continue
            if rc != 0:
                raise GalnError(f"server_result failed ({rc}): {last_error()}")
            del self.n_of[rid]
            return Results(n_aln[:n], rec_off[:n], rec[:int(O.n_rec)], pops[:n],
                           trace[:n] if trace is not None else None, None), lat.value

    def stats(self) -> dict:
        v = np.zeros(8, np.float64)
        hip_lib().galn_hip_server_stats(self.h, _p(v))
        return {"launches": int(v[0]), "reads": int(v[1]), "requests": int(v[2]), "launch_ms": float(v[3]),
                "most_in_flight": int(v[4]), "slots": int(v[5]), "launch_reads": int(v[6])}

    def close(self):
        if self.h:
            hip_lib().galn_hip_server_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cpu_search(words: np.ndarray, meta: BwtMeta | dict, query, qoff, qlen, max_diff, batch, bopts, n_stacks,
               reads_list: np.ndarray | None = None, threads: int = 0, interleave: int = 8) -> Results:
    """galn_cpu_search: the CPU engine over all reads, or over the read ids in `reads_list`
    (results then by list position)."""
    m = meta if isinstance(meta, BwtMeta) else meta_from_dict(meta)
    w = np.ascontiguousarray(words, dtype=np.uint32)
    R, keep = make_reads(query, qoff, qlen, max_diff, batch, bopts, n_stacks)
    lst = None if reads_list is None else np.ascontiguousarray(reads_list, dtype=np.int32)
    n = int(R.n_reads) if lst is None else int(lst.size)
    n_aln = np.zeros(max(n, 1), np.int32)
    rec_off = np.zeros(max(n, 1), np.uint64)
    pops = np.zeros(max(n, 1), np.uint64)
    trace = np.zeros(max(n, 1), np.uint64)
    ob = np.zeros(max(n, 1), np.uint64)
    cap = 2 * n + 16
    while True:
        rec = np.zeros((max(cap, 1), 3), np.uint64)
        O = Out()
        O.n_aln, O.rec_off, O.rec, O.rec_cap = _p(n_aln), _p(rec_off), _p(rec), cap
        O.n_pops, O.trace, O.occ_bytes = _p(pops), _p(trace), _p(ob)
        rc = host_lib().galn_cpu_search(_p(w), C.byref(m), C.byref(R), _p(lst) if lst is not None else None,
                                        n, C.byref(O), int(threads), int(interleave))
        if rc == E_RECCAP:
            cap = int(O.n_rec)
            continue
        if rc != 0:
            raise GalnError(f"cpu_search failed ({rc})")
        return Results(n_aln[:n], rec_off[:n], rec[:int(O.n_rec)], pops[:n], trace[:n], ob[:n])


def sai_bytes(opt: GapOpt, n_aln: np.ndarray, rec_off: np.ndarray, rec: np.ndarray) -> bytes:
    """The .sai byte stream (the same layout galn_sai_write produces)."""
    parts = [b"SAI\x01", gap_opt_bytes(opt)]
    na = np.where(n_aln > 0, n_aln, 0).astype(np.int32)
    for i in range(na.size):
        parts.append(int(na[i]).to_bytes(4, "little", signed=True))
        if na[i]:
            o = int(rec_off[i])
            parts.append(np.ascontiguousarray(rec[o:o + int(na[i])], dtype=np.uint64).tobytes())
    return b"".join(parts)
