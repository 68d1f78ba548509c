"""The search that allocates nothing (galn_hip_search_device_noalloc), on the GPU.

Reads and outputs live in device memory that the test allocates with torch; the library
carves everything else out of its static per-GPU array. Each call's records, pop counts and
trace hashes must equal those of galn_hip_search_plain on the same reads, read for read, and
the golden runs' .sai md5: first on a whole golden read set, then, in the same process, on a
smaller one (the array keeps the first call's contents). Size limits are refused on the host."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import fixtures as fx
from conftest import golden_inputs, load_reads, run_opts
from galn_amd import index as ix
from galn_amd import native as nv

pytestmark = pytest.mark.gpu

E_ARG = -1


def _inputs(name, key, golden_dir, take=None):
    w = golden_inputs(name, golden_dir)
    opts, mate = run_opts(w["meta"], key)
    r = load_reads(w["reads"][mate])
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, r["qlen"], r["batch"])
    q, qoff, qlen, bi = r["query"], r["qoff"], r["qlen"], r["batch"]
    if take is not None:
        qlen, qoff, md, bi = qlen[:take], qoff[:take], md[:take], bi[:take]
        end = int(qoff[-1]) + int(qlen[-1])
        q = q[:end]
    return w, o, (q, qoff, qlen, md, bi, bopts, nst)


def _plain(index, arrays, flags):
    """The existing entry point, host arrays in and out."""
    R, keep = nv.make_reads(*arrays)
    n = int(R.n_reads)
    n_aln = np.empty(max(n, 1), np.int32)
    rec_off = np.empty(max(n, 1), np.uint64)
    pops = np.empty(max(n, 1), np.uint64)
    trace = np.empty(max(n, 1), np.uint64)
    cap = 4 * n + 64
    while True:
        rec = np.empty((cap, 3), np.uint64)
        O = nv.Out()
        O.n_aln, O.rec_off, O.rec, O.rec_cap = nv._p(n_aln), nv._p(rec_off), nv._p(rec), cap
        O.n_pops, O.trace = nv._p(pops), nv._p(trace)
        rc = nv.hip_lib().galn_hip_search_plain(index.h, C.byref(R), C.byref(O), flags)
        if rc == nv.E_RECCAP:
            cap = int(O.n_rec)
            continue
        assert rc == 0, nv.last_error()
        return nv.Results(n_aln[:n], rec_off[:n], rec, pops[:n], trace[:n], None)


def _noalloc(index, arrays, flags, rec_cap=None):
    """The no-allocation entry point: inputs and outputs in device memory (torch)."""
    import torch
    q, qoff, qlen, md, bi, bopts, nst = arrays
    n = int(qlen.size)
    dev = torch.device("cuda", 0)
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)  # noqa: E731
    d_q = t(q if q.size else np.zeros(1, np.uint8), np.uint8)
    d_qoff = t(qoff.view(np.int64) if qoff.size else np.zeros(1, np.int64), np.int64)
    d_qlen, d_md, d_bi = t(qlen, np.int32), t(md, np.int32), t(bi, np.int32)
    cap = rec_cap if rec_cap is not None else 4 * n + 64
    d_n = torch.full((max(n, 1),), -77, dtype=torch.int32, device=dev)
    d_off = torch.full((max(n, 1),), -1, dtype=torch.int64, device=dev)
    d_rec = torch.zeros((max(cap, 1), 3), dtype=torch.int64, device=dev)
    d_pops = torch.zeros(max(n, 1), dtype=torch.int64, device=dev)
    d_trace = torch.zeros(max(n, 1), dtype=torch.int64, device=dev)
    ns = np.ascontiguousarray(nst, dtype=np.int32)
    R = nv.Reads()
    R.n_reads = n
    R.query, R.qoff, R.qlen = d_q.data_ptr(), d_qoff.data_ptr(), d_qlen.data_ptr()
    R.max_diff, R.batch = d_md.data_ptr(), d_bi.data_ptr()
    R.n_batches = int(ns.size)
    R.bopts = C.cast(bopts, C.c_void_p)
    R.n_stacks = nv._p(ns)
    O = nv.Out()
    O.n_aln, O.rec_off, O.rec, O.rec_cap = d_n.data_ptr(), d_off.data_ptr(), d_rec.data_ptr(), cap
    O.n_pops, O.trace = d_pops.data_ptr(), d_trace.data_ptr()
    torch.cuda.synchronize()
    rc = nv.hip_lib().galn_hip_search_device_noalloc(index.h, C.byref(R), C.byref(O), flags)
    res = nv.Results(d_n.cpu().numpy()[:n], d_off.cpu().numpy().view(np.uint64)[:n],
                     d_rec.cpu().numpy().view(np.uint64), d_pops.cpu().numpy().view(np.uint64)[:n],
                     d_trace.cpu().numpy().view(np.uint64)[:n], None)
    return rc, int(O.n_rec), res


def _same(a, b, n):
    np.testing.assert_array_equal(a.n_aln, b.n_aln)
    np.testing.assert_array_equal(a.pops, b.pops)
    np.testing.assert_array_equal(a.trace, b.trace)
    for i in range(n):
        np.testing.assert_array_equal(a.records(i), b.records(i), err_msg=f"records of read {i}")


def _in_order(res, n):
    """Records in read order with offsets = prefix sums (the layout the .sai writer takes)."""
    cnt = np.maximum(res.n_aln, 0).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.uint64)
    rec = np.concatenate([res.records(i) for i in range(n)] + [np.zeros((0, 3), np.uint64)])
    return off, rec


@pytest.mark.parametrize("name,key", [("adna_mixed", "adna_1"), ("plain_small", "default_1"),
                                      ("deep_repeats", "deep_2")])
def test_noalloc_equals_plain_then_smaller_input(name, key, golden_dir):
    print(f"case noalloc {name}/{key}", flush=True)
    w, o, arrays = _inputs(name, key, golden_dir)
    meta, words = ix.read_bwt(w["bwt"])
    n = int(arrays[2].size)
    with nv.Index(words, meta, 0) as index:
        ref = _plain(index, arrays, nv.GALN_HIP_TRACE)
        rc, n_rec, got = _noalloc(index, arrays, nv.GALN_HIP_TRACE)
        assert rc == 0, nv.last_error()
        assert n_rec == int(np.maximum(got.n_aln, 0).sum())
        _same(got, ref, n)
        off, rec = _in_order(got, n)
        assert fx.md5_bytes(nv.sai_bytes(o, got.n_aln, off, rec)) == w["meta"]["runs"][key]["sai_md5"]
        # second call, smaller input: the array still holds the first call's state
        small = n // 3
        _, _, arrays2 = _inputs(name, key, golden_dir, take=small)
        ref2 = _plain(index, arrays2, nv.GALN_HIP_TRACE)
        rc, n_rec, got2 = _noalloc(index, arrays2, nv.GALN_HIP_TRACE)
        assert rc == 0, nv.last_error()
        assert n_rec == int(np.maximum(got2.n_aln, 0).sum())
        _same(got2, ref2, small)


def test_noalloc_record_array_too_small_and_size_limits(golden_dir):
    print("case noalloc limits", flush=True)
    w, o, arrays = _inputs("plain_small", "default_1", golden_dir)
    meta, words = ix.read_bwt(w["bwt"])
    with nv.Index(words, meta, 0) as index:
        ref = _plain(index, arrays, 0)
        need = int(np.maximum(ref.n_aln, 0).sum())
        assert need > 8
        rc, n_rec, _ = _noalloc(index, arrays, 0, rec_cap=8)
        assert rc == nv.E_RECCAP and n_rec > 8, (rc, n_rec, nv.last_error())
        # still correct on the next call, with room
        rc, n_rec, got = _noalloc(index, arrays, 0)
        assert rc == 0 and n_rec == need
        np.testing.assert_array_equal(got.n_aln, ref.n_aln)
        # more reads than a call may take: refused before any device work
        R = nv.Reads()
        R.n_reads = (1 << 20) + 1
        R.n_batches = 1
        R.bopts = C.cast(arrays[5], C.c_void_p)
        ns = np.ascontiguousarray(arrays[6][:1], dtype=np.int32)
        R.n_stacks = nv._p(ns)
        O = nv.Out()
        assert nv.hip_lib().galn_hip_search_device_noalloc(index.h, C.byref(R), C.byref(O), 0) == E_ARG
        assert "at most 1048576" in nv.last_error()
