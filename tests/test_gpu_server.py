"""The serving loop (galn_hip_server_*) on the GPU (pytest -m gpu).

Requests are slices of the serving-mix golden read file, half searched with BWA's default
options and half with the aDNA options, each slice carrying the batch options of the golden
run it comes from. Submitted together to a server whose launches are small, the requests'
reads are split over several launches and mixed with each other; every request must give
the same records, pop counts and trace hashes as the same reads searched alone, and the
pop counts and trace hashes the reference recorded for those reads."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden_inputs, load_reads, run_opts
from galn_amd import index as ix
from galn_amd import native as nv

pytestmark = pytest.mark.gpu

SIZES = [1, 7, 100, 333, 1000, 64, 5, 1490, 1000]


def requests(golden_dir):
    """[(key, lo, hi, arrays)] slices of the golden reads, alternating option sets."""
    w = golden_inputs("serving_mix", golden_dir)
    runs = {}
    for key in ("default_1", "adna_1"):
        opts, mate = run_opts(w["meta"], key)
        r = load_reads(w["reads"][mate])
        o = nv.gap_opt_from_args(opts)
        bopts, nst, md = nv.prepare_batches(o, r["qlen"], r["batch"])
        runs[key] = (r, bopts, nst, md)
    out, lo = [], 0
    for j, size in enumerate(SIZES):
        key = ("default_1", "adna_1")[j % 2]
        r, bopts, nst, md = runs[key]
        hi = min(lo + size, r["n"])
        qo, ql = r["qoff"][lo:hi], r["qlen"][lo:hi]
        q0 = int(qo.min())
        q1 = int((qo + ql).max())
        arrays = (r["query"][q0:q1], qo - np.uint64(q0), ql, md[lo:hi], r["batch"][lo:hi], bopts, nst)
        out.append((key, lo, hi, arrays))
        lo = hi % r["n"]
    return w, out


@pytest.mark.parametrize("slots,heavy_pops", [(1, 0), (4, 0), (3, 1)])
def test_requests_in_a_mix_equal_alone_and_reference(slots, heavy_pops, golden_dir):
    """With one launch slot, and with four (a request's reads then run in several launches at
    once and may finish out of order), every request equals its reads searched alone; also with
    every read that pops sent to the wave engine (cutoff 1)."""
    print(f"case server mix: 9 requests, 2 option sets, {slots} slots, cutoff {heavy_pops}", flush=True)
    w, reqs = requests(golden_dir)
    meta, words = ix.read_bwt(w["bwt"])
    exp = w["expected"]
    with nv.Index(words, meta, 0) as idx:
        alone = []
        for key, lo, hi, arrays in reqs:
            rs = nv.ReadSet(idx, *arrays)
            rs.search(nv.FORM_BATCHED, nv.GALN_HIP_TRACE)
            alone.append(rs.results(want_pops=True, want_trace=True))
            rs.close()
        with nv.Server(idx, max_launch_reads=512, quantum_reads=96, flags=nv.GALN_HIP_TRACE, slots=slots,
                       heavy_pops=heavy_pops) as srv:
            ids = [srv.submit(*arrays) for _, _, _, arrays in reqs]
            assert all(i is not None for i in ids)
            got = []
            for i in ids:
                assert srv.wait(i, 60000)
                got.append(srv.result(i, want_trace=True))
            st = srv.stats()
        assert st["requests"] == len(reqs) and st["reads"] == sum(hi - lo for _, lo, hi, _ in reqs)
        assert st["launches"] >= st["reads"] // 512
        assert st["slots"] == slots and 1 <= st["most_in_flight"] <= slots
        assert st["launch_reads"] == 512  # a configured launch size stays fixed
    for (key, lo, hi, _), a, (g, lat) in zip(reqs, alone, got):
        assert lat >= 0.0
        np.testing.assert_array_equal(g.n_aln, a.n_aln)
        np.testing.assert_array_equal(g.pops, a.pops)
        np.testing.assert_array_equal(g.trace, a.trace)
        for i in range(hi - lo):
            np.testing.assert_array_equal(g.records(i), a.records(i))
        np.testing.assert_array_equal(a.n_aln, exp[key + "_n_aln"][lo:hi])
        np.testing.assert_array_equal(a.pops, exp[key + "_pops"][lo:hi])
        np.testing.assert_array_equal(a.trace, exp[key + "_trace"][lo:hi])


def test_admission_limit_and_release(golden_dir):
    print("case server admission", flush=True)
    w, reqs = requests(golden_dir)
    meta, words = ix.read_bwt(w["bwt"])
    big = reqs[7][3]  # 1490 reads
    small = reqs[0][3]
    with nv.Index(words, meta, 0) as idx:
        with nv.Server(idx, max_queued_bytes=1000) as srv:
            a = srv.submit(*big)  # an empty queue takes any request
            assert a is not None
            assert srv.submit(*small) is None  # over the limit until the first is collected
            assert srv.wait(a, 60000)
            ra, _ = srv.result(a)
            b = srv.submit(*small)
            assert b is not None and srv.wait(b, 60000)
            rb, _ = srv.result(b)
            # the default launch size adapts within its bounds
            assert 262144 <= srv.stats()["launch_reads"] <= 1 << 20
    assert ra.n_aln.size == 1490 and rb.n_aln.size == 1


def test_failed_launch_releases_request_and_server_goes_on(golden_dir, monkeypatch):
    """A launch that fails (injected on the host, before the device sees it) fails every request
    in it; a request spread over several launches is taken out of the queue before its error is
    published, so collecting it while its later reads were still queued is safe, and the server
    keeps serving the requests after it."""
    print("case server failed launch", flush=True)
    w, reqs = requests(golden_dir)
    meta, words = ix.read_bwt(w["bwt"])
    big = reqs[7][3]  # 1490 reads: six launches of 256
    key, lo, hi, small = reqs[4]
    monkeypatch.setenv("GALN_HIP_TEST_SERVER_FAIL", "2")
    with nv.Index(words, meta, 0) as idx:
        with nv.Server(idx, max_launch_reads=256, quantum_reads=64, flags=nv.GALN_HIP_TRACE) as srv:
            a = srv.submit(*big)
            with pytest.raises(nv.GalnError, match="injected"):
                srv.wait(a, 60000)
            with pytest.raises(nv.GalnError, match="injected"):
                srv.result(a)
            b = srv.submit(*small)
            assert srv.wait(b, 60000)
            rb, _ = srv.result(b, want_trace=True)
            st = srv.stats()
    exp = w["expected"]
    np.testing.assert_array_equal(rb.n_aln, exp[key + "_n_aln"][lo:hi])
    np.testing.assert_array_equal(rb.pops, exp[key + "_pops"][lo:hi])
    np.testing.assert_array_equal(rb.trace, exp[key + "_trace"][lo:hi])
    # the failed request's later reads were never launched
    assert st["reads"] < len(big[2]) + (hi - lo)


def test_index_is_busy_while_a_server_holds_it(golden_dir):
    """While a server is open, direct searches on its index are refused (GALN_HIP_E_BUSY): the
    index's stream and stack pool belong to the server's worker. After close they work again."""
    print("case server busy index", flush=True)
    w, reqs = requests(golden_dir)
    meta, words = ix.read_bwt(w["bwt"])
    arrays = reqs[2][3]
    with nv.Index(words, meta, 0) as idx:
        srv = nv.Server(idx)
        with pytest.raises(nv.GalnError, match=r"failed \(-6\)"):
            nv.ReadSet(idx, *arrays)
        rid = srv.submit(*arrays)
        assert srv.wait(rid, 60000)
        srv.result(rid)
        srv.close()
        rs = nv.ReadSet(idx, *arrays)
        rs.search(nv.FORM_BATCHED)
        rs.close()
