"""Genome-scale indexes on the GPU: built by the HIP builder, searched by the HIP search.

* Deep repeats above 2^32 symbols: a 275 Mbp random genome and 8 copies of it with 1-3%
  substitutions and 0.2% indels (2.2 Gbp; the BWT has n ~ 4.4e9 > 2^32 symbols). The .bwt and .sa
  equal the reference builder's (md5s in tests/golden/large_index.json, written by
  make_golden_large.py from oracle/_ref/galn_index_ref). 2 x 150 bp reads of that model are
  searched with `aln -n 0.04 -o 2 -e 6 -k 2`: the plain and batched forms agree bit for bit,
  and a 2,000-read sample equals the oracle read for read (records, pops, traces).
* Human scale: the 3.1 Gbp reference of the human-like model (24 contigs, 41% GC, 45%
  interspersed repeats at 2-20% divergence, tandem repeats; n = 6.2e9). Reads of 30-150 bp with
  deamination are searched with `aln -l 1024 -n 0.01 -o 2`: forms agree, 2,000 reads equal the
  oracle.
SA intervals k, l above 2^32 come out of both searches (checked)."""
from __future__ import annotations

import json
import os
import time

import numpy as np
import pytest

import fixtures as fx
import oracle as orc
from galn_amd import index as ix
from galn_amd import native as nv
from galn_amd import synth

pytestmark = pytest.mark.gpu

LARGE = os.path.join(os.path.dirname(fx.__file__), "large_index.json")


def _search_and_check(meta, words, reads, lens, opts, n_oracle=2000, seed=7):
    """Plain == batched on every read; a sample equals the oracle. Returns the batched results."""
    w_query, w_qoff = synth.encode_queries(reads, lens)
    n = int(lens.size)
    batch = (np.arange(n) // 0x40000).astype(np.int32)
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, lens.astype(np.int32), batch)
    with nv.Index(words, meta, 0) as index:
        rs = nv.ReadSet(index, w_query, w_qoff, lens.astype(np.int32), md, batch, bopts, nst)
        t = time.time()
        rs.search(nv.FORM_BATCHED, nv.GALN_HIP_TRACE)
        res_b = rs.results(want_pops=True, want_trace=True)
        print(f"batched: {n} reads in {time.time() - t:.2f} s", flush=True)
        t = time.time()
        rs.search(nv.FORM_PLAIN, nv.GALN_HIP_TRACE)
        res_p = rs.results(want_pops=True, want_trace=True)
        print(f"plain: {n} reads in {time.time() - t:.2f} s", flush=True)
        rs.close()
    np.testing.assert_array_equal(res_b.n_aln, res_p.n_aln)
    np.testing.assert_array_equal(res_b.pops, res_p.pops)
    np.testing.assert_array_equal(res_b.trace, res_p.trace)
    np.testing.assert_array_equal(res_b.rec_off, res_p.rec_off)
    nr = int(np.maximum(res_b.n_aln, 0).sum())
    np.testing.assert_array_equal(res_b.rec[:nr], res_p.rec[:nr])
    # oracle on a sample, with the same per-batch options (all reads are in batch 0)
    assert n <= 0x40000
    pick = np.sort(np.random.default_rng(seed).choice(n, size=min(n_oracle, n), replace=False))
    q_s, qo_s = synth.encode_queries(reads[pick], lens[pick])
    m = {"primary": int(meta.primary), "L2": [int(meta.L2[i]) for i in range(5)], "seq_len": int(meta.seq_len),
         "bwt_size": int(meta.bwt_size)}
    t = time.time()
    ref = orc.search(words, m, q_s, qo_s, lens[pick].astype(np.int32), md[pick], np.zeros(pick.size, np.int32),
                     bopts, nst, int(o.seed_len))
    print(f"oracle: {pick.size} reads in {time.time() - t:.1f} s", flush=True)
    for j, i in enumerate(pick.tolist()):
        assert res_b.n_aln[i] == ref.n_aln[j], f"read {i}"
        assert res_b.pops[i] == ref.pops[j], f"read {i}"
        assert res_b.trace[i] == ref.trace[j], f"read {i}"
        np.testing.assert_array_equal(res_b.records(i), ref.records(j), err_msg=f"read {i}")
    return res_b


def _pops_histogram(pops):
    h = {}
    for p in pops.astype(np.int64):
        k = "0" if p == 0 else f"{1 << (int(p).bit_length() - 1)}-{(1 << int(p).bit_length()) - 1}"
        h[k] = h.get(k, 0) + 1
    return h


@pytest.mark.skipif(not os.path.exists(LARGE), reason="large_index.json not written yet")
def test_deep_repeats_above_2_32_symbols(tmp_path):
    print("case genome deep repeats > 2^32", flush=True)
    with open(LARGE) as fh:
        exp = json.load(fh)
    t = time.time()
    contigs = fx.reference_contigs(fx.LARGE_SPECS["copies_4g"])
    ref = np.concatenate([c for _, c in contigs])
    del contigs
    print(f"generated {ref.size} bp in {time.time() - t:.1f} s", flush=True)
    assert ref.size == exp["ref_bp"] and 2 * ref.size > 2**32
    t = time.time()
    meta, words, sa, info = nv.build_bwt(ix.encode_forward(ref), 0, want_sa=True)
    print(f"built n={int(meta.seq_len)} in {time.time() - t:.1f} s: {info}", flush=True)
    L2 = [int(meta.L2[i]) for i in range(5)]
    assert int(meta.primary) == exp["primary"] and L2 == exp["L2"]
    prefix = str(tmp_path / "copies_4g")
    ix.write_bwt(prefix + ".bwt", int(meta.primary), L2, words)
    ix.write_sa(prefix + ".sa", int(meta.primary), L2, int(meta.seq_len), sa)
    assert fx.md5_file(prefix + ".bwt") == exp["bwt_md5"]
    assert fx.md5_file(prefix + ".sa") == exp["sa_md5"]
    os.remove(prefix + ".sa")
    m1, l1, _, _ = synth.paired_reads(93, ref, 20_000)
    res = _search_and_check(meta, words, m1, l1, fx.DEEP)
    nr = int(np.maximum(res.n_aln, 0).sum())
    assert int(res.rec[:nr, 2].max()) > 2**32  # intervals past 2^32 reached the .sai records
    print("pops histogram:", _pops_histogram(res.pops), flush=True)
    # positions of records' rows on the GPU equal the oracle's (positions past 2^32 included)
    rows = np.random.default_rng(4).choice(res.rec[:nr, 1], size=3000, replace=False).astype(np.uint64)
    with nv.Index(words, meta, 0) as index:
        index.set_sa(sa)
        pos, st = index.sa_positions(rows, 150)
        pf = index.sa_positions(rows)
    m = {"primary": int(meta.primary), "L2": L2, "seq_len": int(meta.seq_len), "bwt_size": int(meta.bwt_size)}
    np.testing.assert_array_equal(pf, orc.sa_positions(words, m, sa, rows))
    ep, es = orc.sa2pos(words, m, sa, rows, 150)
    np.testing.assert_array_equal(pos, ep)
    np.testing.assert_array_equal(st, es)
    assert int(pf.max()) > 2**32


def test_human_scale_index_and_search(tmp_path):
    print("case genome human scale 3.1 Gbp", flush=True)
    t = time.time()
    contigs = synth.human_like_reference(301, 3_100_000_000)
    ref = np.concatenate([c for _, c in contigs])
    del contigs
    print(f"generated {ref.size} bp in {time.time() - t:.1f} s", flush=True)
    t = time.time()
    meta, words, _, info = nv.build_bwt(ix.encode_forward(ref), 0)
    print(f"built n={int(meta.seq_len)} in {time.time() - t:.1f} s: {info}", flush=True)
    assert int(meta.seq_len) == 2 * ref.size and info["chunks"] >= 3
    reads, lens = synth.uniform_length_reads(302, ref, 50_000, 30, 150)
    res = _search_and_check(meta, words, reads, lens, ["-l", "1024", "-n", "0.01", "-o", "2"])
    nr = int(np.maximum(res.n_aln, 0).sum())
    assert int(res.rec[:nr, 2].max()) > 2**32
