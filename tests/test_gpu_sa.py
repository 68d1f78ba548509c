"""The position step on the GPU (galn_hip_index_set_sa, galn_hip_sa_positions): bwt_sa and
bwa_sa2pos of BWT rows, from the SA samples the HIP builder writes.

* every row of small texts equals the suffix array sorted on the CPU (row 0: bwa's -1);
* on golden fixture indexes, rows of the search's records and random rows equal the oracle's
  bwt_sa / bwa_sa2pos (tests/test_oracle.py pins those against sorted suffixes and exact hits).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import fixtures as fx
import oracle as orc
from conftest import golden_inputs, load_reads, run_opts
from galn_amd import index as ix
from galn_amd import native as nv

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L", [1, 31, 32, 33, 1000, 5003])
def test_positions_of_every_row_equal_sorted_suffixes(L):
    print(f"case sa rows L={L}", flush=True)
    rng = np.random.default_rng(L)
    f2 = rng.integers(0, 4, L).astype(np.uint8)
    if L > 2000:
        f2[3000:4000] = f2[200:1200]
    m, words, sa, _ = nv.build_bwt(f2, 0, want_sa=True)
    f = torch.as_tensor(f2, dtype=torch.uint8)
    ref = ix.suffix_array(torch.cat([f, (3 - f).flip(0)])).numpy().astype(np.uint64)
    n = int(m.seq_len)
    with nv.Index(words, m, 0) as index:
        index.set_sa(sa)
        pos = index.sa_positions(np.arange(n + 1))
    np.testing.assert_array_equal(pos, np.concatenate([[np.uint64(2**64 - 1)], ref]))


@pytest.mark.parametrize("name,key", [("adna_mixed", "adna_1"), ("deep_repeats", "deep_1")])
def test_positions_equal_oracle_on_fixture(name, key, golden_dir):
    print(f"case sa fixture {name}", flush=True)
    w = golden_inputs(name, golden_dir)
    contigs = fx.reference_contigs(fx.SPECS[name])
    f2 = ix.encode_forward(np.concatenate([c for _, c in contigs]))
    m, words, sa, _ = nv.build_bwt(f2, 0, want_sa=True)
    meta = {"primary": int(m.primary), "L2": [int(m.L2[i]) for i in range(5)], "seq_len": int(m.seq_len),
            "bwt_size": int(m.bwt_size)}
    opts, mate = run_opts(w["meta"], key)
    r = load_reads(w["reads"][mate])
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, r["qlen"], r["batch"])
    with nv.Index(words, m, 0) as index:
        rs = nv.ReadSet(index, r["query"], r["qoff"], r["qlen"], md, r["batch"], bopts, nst)
        rs.search(nv.FORM_BATCHED)
        res = rs.results()
        rs.close()
        nr = int(np.maximum(res.n_aln, 0).sum())
        rng = np.random.default_rng(3)
        rows = np.concatenate([res.rec[:nr, 1], res.rec[:nr, 2], rng.integers(0, int(m.seq_len) + 1, 20000)])
        rows = rows.astype(np.uint64)
        index.set_sa(sa)
        pos = index.sa_positions(rows)
        rl = rng.integers(30, 151, rows.size).astype(np.int32)
        p2, st = index.sa_positions(rows, rl)
    np.testing.assert_array_equal(pos, orc.sa_positions(words, meta, sa, rows))
    ep, es = orc.sa2pos(words, meta, sa, rows, rl)
    np.testing.assert_array_equal(p2, ep)
    np.testing.assert_array_equal(st, es)


def test_sa_arguments_are_checked():
    f2 = np.random.default_rng(2).integers(0, 4, 500).astype(np.uint8)
    m, words, sa, _ = nv.build_bwt(f2, 0, want_sa=True)
    with nv.Index(words, m, 0) as index:
        with pytest.raises(nv.GalnError, match="no SA samples"):
            index.sa_positions([1, 2])
        with pytest.raises(nv.GalnError, match="samples"):
            index.set_sa(sa[:-1])
        index.set_sa(sa)
        with pytest.raises(nv.GalnError, match="beyond"):
            index.sa_positions([int(m.seq_len) + 1])
