"""The HIP .bwt builder (galn_hip_build_bwt) on the GPU.

* Every golden fixture's index is byte-identical to the reference builder's (the md5 in the
  fixture JSON comes from oracle/_ref/galn_index_ref, src/index/cpu_ref.c + libsais), also with
  the chunk capacity forced down so that the text is sorted in many chunks.
* Texts that stress the refinement (long runs of one base, short tandem repeats, lengths at the
  word and block boundaries) equal the CPU builder (prefix doubling, itself pinned to the
  reference builder by tests/test_index.py).
* The SA samples equal a suffix array sorted on the CPU.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import fixtures as fx
from galn_amd import index as ix
from galn_amd import native as nv

pytestmark = pytest.mark.gpu

# fixtures with distinct references
FIXTURE_REFS = ["plain_small", "plain_full", "adna_mixed", "human_like", "deep_repeats", "fastq_edge", "beyond_limits"]


def _seq(name):
    contigs = fx.reference_contigs(fx.SPECS[name])
    return b"".join(np.ascontiguousarray(c, dtype=np.uint8).tobytes() for _, c in contigs)


def _cpu_body(f2):
    primary, L = ix.bwt_from_forward(f2, device="cpu")
    L2, words = ix.pack_bwt(L)
    return primary, L2, words


@pytest.mark.parametrize("name", FIXTURE_REFS)
def test_builder_equals_reference_builder(name, tmp_path):
    print(f"case build {name}", flush=True)
    meta, _ = fx.load_expected(name)
    prefix = str(tmp_path / name)
    ix.build_index(_seq(name), prefix, device="cuda")
    assert fx.md5_file(prefix + ".bwt") == meta["bwt_md5"]


@pytest.mark.parametrize("name,cap", [("adna_mixed", 20_000), ("deep_repeats", 50_000), ("plain_small", 3_000)])
def test_builder_in_small_chunks(name, cap, tmp_path, monkeypatch):
    """Chunk boundaries fall inside BWT words and blocks; the bytes do not change."""
    print(f"case build chunks {name} cap={cap}", flush=True)
    monkeypatch.setenv("GALN_HIP_TEST_BUILD_CHUNK", str(cap))
    meta, _ = fx.load_expected(name)
    f2 = ix.encode_forward(_seq(name))
    m, words, _, info = nv.build_bwt(f2, 0)
    assert info["chunks"] > (2 * f2.size) // cap
    prefix = str(tmp_path / name)
    ix.write_bwt(prefix + ".bwt", int(m.primary), [int(m.L2[i]) for i in range(5)], words)
    assert fx.md5_file(prefix + ".bwt") == meta["bwt_md5"]


def _texts():
    rng = np.random.default_rng(5)
    out = {}
    for L in (1, 2, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 1000):
        out[f"random_{L}"] = rng.integers(0, 4, L).astype(np.uint8)
    out["all_A_5000"] = np.zeros(5000, np.uint8)
    out["AC_repeat_3001"] = np.tile(np.array([0, 1], np.uint8), 1501)[:3001]
    out["ACGTT_repeat_4000"] = np.tile(np.array([0, 1, 2, 3, 3], np.uint8), 800)
    # a random text with a 2 kb exact repeat and a palindrome (its reverse complement follows)
    base = rng.integers(0, 4, 20000).astype(np.uint8)
    base[12000:14000] = base[3000:5000]
    base[16000:16500] = 3 - base[15500:16000][::-1]
    out["repeats_20000"] = base
    return out


@pytest.mark.parametrize("name", list(_texts()))
def test_builder_equals_cpu_builder_on_hard_texts(name):
    print(f"case build text {name}", flush=True)
    f2 = _texts()[name]
    m, words, _, info = nv.build_bwt(f2, 0)
    primary, L2, cw = _cpu_body(f2)
    assert int(m.primary) == primary
    assert [int(m.L2[i]) for i in range(5)] == L2
    np.testing.assert_array_equal(words, cw)


@pytest.mark.parametrize("name", ["repeats_20000", "AC_repeat_3001", "random_1000", "random_65"])
def test_sa_samples_equal_cpu_suffix_array(name):
    print(f"case sa samples {name}", flush=True)
    f2 = _texts()[name]
    m, _, sa, _ = nv.build_bwt(f2, 0, want_sa=True)
    f = torch.as_tensor(f2, dtype=torch.uint8)
    text = torch.cat([f, (3 - f).flip(0)])
    ref = ix.suffix_array(text).numpy().astype(np.uint64)
    n = int(text.numel())
    k = (n + 32) // 32 - 1
    np.testing.assert_array_equal(sa, ref[32 * np.arange(1, k + 1) - 1])


def test_builder_index_searches_like_cpu_built_index(golden_dir):
    """An index built on the GPU serves the search: the golden run's .sai md5 on it."""
    from conftest import golden_inputs, load_reads, run_opts
    w = golden_inputs("adna_mixed", golden_dir)
    f2 = ix.encode_forward(_seq("adna_mixed"))
    m, words, _, _ = nv.build_bwt(f2, 0)
    opts, mate = run_opts(w["meta"], "adna_1")
    r = load_reads(w["reads"][mate])
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, r["qlen"], r["batch"])
    with nv.Index(words, m, 0) as index:
        rs = nv.ReadSet(index, r["query"], r["qoff"], r["qlen"], md, r["batch"], bopts, nst)
        rs.search(nv.FORM_BATCHED)
        res = rs.results()
        rs.close()
    assert fx.md5_bytes(nv.sai_bytes(o, res.n_aln, res.rec_off, res.rec)) == w["meta"]["runs"]["adna_1"]["sai_md5"]
