"""The CPU search engine (galn_cpu_search) against the reference's golden runs, on the CPU.

Every committed golden run -- the reference's own C search over seven workloads and thirteen
option sets -- must come out of the engine with the golden .sai bytes, per-read pop counts
and trace hashes, whatever the thread count and the number of reads a thread interleaves.
The engine's algorithmic byte count per read must equal the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

import fixtures as fx
import oracle as orc
from conftest import golden_inputs, golden_runs, load_reads, run_opts
from galn_amd import index as ix
from galn_amd import native as nv


def engine_run(w, key, threads=0, interleave=8, reads_list=None):
    opts, mate = run_opts(w["meta"], key)
    r = load_reads(w["reads"][mate])
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, r["qlen"], r["batch"])
    meta, words = ix.read_bwt(w["bwt"])
    res = nv.cpu_search(words, meta, r["query"], r["qoff"], r["qlen"], md, r["batch"], bopts, nst,
                        reads_list=reads_list, threads=threads, interleave=interleave)
    return o, r, res


@pytest.mark.parametrize("name,key", golden_runs())
def test_cpu_engine_matches_reference_golden(name, key, golden_dir):
    w = golden_inputs(name, golden_dir)
    o, r, res = engine_run(w, key)
    exp = w["expected"]
    np.testing.assert_array_equal(res.n_aln, exp[key + "_n_aln"])
    np.testing.assert_array_equal(res.pops, exp[key + "_pops"])
    np.testing.assert_array_equal(res.trace, exp[key + "_trace"])
    assert fx.md5_bytes(nv.sai_bytes(o, res.n_aln, res.rec_off, res.rec)) == w["meta"]["runs"][key]["sai_md5"]


@pytest.mark.parametrize("threads,interleave", [(1, 1), (1, 16), (3, 5), (2, 0)])  # 0: chosen by index size
def test_cpu_engine_independent_of_threads_and_interleave(threads, interleave, golden_dir):
    w = golden_inputs("adna_mixed", golden_dir)
    o, r, res = engine_run(w, "adna_1", threads=threads, interleave=interleave)
    assert fx.md5_bytes(nv.sai_bytes(o, res.n_aln, res.rec_off, res.rec)) == w["meta"]["runs"]["adna_1"]["sai_md5"]
    np.testing.assert_array_equal(res.trace, w["expected"]["adna_1_trace"])


def test_cpu_engine_read_list_and_bytes(golden_dir):
    """A list of read ids gives those reads' results by list position; the per-read
    algorithmic bytes are the oracle's count."""
    w = golden_inputs("human_like", golden_dir)
    lst = np.arange(0, 6000, 7)[::-1].astype(np.int32)
    o, r, res = engine_run(w, "adna_1", reads_list=lst)
    exp = w["expected"]
    np.testing.assert_array_equal(res.n_aln, exp["adna_1_n_aln"][lst])
    np.testing.assert_array_equal(res.trace, exp["adna_1_trace"][lst])
    meta, words = ix.read_bwt(w["bwt"])
    oo = orc.parse_aln_opts(w["meta"]["runs"]["adna_1"]["opts"])
    bopts, nst, maxd = orc.batch_options(oo, r["qlen"], r["batch"])
    n = 600
    ref = orc.search(words, meta, r["query"], r["qoff"][:n], r["qlen"][:n], maxd[:n], r["batch"][:n], bopts, nst,
                     oo["seed_len"])
    o2, _, full = engine_run(w, "adna_1")
    np.testing.assert_array_equal(full.occ_bytes[:n], ref.occ_bytes)


@pytest.mark.parametrize("name,key", [("plain_small", "default_1"), ("adna_mixed", "adna_1"), ("deep_repeats", "deep_2"),
                                      ("options", "max_entries_m300_1"), ("options", "nonstop_n2_o0_1"),
                                      ("fastq_edge", "adna_1"), ("serving_mix", "default_1"),
                                      ("beyond_limits", "default_1"), ("beyond_limits", "e50_1"),
                                      ("beyond_limits", "n130_1")])
def test_cli_cpu_only_writes_golden_sai(name, key, golden_dir, tmp_path):
    """`galn aln --cpu-only` writes the reference's .sai with no GPU."""
    import subprocess
    w = golden_inputs(name, golden_dir)
    opts, mate = run_opts(w["meta"], key)
    out = tmp_path / "out.sai"
    r = subprocess.run([nv.GALN_BIN, "aln", "--cpu-only", "--cpu-workers", "4", *opts, "-f", str(out), w["prefix"],
                        w["reads"][mate]], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert fx.md5_file(str(out)) == w["meta"]["runs"][key]["sai_md5"]
