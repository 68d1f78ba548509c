"""Parity of the HIP search with the reference, on the GPU (pytest -m gpu).

Every golden run (the reference's own C search, committed in tests/golden/) is searched
through the C ABI in each form -- plain, batched, batched with every read sent to the
wave-per-read engine (also at 5 waves per SIMD), batched without the exact-child register path, batched with one
record slot per lane (forces the re-run path) -- and must give the golden .sai bytes, pop
counts and trace hashes. The `galn` executable must write the golden .sai. Larger inputs
are checked by size-independent properties: the forms agree bit for bit, and a sample
agrees with the oracle."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

import fixtures as fx
import oracle as orc
from conftest import golden_inputs, gpu_golden_runs, load_reads, run_opts
from galn_amd import index as ix
from galn_amd import native as nv

pytestmark = pytest.mark.gpu


def tune(**kw) -> nv.Tune:
    t = nv.Tune()
    t.sort_by_cost = 1
    for k, v in kw.items():
        setattr(t, k, v)
    return t


FORMS = {
    "plain": (nv.FORM_PLAIN, 0, None),
    "batched": (nv.FORM_BATCHED, 0, None),
    "heavy_all": (nv.FORM_BATCHED, 0, dict(heavy_pops=1)),
    "heavy_64": (nv.FORM_BATCHED, 0, dict(heavy_pops=64)),
    "heavy_all_5_waves": (nv.FORM_BATCHED, 0, dict(heavy_pops=1, heavy_waves_per_simd=5)),
    "no_fuse": (nv.FORM_BATCHED, nv.GALN_HIP_NOFUSE, dict(heavy_pops=-1)),
    "rec1": (nv.FORM_BATCHED, 0, dict(rec_per_lane=1, heavy_pops=-1)),
    "unsorted_small_grid": (nv.FORM_BATCHED, 0, dict(sort_by_cost=0, waves_per_cu=1)),
}

_index_cache: dict = {}


def readset_for(name, key, golden_dir):
    w = golden_inputs(name, golden_dir)
    opts, mate = run_opts(w["meta"], key)
    r = load_reads(w["reads"][mate])
    if name not in _index_cache:
        meta, words = ix.read_bwt(w["bwt"])
        _index_cache[name] = nv.Index(words, meta, 0)
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, r["qlen"], r["batch"])
    rs = nv.ReadSet(_index_cache[name], r["query"], r["qoff"], r["qlen"], md, r["batch"], bopts, nst)
    return w, o, rs


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("name,key", gpu_golden_runs())
def test_golden_sai_pops_trace(name, key, form, golden_dir):
    print(f"case {name}/{key} form={form}", flush=True)
    w, o, rs = readset_for(name, key, golden_dir)
    f, flags, tk = FORMS[form]
    rs.search(f, flags | nv.GALN_HIP_TRACE, tune(**tk) if tk is not None else None)
    res = rs.results(want_pops=True, want_trace=True)
    exp = w["expected"]
    np.testing.assert_array_equal(res.n_aln, exp[key + "_n_aln"])
    np.testing.assert_array_equal(res.pops, exp[key + "_pops"])
    np.testing.assert_array_equal(res.trace, exp[key + "_trace"])
    assert fx.md5_bytes(nv.sai_bytes(o, res.n_aln, res.rec_off, res.rec)) == w["meta"]["runs"][key]["sai_md5"]


@pytest.mark.parametrize("name,key,must_cut", [("deep_repeats", "deep_1", True), ("adna_mixed", "adna_1", True),
                                               ("options", "max_entries_m300_1", False)])
def test_push_cap_cuts_rounds_without_changing_results(name, key, must_cut, golden_dir, monkeypatch):
    """With the wave engine's per-round push cap lowered to 128 (every read forced through the
    engine), rounds end at the cap and the deferred chains run in the next round: records, pops
    and traces stay the reference's."""
    print(f"case push cap {name}/{key}", flush=True)
    w, o, rs = readset_for(name, key, golden_dir)
    t = tune(heavy_pops=1)
    rs.search(nv.FORM_BATCHED, nv.GALN_HIP_TRACE | nv.GALN_HIP_STATS, t)
    base_rounds = rs.engine_stats()["rounds"]
    monkeypatch.setenv("GALN_HIP_TEST_PUSH_CAP", "128")
    rs.search(nv.FORM_BATCHED, nv.GALN_HIP_TRACE | nv.GALN_HIP_STATS, t)
    rounds = rs.engine_stats()["rounds"]
    res = rs.results(want_pops=True, want_trace=True)
    exp = w["expected"]
    np.testing.assert_array_equal(res.n_aln, exp[key + "_n_aln"])
    np.testing.assert_array_equal(res.pops, exp[key + "_pops"])
    np.testing.assert_array_equal(res.trace, exp[key + "_trace"])
    assert fx.md5_bytes(nv.sai_bytes(o, res.n_aln, res.rec_off, res.rec)) == w["meta"]["runs"][key]["sai_md5"]
    assert rounds > base_rounds if must_cut else rounds >= base_rounds


@pytest.mark.parametrize("form", ["plain", "batched"])
@pytest.mark.parametrize("name,key", [("plain_small", "default_1"), ("adna_mixed", "adna_1"), ("deep_repeats", "deep_2"),
                                      ("options", "max_entries_m300_1"), ("fastq_edge", "adna_1"),
                                      ("serving_mix", "default_1"), ("beyond_limits", "default_1"),
                                      ("beyond_limits", "default_2"), ("beyond_limits", "e50_1"),
                                      ("beyond_limits", "n130_1")])
def test_cli_writes_golden_sai(name, key, form, golden_dir, tmp_path):
    """`galn aln` on the GPU writes the reference's .sai; reads outside the kernels' limits (long
    reads, more than 256 score buckets, max_diff above 126) are finished by the CPU engine."""
    print(f"case cli {name}/{key} form={form}", flush=True)
    w = golden_inputs(name, golden_dir)
    opts, mate = run_opts(w["meta"], key)
    out = tmp_path / "out.sai"
    r = subprocess.run([nv.GALN_BIN, "aln", "--form", form, *opts, "-f", str(out), w["prefix"], w["reads"][mate]],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert fx.md5_file(str(out)) == w["meta"]["runs"][key]["sai_md5"]


@pytest.mark.parametrize("key", ["adna_1", "default_1"])
def test_cli_streams_blocks_of_one_batch(key, golden_dir, tmp_path):
    """The edge FASTQ has two bwa batches (a malformed record ends the first): streamed one batch
    per block, the .sai is still the reference's."""
    print(f"case cli stream fastq_edge/{key}", flush=True)
    w = golden_inputs("fastq_edge", golden_dir)
    opts, mate = run_opts(w["meta"], key)
    out = tmp_path / "out.sai"
    r = subprocess.run([nv.GALN_BIN, "aln", "--block-batches", "1", *opts, "-f", str(out), w["prefix"], w["reads"][mate]],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "(2 blocks)" in r.stderr
    assert fx.md5_file(str(out)) == w["meta"]["runs"][key]["sai_md5"]


@pytest.mark.parametrize("name,key", [("adna_mixed", "adna_1"), ("plain_small", "default_1")])
def test_cli_verify_reruns_reads_on_cpu_engine(name, key, golden_dir, tmp_path):
    """`galn aln --verify` re-runs every GPU read on the CPU engine: records, pops and pop traces
    agree, and the .sai is still the reference's."""
    print(f"case cli verify {name}/{key}", flush=True)
    w = golden_inputs(name, golden_dir)
    opts, mate = run_opts(w["meta"], key)
    out = tmp_path / "out.sai"
    r = subprocess.run([nv.GALN_BIN, "aln", "--verify", *opts, "-f", str(out), w["prefix"], w["reads"][mate]],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "differ -> PASS" in r.stderr, r.stderr
    assert fx.md5_file(str(out)) == w["meta"]["runs"][key]["sai_md5"]


def test_cli_streamed_equals_whole_input_on_many_batches(golden_dir, tmp_path):
    """300k reads (two bwa batches, gzip input): streamed in blocks of one batch, of four, and
    read whole first, the .sai bytes are the same."""
    import gzip
    print("case cli stream 300k reads", flush=True)
    w = golden_inputs("adna_mixed", golden_dir)
    src = open(w["reads"][0], "rb").read()
    fq = tmp_path / "many.fq.gz"
    with gzip.open(fq, "wb", compresslevel=1) as fh:
        for _ in range(38):  # 38 x 8000 reads
            fh.write(src)
    outs = {}
    for bb in ("1", "4", "0"):
        out = tmp_path / f"out{bb}.sai"
        r = subprocess.run([nv.GALN_BIN, "aln", "--block-batches", bb, "-l", "1024", "-n", "0.01", "-o", "2", "-f",
                            str(out), w["prefix"], str(fq)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        outs[bb] = fx.md5_file(str(out))
    assert outs["1"] == outs["4"] == outs["0"]


def test_algorithmic_bytes_match_oracle(golden_dir):
    """The kernel's per-read algorithmic byte count is the oracle's (same definition)."""
    w, o, rs = readset_for("adna_mixed", "adna_1", golden_dir)
    rs.search(nv.FORM_BATCHED)
    res = rs.results(want_bytes=True)
    r = load_reads(w["reads"][0])
    meta, words = ix.read_bwt(w["bwt"])
    oo = orc.parse_aln_opts(w["meta"]["runs"]["adna_1"]["opts"])
    bopts, nst, maxd = orc.batch_options(oo, r["qlen"], r["batch"])
    ref = orc.search(words, meta, r["query"], r["qoff"], r["qlen"], maxd, r["batch"], bopts, nst, oo["seed_len"])
    np.testing.assert_array_equal(res.occ_bytes, ref.occ_bytes)
    launches = rs.launches()
    assert sum(l["algo_bytes"] for l in launches) == int(ref.occ_bytes.sum())


def test_autotuned_search_gives_the_golden_sai(golden_dir):
    """The autotuner only changes how the batched form runs, never its records."""
    print("case autotune adna_mixed/adna_1", flush=True)
    w, o, rs = readset_for("adna_mixed", "adna_1", golden_dir)
    tuned = rs.autotune()
    assert tuned["heavy_pops"] in (1, 16, 64, 128, 256, 512, 1024, 4096) and tuned["block_threads"] in (64, 128)
    rs.search(nv.FORM_BATCHED)
    res = rs.results()
    assert fx.md5_bytes(nv.sai_bytes(o, res.n_aln, res.rec_off, res.rec)) == w["meta"]["runs"]["adna_1"]["sai_md5"]


@pytest.mark.parametrize("name,key", [("adna_mixed", "adna_1"), ("serving_mix", "default_1")])
def test_graph_replays_equal_eager_and_golden(name, key, golden_dir):
    """The batched form run as a captured graph: every replay, with eager searches between them,
    gives the eager result and the reference's .sai, pops and traces."""
    print(f"case graph {name}/{key}", flush=True)
    w, o, rs = readset_for(name, key, golden_dir)
    exp = w["expected"]
    rs.search(nv.FORM_BATCHED, nv.GALN_HIP_TRACE)
    eager = rs.results(want_pops=True, want_trace=True)
    for rep in range(3):
        rs.search(nv.FORM_BATCHED, nv.GALN_HIP_TRACE | nv.GALN_HIP_GRAPH)
        g = rs.results(want_pops=True, want_trace=True)
        np.testing.assert_array_equal(g.n_aln, eager.n_aln)
        np.testing.assert_array_equal(g.rec, eager.rec)
        np.testing.assert_array_equal(g.pops, exp[key + "_pops"])
        np.testing.assert_array_equal(g.trace, exp[key + "_trace"])
        assert fx.md5_bytes(nv.sai_bytes(o, g.n_aln, g.rec_off, g.rec)) == w["meta"]["runs"][key]["sai_md5"]
        assert [l["tier"] for l in rs.launches()][:2] == [0, 10]
        rs.search(nv.FORM_PLAIN)  # eager work between replays
        rs.results()


def test_results_into_reused_buffers(golden_dir):
    """results(out=...) fills caller-owned arrays (reused, grown when the records outnumber
    them) with what a fresh call returns."""
    print("case results into reused buffers adna_mixed/adna_1", flush=True)
    w, o, rs = readset_for("adna_mixed", "adna_1", golden_dir)
    rs.search(nv.FORM_BATCHED)
    fresh = rs.results(want_pops=True, want_bytes=True)
    bufs = rs.result_buffers(want_pops=True, want_bytes=True, rec_cap=8)  # too small: grows
    for rep in range(2):
        rs.search(nv.FORM_BATCHED)
        got = rs.results(want_pops=True, want_bytes=True, out=bufs)
        np.testing.assert_array_equal(got.n_aln, fresh.n_aln)
        np.testing.assert_array_equal(got.rec_off, fresh.rec_off)
        np.testing.assert_array_equal(got.rec, fresh.rec)
        np.testing.assert_array_equal(got.pops, fresh.pops)
        np.testing.assert_array_equal(got.occ_bytes, fresh.occ_bytes)
        assert bufs["rec"].shape[0] >= fresh.rec.shape[0]
    assert fx.md5_bytes(nv.sai_bytes(o, got.n_aln, got.rec_off, got.rec)) == w["meta"]["runs"]["adna_1"]["sai_md5"]
    with pytest.raises(nv.GalnError):
        rs.results(want_trace=True, out=bufs)  # no trace array in these buffers


def test_repeated_searches_identical(golden_dir):
    w, o, rs = readset_for("human_like", "adna_1", golden_dir)
    outs = []
    for form in (nv.FORM_BATCHED, nv.FORM_PLAIN, nv.FORM_BATCHED):
        rs.search(form)
        res = rs.results()
        outs.append(nv.sai_bytes(o, res.n_aln, res.rec_off, res.rec))
    assert outs[0] == outs[1] == outs[2]


def test_reads_outside_limits_are_reported():
    """A read longer than the kernel's limit gets the unsupported code, not a wrong answer."""
    from galn_amd import synth
    ref = synth.plain_reference(5, 20000)
    meta = ix.build_index(ref.tobytes(), "/tmp/galn_lim")
    meta, words = ix.read_bwt("/tmp/galn_lim.bwt")
    reads = np.zeros((3, 1100), np.uint8)
    lens = np.array([50, 1100, 80])
    for i, L in enumerate(lens):
        reads[i, :L] = ref[100 * i:100 * i + L]
    q, qo = synth.encode_queries(reads, lens)
    o = nv.gap_opt_from_args([])
    batch = np.zeros(3, np.int32)
    bopts, nst, md = nv.prepare_batches(o, lens.astype(np.int32), batch)
    with nv.Index(words, meta, 0) as idx:
        rs = nv.ReadSet(idx, q, qo, lens.astype(np.int32), md, batch, bopts, nst)
        for form in (nv.FORM_PLAIN, nv.FORM_BATCHED):
            rs.search(form)
            res = rs.results()
            assert res.n_aln[1] == -5
            assert res.n_aln[0] >= 1 and res.n_aln[2] >= 1


def test_larger_reference_forms_agree_and_sample_matches_oracle(tmp_path):
    """100 Mbp aDNA-like reference (index built on the GPU), 100k reads: plain and batched
    agree bit for bit, and 1500 reads agree with the oracle."""
    from galn_amd import workloads
    print("case large: building 100 Mbp workload", flush=True)
    w = workloads.adna_mixed(str(tmp_path), ref_bp=100_000_000, n_reads=100_000, device="cuda", workers=8)
    with nv.Index(w.words, w.meta, 0) as idx:
        rs = w.readset(idx)
        rs.search(nv.FORM_BATCHED)
        rb = rs.results()
        rs.search(nv.FORM_PLAIN)
        rp = rs.results()
    np.testing.assert_array_equal(rb.n_aln, rp.n_aln)
    np.testing.assert_array_equal(rb.rec, rp.rec)
    o = orc.parse_aln_opts(w.opts)
    n = 1500
    bid = orc.batch_ids(n)
    bopts, nst, maxd = orc.batch_options(o, w.lens[:n].astype(np.int32), bid)
    ref = orc.search(w.words, w.meta, w.query, w.qoff[:n], w.lens[:n].astype(np.int32), maxd, bid, bopts, nst,
                     o["seed_len"], 16)
    np.testing.assert_array_equal(rb.n_aln[:n], ref.n_aln)
    for i in range(n):
        np.testing.assert_array_equal(rb.records(i), ref.records(i))


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_exhausted_pool_reruns_reads_with_fewer_lanes(form, golden_dir, monkeypatch):
    """With the stack pool capped at 2048 granules (GALN_HIP_TEST_POOL_GRANULES) the first launch
    runs out of granules: the reads that found the pool empty are re-run with a quarter of the
    lanes (and of the wave engine's waves) until none is left, and records, pops and traces are
    still the reference's."""
    print(f"case pool cap adna_mixed/adna_1 form={form}", flush=True)
    monkeypatch.setenv("GALN_HIP_TEST_POOL_GRANULES", "2048")
    w = golden_inputs("adna_mixed", golden_dir)
    opts, mate = run_opts(w["meta"], "adna_1")
    r = load_reads(w["reads"][mate])
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, r["qlen"], r["batch"])
    meta, words = ix.read_bwt(w["bwt"])
    with nv.Index(words, meta, 0) as idx:  # a fresh index: its pool starts at the cap
        rs = nv.ReadSet(idx, r["query"], r["qoff"], r["qlen"], md, r["batch"], bopts, nst)
        rs.search(FORMS[form][0], nv.GALN_HIP_TRACE)
        res = rs.results(want_pops=True, want_trace=True)
        tiers = [l["tier"] for l in rs.launches()]
        rs.close()
    exp = w["expected"]
    np.testing.assert_array_equal(res.n_aln, exp["adna_1_n_aln"])
    np.testing.assert_array_equal(res.pops, exp["adna_1_pops"])
    np.testing.assert_array_equal(res.trace, exp["adna_1_trace"])
    assert fx.md5_bytes(nv.sai_bytes(o, res.n_aln, res.rec_off, res.rec)) == w["meta"]["runs"]["adna_1"]["sai_md5"]
    assert any(t not in (0, 10) for t in tiers), f"no re-run launch: {tiers}"


def test_pool_grows_on_demand_and_results_stay_golden(golden_dir, monkeypatch):
    """A context whose first pool holds only 1024 granules (GALN_HIP_TEST_POOL_START) runs out in
    the first search: the pool doubles before the reads that found it empty are re-run, and again
    after each search that used up its never-used granules, so later searches need fewer re-runs.
    Every search gives the reference's records, pops and traces."""
    print("case pool growth adna_mixed/adna_1", flush=True)
    monkeypatch.setenv("GALN_HIP_TEST_POOL_START", "1024")
    w = golden_inputs("adna_mixed", golden_dir)
    opts, mate = run_opts(w["meta"], "adna_1")
    r = load_reads(w["reads"][mate])
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, r["qlen"], r["batch"])
    meta, words = ix.read_bwt(w["bwt"])
    exp = w["expected"]
    launches = []
    with nv.Index(words, meta, 0) as idx:  # a fresh index: a fresh pool
        rs = nv.ReadSet(idx, r["query"], r["qoff"], r["qlen"], md, r["batch"], bopts, nst)
        for _ in range(3):
            rs.search(nv.FORM_BATCHED, nv.GALN_HIP_TRACE, tune(heavy_pops=1))
            res = rs.results(want_pops=True, want_trace=True)
            launches.append(len(rs.launches()))
            np.testing.assert_array_equal(res.n_aln, exp["adna_1_n_aln"])
            np.testing.assert_array_equal(res.pops, exp["adna_1_pops"])
            np.testing.assert_array_equal(res.trace, exp["adna_1_trace"])
        rs.close()
    assert launches[0] > 2, f"the first search did not run out of granules: {launches}"
    assert launches[-1] <= launches[0], launches


def test_index_close_is_refused_while_a_read_set_is_open(golden_dir):
    """galn_hip_index_close returns GALN_HIP_E_BUSY and keeps the index while one of its read
    sets is open (freeing the read set afterwards used its index's stream), and succeeds once
    the read set is freed."""
    print("case index close order", flush=True)
    w = golden_inputs("plain_small", golden_dir)
    opts, mate = run_opts(w["meta"], "default_1")
    r = load_reads(w["reads"][mate])
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, r["qlen"], r["batch"])
    meta, words = ix.read_bwt(w["bwt"])
    idx = nv.Index(words, meta, 0)
    rs = nv.ReadSet(idx, r["query"], r["qoff"], r["qlen"], md, r["batch"], bopts, nst)
    L = nv.hip_lib()
    assert L.galn_hip_index_close(idx.h) == -6
    rs.search(nv.FORM_BATCHED)  # the index is still usable
    res = rs.results()
    assert fx.md5_bytes(nv.sai_bytes(o, res.n_aln, res.rec_off, res.rec)) == w["meta"]["runs"]["default_1"]["sai_md5"]
    rs.close()
    assert L.galn_hip_index_close(idx.h) == 0
    idx.h = None
