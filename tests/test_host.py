"""Host side of `galn aln` (libgaln_host.so, the `galn` executable) and the C ABI surface.
No GPU is used: option parsing, batch options, FASTQ input, .bwt loading, .sai bytes,
exit codes, and the exported symbols of both libraries."""
from __future__ import annotations

import ctypes as C
import gzip
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import fixtures as fx
import oracle as orc
from conftest import ROOT, golden_inputs
from galn_amd import index as ix
from galn_amd import native as nv


def header_functions(path: str) -> list[str]:
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(galn_[a-z_0-9]+)\s*\(", txt, flags=re.M)))


HEADERS = {"galn_hip.h": nv.HIP_LIB, "galn_host.h": nv.HOST_LIB, "galn_cpu.h": nv.HOST_LIB}


@pytest.mark.parametrize("header,lib", list(HEADERS.items()))
def test_libraries_load_and_export_every_declared_function(header, lib):
    names = header_functions(os.path.join(ROOT, "include", header))
    assert len(names) >= 1
    L = C.CDLL(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("lib,mirrored", [(nv.HIP_LIB, nv.HIP_SYMBOLS), (nv.HOST_LIB, nv.HOST_SYMBOLS)])
def test_ctypes_mirror_binds_every_declared_function(lib, mirrored):
    names = sorted(n for h, l in HEADERS.items() if l == lib
                   for n in header_functions(os.path.join(ROOT, "include", h)))
    assert sorted(mirrored) == names


def test_abi_version_and_limits():
    assert nv.hip_lib().galn_hip_abi_version() == 5
    lim = nv.limits()
    assert lim["max_read_len"] == 1023 and lim["max_n_stacks"] == 256 and lim["max_seq_len"] == (1 << 38) - 1


OPTION_SETS = [
    [], ["-n", "0.01"], ["-n", "3"], ["-o", "2"], ["-e", "6"], ["-e", "-1"], ["-i", "2"], ["-d", "3"], ["-l", "1024"],
    ["-k", "0"], ["-m", "500"], ["-dt", "16"], ["-t", "4"], ["-dt", "0"], ["-M", "5"], ["-O", "15"], ["-E", "2"],
    ["-R", "7"], ["-L"], ["-N"], ["-q", "0"], ["-l", "1024", "-n", "0.01", "-o", "2"],
    ["-n", "0.04", "-o", "2", "-e", "6", "-k", "2"], ["-N", "-L", "-e", "3", "-n", "4", "-R", "2"],
]


@pytest.mark.parametrize("opts", OPTION_SETS, ids=lambda o: " ".join(o) or "defaults")
def test_option_header_matches_reference_semantics(opts):
    """The 64-byte gap_opt_t written after "SAI\\1" (src/galn.cu:185-250, src/opt_build.h:23-36)."""
    # options before and after the positional arguments
    for args in (opts + ["p", "r.fq"], ["p"] + opts + ["r.fq"], ["p", "r.fq"] + opts):
        rc, a = nv.parse_aln_args(args)
        assert rc == 0
        assert b"SAI\x01" + nv.gap_opt_bytes(a.opt) == orc.sai_header(orc.parse_aln_opts(opts))
        assert a.prefix == b"p" and a.reads == b"r.fq"


def test_defaults_are_bwa_defaults():
    o = nv.gap_opt_from_args([])
    assert (o.s_mm, o.s_gapo, o.s_gape, o.max_diff, o.max_gapo, o.max_gape) == (3, 11, 4, -1, 1, 6)
    assert (o.indel_end_skip, o.max_del_occ, o.max_entries, o.mode, o.seed_len, o.max_seed_diff) == (5, 10, 2000000, 3, 32, 2)
    assert abs(o.fnr - 0.04) < 1e-7 and o.n_threads == 1 and o.max_top2 == 30 and o.trim_qual == 0


@pytest.mark.parametrize("args", [["p"], ["p", "r", "x"], ["-q", "5", "p", "r"], ["--bogus", "p", "r"], ["p", "r", "-n"],
                                  ["--cpu-workers", "p", "r"], ["--form", "fast", "p", "r"]])
def test_option_errors(args):
    rc, _ = nv.parse_aln_args(args)
    assert rc != 0


def test_engine_options_parse():
    rc, a = nv.parse_aln_args(["--cpu-only", "--verify", "--cpu-workers", "3", "--cpu-interleave", "5", "--gpu-only",
                               "p", "r"])
    assert rc == 0 and (a.cpu_only, a.verify, a.cpu_threads, a.cpu_interleave) == (1, 1, 3, 5)
    rc, a = nv.parse_aln_args(["p", "r"])
    assert rc == 0 and (a.cpu_only, a.verify, a.cpu_threads, a.cpu_interleave) == (0, 0, 0, 0)


def test_reference_tuning_flags_accepted():
    rc, a = nv.parse_aln_args(["--light-cap", "64", "--warps", "8", "--gpu-only", "--handoff", "5", "p", "r"])
    assert rc == 0 and a.prefix == b"p"


def test_cal_maxdiff_same_as_oracle():
    L = nv.host_lib()
    for fnr in (0.04, 0.01, 0.02, 0.001, 0.1):
        t = float(np.float32(fnr))
        for n in range(1, 1025, 7):
            assert L.galn_cal_maxdiff(n, 0.02, t) == orc.cal_maxdiff(n, 0.02, t)


@pytest.mark.parametrize("opts", [[], ["-l", "1024", "-n", "0.01", "-o", "2"], ["-n", "5", "-o", "9"], ["-e", "3"]])
def test_prepare_batches_same_as_oracle(opts):
    rng = np.random.default_rng(3)
    n = 3000
    lens = rng.integers(20, 250, size=n).astype(np.int32)
    batch = np.repeat(np.arange(4), [700, 900, 1, 1399]).astype(np.int32)
    o = nv.gap_opt_from_args(opts)
    bopts, nst, md = nv.prepare_batches(o, lens, batch)
    ob, onst, omd = orc.batch_options(orc.parse_aln_opts(opts), lens, batch)
    np.testing.assert_array_equal(nst, onst)
    np.testing.assert_array_equal(md, omd)
    for b in range(4):
        assert bytes(memoryview(bopts[b])) == bytes(memoryview(ob[b]))


def test_fastq_edge_rules(tmp_path):
    """Empty records are skipped, a malformed record ends the batch, CRLF / lower case /
    multi-line / FASTA records are read as bwa reads them; gzip gives the same."""
    meta, _ = fx.load_expected("fastq_edge")
    path = os.path.join(ROOT, "tests", "golden", "edge.fq")
    r = nv.read_fastq(path)
    assert r["n"] == meta["runs"]["adna_1"]["n_reads"] == 32
    assert r["n_batches"] == 2
    names = r["names"]
    assert "empty" not in names and "bad_quality" not in names
    assert names.index("b1_0") > names.index("b0_4")
    assert all(r["batch"][i] == (1 if names[i].startswith("b1_") else 0) for i in range(r["n"]))
    i = names.index("crlf")
    assert r["qlen"][i] == 70
    i = names.index("multiline")
    assert r["qlen"][i] == 90
    i = names.index("dash")
    q = r["query"][int(r["qoff"][i]):int(r["qoff"][i]) + int(r["qlen"][i])]
    assert r["qlen"][i] == 41 and int(q[20]) == 4  # '-' is not a base
    i = names.index("lower")
    assert r["query"][int(r["qoff"][i]):int(r["qoff"][i]) + 60].max() <= 3
    gz = tmp_path / "edge.fq.gz"
    with open(path, "rb") as fi, gzip.open(gz, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    r2 = nv.read_fastq(str(gz))
    for k in ("qlen", "qoff", "query", "batch"):
        np.testing.assert_array_equal(r[k], r2[k])


def test_query_encoding():
    """query = complement(nt4(reverse(read))), anything but ACGT -> 4 (src/host_io.c:178-182)."""
    from galn_amd import synth
    rng = np.random.default_rng(1)
    reads = np.frombuffer(b"ACGTNacgtRY", np.uint8)[rng.integers(0, 11, size=(50, 40))]
    lens = rng.integers(1, 41, size=50)
    q, qo = synth.encode_queries(reads, lens)
    comp = {ord("A"): 3, ord("C"): 2, ord("G"): 1, ord("T"): 0, ord("a"): 3, ord("c"): 2, ord("g"): 1, ord("t"): 0}
    for i in range(50):
        s = reads[i, :lens[i]][::-1]
        exp = [comp.get(int(c), 4) for c in s]
        assert list(q[int(qo[i]):int(qo[i]) + int(lens[i])]) == exp


def test_load_bwt_matches_file(golden_dir):
    w = golden_inputs("plain_small", golden_dir)
    m, words = nv.load_bwt(w["bwt"])
    meta, ref_words = ix.read_bwt(w["bwt"])
    assert m.primary == meta["primary"] and list(m.L2) == meta["L2"] and m.bwt_size == ref_words.size
    np.testing.assert_array_equal(words, ref_words)


def test_sai_writer_bytes(tmp_path):
    """galn_sai_write: header, then per read int32 n_aln and 24-byte records; errors -> 0 records."""
    o = nv.gap_opt_from_args(["-dt", "8"])
    n_aln = np.array([0, 2, -3, 1], np.int32)
    rec_off = np.array([0, 0, 2, 2], np.uint64)
    rec = np.arange(9, dtype=np.uint64).reshape(3, 3)
    p = tmp_path / "x.sai"
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen(str(p).encode(), b"wb")
    L = nv.host_lib()
    L.galn_sai_write.argtypes = [C.c_void_p, C.POINTER(nv.GapOpt), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
    rc = L.galn_sai_write(fp, C.byref(o), 4, n_aln.ctypes.data, rec_off.ctypes.data, rec.ctypes.data)
    libc.fclose(fp)
    assert rc == 0
    expect = orc.sai_header(orc.parse_aln_opts(["-dt", "8"]))
    expect += (0).to_bytes(4, "little") + (2).to_bytes(4, "little") + rec[:2].tobytes()
    expect += (0).to_bytes(4, "little") + (1).to_bytes(4, "little") + rec[2:3].tobytes()
    assert p.read_bytes() == expect
    assert nv.sai_bytes(o, n_aln, rec_off, rec) == expect


def _galn(*args, timeout=60):
    return subprocess.run([nv.GALN_BIN, *args], capture_output=True, text=True, timeout=timeout)


def test_cli_version_and_usage():
    r = _galn("--version")
    assert r.returncode == 0 and r.stdout.strip() == "G-ALN 0.2.7"
    assert _galn().returncode == 1
    assert _galn("aln").returncode == 1
    assert _galn("index").returncode == 1


def test_cli_exit_codes(tmp_path, golden_dir):
    w = golden_inputs("plain_small", golden_dir)
    assert _galn("aln", "--cpu-workers", w["prefix"], w["reads"][0]).returncode == 1
    assert _galn("aln", str(tmp_path / "nope"), w["reads"][0]).returncode == 2
    assert _galn("aln", w["prefix"], str(tmp_path / "nope.fq")).returncode == 3
    assert _galn("aln", "-f", "/nonexistent/dir/x.sai", w["prefix"], w["reads"][0]).returncode == 1


def test_cli_without_gpu_fails_loudly(golden_dir):
    """No silent CPU fallback: with no HIP device the GPU search is refused (exit 10); only
    --cpu-only searches on the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    w = golden_inputs("plain_small", golden_dir)
    r = _galn("aln", w["prefix"], w["reads"][0])
    assert r.returncode == 10 and "GPU" in r.stderr


def test_library_without_gpu_refuses(golden_dir):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    w = golden_inputs("plain_small", golden_dir)
    meta, words = ix.read_bwt(w["bwt"])
    with pytest.raises(nv.GalnError):
        nv.Index(words, meta, 0)
