#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own C code.

Run in a container where the reference tree and oracle/_ref/ are present
(`make -C oracle`): for every spec in fixtures.SPECS it regenerates the inputs,
builds the index with both this repository's builder and the reference's CPU
builder (oracle/_ref/galn_index_ref: src/index/cpu_ref.c + libsais) and checks
they agree, then runs the reference's CPU search (oracle/_ref/galn_ref:
src/aln_host.c, src/bwtgap_hd.h, src/host_io.c) for each option list and stores
the .sai md5 and the per-read n_aln / pops / trace hash.

    python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import fixtures as fx  # noqa: E402

REF_BIN = os.path.join(fx.ROOT, "oracle", "_ref", "galn_ref")
REF_INDEX = os.path.join(fx.ROOT, "oracle", "_ref", "galn_index_ref")


def write_edge_fastq(path: str) -> None:
    """Hand-made FASTQ exercising the record rules (kseq_read / bwa_read_seq): lower case, Ns,
    an empty record, CRLF line endings, a multi-line sequence, a FASTA record, a record whose
    quality is longer than its sequence (ends the batch), and reads after it."""
    from galn_amd import synth
    ref = synth.plain_reference(71, 200_000)
    rng = np.random.default_rng(72)

    def sub(L):
        p = int(rng.integers(0, ref.size - L))
        s = bytearray(ref[p:p + L].tobytes())
        for _ in range(int(rng.integers(0, 3))):
            j = int(rng.integers(0, L))
            s[j] = b"ACGT"[(b"ACGT".index(s[j]) + 1) % 4]
        return bytes(s)

    out = []
    for i in range(12):
        s = sub(int(rng.integers(30, 101)))
        out.append(b"@r%d extra comment\n%s\n+\n%s\n" % (i, s, b"I" * len(s)))
    s = sub(60).lower()
    out.append(b"@lower\n%s\n+\n%s\n" % (s, b"#" * len(s)))
    s = bytearray(sub(80))
    s[10] = ord("N")
    s[40] = ord("N")
    out.append(b"@two_n\n%s\n+\n%s\n" % (bytes(s), b"I" * len(s)))
    s = b"N" * 50
    out.append(b"@all_n\n%s\n+\n%s\n" % (s, b"I" * 50))
    out.append(b"@empty\n\n+\n\n")
    s = sub(70)
    out.append(b"@crlf\r\n%s\r\n+\r\n%s\r\n" % (s, b"I" * len(s)))
    s = sub(90)
    out.append(b"@multiline\n%s\n%s\n+\n%s\n" % (s[:45], s[45:], b"I" * len(s)))
    s = sub(55)
    out.append(b">fasta_record\n%s\n" % s)
    s = sub(40)
    out.append(b"@dash\n%s-%s\n+\n%s\n" % (s[:20], s[20:], b"I" * (len(s) + 1)))
    for i in range(5):
        s = sub(int(rng.integers(30, 150)))
        out.append(b"@b0_%d\n%s\n+\n%s\n" % (i, s, b"I" * len(s)))
    s = sub(50)
    out.append(b"@bad_quality\n%s\n+\n%s\n" % (s, b"I" * (len(s) + 7)))
    for i in range(8):
        s = sub(int(rng.integers(30, 150)))
        out.append(b"@b1_%d\n%s\n+\n%s\n" % (i, s, b"I" * len(s)))
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


def run_ref(prefix: str, reads: str, opts: list[str], sai: str, stats: str) -> None:
    r = subprocess.run([REF_BIN, *opts, "--stats", stats, "-f", sai, prefix, reads], capture_output=True, text=True)
    if r.returncode not in (0, 4):
        raise RuntimeError(f"reference search failed ({r.returncode}): {r.stderr}")


def read_stats(path: str):
    n_aln, pops, trace = [], [], []
    with open(path) as fh:
        next(fh)
        for line in fh:
            f = line.split("\t")
            n_aln.append(int(f[1]))
            pops.append(int(f[2]))
            trace.append(int(f[3], 16))
    return np.array(n_aln, np.int32), np.array(pops, np.uint64), np.array(trace, np.uint64)


def make(name: str) -> None:
    edge = os.path.join(HERE, "edge.fq")
    if name == "fastq_edge" and not os.path.exists(edge):
        write_edge_fastq(edge)
    with tempfile.TemporaryDirectory() as tmp:
        w = fx.build(name, tmp)
        ref_prefix = os.path.join(tmp, "refbuilt")
        r = subprocess.run([REF_INDEX, "-f", w["fa"], ref_prefix], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        bwt_md5 = fx.md5_file(w["bwt"])
        if fx.md5_file(ref_prefix + ".bwt") != bwt_md5:
            raise RuntimeError(f"{name}: this repository's .bwt differs from the reference builder's")
        meta = {"name": name, "spec": w["spec"], "bwt_md5": bwt_md5,
                "reads_md5": [fx.md5_file(p) for p in w["reads"]], "runs": {}}
        arrs = {}
        for oname, v in w["spec"]["opts"].items():
            opts, files = fx.opts_and_files(v)
            for mi, rp in enumerate(w["reads"]):
                if files is not None and mi not in files:
                    continue
                key = f"{oname}_{mi + 1}"
                sai = os.path.join(tmp, key + ".sai")
                st = os.path.join(tmp, key + ".tsv")
                run_ref(w["prefix"], rp, opts, sai, st)
                n_aln, pops, trace = read_stats(st)
                meta["runs"][key] = {"opts": opts, "reads": mi, "sai_md5": fx.md5_file(sai),
                                     "n_reads": int(n_aln.size), "records": int(np.maximum(n_aln, 0).sum())}
                arrs[key + "_n_aln"] = n_aln
                arrs[key + "_pops"] = pops
                arrs[key + "_trace"] = trace
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(meta, fh, indent=1, sort_keys=True)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrs)
        print(name, {k: (v["n_reads"], v["records"], v["sai_md5"][:8]) for k, v in meta["runs"].items()})


if __name__ == "__main__":
    if not (os.path.exists(REF_BIN) and os.path.exists(REF_INDEX)):
        sys.exit("oracle/_ref is not built (make -C oracle in a container with the reference tree)")
    for n in (sys.argv[1:] or list(fx.SPECS)):
        make(n)
