"""Golden fixture specs: seeded workloads whose reference outputs are committed here.

Each spec names a reference generator, a read generator and one or more `aln`
option lists. ``build(spec, workdir)`` regenerates the inputs (FASTA, .bwt,
FASTQ) deterministically; the committed JSON holds the md5 of every generated
input (so generator drift is caught) and of the reference's .sai for each option
list, and the .npz the reference's per-read n_aln / pop count / trace hash.

The expected values were written by make_golden.py from the reference's own C
search (G-ALN 0.2.7 src/aln_host.c + src/bwtgap_hd.h, built by oracle/Makefile)
and the reference's own CPU index builder (src/index/cpu_ref.c with libsais).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from galn_amd import index as ix  # noqa: E402
from galn_amd import synth  # noqa: E402

ADNA = ["-l", "1024", "-n", "0.01", "-o", "2"]
DEEP = ["-n", "0.04", "-o", "2", "-e", "6", "-k", "2"]

SPECS = {
    "plain_small": {
        "ref": ("plain", 11, 500_000), "reads": ("plain", 12, 3000, 100),
        "opts": {"default": []},
    },
    # the plain small configuration at full size: 5 Mbp i.i.d. reference, 100k x 100 bp reads
    "plain_full": {
        "ref": ("plain", 101, 5_000_000), "reads": ("plain", 102, 100_000, 100),
        "opts": {"default": []},
    },
    "adna_mixed": {
        "ref": ("adna", 21, 2_000_000), "reads": ("adna", 22, 8000),
        "opts": {"adna": ADNA},
    },
    "human_like": {
        "ref": ("human", 31, 2_000_000), "reads": ("uniform", 32, 6000, 30, 150),
        "opts": {"adna": ADNA},
    },
    "serving_mix": {
        "ref": ("human", 31, 2_000_000), "reads": ("uniform", 42, 4000, 30, 250),
        "opts": {"default": [], "adna": ADNA},
    },
    "deep_repeats": {
        "ref": ("copies", 51, 250_000, 8), "reads": ("paired", 52, 1500),
        "opts": {"deep": DEEP},
    },
    "options": {
        "ref": ("plain", 11, 500_000), "reads": ("adna", 62, 1500),
        "opts": {
            "nonstop_n2_o0": ["-N", "-n", "2", "-o", "0"],
            "loggap_e3": ["-L", "-e", "3"],
            "top2_R2": ["-R", "2"],
            "max_entries_m300": ["-m", "300"],
            "penalties": ["-M", "5", "-O", "15", "-E", "2"],
            "n3_k0_l20": ["-n", "3", "-k", "0", "-l", "20"],
            "i2_d2_dt16": ["-i", "2", "-d", "2", "-dt", "16"],
            "exact_n0": ["-n", "0"],
        },
    },
    "fastq_edge": {
        "ref": ("plain", 71, 200_000), "reads": ("file", "edge.fq"),
        "opts": {"adna": ADNA, "default": []},
    },
    # beyond the GPU kernels' limits: more than 256 score buckets (-e 50), max_diff above 126
    # (-n 130), reads of 1024-2000 bp (the second read file); an option list given with read
    # file numbers runs on those files only
    "beyond_limits": {
        "ref": ("plain", 81, 500_000), "reads": ("two_uniform", 82, 300, 30, 150, 24, 1024, 2000),
        "opts": {"default": [], "e50": (["-o", "2", "-e", "50"], [0]), "n130": (["-n", "130"], [0])},
    },
}


# genome-scale texts (no reads): the index builder's parity above 2^32 symbols
LARGE_SPECS = {
    "copies_4g": {"ref": ("copies", 91, 275_000_000, 8)},
}


def opts_and_files(v) -> tuple[list[str], list[int] | None]:
    """An option list of a spec, and the read files it runs on (None: all)."""
    return (list(v), None) if isinstance(v, list) else (list(v[0]), list(v[1]))


def md5_file(path: str) -> str:
    h = hashlib.md5()
    with open(path, "rb") as fh:
        for b in iter(lambda: fh.read(1 << 20), b""):
            h.update(b)
    return h.hexdigest()


def md5_bytes(b: bytes) -> str:
    return hashlib.md5(b).hexdigest()


def reference_contigs(spec) -> list[tuple[str, np.ndarray]]:
    kind = spec["ref"][0]
    if kind == "plain":
        return [("chr1", synth.plain_reference(spec["ref"][1], spec["ref"][2]))]
    if kind == "adna":
        return [("chr1", synth.adna_reference(spec["ref"][1], spec["ref"][2]))]
    if kind == "human":
        return synth.human_like_reference(spec["ref"][1], spec["ref"][2])
    if kind == "copies":
        return synth.diverged_copies_reference(spec["ref"][1], spec["ref"][2], spec["ref"][3])
    raise ValueError(kind)


def build(name: str, workdir: str, device: str = "cpu") -> dict:
    """Regenerate a fixture's inputs under workdir: <name>.fa, <name>.bwt, read files."""
    spec = SPECS[name]
    os.makedirs(workdir, exist_ok=True)
    contigs = reference_contigs(spec)
    fa = os.path.join(workdir, name + ".fa")
    synth.write_fasta(fa, contigs)
    prefix = os.path.join(workdir, name)
    seq = b"".join(np.ascontiguousarray(c, dtype=np.uint8).tobytes() for _, c in contigs)
    ix.build_index(seq, prefix, device=device)
    ref = np.frombuffer(seq, dtype=np.uint8)
    kind = spec["reads"][0]
    fqs = []
    if kind == "plain":
        r, L = synth.plain_reads(spec["reads"][1], ref, spec["reads"][2], spec["reads"][3])
        fqs.append((r, L))
    elif kind == "adna":
        r, L = synth.adna_reads(spec["reads"][1], ref, spec["reads"][2])
        fqs.append((r, L))
    elif kind == "uniform":
        r, L = synth.uniform_length_reads(spec["reads"][1], ref, spec["reads"][2], spec["reads"][3], spec["reads"][4])
        fqs.append((r, L))
    elif kind == "two_uniform":
        sd, n1, lo1, hi1, n2, lo2, hi2 = spec["reads"][1:]
        fqs.append(synth.uniform_length_reads(sd, ref, n1, lo1, hi1, sub_rate=0.01, deam=0.0))
        fqs.append(synth.uniform_length_reads(sd + 1, ref, n2, lo2, hi2, sub_rate=0.01, deam=0.0))
    elif kind == "paired":
        m1, l1, m2, l2 = synth.paired_reads(spec["reads"][1], ref, spec["reads"][2])
        fqs += [(m1, l1), (m2, l2)]
    files = []
    if kind == "file":
        files.append(os.path.join(HERE, spec["reads"][1]))
    else:
        for i, (r, L) in enumerate(fqs):
            p = os.path.join(workdir, f"{name}_{i + 1}.fq")
            synth.write_fastq(p, r, L)
            files.append(p)
    return {"name": name, "spec": spec, "fa": fa, "prefix": prefix, "bwt": prefix + ".bwt", "reads": files}


def load_expected(name: str) -> tuple[dict, dict]:
    with open(os.path.join(HERE, name + ".json")) as fh:
        meta = json.load(fh)
    arrs = dict(np.load(os.path.join(HERE, name + ".npz")))
    return meta, arrs
