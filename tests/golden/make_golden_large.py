#!/usr/bin/env python3
"""Writes tests/golden/large_index.json: the reference builder's .bwt for a text above 2^32
symbols.

The text is the deep-repeats model at genome scale -- a random 275 Mbp base genome and 8 copies
of it with 1-3% substitutions and 0.2% indels (synth.diverged_copies_reference, seed 91), about
2.2 Gbp, so the BWT of F . revcomp(F) has n ~ 4.4e9 > 2^32 symbols. The reference's CPU builder
(oracle/_ref/galn_index_ref: src/index/cpu_ref.c + libsais64, built by `make -C oracle`) writes
its .bwt and .sa; the md5s, primary and L2 go into the JSON, which the GPU builder's test
compares against. Needs ~48 GB of host memory and the reference tree.

    python tests/golden/make_golden_large.py <workdir>
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import fixtures as fx  # noqa: E402

REF_INDEX = os.path.join(fx.ROOT, "oracle", "_ref", "galn_index_ref")


def main(workdir: str) -> None:
    from galn_amd import synth
    os.makedirs(workdir, exist_ok=True)
    spec = fx.LARGE_SPECS["copies_4g"]
    t = time.time()
    contigs = fx.reference_contigs(spec)
    total = sum(int(c.size) for _, c in contigs)
    print(f"generated {total} bp in {time.time() - t:.1f} s", flush=True)
    fa = os.path.join(workdir, "copies_4g.fa")
    synth.write_fasta(fa, contigs)
    del contigs
    print(f"fasta written ({time.time() - t:.1f} s)", flush=True)
    prefix = os.path.join(workdir, "copies_4g")
    t = time.time()
    subprocess.run([REF_INDEX, "-t", str(os.cpu_count() or 1), "-v", "-f", fa, prefix], check=True)
    print(f"reference builder: {time.time() - t:.1f} s", flush=True)
    import struct
    with open(prefix + ".bwt", "rb") as fh:
        hdr = struct.unpack("<5Q", fh.read(40))
    out = {"spec": list(spec["ref"]), "ref_bp": total, "seq_len": 2 * total, "primary": hdr[0],
           "L2": [0] + list(hdr[1:5]), "bwt_md5": fx.md5_file(prefix + ".bwt"),
           "sa_md5": fx.md5_file(prefix + ".sa"), "fasta_md5": fx.md5_file(fa),
           "builder": "oracle/_ref/galn_index_ref (src/index/cpu_ref.c, libsais64)"}
    with open(os.path.join(HERE, "large_index.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
