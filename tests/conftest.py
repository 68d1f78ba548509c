"""Shared test set-up: the `gpu` marker, paths, and session-cached golden workloads."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def golden_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("golden_inputs"))


_built: dict = {}


def golden_inputs(name: str, workdir: str, device: str = "cpu") -> dict:
    """A golden fixture's inputs (built once per session), checked against the recorded md5s."""
    import fixtures as fx
    if name not in _built:
        w = fx.build(name, os.path.join(workdir, name), device=device)
        meta, arrs = fx.load_expected(name)
        assert fx.md5_file(w["bwt"]) == meta["bwt_md5"], f"{name}: generated index drifted from the fixture"
        assert [fx.md5_file(p) for p in w["reads"]] == meta["reads_md5"], f"{name}: generated reads drifted"
        w["meta"], w["expected"] = meta, arrs
        _built[name] = w
    return _built[name]


def load_reads(path: str) -> dict:
    from galn_amd import native as nv
    return nv.read_fastq(path)


def golden_runs():
    """(fixture, run key) for every committed golden run."""
    import fixtures as fx
    out = []
    for name in fx.SPECS:
        meta, _ = fx.load_expected(name)
        for key in meta["runs"]:
            out.append((name, key))
    return out


# fixtures whose reads or options lie outside the GPU kernels' limits: the kernels report those
# reads as unsupported and `galn aln` finishes them on the CPU engine
BEYOND_GPU_LIMITS = ("beyond_limits",)


def gpu_golden_runs():
    """(fixture, run key) for every golden run the GPU kernels search whole."""
    return [(n, k) for n, k in golden_runs() if n not in BEYOND_GPU_LIMITS]


def run_opts(meta: dict, key: str) -> tuple[list[str], int]:
    r = meta["runs"][key]
    return list(r["opts"]), int(r["reads"])


def sai_md5(opt_struct, n_aln, rec_off, rec) -> str:
    import fixtures as fx
    from galn_amd import native as nv
    return fx.md5_bytes(nv.sai_bytes(opt_struct, np.asarray(n_aln), np.asarray(rec_off), np.asarray(rec)))
