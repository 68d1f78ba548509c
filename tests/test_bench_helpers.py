"""The bench's own helpers, on the CPU: .sai parsing (the CPU baseline's records are read back with
it), the lookup of the profiled HBM traffic, the CPU thread count and the one-GPU guard."""
from __future__ import annotations

import os
import subprocess
import sys
import types

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from galn_amd import native as nv  # noqa: E402


def test_parse_sai_reads_back_what_the_writer_wrote(tmp_path):
    opt = nv.gap_opt_from_args(["-l", "1024", "-n", "0.01", "-o", "2"])
    n_aln = np.array([0, 2, -5, 1], np.int32)  # -5, a per-read error code, is written as no records
    rec = np.arange(9, dtype=np.uint64).reshape(3, 3)
    rec_off = np.array([0, 0, 2, 2], np.uint64)
    path = tmp_path / "x.sai"
    path.write_bytes(nv.sai_bytes(opt, n_aln, rec_off, rec))
    got_n, got_recs = bench.parse_sai(str(path), 4)
    assert got_n.tolist() == [0, 2, 0, 1]
    assert got_recs[0].shape == (0, 3) and got_recs[2].shape == (0, 3)
    assert np.array_equal(got_recs[1], rec[0:2]) and np.array_equal(got_recs[3], rec[2:3])


def _args(**kw):
    a = dict(workload="adna", reads=10_000_000, ref_bp=100_000_000)
    a.update(kw)
    return types.SimpleNamespace(**a)


def test_measured_traffic_only_for_the_profiled_workload_and_tuning():
    tuned = {"heavy_pops": 1, "block_threads": 64}
    heavy = bench.measured_traffic("k_heavy", _args(), tuned)
    # the profile's figure: above the launch's algorithmic bytes (1.66e12), each 64-byte block a 128-byte line
    assert heavy is not None and heavy > 1_657_419_507_200
    assert bench.measured_traffic("k_search", _args(), tuned) is not None
    assert bench.measured_traffic("k_heavy", _args(), {"heavy_pops": 512}) is None  # another tuning
    assert bench.measured_traffic("k_heavy", _args(workload="human"), tuned) is None
    assert bench.measured_traffic("k_heavy", _args(reads=1_000_000), tuned) is None
    assert bench.measured_traffic("k_widths", _args(), tuned) is None  # not profiled


def test_cpu_threads_within_the_hosts_share(monkeypatch):
    n = len(os.sched_getaffinity(0))
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.cpu_threads() == 1
    monkeypatch.setenv("OMP_NUM_THREADS", str(n + 100))
    assert bench.cpu_threads() == n
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.cpu_threads() == n


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
