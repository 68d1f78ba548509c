#!/usr/bin/env python3
"""bench.py -- reads/s of the MI355X BWA-aln search, batched vs plain form, on one GPU.

Workload (default): the batched aDNA mixed-length configuration -- a 100 Mbp synthetic
reference (41% GC, 10% duplicated 300 bp-5 kb segments at 0-5% divergence, 1% N runs)
and 10M reads of log-normal length (median 45, clipped 30-150) with 5' C->T / 3' G->A
deamination, 0.5% substitutions and 5% reads from an unrelated random sequence,
searched with `aln -l 1024 -n 0.01 -o 2`.

A step is one search of the whole read set with the reads and the index resident in
HBM, from the call into the library until every record is in host memory. The bench
times K steps of the batched form (the reported value) and --plain-steps steps of the
plain form on the first --plain-reads reads, checks that both forms give identical records, runs the serving loop under
Poisson arrivals at 25/50/75/90% of its peak (p50/p99 request latency), times the reference's CPU search
(oracle/_ref, built from the reference's C sources) on a bounded sample of the same
reads and checks the GPU records against it, and reports the dominant kernel's
achieved algorithmic bytes/s against MI355X's HBM peak and the measured
random-access (pointer-chase) ceiling.

Other workloads (--workload): `plain` (5 Mbp i.i.d. reference, 100k x 100 bp reads, default
options), `human` (3.1 Gbp human-like reference, 50M reads of 30-150 bp, aDNA options, and the
serving sweep on that resident index with requests of 30-250 bp reads), `deep` (20 Gbp: a 2.5 Gbp
genome and 8 diverged copies, n = 4e10 > 2^32; 2 x 20M x 150 bp pairs, aln -n 0.04 -o 2 -e 6 -k 2).
Their indexes are built on the GPU by the library's builder.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import struct
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "reads/s aligned, bit-exact .sai, 1 GPU, batched vs plain; p50/p99 request ms"
WORKLOADS = {
    "adna": "batched aDNA mixed length: 100 Mbp reference (41% GC, 10% duplicated segments, 1% N), "
            "10M reads 30-150 bp (median 45), aln -l 1024 -n 0.01 -o 2",
    "plain": "plain small: 5 Mbp i.i.d. reference, 100k x 100 bp reads (1% substitutions, 0.1% indels), "
             "default aln options",
    "human": "human scale: 3.1 Gbp reference (24 contigs, 41% GC, 45% interspersed repeats, tandem repeats), "
             "50M reads 30-150 bp with deamination, aln -l 1024 -n 0.01 -o 2; serving on the same resident index "
             "with requests of 1k-1M reads x 30-250 bp, 50/50 default / aDNA options",
    "deep": "deep repeats: 20 Gbp (a 2.5 Gbp genome and 8 copies at 1-3% substitutions, 0.2% indels; n = 4e10), "
            "2 x 20M x 150 bp pairs (both mates' files as one read set), aln -n 0.04 -o 2 -e 6 -k 2",
}
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    sys.stderr.write(f"[bench {time.strftime('%H:%M:%S')}] {msg}\n")
    sys.stderr.flush()


def parse_sai(path: str, n: int):
    """(n_aln[n], list of record arrays) from a .sai file."""
    raw = open(path, "rb").read()
    off = 4 + 64
    n_aln = np.zeros(n, np.int32)
    recs = []
    for i in range(n):
        (na,) = struct.unpack_from("<i", raw, off)
        off += 4
        n_aln[i] = na
        recs.append(np.frombuffer(raw, dtype=np.uint64, count=3 * na, offset=off).reshape(na, 3))
        off += 24 * na
    return n_aln, recs


def host_cpu_info() -> dict:
    """The host's CPUs: logical CPUs, those this process may run on, model name and sockets."""
    model, sockets = None, set()
    try:
        for ln in open("/proc/cpuinfo"):
            k, _, v = ln.partition(":")
            k = k.strip()
            if k == "model name" and model is None:
                model = v.strip()
            elif k == "physical id":
                sockets.add(v.strip())
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model,
            "sockets": len(sockets) or None}


def cpu_threads() -> int:
    """Every CPU this process may run on, within the CPU share its host grants (OMP_NUM_THREADS,
    when the host sets it)."""
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def cpu_baseline(w, res, sample: int, threads: int, workdir: str) -> dict:
    """The reference's CPU search on the first `sample` reads, and its agreement with the GPU."""
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "galn_ref")
    fq = w.fastq(os.path.join(workdir, "sample.fq"), 0, sample)
    sai = os.path.join(workdir, "sample.sai")
    if os.path.exists(ref_bin):
        kind = "reference"
        cmd = [ref_bin, *w.opts, "--il", "8", "--threads", str(threads), "-f", sai, w.prefix, fq]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode not in (0, 4):
            raise RuntimeError(f"reference CPU search failed: {r.stderr}")
        m = re.search(r"(\d+) reads searched in ([0-9.]+) s", r.stderr)
        secs = float(m.group(2))
        n_aln, recs = parse_sai(sai, sample)
    else:
        # the oracle's restatement, when the reference build is absent
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        kind = "port"
        o = orc.parse_aln_opts(w.opts)
        q = w.query
        t = time.time()
        bid = orc.batch_ids(sample)
        bopts, nst, maxd = orc.batch_options(o, w.lens[:sample].astype(np.int32), bid)
        rr = orc.search(w.words, w.meta, q, w.qoff[:sample], w.lens[:sample].astype(np.int32), maxd, bid, bopts,
                        nst, o["seed_len"], threads)
        secs = time.time() - t
        n_aln = rr.n_aln
        recs = [rr.records(i) for i in range(sample)]
    agree = bool((np.where(n_aln > 0, n_aln, 0) == np.where(res.n_aln[:sample] > 0, res.n_aln[:sample], 0)).all())
    if agree:
        for i in range(sample):
            if n_aln[i] > 0 and not np.array_equal(recs[i], res.records(i)):
                agree = False
                break
    return {"value": sample / secs, "unit": "reads/s", "cores": threads, "kind": kind, "seconds": secs,
            "sample": f"first {sample} reads of the same workload, {threads} host threads, interleaved search",
            "host": host_cpu_info(), "records_equal_gpu": agree}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_search_k_heavy_adna10M_traffic.md")


def measured_traffic(kernel: str, args, tuned: dict) -> int | None:
    """HBM bytes per launch of `kernel` measured by the FETCH_SIZE / WRITE_SIZE passes recorded
    in profiles/ (scaled by a device copy of known size): for the default workload run with the
    tuning that was profiled, else None."""
    if (args.workload != "adna" or args.reads != 10_000_000 or args.ref_bp != 100_000_000
            or not os.path.exists(TRAFFIC_PROFILE)):
        return None
    value, config = None, None
    for ln in open(TRAFFIC_PROFILE):
        f = ln.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            value = int(f[2])
        if len(f) >= 3 and len(f) % 2 == 1 and f[0] == "traffic_config":
            config = {f[i]: int(f[i + 1]) for i in range(1, len(f), 2)}
    if config is None or any(tuned.get(k) != v for k, v in config.items()):
        return None
    return value


def serving_sweep(w, index, nv, loads, seconds: float, max_req: int, seed: int = 7, slots: int = 0,
                  launch_reads: int = 0) -> dict:
    """Serving loop under Poisson arrivals. Requests are slices of the workload's reads of
    log-uniform size (1000 .. max_req reads), each searched with BWA's default options or the
    aDNA options (50/50), batch options computed per request as `galn aln` would. The
    server's peak rate is measured first (requests submitted back to back); each offered
    load then submits requests at that fraction of the peak for `seconds`, and reports the
    p50/p99 latency (submit to the request's records ready) and the reads/s served."""
    import threading

    rng = np.random.default_rng(seed)
    opt_sets = {"default": nv.gap_opt_from_args([]), "adna": nv.gap_opt_from_args(list(w.opts))}
    lens32 = w.lens.astype(np.int32)

    def make_request():
        size = int(np.exp(rng.uniform(np.log(1000), np.log(max_req))))
        size = min(size, w.n)
        lo = int(rng.integers(0, w.n - size + 1))
        hi = lo + size
        o = opt_sets["default" if rng.random() < 0.5 else "adna"]
        ql = lens32[lo:hi]
        bid = (np.arange(size) // 0x40000).astype(np.int32)
        bopts, nst, md = nv.prepare_batches(o, ql, bid)
        qo = w.qoff[lo:hi]
        q0 = int(qo.min())
        q1 = int((qo + ql).max())
        return (w.query[q0:q1], qo - np.uint64(q0), ql, md, bid, bopts, nst)

    def run(srv, reqs, rate):
        """Submit reqs (at `rate` requests/s, or back to back if None); latencies in ms."""
        ids, lats, t_sub = [], [], []
        done = threading.Event()

        def collect():
            k = 0
            while k < len(reqs):
                while k >= len(ids):
                    time.sleep(0.0005)
                srv.wait(ids[k], -1)
                _, lat = srv.result(ids[k])
                lats.append(lat)
                k += 1
                if k % 8 == 0 and k < len(reqs):
                    log(f"  serving: {k} of {len(reqs)} requests done")
            done.set()

        th = threading.Thread(target=collect, daemon=True)
        th.start()
        t0 = time.perf_counter()
        nxt = t0
        for r in reqs:
            if rate:
                nxt += float(rng.exponential(1.0 / rate))
                d = nxt - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            while True:
                rid = srv.submit(*r)
                if rid is not None:
                    break
                time.sleep(0.001)  # admission limit: retry shortly
            ids.append(rid)
            t_sub.append(time.perf_counter())
        done.wait()
        return lats, time.perf_counter() - t0

    out = {"request_reads": f"log-uniform 1000..{max_req}", "options": "50/50 default / aln -l 1024 -n 0.01 -o 2",
           "loads": []}
    with nv.Server(index, slots=slots, max_launch_reads=launch_reads) as srv:
        warm = [make_request() for _ in range(2)]
        run(srv, warm, None)
        # the sustained rate over enough requests to keep every launch slot busy
        peak_reqs = [make_request() for _ in range(48)]
        lats, el = run(srv, peak_reqs, None)
        n_reads = sum(int(r[2].size) for r in peak_reqs)
        peak = n_reads / el
        mean_size = n_reads / len(peak_reqs)
        out["peak_reads_per_s"] = peak
        log(f"serving: peak {peak:.3e} reads/s over {len(peak_reqs)} requests")
        for load in loads:
            rate = load * peak / mean_size
            n_req = max(4, int(rate * seconds))
            reqs = [make_request() for _ in range(n_req)]
            lats, el = run(srv, reqs, rate)
            reads = sum(int(r[2].size) for r in reqs)
            rec = {"offered": load, "requests": n_req, "reads_per_s": reads / el,
                   "p50_ms": float(np.percentile(lats, 50)), "p99_ms": float(np.percentile(lats, 99))}
            out["loads"].append(rec)
            log(f"serving: load {load:.2f}: {n_req} requests, p50 {rec['p50_ms']:.1f} ms, p99 {rec['p99_ms']:.1f} ms, "
                f"{rec['reads_per_s']:.3e} reads/s")
        out["server_stats"] = srv.stats()
        log(f"serving: server counters {out['server_stats']}")
    # the batched form on exactly the peak phase's requests (one read set, each request's reads
    # with its own batch options): the throughput the serving peak is to be compared with
    parts = list(zip(*[r for r in peak_reqs]))
    qs, qos, qls, mds, bids, bopt_l, nst_l = parts
    qoff_all, q_all, batch_all, base, b0 = [], [], [], 0, 0
    bopts_all = (nv.SearchOpt * sum(len(b) for b in bopt_l))()
    nst_all = np.concatenate(nst_l).astype(np.int32)
    k = 0
    for q, qo, bid, bo in zip(qs, qos, bids, bopt_l):
        q_all.append(q)
        qoff_all.append(qo + np.uint64(base))
        batch_all.append(bid + b0)
        base += int(q.size)
        for x in bo:
            bopts_all[k] = x
            k += 1
        b0 += len(bo)
    rs_mix = nv.ReadSet(index, np.concatenate(q_all), np.concatenate(qoff_all), np.concatenate(qls),
                        np.concatenate(mds), np.concatenate(batch_all).astype(np.int32), bopts_all, nst_all)
    # one untimed search when a search of these requests is short (first-use costs); on a slow
    # workload one timed search, whose first-use costs are a few ms of it
    log(f"serving: the batched form on the same {len(peak_reqs)} requests ({n_reads} reads)")
    if n_reads / peak < 20:
        rs_mix.search(nv.FORM_BATCHED)
    t0 = time.perf_counter()
    rs_mix.search(nv.FORM_BATCHED)
    rs_mix.results()
    el = time.perf_counter() - t0
    rs_mix.close()
    out["batched_same_requests_reads_per_s"] = n_reads / el
    out["peak_over_batched_same_requests"] = peak / (n_reads / el)
    log(f"serving: the batched form on the same {len(peak_reqs)} requests: {n_reads / el:.3e} reads/s")
    return out


def end_to_end(w, workdir: str) -> dict:
    """`galn aln` on the workload's reads as a gzip FASTQ, from the command line to the .sai on disk:
    streamed (blocks of four bwa batches read, searched and written at the same time) and with the
    whole input read first; both .sai files must be the same bytes."""
    import gzip
    from galn_amd import native as nv
    fq = os.path.join(workdir, "reads.fq.gz")
    t = time.time()
    step = 1 << 20
    with gzip.open(fq, "wb", compresslevel=1) as fh:
        for s in range(0, w.n, step):
            e = min(w.n, s + step)
            L = w.lens[s:e]
            rows = w.reads[s:e]
            fh.write(b"".join(b"@r%d\n%s\n+\n%s\n" % (s + i, rows[i, :L[i]].tobytes(), b"I" * int(L[i]))
                              for i in range(e - s)))
    log(f"end to end: wrote {w.n} reads as {os.path.getsize(fq) / 1e6:.1f} MB of gzip FASTQ in {time.time() - t:.1f} s")
    res = {"reads": w.n, "fq_gz_bytes": os.path.getsize(fq)}
    md5 = {}
    for name, extra in (("streamed", []), ("whole_input", ["--block-batches", "0"])):
        sai = os.path.join(workdir, f"e2e_{name}.sai")
        t = time.perf_counter()
        r = subprocess.run([nv.GALN_BIN, "aln", *extra, *w.opts, "-f", sai, w.prefix, fq], capture_output=True, text=True)
        el = time.perf_counter() - t
        if r.returncode != 0:
            raise RuntimeError(f"galn aln failed ({r.returncode}): {r.stderr}")
        import hashlib
        h = hashlib.md5()
        with open(sai, "rb") as fh:
            for b in iter(lambda: fh.read(1 << 24), b""):
                h.update(b)
        md5[name] = h.hexdigest()
        res[name] = {"seconds": el, "reads_per_s": w.n / el, "galn_stderr": r.stderr.strip().splitlines()[-1]}
        log(f"end to end, {name}: {el:.2f} s ({w.n / el:.3e} reads/s): {res[name]['galn_stderr']}")
    res["sai_equal"] = md5["streamed"] == md5["whole_input"]
    return res


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=list(WORKLOADS), default="adna")
    ap.add_argument("--reads", type=int, default=0, help="reads (pairs for `deep`); 0: the workload's size")
    ap.add_argument("--ref-bp", type=int, default=0, help="reference bp (base genome for `deep`); 0: the workload's")
    ap.add_argument("--cpu-sample", type=int, default=0, help="reads for the CPU baseline (0: ~15 s of CPU work)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=1, help="timed steps of the plain form")
    ap.add_argument("--plain-reads", type=int, default=-1,
                    help="reads the plain form is timed on, the first ones of the workload (0: all; it is ~60x "
                         "slower; -1: 1M, or 200k for the genome-scale workloads)")
    ap.add_argument("--serve-seconds", type=float, default=8.0, help="arrival time per offered load (0: no serving)")
    ap.add_argument("--serve-max-request", type=int, default=1_000_000)
    ap.add_argument("--serve-slots", type=int, default=0, help="serving launches on the device at once (0: library default)")
    ap.add_argument("--serve-launch-reads", type=int, default=0, help="reads per serving launch (0: the library's adaptive size)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also copy a buffer of this many MiB device to device three times (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--tune", default="", metavar="JSON",
                    help='use this tuning instead of autotuning, e.g. \'{"heavy_pops": 1, "block_threads": 128}\'')
    ap.add_argument("--autotune-reads", type=int, default=0,
                    help="reads the autotuner searches (the first ones; 0: all the timed reads)")
    ap.add_argument("--end-to-end", action="store_true",
                    help="also time `galn aln` on the reads as a gzip FASTQ, streamed and whole-input")
    ap.add_argument("--workdir", default="")
    args = ap.parse_args()
    # the library reports autotune trials and index-build chunks on stderr (long, otherwise silent calls)
    os.environ.setdefault("GALN_HIP_VERBOSE", "1")
    if args.gpus != 1:
        sys.stderr.write("bench.py: measures one GPU only (reads shard across GPUs by read, each GPU holding "
                         "the whole index, with no exchange during the search); run with --gpus 1\n")
        return 2

    import torch

    from galn_amd import native as nv
    from galn_amd import synth, workloads

    if not torch.cuda.is_available():
        sys.stderr.write("bench.py: no GPU visible\n")
        return 3
    workdir = args.workdir or tempfile.mkdtemp(prefix="galn_bench_")
    os.makedirs(workdir, exist_ok=True)
    threads = args.cpu_threads or cpu_threads()

    # (reference bp, reads timed, reads of the configuration): the genome-scale configurations are
    # timed on the first reads of their generators (reads are i.i.d. samples of the same model;
    # their repeat-rich reads backtrack ~70x more than the aDNA workload's)
    sizes = {"adna": (100_000_000, 10_000_000, 10_000_000), "plain": (5_000_000, 100_000, 100_000),
             "human": (3_100_000_000, 2_000_000, 50_000_000), "deep": (2_500_000_000, 2_500_000, 20_000_000)}[args.workload]
    args.ref_bp = args.ref_bp or sizes[0]
    args.reads = args.reads or sizes[1]
    log(f"building workload {args.workload}: {args.ref_bp} bp reference, {args.reads} reads (index on the GPU)")
    t = time.time()
    serve_pool = None
    if args.workload == "adna":
        w = workloads.adna_mixed(workdir, ref_bp=args.ref_bp, n_reads=args.reads, device="cuda", workers=threads)
    elif args.workload == "plain":
        w = workloads.plain_small(workdir, ref_bp=args.ref_bp, n_reads=args.reads, device="cuda")
    elif args.workload == "human":
        w = workloads.human_scale(workdir, ref_bp=args.ref_bp, n_reads=args.reads, workers=threads)
        if args.serve_seconds > 0:
            # serving requests: reads of 30-250 bp from the same reference
            sr, sl = synth.uniform_length_reads(303, w.ref, 2_000_000, 30, 250, workers=threads)
            serve_pool = workloads.Workload("serving_pool", w.prefix, w.meta, w.words, sr, sl,
                                            list(workloads.ADNA_OPTS)).prepare()
    else:
        w = workloads.deep_copies(workdir, base_bp=args.ref_bp, n_pairs=args.reads, workers=threads, progress=log)
    w.ref = None
    log(f"workload ready in {time.time() - t:.1f} s ({w.timings}); n={w.n}, index {w.words.nbytes / 1e6:.1f} MB")
    torch.cuda.synchronize()
    torch.cuda.empty_cache()

    # random-access ceiling at the index's working-set size
    ws = int(w.words.nbytes)
    best = 0.0
    for waves in (8, 16, 32):
        lps, _ = nv.gather_ceiling(max(ws, 1 << 20), 64, waves, 1000)
        best = max(best, lps)
    ceiling_gbs = best * 64 / 1e9
    log(f"pointer-chase ceiling at {ws / 1e6:.1f} MB: {ceiling_gbs:.1f} GB/s of 64-byte lines")

    index = nv.Index(w.words, w.meta, 0)
    rs = w.readset(index)
    # the batched form: persistent kernel + device work queue, launches replayed as a captured
    # hipGraph, tuning chosen by the library's autotuner on this read set
    BATCHED_FLAGS = nv.GALN_HIP_GRAPH
    # tuned on the read set that is timed: on a smaller sample of the repeat-rich workloads the
    # longest single searches decide every trial (on the first 200k human-scale reads cutoff 1 won
    # by 8%; on the 2M timed reads it was 1.8x slower than cutoff 128)
    at_n = args.autotune_reads or w.n
    at_n = min(at_n, w.n)
    if args.tune:
        # a given tuning (profiling passes that must all run the same one)
        tuned = {k: int(v) for k, v in json.loads(args.tune).items()}
        log("tuning given on the command line")
    elif at_n < w.n:
        # tune on the first at_n reads, then search the whole read set with that tuning
        log(f"autotuning the batched form on the first {at_n} reads")
        hi_q = int((w.qoff[:at_n] + w.lens[:at_n].astype(np.uint64)).max())
        lo_q = int(w.qoff[:at_n].min())
        rs_t = nv.ReadSet(index, w.query[lo_q:hi_q], w.qoff[:at_n] - np.uint64(lo_q), w.lens[:at_n].astype(np.int32),
                          w.maxdiff[:at_n], w.batch[:at_n], w.bopts, w.nstacks)
        tuned = rs_t.autotune(BATCHED_FLAGS)
        rs_t.close()
    else:
        log("autotuning the batched form on this read set")
        tuned = rs.autotune(BATCHED_FLAGS)
    TUNE = nv.Tune()
    for k, v in tuned.items():
        setattr(TUNE, k, int(v))
    log(f"tuning: {tuned}")
    # host arrays for the records, allocated once in pinned memory and reused by every step (as a
    # caller that searches batch after batch would): the copies then run at full PCIe speed
    def pinned(shape, dtype):
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return torch.empty(nbytes, dtype=torch.uint8, pin_memory=True).numpy().view(dtype).reshape(shape)

    bufs = rs.result_buffers(want_pops=True, want_bytes=True, alloc=pinned)
    log("warm-up")
    for _ in range(max(args.warmup, 0)):
        t0 = time.perf_counter()
        rs.search(nv.FORM_BATCHED, BATCHED_FLAGS, TUNE)
        rs.results(want_pops=False, out=bufs)
        log(f"  warm-up search: {(time.perf_counter() - t0) * 1e3:.1f} ms")

    def timed(rset, n: int, form: int, steps: int, out):
        times, launches = [], []
        res = None
        for s in range(steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if form == nv.FORM_BATCHED:
                rset.search(form, BATCHED_FLAGS, TUNE)
            else:
                rset.search(form, 0)
            res = rset.results(want_pops=(s == steps - 1), want_bytes=(s == steps - 1), out=out)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            launches.append(rset.launches())
            log(f"  form {form} step {s}: {times[-1] * 1e3:.1f} ms, {n / times[-1]:.3e} reads/s")
        return times, launches, res

    log(f"timing batched form, {args.steps} steps")
    tb, lb, res_b = timed(rs, w.n, nv.FORM_BATCHED, args.steps, bufs)
    if args.plain_reads < 0:
        args.plain_reads = 200_000 if args.workload in ("human", "deep") else 1_000_000
    n_plain = w.n if args.plain_reads == 0 else min(args.plain_reads, w.n)
    if not args.no_plain:
        # the plain form (~60x slower) on the first n_plain reads, a sample of the same generator;
        # its records must equal the batched form's for those reads
        lo_q = int(w.qoff[:n_plain].min())
        hi_q = int((w.qoff[:n_plain] + w.lens[:n_plain].astype(np.uint64)).max())
        rs_p = nv.ReadSet(index, w.query[lo_q:hi_q], w.qoff[:n_plain] - np.uint64(lo_q),
                          w.lens[:n_plain].astype(np.int32), w.maxdiff[:n_plain], w.batch[:n_plain], w.bopts,
                          w.nstacks)
        log(f"timing plain form on the first {n_plain} reads, {args.plain_steps} steps")
        tp, lp, res_p = timed(rs_p, n_plain, nv.FORM_PLAIN, args.plain_steps, None)
        nb = n_plain
        end_b = int(res_b.rec_off[nb - 1]) + max(int(res_b.n_aln[nb - 1]), 0)
        same = (np.array_equal(res_p.n_aln, res_b.n_aln[:nb]) and np.array_equal(res_p.rec_off, res_b.rec_off[:nb])
                and np.array_equal(res_p.rec, res_b.rec[:end_b]))
        rs_p.close()
    else:
        tp, lp, same = [], [], None
    log(f"plain == batched: {same}")

    ms_b = 1e3 * sum(tb) / len(tb)
    value = w.n / (ms_b / 1e3)
    # per kernel of the last timed search: k_search (tiers 0-9) and k_heavy (tiers 10-)
    per = {}
    for l in lb[-1]:
        kname = "k_search" if l["tier"] < 10 else "k_heavy"
        d = per.setdefault(kname, {"ms": 0.0, "algo_bytes": 0, "launches": 0})
        d["ms"] += l["ms"]
        d["algo_bytes"] += l["algo_bytes"]
        d["launches"] += 1
    # the dominant kernel: the one with the most time; its first-tier launch is the unit
    dom = max(per, key=lambda k: per[k]["ms"])
    first = [l for l in lb[-1] if l["tier"] == (0 if dom == "k_search" else 10)]
    k_ms = first[0]["ms"]
    k_bytes = first[0]["algo_bytes"]
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    total_algo = int(res_b.occ_bytes.sum())
    traffic = measured_traffic(dom, args, tuned)
    # per-read backtracking (popped stack entries), log2 bins: [0], [1], [2,3], [4,7], ...
    pops = res_b.pops.astype(np.int64)
    bins = np.bincount(np.where(pops > 0, np.floor(np.log2(np.maximum(pops, 1))).astype(np.int64) + 1, 0))
    pops_hist = {("0" if b == 0 else f"{1 << (b - 1)}-{(1 << b) - 1}"): int(c) for b, c in enumerate(bins) if c}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_b,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded generator; index built on the GPU from the generated reference)",
        "config": {"workload": WORKLOADS[args.workload], "name": args.workload,
                   "ref_bp": args.ref_bp, "reads": w.n, "config_reads": sizes[2] * (2 if args.workload == "deep" else 1), "aln_opts": " ".join(w.opts), "form": "batched",
                   "autotuned": tuned,
                   "index_bytes": ws},
        "plain": {"value": n_plain / (sum(tp) / len(tp)) if tp else None,
                  "ms_per_step": 1e3 * sum(tp) / len(tp) if tp else None, "steps": len(tp), "reads": n_plain},
        "batched_over_plain": (value / (n_plain / (sum(tp) / len(tp)))) if tp else None,
        "parity": {"plain_eq_batched": same},
        "pops_per_read": {"mean": float(pops.mean()), "max": int(pops.max()), "log2_histogram": pops_hist},
        "kernel": {"dominant": dom, "launch_ms": k_ms, "launch_algo_bytes": k_bytes, "per_kernel": per,
                   "all_launches_ms": sum(l["ms"] for l in lb[-1]), "launches": len(lb[-1]),
                   "algo_bytes_total": total_algo},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "achieved_is": "algorithmic index bytes (64 per distinct 128-symbol block per occurrence "
                                    "query) of the dominant kernel's launch / its duration",
                     "traffic_source": (f"profiled: {os.path.relpath(TRAFFIC_PROFILE, ROOT)} (rocprofv3 --pmc "
                                        "FETCH_SIZE, WRITE_SIZE of the same workload and tuning)")
                     if traffic is not None else None},
        "ceiling": {"kind": "pointer chase, 64-byte random lines", "ws_bytes": ws, "GB/s": ceiling_gbs,
                    "frac": achieved / ceiling_gbs if ceiling_gbs else None},
    }
    # the wave engine's counters (rounds, chains, clock ticks per phase) from one more search, not timed
    rs.search(nv.FORM_BATCHED, BATCHED_FLAGS | nv.GALN_HIP_STATS, TUNE)
    out["engine_stats"] = rs.engine_stats()
    if args.copy_calibration > 0:
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 device copies of {n} bytes")
        del src, dst
    if args.serve_seconds > 0 and args.workload in ("adna", "human"):
        pool = serve_pool if serve_pool is not None else w
        out["serving"] = serving_sweep(pool, index, nv, (0.25, 0.5, 0.75, 0.9), args.serve_seconds,
                                       min(args.serve_max_request, pool.n), slots=args.serve_slots,
                                       launch_reads=args.serve_launch_reads)
        out["serving"]["peak_over_batched"] = out["serving"]["peak_reads_per_s"] / value
    rs.close()
    index.close()
    if not args.no_cpu:
        sample = args.cpu_sample
        if sample <= 0:
            # calibrate: ~15 s of CPU work
            cal = min(w.n, 20000)
            c0 = cpu_baseline(w, res_b, cal, threads, workdir)
            sample = int(min(w.n, max(cal, 15.0 * c0["value"])))
        log(f"CPU baseline on {sample} reads, {threads} threads")
        cb = cpu_baseline(w, res_b, sample, threads, workdir)
        out["cpu_baseline"] = cb
        out["parity"]["cpu_sample_eq_gpu"] = cb["records_equal_gpu"]
        out["parity"]["cpu_sample_reads"] = sample
        log(f"CPU baseline: {cb['value']:.3e} reads/s ({cb['kind']}), records equal: {cb['records_equal_gpu']}")
    if args.end_to_end:
        out["end_to_end"] = end_to_end(w, workdir)
    if not args.workdir:
        shutil.rmtree(workdir, ignore_errors=True)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
