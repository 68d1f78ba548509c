"""The C ABI as a caller outside this repository sees it: a C11 and a C++17 caller, written the way a
maintainer would bind the library inside the reference's driver (search, record-array regrowth,
serving, the position step), compile against include/ with warnings as errors and link against the
built libraries. Nothing is run (no GPU here); the loader test in test_host.py covers the exports."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "galn_amd", "lib")

BODY = r"""
    galn_hip_index *idx;
    if (galn_hip_index_open(bwt, &meta, dev, &idx)) {
        fprintf(stderr, "[galn] GPU: %s\n", galn_hip_last_error());
        return 10;
    }
    int r = galn_hip_search_batched(idx, &R, &O, 0, NULL);
    if (r == GALN_HIP_E_RECCAP) {
        /* the caller grows its record array to O.n_rec and searches again */
        r = galn_hip_search_batched(idx, &R, &O, 0, NULL);
    }
    galn_hip_index_set_sa(idx, sa_samples, n_sa, 32);
    uint64_t pos[1];
    int32_t ref_len[1] = {0};
    int32_t strand[1];
    galn_hip_sa_positions(idx, rows, 1, pos, ref_len, strand);
    galn_hip_server_cfg_t cfg;
    memset(&cfg, 0, sizeof cfg);
    galn_hip_server *srv;
    if (!galn_hip_server_open(idx, &cfg, &srv)) {
        uint64_t id;
        double ms;
        if (galn_hip_server_submit(srv, &R, &id) == 0 && galn_hip_server_wait(srv, id, -1) == 1)
            galn_hip_server_result(srv, id, &O, &ms);
        galn_hip_server_close(srv);
    }
    galn_hip_index_close(idx);
    return r;
"""

C_CALLER = r"""
#include <stdio.h>
#include <string.h>
#include "galn_hip.h"
#include "galn_host.h"
#include "galn_cpu.h"

int search(const uint32_t *bwt, galn_bwt_meta_t meta, int dev, const uint8_t *h_q, const uint64_t *h_off,
           const int32_t *h_len, const int32_t *h_maxdiff, const int32_t *h_bid, int nreads,
           const galn_search_opt_t *bopts, const int32_t *bstacks, int nbatch, int32_t *n_aln, uint64_t *off,
           galn_aln_rec_t *rec, uint64_t rec_cap, const uint64_t *sa_samples, uint64_t n_sa, const uint64_t *rows)
{
    galn_hip_reads R = {.n_reads = nreads, .query = h_q, .qoff = h_off, .qlen = h_len, .max_diff = h_maxdiff,
                        .batch = h_bid, .n_batches = nbatch, .bopts = bopts, .n_stacks = bstacks};
    galn_hip_out O = {n_aln, off, rec, rec_cap, 0, NULL, NULL, NULL};
""" + BODY + r"""
}

int main(int argc, char **argv) { return argc > 1 ? galn_aln_main(argc - 1, argv + 1) : 0; }
"""

CXX_CALLER = r"""
#include <cstdio>
#include <cstring>
#include <vector>
#include "galn_hip.h"
#include "galn_host.h"
#include "galn_cpu.h"

int search(const uint32_t *bwt, const galn_bwt_meta_t &meta, int dev, const std::vector<uint8_t> &h_q,
           const std::vector<uint64_t> &h_off, const std::vector<int32_t> &h_len,
           const std::vector<int32_t> &h_maxdiff, const std::vector<int32_t> &h_bid,
           const std::vector<galn_search_opt_t> &bopts, const std::vector<int32_t> &bstacks,
           const std::vector<uint64_t> &sa, const uint64_t *rows)
{
    const int nreads = (int)h_len.size();
    const uint64_t *sa_samples = sa.data();
    const uint64_t n_sa = sa.size();
    galn_hip_reads R = {nreads, h_q.data(), h_off.data(), h_len.data(), h_maxdiff.data(), h_bid.data(),
                        (int32_t)bopts.size(), bopts.data(), bstacks.data()};
    std::vector<int32_t> n_aln(nreads);
    std::vector<uint64_t> off(nreads);
    std::vector<galn_aln_rec_t> rec(2 * (size_t)nreads);
    galn_hip_out O = {n_aln.data(), off.data(), rec.data(), rec.size(), 0, nullptr, nullptr, nullptr};
""" + BODY + r"""
}

int main() { return 0; }
"""


def _compile(compiler, std, src, tmp_path):
    exe = str(tmp_path / "caller")
    cmd = [compiler, f"-std={std}", "-Wall", "-Wextra", "-Werror", f"-I{INC}", "-o", exe, str(src),
           f"-L{LIB}", "-lgaln_hip", "-lgaln_host", f"-Wl,-rpath-link,{LIB}", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(exe)


@pytest.mark.skipif(not shutil.which("gcc"), reason="no C compiler")
def test_c11_caller_compiles_and_links(tmp_path):
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    _compile("gcc", "c11", src, tmp_path)


@pytest.mark.skipif(not shutil.which("g++"), reason="no C++ compiler")
def test_cxx17_caller_compiles_and_links(tmp_path):
    src = tmp_path / "caller.cpp"
    src.write_text(CXX_CALLER)
    _compile("g++", "c++17", src, tmp_path)
