"""The .bwt builder: BWT properties on tiny texts, and byte identity with the reference's own
builder (every golden fixture records the reference builder's md5, checked when the fixture's
inputs are regenerated in tests/conftest.py)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import fixtures as fx
from galn_amd import index as ix


def inverse_bwt(B: np.ndarray) -> np.ndarray:
    """T from the $-inclusive BWT (4 = $) by LF walking."""
    n = B.size
    key = np.where(B == 4, -1, B)  # $ sorts first
    order = np.argsort(key, kind="stable")  # F column: row i of F holds B[order[i]]
    lf = np.empty(n, np.int64)
    lf[order] = np.arange(n)
    out = np.empty(n - 1, np.uint8)
    r = 0  # row 0 = the $ suffix
    for i in range(n - 2, -1, -1):
        out[i] = B[r]
        r = lf[r]
    return out


@pytest.mark.parametrize("seed,n", [(1, 1), (2, 7), (3, 128), (4, 1000), (5, 4097)])
def test_bwt_inverts_to_the_text(seed, n):
    rng = np.random.default_rng(seed)
    f = rng.integers(0, 4, size=n).astype(np.uint8)
    primary, L = ix.bwt_from_forward(f)
    B = np.insert(L.astype(np.int64), primary, 4)
    text = np.concatenate([f, 3 - f[::-1]])
    np.testing.assert_array_equal(inverse_bwt(B), text)


def test_pack_layout_and_counts():
    rng = np.random.default_rng(9)
    L = rng.integers(0, 4, size=1000).astype(np.uint8)
    L2, words = ix.pack_bwt(L)
    assert words.size == 8 * (1000 // 128 + 2) + (1000 + 15) // 16
    assert L2 == [0] + list(np.cumsum([int((L == c).sum()) for c in range(4)]))
    meta = {"primary": 0, "L2": L2, "seq_len": 1000, "bwt_size": int(words.size)}
    np.testing.assert_array_equal(ix.decode_bwt(meta, words), L)
    # each block starts with the exclusive counts before it
    for b in range(0, 8):
        cnt = words[16 * b:16 * b + 8].view(np.uint64)
        np.testing.assert_array_equal(cnt, [int((L[:128 * b] == c).sum()) for c in range(4)])


def test_suffix_array_sorted():
    rng = np.random.default_rng(3)
    t = rng.integers(0, 3, size=500).astype(np.uint8)
    sa = ix.suffix_array(torch.as_tensor(t)).numpy()
    sufs = [t[i:].tobytes() for i in sa]
    assert sufs == sorted(sufs) and sorted(sa.tolist()) == list(range(t.size))


def test_ambiguous_bases_follow_lrand48():
    """N and other ambiguous bases become lrand48() & 3 after srand48(11), in order."""
    seq = b"ACGTNNRACGTN"
    f = ix.encode_forward(seq)
    draws = ix.lrand48_low2(4)
    np.testing.assert_array_equal(f[[4, 5, 6, 11]], draws)
    np.testing.assert_array_equal(f[[0, 1, 2, 3]], [0, 1, 2, 3])


@pytest.mark.parametrize("name", ["adna_mixed", "human_like", "deep_repeats"])
def test_builder_matches_reference_builder_md5(name, golden_dir):
    """The recorded md5 is the reference builder's (make_golden.py refuses to write a fixture
    on which the two builders disagree); regenerating must reproduce it."""
    from conftest import golden_inputs
    w = golden_inputs(name, golden_dir)
    meta, _ = fx.load_expected(name)
    assert fx.md5_file(w["bwt"]) == meta["bwt_md5"]
