"""The triage oracle (oracle/triage_ref.py) pinned by the mathematics: the
canonical 6-mer classes in closed form, profiles against a brute-force count,
and the backward pass against central finite differences in float64 on every
parameter block.  A mistake planted in a copy of the oracle's backward must be
caught by the same check."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import triage_ref as T  # noqa: E402


def test_canonical_classes_closed_form():
    idx = T.canonical_index()
    assert idx.max() + 1 == T.KMER_DIM == (4 ** 6 + 4 ** 3) // 2
    sizes = np.bincount(idx)
    # 4^3 reverse-complement palindromes form singleton classes, the rest pairs
    assert (sizes == 1).sum() == 64 and (sizes == 2).sum() == T.KMER_DIM - 64

    def rc(x):
        r = 0
        for i in range(6):
            r = (r << 2) | (3 - ((x >> (2 * i)) & 3))
        return r
    for x in (0, 1, 27, 1000, 4095, 0b000110110011):
        assert idx[x] == idx[rc(x)]


def test_profile_matches_brute_force():
    rng = np.random.default_rng(5)
    seq = rng.integers(0, 4, 700).astype(np.uint8)
    seq[[10, 11, 300, 650]] = 14  # N: words through them are skipped
    start, length = 37, 600
    counts = {}
    for p in range(start, start + length - 5):
        w = seq[p:p + 6]
        if (w > 3).any():
            continue
        code = 0
        for b in w:
            code = code * 4 + int(b)
        r = 0
        for i in range(6):
            r = (r << 2) | (3 - ((code >> (2 * i)) & 3))
        counts[min(code, r)] = counts.get(min(code, r), 0) + 1
    prof = T.kmer_profile(seq, start, length)
    idx = T.canonical_index()
    expect = np.zeros(T.KMER_DIM)
    for c, n in counts.items():
        expect[idx[c]] = n
    np.testing.assert_allclose(prof, np.log1p(expect), rtol=1e-6)


def _batch(seed, nw=3, nf=2):
    rng = np.random.default_rng(seed)
    xw = np.log1p(rng.poisson(1.0, (nw, T.KMER_DIM))).astype(np.float64)
    xf = np.log1p(rng.poisson(1.0, (nf, T.KMER_DIM))).astype(np.float64)
    lab = (rng.random(nw * nf) < 0.4).astype(np.uint8)
    lab[0] = 1
    lab[-1] = 2  # ignored
    flat = T.init_params(seed).astype(np.float64)
    flat[T.layout()[0]["a"][0]] = 0.01
    flat[T.layout()[0]["c"][0]] = 0.3
    return flat, xw, xf, lab


def _fd_check(fb, flat, xw, xf, lab, picks, h=1e-6):
    _, loss, g = fb(flat, xw, xf, lab, pos_weight=3.0, bf16=False, dtype=np.float64)
    worst = 0.0
    for i in picks:
        p, m = flat.copy(), flat.copy()
        p[i] += h
        m[i] -= h
        lp = fb(p, xw, xf, lab, pos_weight=3.0, bf16=False, dtype=np.float64)[1]
        lm = fb(m, xw, xf, lab, pos_weight=3.0, bf16=False, dtype=np.float64)[1]
        fd = (lp - lm) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(1e-7, abs(fd) + abs(g[i])))
    return worst


def _picks(seed):
    rng = np.random.default_rng(seed)
    off, _ = T.layout()
    picks = []
    for k, (o, shape) in off.items():
        n = int(np.prod(shape)) if shape else 1
        picks += list(o + rng.choice(n, size=min(n, 4), replace=False))
    return picks


def test_backward_matches_finite_differences():
    flat, xw, xf, lab = _batch(7)
    assert _fd_check(T.forward_backward, flat, xw, xf, lab, _picks(7)) < 1e-5


def test_finite_differences_catch_a_planted_mistake():
    src = open(os.path.join(ROOT, "oracle", "triage_ref.py")).read()
    bad = src.replace('G["a"] = (g * dot).sum()', 'G["a"] = (g * dot).sum() * 0.5')
    assert bad != src
    ns = {}
    exec(compile(bad, "triage_ref_planted", "exec"), ns)
    flat, xw, xf, lab = _batch(8)
    a_idx = T.layout()[0]["a"][0]
    assert _fd_check(ns["forward_backward"], flat, xw, xf, lab, [a_idx]) > 1e-2


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], dtype=np.float32)
    r = T.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -6) and r[3] == -2.5
    np.testing.assert_array_equal(T.from_bf16_bits(T.bf16_bits(x)), r)
