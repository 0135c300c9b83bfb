"""The stage taps against the oracle and the reference's goldens: the query
lookup table (rmb_tap_lookup: cells and entries in order, small and
megablast-width tables, ambiguity codes), the post-dedup seeds
(rmb_tap_seeds: content and scan order against the oracle; the reference's
chr22 dedup-seeds golden, rmblast-lib/tests/dedup_seeds_regression.rs, on
the two chunks of the committed 3 Mbp fixture) and Phase 2a's preliminary
HSPs (rmb_tap_prelims: content and order); the library as the device
prepares it (rmb_tap_subject: ambiguity codes resolved by makeblastdb's
generator, reverse complement) against the oracle's generator."""
import ctypes
import gzip

import numpy as np
import pytest

import rmblast_amd
from rmblast_amd import synth
from conftest import CHR22_3MB, golden, ptr
from test_oracle_chr22 import read_2bit

pytestmark = pytest.mark.gpu


def comparison():
    m = rmblast_amd.load_matrix(golden("comparison.matrix"))
    mat = np.ascontiguousarray(np.array([[m.scores[i][j] for j in range(16)] for i in range(16)], dtype=np.int32))
    return m, mat


@pytest.mark.parametrize("ws,qlen", [(7, 3000), (11, 20000), (14, 60000), (16, 120000), (28, 9000), (7, 5)])
def test_tap_lookup_matches_oracle(oracle_lib, ws, qlen):
    m, _ = comparison()
    rng = np.random.default_rng(ws * 1000 + qlen)
    q = rng.integers(0, 4, qlen).astype(np.uint8)
    q[rng.random(qlen) < 0.01] = 14  # N runs break words
    if qlen > 100:
        q[qlen // 3:qlen // 3 + 40] = np.tile(np.array([0, 1], np.uint8), 20)  # a repeat: long cells
    s = rmblast_amd.Searcher(m, word_size=ws)
    lw, cells, entries = s.tap_lookup(q)
    olw = ctypes.c_int(0)
    n = oracle_lib.orc_lookup(ptr(q), len(q), ws, ctypes.byref(olw), None, 0, None, 0)
    ocells = np.zeros(4 ** olw.value + 1, dtype=np.uint32)
    oent = np.zeros(max(n, 1), dtype=np.uint32)
    oracle_lib.orc_lookup(ptr(q), len(q), ws, ctypes.byref(olw), ptr(ocells, ctypes.c_uint32), len(ocells),
                          ptr(oent, ctypes.c_uint32), len(oent))
    assert lw == olw.value
    if qlen >= ws:
        assert np.array_equal(cells, ocells)
        assert np.array_equal(entries, oent[:n])
    else:
        assert len(entries) == 0


@pytest.mark.parametrize("ws", [7, 11, 14])
def test_tap_seeds_match_oracle_in_order(oracle_lib, ws):
    w = synth.parity_config(seed=23, genome_len=120_000, n_families=4, n_copies=120)
    m, mat = comparison()
    s = rmblast_amd.Searcher(m, word_size=ws, xdrop_ungap=20, min_raw_gapped_score=60)
    s.load_library([f[1] for f in w.families])
    q = w.records[0][1]
    for f in range(len(w.families)):
        got = s.tap_seeds(q, f)
        subj = w.families[f][1]
        out = np.zeros((400000, 3), dtype=np.uint32)
        n = oracle_lib.orc_seeds(ptr(q), len(q), ptr(subj), len(subj), ws, 20, 60, ptr(mat, ctypes.c_int32),
                                 ptr(out, ctypes.c_uint32), 400000)
        assert n > 0 and np.array_equal(got, out[:n]), (ws, f, len(got), n)


def test_chr22_dedup_seeds_golden_on_gpu():
    want = {}
    for line in gzip.open(golden("chr22_seeds_chunks30_31.txt.gz"), "rt"):
        if line.startswith("#"):
            continue
        c, l = line.rstrip("\n").split("\t")
        want.setdefault(c, []).append(l)
    seq = "".join(l.strip() for l in gzip.open(golden(CHR22_3MB + ".gz"), "rt") if not l.startswith(">"))
    # chunks 30 and 31 of chr22's NCBI query chunks, in the fixture's coordinates
    bounds = {"c30": (491_040, 1_507_508), "c31": (1_507_408, 2_523_876)}
    m = rmblast_amd.load_matrix(golden("comparison.matrix"))
    s = rmblast_amd.Searcher(m, gap_open=20, gap_extend=5, word_size=14, xdrop_ungap=400, xdrop_gap=100,
                             xdrop_gap_final=200, min_raw_gapped_score=200, complexity_adjust=True)
    alu = read_2bit(golden("alu.2bit"))
    s.load_library([rmblast_amd.encode(x) for _, x in alu])
    total = 0
    for c, (a, b) in bounds.items():
        q = rmblast_amd.encode(seq[a:b])
        got = []
        for f in range(len(alu)):
            for qo, so, minus in s.tap_seeds(q, f):
                got.append(f"SEED strand={'-' if minus else '+'} q={qo} s={so}")
        assert sorted(got) == want[c], c
        total += len(got)
    assert total == 20860


@pytest.mark.parametrize("ws,gaps", [(7, (20, 5)), (11, (4, 4)), (14, (20, 5))])
def test_tap_prelims_match_oracle_in_order(oracle_lib, ws, gaps):
    w = synth.parity_config(seed=29, genome_len=150_000, n_families=5, n_copies=160)
    m, mat = comparison()
    s = rmblast_amd.Searcher(m, gap_open=gaps[0], gap_extend=gaps[1], word_size=ws, xdrop_ungap=20, xdrop_gap=30,
                             min_raw_gapped_score=60)
    s.load_library([f[1] for f in w.families])
    q = w.records[0][1]
    seen = 0
    for f in range(len(w.families)):
        got = s.tap_prelims(q, f)
        subj = w.families[f][1]
        out = np.zeros((200000, 8), dtype=np.uint32)
        n = oracle_lib.orc_prelims(ptr(q), len(q), ptr(subj), len(subj), ws, 20, 30, gaps[0], gaps[1], 60,
                                   ptr(mat, ctypes.c_int32), ptr(out, ctypes.c_uint32), 200000)
        assert np.array_equal(got, out[:n]), (ws, f, len(got), n)
        seen += n
    assert seen > 20


COMPLEMENT = np.array([3, 2, 1, 0, 5, 4, 7, 6, 8, 9, 13, 12, 11, 10, 14, 15], dtype=np.uint8)


def test_device_library_prep_matches_oracle_generator(oracle_lib):
    """Subjects with every ambiguity code (scattered, in runs, at the ends,
    none at all, and one of 20 kbp mostly N) resolved on the device: the plus
    strand equals the oracle's NcbiRandom seeded with the length, drawn once
    per ambiguous base 5' to 3', and the minus strand is its reverse
    complement."""
    rng = np.random.default_rng(31)
    subs = []
    for k in range(40):
        n = int(rng.integers(1, 3000))
        x = rng.integers(0, 4, n).astype(np.uint8)
        if k % 4:
            x[rng.random(n) < 0.05] = rng.integers(4, 15, int((rng.random(n) < 0.05).sum()) or 1)[0]
            amb = rng.random(n) < 0.03
            x[amb] = rng.integers(4, 15, int(amb.sum()))
        if k % 5 == 0 and n > 10:
            x[:3] = 14
            x[-3:] = 11
        subs.append(x)
    big = rng.integers(4, 15, 20000).astype(np.uint8)
    big[::7] = 2
    subs.append(big)
    m, _ = comparison()
    s = rmblast_amd.Searcher(m, word_size=11)
    s.load_library(subs)
    for i, x in enumerate(subs):
        plus, rc = s.tap_subject(i)
        want = x.copy()
        amb = np.nonzero((x >= 4) & (x <= 14))[0]
        if len(amb):
            out = np.zeros(len(amb), dtype=np.uint8)
            codes = np.ascontiguousarray(x[amb])
            oracle_lib.orc_ncbi_random(len(x), ptr(codes), len(amb), ptr(out))
            want[amb] = out
        assert np.array_equal(plus, want), i
        assert np.array_equal(rc, COMPLEMENT[want[::-1]]), i
