"""The oracle's split-query path (queries of 1,999,800 bases or more).

Pinned by: NCBI's chunk plan worked by hand and its covering invariants
(rmblastn/src/main.rs:1150-1193), the cross-chunk prelim merge against the
reference's own unit cases (main.rs:1472-1551, restated here), and the
NCBI RMBlast 2.17.1 output for chr22:30,000,000-33,000,000 at word size 7
(rmblastn/tests/cli.rs:244-274), which runs three chunks.
"""
import ctypes

import numpy as np
import pytest

from conftest import CHR22_3MB, CHUNKED_SEARCH, ORACLE_BIN, golden, golden_unzipped, ptr, run


def plan(L, n):
    out = np.zeros(3, dtype=np.uint64)
    L.orc_chunk_plan(n, ptr(out, ctypes.c_uint64))
    return tuple(int(x) for x in out)


def chunks(L, n):
    size, count, step = plan(L, n)
    res = []
    for k in range(count):
        s = k * step
        if s >= n:
            break
        res.append((s, min(s + size, n)))
    return res


@pytest.mark.parametrize("n,want", [
    (1_000_000, (1_000_000, 1)),
    (1_999_799, (1_999_799, 1)),            # 1,999,799 // 999,900 = 1: one chunk
    (1_999_800, (999_951, 2)),              # (1,999,800 + 100) // 2 = 999,950, +1 as 2 < 999,850
    (3_000_000, (1_000_067, 3)),            # (3,000,000 + 200) // 3 = 1,000,066, +1
    (50_818_468, (1_016_468, 50)),          # chr22: 50 chunks, (50,818,468 + 4,900) // 50 = 1,016,467, +1
])
def test_chunk_plan_known_answers(oracle_lib, n, want):
    size, count, step = plan(oracle_lib, n)
    assert (size, count) == want
    assert step == size - 100


@pytest.mark.parametrize("seed", range(4))
def test_chunks_cover_the_query_with_overlap_100(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    for n in rng.integers(1_999_800, 40_000_000, size=50):
        c = chunks(oracle_lib, int(n))
        assert len(c) >= 2
        assert c[0][0] == 0 and c[-1][1] == n
        for (a0, a1), (b0, b1) in zip(c, c[1:]):
            assert a1 - b0 == 100
        sizes = {b - a for a, b in c[:-1]}
        assert len(sizes) == 1


PLUS, MINUS = 0, 1


def hsp(strand, score, q, s, seed):
    # {q_seed, s_seed, q_start, q_end, s_start, s_end, score, strand}
    return [seed[0], seed[1], q[0], q[1], s[0], s[1], score, strand]


def merge(L, acc, new, split):
    a = np.array(acc, dtype=np.uint32).reshape(-1, 8)
    b = np.array(new, dtype=np.uint32).reshape(-1, 8)
    out = np.zeros((len(a) + len(b), 8), dtype=np.uint32)
    n = L.orc_merge_chunk_prelims(ptr(a, ctypes.c_uint32), len(a), ptr(b, ctypes.c_uint32), len(b), split,
                                  ptr(out, ctypes.c_uint32))
    return [tuple(int(v) for v in r) for r in out[:n]]


def pieces(strand):
    # one alignment cut at 1000, the two pieces overlapping by 50
    return (hsp(strand, 300, (800, 1050), (100, 350), (900, 200)),
            hsp(strand, 200, (1000, 1200), (300, 500), (1100, 400)))


def test_merged_score_is_density_times_merged_length(oracle_lib):
    a, b = pieces(PLUS)
    m = merge(oracle_lib, [a], [b], 1000)
    assert len(m) == 1
    assert m[0][2:6] == (800, 1200, 100, 500)
    # (300 + 200) / (250 + 200) * 400 = 444.4, truncated
    assert m[0][6] == 444
    assert m[0][:2] == (900, 200)


def test_higher_scoring_new_piece_supplies_the_seed(oracle_lib):
    a, b = pieces(PLUS)
    a[6], b[6] = 200, 300
    m = merge(oracle_lib, [a], [b], 1000)
    assert m[0][:2] == (1100, 400)
    assert m[0][6] == 444


def test_growth_is_seen_by_later_candidates(oracle_lib):
    a, b = pieces(PLUS)
    c = hsp(PLUS, 100, (1090, 1300), (390, 600), (1200, 500))   # inside a only after a has grown by b
    m = merge(oracle_lib, [a], [c, b], 1000)
    assert len(m) == 1 and m[0][3] == 1300


def test_best_accumulated_piece_takes_the_merge(oracle_lib):
    a, b = pieces(PLUS)
    weaker = hsp(PLUS, 250, (810, 1040), (110, 340), (850, 150))
    m = merge(oracle_lib, [weaker, a], [b], 1000)
    assert len(m) == 2
    assert m[0][3] == 1040 and m[1][3] == 1200


def test_minus_strand_and_mixed_strands_never_merge(oracle_lib):
    a, b = pieces(MINUS)
    assert len(merge(oracle_lib, [a], [b], 1000)) == 2
    a, _ = pieces(PLUS)
    assert len(merge(oracle_lib, [a], [b], 1000)) == 2


def test_far_diagonal_or_outside_the_split_window_does_not_merge(oracle_lib):
    a, b = pieces(PLUS)
    far = list(b)
    far[4] -= 10
    far[5] -= 10          # start diagonal 10 away from a's end diagonal
    assert len(merge(oracle_lib, [a], [far], 1000)) == 2
    near = list(b)
    near[4] -= 9
    near[5] -= 9
    assert len(merge(oracle_lib, [a], [near], 1000)) == 1
    # accumulated pieces ending at or before the split are not candidates
    assert len(merge(oracle_lib, [a], [b], 1050)) == 2


def test_unabsorbed_new_pieces_follow_in_their_order(oracle_lib):
    a, b = pieces(PLUS)
    x = hsp(PLUS, 50, (5000, 5100), (900, 1000), (5050, 950))
    y = hsp(PLUS, 70, (6000, 6100), (900, 1000), (6050, 950))
    m = merge(oracle_lib, [a], [x, b, y], 1000)
    assert [r[0] for r in m] == [900, 5050, 6050]


def test_chr22_3mb_ws7_matches_ncbi(tmp_path):
    # three chunks: [0, 1,000,067), [999,967, 2,000,034), [1,999,934, 3,000,000)
    q = golden_unzipped(CHR22_3MB, tmp_path)
    r = run([ORACLE_BIN, "--db", golden("db.fa"), "--query", q, *CHUNKED_SEARCH, "--num-threads", "8"])
    assert r.returncode == 0, r.stderr
    want = open(golden("expected_chr22_30M-33M_ws7.ncbi-2.17.1.tsv")).read()
    assert r.stdout.count("\n") == 4236
    assert r.stdout == want
