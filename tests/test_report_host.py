"""The library's host tail without a GPU: mask level (rmb_mask_level, the
reference's apply_mask_level, engine.rs:863-974) against the reference's
known answers (engine.rs:3291-3328) and against the oracle on random
cross-subject HSP sets, including ties, touching ranges and both strands."""
import ctypes

import numpy as np
import pytest

import rmblast_amd
from conftest import ptr


def hsps_of(rows):
    """rows: (score, q_start, q_end, s_start, s_end, strand, subject)"""
    h = np.zeros(len(rows), dtype=rmblast_amd.HSP_DTYPE)
    for i, (sc, qs, qe, ss, se, st, sj) in enumerate(rows):
        h[i] = (0, sj, sc, qs, qe, ss, se, st, 0, 0)
    return h


def oracle_keep(L, rows, q_len, n_subj, level):
    L.orc_mask_level_xs.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                    ctypes.c_int, ctypes.POINTER(ctypes.c_uint8)]
    a = np.ascontiguousarray(np.array(rows, dtype=np.uint32).reshape(-1, 7))
    keep = np.zeros(len(a), dtype=np.uint8)
    L.orc_mask_level_xs(ptr(a, ctypes.c_uint32), len(a), q_len, n_subj, level, ptr(keep))
    return keep.astype(bool)


def test_known_answers():
    # engine.rs:3291-3303: 26 HSPs, level 80 keeps 17; of those level 25 keeps
    # the nine 17-base ones
    so = [10, 382, 14, 83, 4, 382, 1000, 203, 54, 32, 64, 382, 89, 183, 813, 132, 14, 1344, 321, 224, 34, 8341, 1344,
          254, 861, 834]
    qo = [1, 4, 14, 21, 24, 34, 41, 44, 54, 61, 64, 74, 81, 84, 94, 101, 104, 114, 121, 124, 134, 141, 144, 154, 161,
          164]
    ln = [17, 9, 9] * 8 + [17, 9]
    sc = [93, 85, 85] * 8 + [93, 85]
    rows = [(sc[i], qo[i], qo[i] + ln[i], so[i], so[i] + ln[i], 0, 0) for i in range(26)]
    k80 = rmblast_amd.mask_level(hsps_of(rows), 10000, 80)
    assert k80.sum() == 17
    r17 = [r for r, k in zip(rows, k80) if k]
    k25 = rmblast_amd.mask_level(hsps_of(r17), 10000, 25)
    assert sorted(r[1] for r, k in zip(r17, k25) if k) == [1, 21, 41, 61, 81, 101, 121, 141, 161]
    # engine.rs:3311-3328: 29 of 50 bases overlapped is 57.99...%, not 58%
    pair = [(100, 0, 50, 0, 50, 0, 0), (90, 21, 71, 21, 71, 0, 0)]
    assert rmblast_amd.mask_level(hsps_of(pair), 1000, 58).sum() == 2
    assert rmblast_amd.mask_level(hsps_of(pair), 1000, 57).sum() == 1
    touch = [(100, 0, 50, 0, 50, 0, 0), (90, 50, 100, 50, 100, 0, 0)]
    assert rmblast_amd.mask_level(hsps_of(touch), 1000, 0).sum() == 1
    assert rmblast_amd.mask_level(hsps_of(touch), 1000, 1).sum() == 2
    # levels at or above 100 (and below 0) keep everything
    assert rmblast_amd.mask_level(hsps_of(pair), 1000, 100).all()


@pytest.mark.parametrize("seed", range(12))
def test_random_sets_match_oracle(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    q_len = 5000
    n_subj = int(rng.integers(1, 6))
    rows = []
    for _ in range(int(rng.integers(5, 400))):
        qs = int(rng.integers(0, q_len - 1))
        qe = min(q_len, qs + int(rng.integers(0, 900)))   # empty ranges included
        ss = int(rng.integers(0, 3000))
        rows.append((int(rng.integers(40, 120)), qs, qe, ss, ss + (qe - qs), int(rng.integers(0, 2)),
                     int(rng.integers(0, n_subj))))
    for k in range(len(rows) // 10):                       # exact duplicates on other subjects/strands
        r = list(rows[int(rng.integers(0, len(rows)))])
        r[5] = int(rng.integers(0, 2))
        r[6] = int(rng.integers(0, n_subj))
        rows.append(tuple(r))
    for level in (0, 1, 29, 57, 58, 80, 99):
        want = oracle_keep(oracle_lib, rows, q_len, n_subj, level)
        got = rmblast_amd.mask_level(hsps_of(rows), q_len, level)
        assert (got == want).all(), (seed, level)
