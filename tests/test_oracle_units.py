"""Known answers that pin the oracle's pieces, ported from the reference's own
unit tests (file:line of each is in the test's comment), plus brute-force
checks where the mathematics fixes the answer."""
import ctypes
import math
import os

import numpy as np
import pytest

from conftest import golden, ptr, seq_from_fasta, simple_matrix

INT_MIN_HALF = -(2 ** 31) // 2
CODE = {c: i for i, c in enumerate("ACGTRYMKWSBDHVN")}
COMP = [3, 2, 1, 0, 5, 4, 7, 6, 8, 9, 13, 12, 11, 10, 14, 15]


def enc(s, wrap=True):
    v = [CODE.get(c, 14) for c in s.upper()]
    return np.array(([15] + v + [15]) if wrap else v, dtype=np.uint8)


def revcomp(a):
    return np.array([COMP[x] for x in a[::-1]], dtype=np.uint8)


MAT54 = simple_matrix(5, -4, other=-32768, sentinel=INT_MIN_HALF)     # ungapped.rs:315-322
BLASTN_1M3 = simple_matrix(1, -3, other=0)                             # engine.rs:3072-3080


def ungapped(L, q, s, qo, so, x, mat=MAT54):
    out = (ctypes.c_int32 * 5)()
    L.orc_extend_ungapped(ptr(q), len(q), ptr(s), len(s), qo, so, ptr(mat, ctypes.c_int32), x, out)
    return tuple(out)


# ---- ungapped.rs:328-428 -------------------------------------------------------
def test_ungapped_known_answers(oracle_lib):
    L = oracle_lib
    assert ungapped(L, enc("ACGTACGT"), enc("ACGTACGT"), 0, 0, 20) == (40, 0, 8, 0, 8)
    assert ungapped(L, enc("AAAAAAAA"), enc("CCCCCCCC"), 0, 0, 20)[0] < 1
    assert ungapped(L, enc("ACGTACGT"), enc("ACGTACGT"), 4, 4, 20)[:3] == (40, 0, 8)
    assert ungapped(L, enc("AAAAAAAA"), enc("AAAAAAAA"), 4, 4, 20)[:3] == (40, 0, 8)


def minus_via_ncbi_frame(L, qs, ss, q_off, s_off, x):
    q, s = enc(qs), enc(ss)
    q_len, s_len = len(q) - 2, len(s) - 2
    qn, sn = q_len - 1 - q_off, s_len - 1 - s_off
    sc, a, b, c, d = ungapped(L, revcomp(q), revcomp(s), qn, sn, x)
    return sc, q_len - b, q_len - a, s_len - d, s_len - c


def test_ungapped_minus_frame_bug36(oracle_lib):
    # ungapped.rs:403-417: NCBI's frame stops at the mismatch stretch
    qs = "GGGGGGGGGGGGGGGGGGGGGGGGACACACACACACACACACACACGTACGTACGT"
    ss = "GGGGGGGGGGGGGGGGGGGGGGGGCACACACACACACACACACAACGTACGTACGT"
    assert minus_via_ncbi_frame(oracle_lib, qs, ss, 44, 44, 200) == (60, 44, 56, 44, 56)
    r = minus_via_ncbi_frame(oracle_lib, "ACGTACGTACGTACGT", "ACGTACGTACGTACGT", 8, 8, 30)
    assert r[1:] == (0, 16, 0, 16) and r[0] == 80


# ---- diag_hash.rs:127-184 --------------------------------------------------------
def hash_run(L, ops, window):
    a = np.ascontiguousarray(np.array(ops, dtype=np.int32).reshape(-1, 4))
    out = np.zeros(len(a), dtype=np.uint32)
    L.orc_diag_hash_run(ptr(a, ctypes.c_int32), len(a), window, ptr(out, ctypes.c_uint32))
    return out


def test_diag_hash_is_a_map_when_nothing_is_stale(oracle_lib):
    ops, want = [], {}
    x = 0x9E3779B97F4A7C15
    M = (1 << 64) - 1
    for i in range(20000):
        x ^= (x << 13) & M
        x ^= x >> 7
        x ^= (x << 17) & M
        d = x % 3000 - 1500
        ops.append((1, d, 1 + i, 0))
        want[d] = 1 + i
    gets = list(range(-1600, 1600))
    ops += [(0, d, 0, 0) for d in gets] + [(2, 0, 0, 0)] + [(0, d, 0, 0) for d in gets]
    ops += [(1, 7, 42, 0), (0, 7, 0, 0)]
    out = hash_run(oracle_lib, ops, 2 ** 31 - 1)
    got = out[20000:20000 + len(gets)]
    assert list(got) == [want.get(d, 0) for d in gets]
    assert not out[20000 + len(gets) + 1:20000 + 2 * len(gets) + 1].any()
    assert out[-1] == 42


def test_diag_hash_stale_rule_boundaries(oracle_lib):
    d1 = 5
    d2 = next(d for d in range(6, 5000) if ((d * 0x9E370001) & 0xFFFFFFFF) % 512 == ((d1 * 0x9E370001) % 512))
    w = 1 - 14
    L = oracle_lib
    get2 = [(0, d1, 0, 0), (0, d2, 0, 0)]
    assert list(hash_run(L, [(1, d1, 100, 50), (1, d2, 300, 80)] + get2, w)[-2:]) == [100, 300]
    assert list(hash_run(L, [(1, d1, 100, 50), (1, d2, 300, 200)] + get2, w)[-2:]) == [0, 300]
    assert list(hash_run(L, [(1, d1, 100, 50), (1, d2, 300, 87)] + get2, w)[-2:]) == [100, 300]
    assert hash_run(L, [(1, d1, 100, 50), (1, d1, 500, 10000), (0, d1, 0, 0)], w)[-1] == 500


# ---- seq/twobit.rs:584-600 ---------------------------------------------------------
def test_ncbi_random_known_draws(oracle_lib):
    out = np.zeros(3, dtype=np.uint8)
    oracle_lib.orc_ncbi_random(23, ptr(np.array([14, 14, 14], dtype=np.uint8)), 3, ptr(out))
    assert list(out) == [3, 0, 2]
    oracle_lib.orc_ncbi_random(428, ptr(np.array([14], dtype=np.uint8)), 1, ptr(out))
    assert out[0] == 1


# ---- complexity_adjust.rs:159-207 --------------------------------------------------
CA_MAT = """# FREQS A 0.255 C 0.245 G 0.245 T 0.255
   A   C   G   T
A  9  -6  -5 -13
C -6   9 -13  -5
G -5 -13   9  -6
T -13  -5  -6   9
"""


def test_complexity_adjust_closed_forms(oracle_lib, tmp_path):
    p = tmp_path / "ca.matrix"
    p.write_text(CA_MAT)
    L = oracle_lib
    scores = np.zeros(256, dtype=np.int32)
    freqs = np.zeros(16)
    lam = ctypes.c_double()
    L.orc_matrix_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]
    assert L.orc_matrix_load(str(p).encode(), ptr(scores, ctypes.c_int32), ptr(freqs, ctypes.c_double),
                             ctypes.byref(lam)) == 0
    lam = lam.value

    def adj(raw, bases):
        b = np.array(bases, dtype=np.uint8)
        out = ctypes.c_double()
        assert L.orc_adjusted_score(str(p).encode(), raw, ptr(b), len(b), ctypes.byref(out)) == 0
        return out.value

    bal = [0, 1, 2, 3] * 16
    t_sum = 32.0 * (math.log(0.255) + math.log(0.245)) - 64.0 * math.log(0.25)
    assert abs(adj(100, bal) - (100 + t_sum / lam + 0.999)) < 1e-9
    assert abs(adj(100, bal) - 100.8808) < 1e-3
    want = 1000.0 + 64.0 * math.log(0.255) / lam + 0.999
    assert abs(adj(1000, [0] * 64) - want) < 1e-9
    assert adj(200, [0] * 64) == 0.0
    a, c, t = 30.0, 10.0, 20.0
    n = a + c + t
    tf = a * math.log(a) + c * math.log(c) + t * math.log(t) - n * math.log(n)
    ts = a * math.log(0.255) + c * math.log(0.245) + t * math.log(0.255) - tf
    assert abs(adj(500, [0] * 30 + [1] * 10 + [3] * 20 + [14] * 5) - (500 + ts / lam + 0.999)) < 1e-9
    assert math.isnan(adj(500, [14] * 8))


# ---- gapped.rs:869-944: X-drop DP = Gotoh optimum when X cannot prune ---------------
def gotoh_anchored(a, b, mat, go, ge):
    NEG = -(1 << 60)
    la, lb = len(a), len(b)
    H = np.full((la + 1, lb + 1), NEG, dtype=np.int64)
    GX = H.copy()
    GY = H.copy()
    H[0, 0] = 0
    best = 0
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            if i:
                GX[i, j] = max(H[i - 1, j] - go - ge, GX[i - 1, j] - ge)
            if j:
                GY[i, j] = max(H[i, j - 1] - go - ge, GY[i, j - 1] - ge)
            d = H[i - 1, j - 1] + mat[a[i - 1], b[j - 1]] if i and j else NEG
            H[i, j] = max(d, GX[i, j], GY[i, j])
            best = max(best, H[i, j])
    return int(best)


def rescore(ops, a, b, mat, go, ge):
    i = j = sc = 0
    prev = None
    for k, n in ops:
        for _ in range(n):
            if k == 0:
                sc += mat[a[i], b[j]]; i += 1; j += 1
            elif k == 2:
                sc -= ge if prev == 2 else go + ge; i += 1
            else:
                sc -= ge if prev == 1 else go + ge; j += 1
            prev = k
    return sc, i, j


def test_xdrop_dp_equals_gotoh_without_pruning(oracle_lib):
    L = oracle_lib
    mat = MAT54
    rng = np.random.default_rng(7)
    for trial in range(300):
        a = rng.integers(0, 4, 1 + rng.integers(0, 40)).astype(np.uint8)
        b = []
        for c in a:
            r = rng.integers(0, 12)
            if r == 1:
                b += [c, rng.integers(0, 4)]
            elif r == 2:
                b.append(rng.integers(0, 4))
            elif r != 0:
                b.append(c)
        b = np.array(b or [0], dtype=np.uint8)
        go, ge = [(4, 4), (20, 5), (8, 2), (1, 1)][trial % 4]
        for rev in (0, 1):
            aa, bb = (a[::-1].copy(), b[::-1].copy()) if rev else (a, b)
            want = gotoh_anchored(aa, bb, mat, go, ge)
            out = (ctypes.c_int64 * 4)()
            ops = np.zeros(256, dtype=np.uint32)
            n = L.orc_align_ex(ptr(a), ptr(b), len(a), len(b), go, ge, 10 ** 6, ptr(mat, ctypes.c_int32), rev, out,
                               ptr(ops, ctypes.c_uint32), 128)
            assert out[0] == want, (trial, rev)
            if not rev:
                pairs = [(int(ops[2 * k]), int(ops[2 * k + 1])) for k in range(n)]
                assert rescore(pairs, a, b, mat, go, ge) == (want, out[1], out[2])
            so = (ctypes.c_int64 * 3)()
            L.orc_align_score_only(ptr(a), ptr(b), len(a), len(b), go, ge, 10 ** 6, ptr(mat, ctypes.c_int32), rev,
                                   so)
            assert so[0] == want


# ---- engine.rs:3431-3490: NCBI testNuclWordFinder (word size 19, stride 12) --------
def test_word_finder_five_hits(oracle_lib):
    q = enc(seq_from_fasta(golden("gi3090_xylanase_query.fa")), wrap=False)
    s = enc(seq_from_fasta(golden("gi33383640_xylanase_subject.fa")), wrap=False)
    out = np.zeros((4096, 8), dtype=np.uint32)
    n = oracle_lib.orc_ungapped_hits(ptr(q), len(q), ptr(s), len(s), 19, 11, 28, ptr(BLASTN_1M3, ctypes.c_int32),
                                     ptr(out, ctypes.c_uint32), 4096)
    plus = sorted((int(h[2]), int(h[4]), int(h[3] - h[2]), int(h[6])) for h in out[:n] if h[7] == 0)
    assert plus == sorted([(233, 0, 31, 31), (1037, 66, 945, 853), (263, 27, 685, 541), (1911, 811, 101, 73),
                           (1782, 940, 63, 51)])


# ---- engine.rs:3512-3549: NCBI testGapAlignment seed 7 ----------------------------
def test_gapped_extension_seed7(oracle_lib):
    q = enc(seq_from_fasta(golden("gi2655203_unp_20001_35000.fa")), wrap=False)
    s = enc(seq_from_fasta(golden("gi2516238_rab33b.fa")), wrap=False)
    o2 = np.zeros(2, dtype=np.uint32)
    oracle_lib.orc_improve_seed(ptr(q), len(q), ptr(s), len(s),
                                ptr(np.array([3212, 3640, 3201, 3221, 3629, 3649], dtype=np.uint32), ctypes.c_uint32),
                                0, ptr(o2, ctypes.c_uint32))
    out = (ctypes.c_int64 * 6)()
    oracle_lib.orc_gapped_bidir(ptr(q), len(q), ptr(s), len(s), int(o2[0]), int(o2[1]), 5, 2, 30,
                                ptr(BLASTN_1M3, ctypes.c_int32), out)
    assert out[5] == 1
    assert tuple(out[1:5]) == (3201, 3221, 3629, 3649)


# ---- engine.rs:3213-3234: NCBI testCheckHSPCommonEndpoints -------------------------
def test_purge_prelim_common_endpoints(oracle_lib):
    scores = [1044, 995, 965, 219, 160, 125, 110, 107, 103]
    qs = [2, 2, 2, 236, 88, 259, 278, 259, 278]
    qe = [322, 336, 300, 322, 182, 322, 341, 341, 341]
    ss = [7, 7, 7, 194, 2, 194, 197, 194, 197]
    se = [292, 293, 301, 292, 96, 292, 260, 260, 266]
    h = np.ascontiguousarray(np.array([scores, qs, qe, ss, se, [0] * 9], dtype=np.uint32).T)
    keep = np.zeros(9, dtype=np.uint8)
    assert oracle_lib.orc_purge_prelims(ptr(h, ctypes.c_uint32), 9, 400, ptr(keep)) == 3
    assert sorted(np.array(scores)[keep.astype(bool)], reverse=True) == [1044, 160, 110]


# ---- engine.rs:3270-3328: NCBI testHSPResultsApplyMasklevel + 57/58 and touching ----
def mask(L, hsps, level, q_len=10000):
    h = np.ascontiguousarray(np.array(hsps, dtype=np.uint32).reshape(-1, 6))
    keep = np.zeros(len(h), dtype=np.uint8)
    L.orc_mask_level(ptr(h, ctypes.c_uint32), len(h), q_len, level, ptr(keep))
    return keep.astype(bool)


def test_mask_level_known_answers(oracle_lib):
    L = oracle_lib
    so = [10, 382, 14, 83, 4, 382, 1000, 203, 54, 32, 64, 382, 89, 183, 813, 132, 14, 1344, 321, 224, 34, 8341, 1344,
          254, 861, 834]
    qo = [1, 4, 14, 21, 24, 34, 41, 44, 54, 61, 64, 74, 81, 84, 94, 101, 104, 114, 121, 124, 134, 141, 144, 154, 161,
          164]
    ln = [17, 9, 9, 17, 9, 9, 17, 9, 9, 17, 9, 9, 17, 9, 9, 17, 9, 9, 17, 9, 9, 17, 9, 9, 17, 9]
    sc = [93, 85, 85, 93, 85, 85, 93, 85, 85, 93, 85, 85, 93, 85, 85, 93, 85, 85, 93, 85, 85, 93, 85, 85, 93, 85]
    hs = [(sc[i], qo[i], qo[i] + ln[i], so[i], so[i] + ln[i], 0) for i in range(26)]
    k80 = mask(L, hs, 80)
    assert k80.sum() == 17
    hs17 = [h for h, k in zip(hs, k80) if k]
    k25 = mask(L, hs17, 25)
    assert sorted(h[1] for h, k in zip(hs17, k25) if k) == [1, 21, 41, 61, 81, 101, 121, 141, 161]
    pair = [(100, 0, 50, 0, 50, 0), (90, 21, 71, 21, 71, 0)]
    assert mask(L, pair, 58).sum() == 2 and mask(L, pair, 57).sum() == 1
    touch = [(100, 0, 50, 0, 50, 0), (90, 50, 100, 50, 100, 0)]
    assert mask(L, touch, 0).sum() == 1 and mask(L, touch, 1).sum() == 2
