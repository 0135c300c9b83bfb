"""E-value statistics (rmblast-lib/src/ka_stats.rs, rmblast-stats/rmstats) in
the library's C ABI (host code, no GPU call) and in the oracle.

Pinned by: the reference's own known answers (ka_stats.rs:398-597,
rmstats/tests/stats_tests.rs:84-362, output.rs:768-790), the ALP fit summary
the reference ships (tests/golden/alp_fits_summary.tsv: every table entry's
gap pair, lambda, K and H), and the length adjustment's defining inequality
(BLAST_ComputeLengthAdjustment returns floor of the fixed point of
ell = alpha/lambda * ln(K (m - ell)(n - N ell)) + beta), checked by brute
force on random geometries."""
import csv
import ctypes
import math

import numpy as np
import pytest

import rmblast_amd
from conftest import ORACLE_LIB, golden

COMPARISON_LIKE = """# FREQS A 0.265 C 0.235 G 0.235 T 0.265
   A   C   G   T   N
A  9 -15  -6 -17  -1
C -15  10 -15  -6  -1
G  -6 -15  10 -15  -1
T -17  -6 -15   9  -1
N  -1  -1  -1  -1  -1
"""


@pytest.fixture(scope="module")
def orc():
    L = ctypes.CDLL(ORACLE_LIB)
    L.orc_ka_stats.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                               ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.orc_length_adjustment.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int64, ctypes.c_int]
    L.orc_format_scores.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p]
    return L


def matrix_file(tmp_path, name, text=COMPARISON_LIKE):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def both(orc, path, go, ge, cli=(0.0, 0.0, 0.0, 0.0), compat=False):
    """(source, lambda, k, log_k, h, alpha, beta) from the library and the oracle, which must agree."""
    s = rmblast_amd.ka_stats(path, go, ge, *cli, ncbi_compat=compat)
    got = (rmblast_amd.KA_SOURCES[s.source], s.lambda_, s.k, s.log_k, s.h, s.alpha, s.beta)
    out = (ctypes.c_double * 6)()
    src = orc.orc_ka_stats(path.encode(), go, ge, (ctypes.c_double * 4)(*cli), int(compat), out)
    assert (rmblast_amd.KA_SOURCES[src],) + tuple(out) == got
    return got, s


def test_table_matches_the_alp_fits(tmp_path, orc):
    rows = list(csv.DictReader(open(golden("alp_fits_summary.tsv")), delimiter="\t"))
    assert len(rows) == 42
    for r in rows:
        path = matrix_file(tmp_path, r["matrix"] + ".matrix")
        go, ge = int(r["gap_open"]), int(r["gap_extend"])
        (src, lam, k, logk, h, alpha, beta), _ = both(orc, path, go, ge)
        assert src == "table", r["matrix"]
        for got, want in ((lam, r["lambda"]), (k, r["K"]), (h, r["H"])):
            assert abs(got - float(want)) <= 5e-10, (r["matrix"], got, want)  # NCBI rounds H in its last digit
        assert logk == math.log(k) and alpha == lam / h and beta == 0.0
        assert both(orc, path, go + 1, ge)[0][0] == "comment" or both(orc, path, go + 1, ge)[0][0] == "sentinel"


def test_baked_table_hit_wins(tmp_path, orc):
    # ka_stats.rs:417-433: 20p41g at 25/5 from the table in both modes; other
    # gap costs miss the table (no ALP fit here: sentinel)
    path = matrix_file(tmp_path, "20p41g.matrix")
    for compat in (False, True):
        assert both(orc, path, 25, 5, compat=compat)[0][0] == "table"
    assert both(orc, path, 11, 1)[0][0] == "sentinel"
    assert both(orc, path, 11, 1, compat=True)[0][0] == "sentinel"


def test_cli_overrides_used_on_table_miss(tmp_path, orc):
    # ka_stats.rs:436-452
    path = matrix_file(tmp_path, "nosuch.matrix")
    for compat in (False, True):
        (src, lam, k, logk, h, alpha, beta), _ = both(orc, path, 20, 5, (0.1, 0.02, 0.3, -25.0), compat)
        assert src == "cli" and abs(lam - 0.1) < 1e-15 and abs(h - 0.1 / 0.3) < 1e-15
        assert (alpha, beta) == (0.3, -25.0)


def test_karlin_comment_used_when_table_and_cli_miss(tmp_path, orc):
    # ka_stats.rs:455-474: alpha derived from H, then H = lambda / alpha recovers it
    path = matrix_file(tmp_path, "nosuch.matrix", "# KARLIN lambda 0.1276 k 0.0179 h 0.3562\n" + COMPARISON_LIKE)
    (src, lam, k, logk, h, alpha, beta), s = both(orc, path, 20, 5)
    assert src == "comment" and abs(lam - 0.1276) < 1e-15 and abs(h - 0.3562) < 1e-12
    sp, la = rmblast_amd.ka_search_space(s, 1_000_000, 50_000, 25)
    e, b = rmblast_amd.ka_evalue(s, sp, 300)
    assert e > 0 and b > 0
    assert both(orc, path, 20, 5, compat=True)[0][0] == "sentinel"
    # matrix.rs:409-428: an unknown key is a parse error of the matrix
    bad = matrix_file(tmp_path, "bad.matrix", "# KARLIN lambdo 0.1\n" + COMPARISON_LIKE)
    with pytest.raises(rmblast_amd.RmbError):
        rmblast_amd.load_matrix(bad)
    with pytest.raises(rmblast_amd.RmbError):
        rmblast_amd.ka_stats(bad, 20, 5)


def test_sentinel_when_no_source(tmp_path, orc):
    # ka_stats.rs:477-490 and stats_tests.rs:136-145
    (src, lam, k, logk, h, _, _), s = both(orc, matrix_file(tmp_path, "nosuch.matrix"), 20, 5)
    assert src == "sentinel" and (lam, k, logk, h) == (-1.0, -1.0, -1.0, -1.0)
    assert rmblast_amd.ka_search_space(s, 1000, 1000, 1) == (0, 0)
    assert rmblast_amd.ka_evalue(s, 0, 300) == (1.0, 0.0)


def test_comparison_matrix_modes(tmp_path, orc):
    # ka_stats.rs:493-523
    path = matrix_file(tmp_path, "comparison.matrix")
    (src, lam, *_), _ = both(orc, path, 20, 5)
    assert src == "table" and abs(lam - 0.09836223) < 1e-12
    assert both(orc, path, 20, 5, compat=True)[0][0] == "sentinel"
    cli = (0.09836223, 0.0789918084, 0.881279665, -26.8400935093)
    assert both(orc, path, 20, 5, cli, compat=True)[0][0] == "cli"


def test_30p53g_placeholder_in_compat_mode(tmp_path, orc):
    # ka_stats.rs:526-546
    path = matrix_file(tmp_path, "30p53g.matrix")
    (src, lam, k, logk, h, alpha, beta), _ = both(orc, path, 22, 5, compat=True)
    assert src == "table" and (lam, k, h, alpha, beta) == (0.1, 0.1, 0.1, 1.0, 0.0)
    assert abs(both(orc, path, 22, 5)[0][1] - 0.1204393447) < 1e-12


def test_table_14p35g_and_mode2(tmp_path, orc):
    # stats_tests.rs:84-134
    (_, lam, k, logk, h, _, _), _ = both(orc, matrix_file(tmp_path, "14P35G.MATRIX"), 29, 6)
    assert (lam, k, h, logk) == (0.1241860392, 0.2565118357, 0.3349210581, math.log(0.2565118357))
    (src, lam, k, logk, h, _, _), _ = both(orc, matrix_file(tmp_path, "x.matrix"), 20, 5, (0.1227, 0.0919, 1.5, 0))
    assert src == "cli" and (lam, k, logk, h) == (0.1227, 0.0919, math.log(0.0919), 0.1227 / 1.5)


def fixed_point(alpha_d_lambda, beta, logk, m, n, nseq):
    return lambda ell: alpha_d_lambda * (logk + math.log((m - ell) * (n - nseq * ell))) + beta


def test_length_adjustment_is_the_floor_of_the_fixed_point(tmp_path, orc):
    # stats_tests.rs:226-258: 25p41g, a chr20-sized query against a 1.8 Mbp library
    _, s = both(orc, matrix_file(tmp_path, "25p41g.matrix"), 22, 5)
    m, n, nseq = 64_444_167, 1_838_298, 1_580
    sp, la = rmblast_amd.ka_search_space(s, m, n, nseq)
    f = fixed_point(s.alpha / s.lambda_, s.beta, s.log_k, m, n, nseq)
    assert la > 0 and f(la) >= la and f(la + 1) < la + 1 + 1.0
    assert sp == (n - nseq * la) * (m - la)
    assert orc.orc_length_adjustment(s.k, s.log_k, s.alpha / s.lambda_, s.beta, m, n, nseq) == la
    # degenerate: c < 0 -> 0
    assert orc.orc_length_adjustment(1e-9, math.log(1e-9), 2.0, 0.0, 10, 10, 1) == 0
    # random geometries: the library and the oracle agree, and where the
    # iteration converges the result is the fixed point's floor
    rng = np.random.default_rng(3)
    for _ in range(300):
        m = int(rng.integers(100, 5_000_000))
        n = int(rng.integers(1_000, 50_000_000))
        nseq = int(rng.integers(1, 20_000))
        _, la = rmblast_amd.ka_search_space(s, m, n, nseq)
        assert orc.orc_length_adjustment(s.k, s.log_k, s.alpha / s.lambda_, s.beta, m, n, nseq) == la
        f = fixed_point(s.alpha / s.lambda_, s.beta, s.log_k, m, n, nseq)
        if 0 < la and (m - la - 1) > 0 and (n - nseq * (la + 1)) > 0:
            assert f(la) >= la - 1e-9


def test_rmstats_end_to_end(tmp_path, orc):
    # stats_tests.rs:296-338
    _, s = both(orc, matrix_file(tmp_path, "20p41g.matrix"), 25, 5)
    sp, la = rmblast_amd.ka_search_space(s, 1_000_000, 1_838_298, 1_580)
    assert la > 0 and sp == (1_838_298 - 1_580 * la) * (1_000_000 - la)
    e, b = rmblast_amd.ka_evalue(s, sp, 250)
    assert e == sp * math.exp(-s.lambda_ * 250 + s.log_k)
    assert abs(b - (250 * s.lambda_ - s.log_k) / math.log(2)) < 1e-12
    _, s2 = both(orc, matrix_file(tmp_path, "comparison.matrix"), 20, 6)
    assert rmblast_amd.ka_search_space(s2, 1_000_000, 1_838_298, 1_580) == (0, 0)
    assert rmblast_amd.ka_evalue(s2, 0, 250) == (1.0, 0.0)


@pytest.mark.parametrize("value,text", [(1.0e-200, "0.0"), (2.5e-105, "2.50e-105"), (1.234e-30, "1.23e-30"),
                                        (5.0e-5, "5.00e-05"), (8.9e-4, "8.90e-04"), (0.0009, "0.001"),
                                        (0.05, "0.050"), (0.55, "0.55"), (1.0, "1.0"), (3.7, "3.7"),
                                        (15.0, "15"), (2000.0, "2000")])
def test_evalue_format_known_answers(orc, value, text):
    # output.rs:768-781
    e, b = ctypes.create_string_buffer(64), ctypes.create_string_buffer(64)
    orc.orc_format_scores(value, 0.0, e, b)
    assert e.value.decode() == text


@pytest.mark.parametrize("value,text", [(0.0, " 0.0"), (7.25, " 7.2"), (45.67, "45.7"), (99.9, "99.9"),
                                        (100.4, "100"), (12345.9, "12345"), (123456.0, "1.235e+05")])
def test_bit_score_format_known_answers(orc, value, text):
    # output.rs:784-790 (and the truncating and scientific arms, output.rs:279-288)
    e, b = ctypes.create_string_buffer(64), ctypes.create_string_buffer(64)
    orc.orc_format_scores(1.0, value, e, b)
    assert b.value.decode() == text
