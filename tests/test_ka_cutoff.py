"""Karlin-Altschul ungapped cutoff (rmblast-lib/src/search/ka_cutoff.rs) in the
library's C ABI (host code, no GPU call) and in the oracle.

Pinned by the reference's own known answers (ka_cutoff.rs:159-191), by the
ALP fit summary the reference ships (rmblast-stats/fits/SUMMARY.tsv, copied to
tests/golden/alp_fits_summary.tsv: each table entry's gap pair, lambda and K)
and by the edge cases the formula fixes (2 * query length below the average
subject length, empty query, saturating search space)."""
import csv
import ctypes
import math

import pytest

import rmblast_amd
from conftest import ORACLE_LIB, golden


@pytest.fixture(scope="module")
def orc():
    L = ctypes.CDLL(ORACLE_LIB)
    L.orc_ungapped_cutoff.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)]
    L.orc_ungapped_cutoff.restype = ctypes.c_int
    return L


def oracle_cutoff(L, name, go, ge, min_raw, qlen, avg):
    out = (ctypes.c_int * 2)()
    sp = ctypes.c_uint64(0)
    c = L.orc_ungapped_cutoff(name.encode(), go, ge, min_raw, qlen, avg, out, ctypes.byref(sp))
    return c, (out[1], sp.value) if out[0] else None


def both_cutoffs(L, *args):
    lib_c, info = rmblast_amd.ungapped_cutoff(*args)
    orc_c, oinfo = oracle_cutoff(L, *args)
    assert lib_c == orc_c
    assert (info is None) == (oinfo is None)
    if info is not None:
        assert (info["es"], info["searchsp"]) == oinfo
    return lib_c, info


def test_lookup_matches_basename_and_gaps_case_insensitively():
    # ka_cutoff.rs:159-168
    want = (0.1179924113, 0.1957371364)
    assert rmblast_amd.ka_lookup("18p43g.matrix", 28, 5) == want
    assert rmblast_amd.ka_lookup("/some/path/18p43g.matrix", 28, 5) == want
    assert rmblast_amd.ka_lookup("18P43G.MATRIX", 28, 5) == want
    assert rmblast_amd.ka_lookup("18p43g.matrix", 20, 5) is None
    assert rmblast_amd.ka_lookup("comparison.matrix", 28, 5) is None


def test_cutoff_falls_back_without_ka(orc):
    # ka_cutoff.rs:170-176
    c, info = both_cutoffs(orc, "comparison.matrix", 20, 5, 200, 1_000_000, 500)
    assert c == 100 and info is None
    # cutoff_score_max is 1 when min_raw_gapped_score is 0 (options.rs:80-82)
    assert both_cutoffs(orc, "comparison.matrix", 20, 5, 0, 1000, 500)[0] == 1


def test_cutoff_ka_based_18p43g_below_fixed(orc):
    # ka_cutoff.rs:178-191, the bug #35 case: 18p43g {28,5}, min_raw 300,
    # query 49,802 bp, average subject 919 bp -> searchsp = 919 * 919
    c, info = both_cutoffs(orc, "18p43g.matrix", 28, 5, 300, 49802, 919)
    assert info["searchsp"] == 919 * 919
    assert info["es"] == 128          # ceil(ln(0.1957371364 * 844561 / 0.05) / 0.1179924113) = ceil(127.06)
    assert c == 128 and c < 150 and c <= 140


def test_short_query_limits_the_search_space(orc):
    # 2 * q_len below the average subject length: searchsp = 2 q_len * avg
    c, info = both_cutoffs(orc, "18p43g.matrix", 28, 5, 300, 100, 919)
    assert info["searchsp"] == 200 * 919
    assert info["es"] == math.ceil(math.log(0.1957371364 * 200 * 919 / 0.05) / 0.1179924113)
    assert c == info["es"] < 128


def test_degenerate_inputs(orc):
    # empty query: searchsp 0, ln(0) = -inf, es saturates to i32::MIN, cutoff 1
    c, info = both_cutoffs(orc, "18p43g.matrix", 28, 5, 300, 0, 919)
    assert info["searchsp"] == 0 and info["es"] == -2 ** 31 and c == 1
    # no subjects: the KA branch does not apply (ka_cutoff.rs:137)
    c, info = both_cutoffs(orc, "18p43g.matrix", 28, 5, 300, 1000, 0)
    assert info is None and c == 150
    # the search space saturates at u64::MAX
    c, info = both_cutoffs(orc, "18p43g.matrix", 28, 5, 10_000, 2 ** 62, 2 ** 40)
    assert info["searchsp"] == 2 ** 64 - 1 and c == info["es"]


def test_table_matches_the_alp_fits(orc):
    """Every table entry equals the reference's ALP fit (gap pair, lambda, K),
    except 30p53g, which carries NCBI's placeholder; comparison.matrix is not
    in NCBI's table."""
    rows = list(csv.DictReader(open(golden("alp_fits_summary.tsv")), delimiter="\t"))
    assert len(rows) == 42
    for r in rows:
        name, go, ge = r["matrix"] + ".matrix", int(r["gap_open"]), int(r["gap_extend"])
        got = rmblast_amd.ka_lookup(name, go, ge)
        if r["matrix"] == "comparison":
            assert got is None
            continue
        want = (0.1, 0.1) if r["matrix"] == "30p53g" else (float(r["lambda"]), float(r["K"]))
        assert got == want, name
        assert rmblast_amd.ka_lookup(name, go + 1, ge) is None
        # the oracle's table gives the same cutoff on a query where KA applies
        lib_c, info = rmblast_amd.ungapped_cutoff(name, go, ge, 400, 5000, 1000)
        assert info is not None and (lib_c, (info["es"], info["searchsp"])) == oracle_cutoff(
            orc, name, go, ge, 400, 5000, 1000)
