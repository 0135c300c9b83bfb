"""The CPU oracle against the reference's NCBI RMBlast 2.17.1 goldens.

rmblastn/tests/cli.rs:107-300 produced expected_*.ncbi-2.17.1.tsv with real
NCBI rmblastn: mask level off and dust off, RepeatMasker's options with DUST,
'-' in either sequence, a matrix with asymmetric frequencies, and NCBI's
defaults (word size 11).  The multi-chunk chr22 golden is checked in
test_oracle_chunks.py.
"""
import os

import pytest

from conftest import ORACLE_BIN, RM_FIELDS, RM_SEARCH, golden, run


def oracle(args):
    r = run([ORACLE_BIN] + args)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_nomask_nodust_matches_ncbi_byte_for_byte():
    # cli.rs:167-171: -mask_level 101 -dust no, word size 14 (megablast table)
    out = oracle(["--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"),
                  "--word-size", "14", *RM_SEARCH, "--outfmt", RM_FIELDS, "--mask-level", "101", "--dust", "no"])
    want = open(golden("expected_nomask_nodust.ncbi-2.17.1.tsv")).read()
    assert out.count("\n") == 192
    assert out == want


def test_ncbi_single_dash_spelling_is_accepted():
    a = oracle(["-db", golden("db.fa"), "-query", golden("query.fa"), "-matrix", golden("comparison.matrix"),
                "-word_size", "14", "-gapopen", "20", "-gapextend", "5", "-xdrop_ungap", "400",
                "-xdrop_gap_final", "200", "-xdrop_gap", "100", "-min_raw_gapped_score", "200",
                "-outfmt", RM_FIELDS, "-mask_level", "101", "-dust", "no"])
    assert a == open(golden("expected_nomask_nodust.ncbi-2.17.1.tsv")).read()


def test_identity_matrix_default_cutoff_has_no_score_zero_hits():
    # cli.rs:217-230: with --min-raw-gapped-score 0 no hit below NCBI's cutoff of 1
    out = oracle(["--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", golden("identity.matrix"),
                  "--gapopen", "20", "--gapextend", "5", "--word-size", "14", "--dust", "no",
                  "--outfmt", "6 score qstart qend"])
    lines = out.splitlines()
    assert len(lines) > 300
    for line in lines:
        s, qs, qe = map(int, line.split("\t"))
        assert s >= 1 and qe >= qs


def test_thread_count_does_not_change_output():
    base = ["--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"),
            "--word-size", "7", *RM_SEARCH, "--outfmt", RM_FIELDS, "--mask-level", "80", "--complexity-adjust",
            "--dust", "no"]
    assert oracle(base + ["--num-threads", "1"]) == oracle(base + ["--num-threads", "4"])


# cli.rs:147-299: the goldens that run with DUST on
DUST_GOLDENS = [
    ("rm_defaults", "query.fa", "db.fa", "comparison.matrix", ["--mask-level", "80", "--complexity-adjust"], 33),
    ("query_dash", "query_dash.fa", "db.fa", "comparison.matrix", ["--mask-level", "80", "--complexity-adjust"], 24),
    ("db_dash", "query.fa", "db_dash.fa", "comparison.matrix", ["--mask-level", "80", "--complexity-adjust"], 34),
    ("asym_freqs", "query.fa", "db.fa", "asym_freqs.matrix", ["--mask-level", "101", "--complexity-adjust"], 181),
]


def dust_golden_args(query, db, matrix, extra):
    return ["--db", golden(db), "--query", golden(query), "--matrix", golden(matrix), "--word-size", "14",
            *RM_SEARCH, "--outfmt", RM_FIELDS, *extra, "--dust", "yes"]


@pytest.mark.parametrize("name,query,db,matrix,extra,rows", DUST_GOLDENS)
def test_dust_goldens_match_ncbi(name, query, db, matrix, extra, rows):
    out = oracle(dust_golden_args(query, db, matrix, extra))
    want = open(golden(f"expected_{name}.ncbi-2.17.1.tsv")).read()
    assert out.count("\n") == rows
    assert out == want


def test_ncbi_defaults_match():
    # cli.rs:293-300: only the gap costs given; word size 11, dust on, no mask level
    out = oracle(["--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"),
                  "--gapopen", "20", "--gapextend", "5", "--outfmt", RM_FIELDS])
    assert out.count("\n") == 350
    assert out == open(golden("expected_defaults.ncbi-2.17.1.tsv")).read()
