"""The CPU oracle against the reference's NCBI RMBlast 2.17.1 goldens.

rmblastn/tests/cli.rs:107-300 produced expected_*.ncbi-2.17.1.tsv with real
NCBI rmblastn.  The oracle has no DUST filter, so only the dust-free golden
(nomask_nodust) is an end-to-end pin; the others run with `-dust yes`.
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
