"""E-value and bit-score columns and the --evalue reap through the product
front end (HIP search + the library's host tail) against the oracle: the
table's statistics (comparison.matrix is in rmstats' table at 20/5, a
RepeatMasker matrix name at its gap costs), -matrix_* overrides, a "# KARLIN"
comment, NCBI 2.17.1 emulation (--ncbi-compat: sentinel columns and the reap
at 10), and the reap before mask level (rmblastn/src/main.rs:1224-1234)."""
import pytest

from conftest import CLI, ORACLE_BIN, RM_FIELDS, RM_SEARCH, golden, run

pytestmark = pytest.mark.gpu

FIELDS = RM_FIELDS + " evalue bitscore"


def both(args):
    g = run([CLI] + args)
    assert g.returncode == 0, g.stderr
    o = run([ORACLE_BIN] + args)
    assert o.returncode == 0, o.stderr
    return g, o


def fixture_args(matrix, extra=(), ws="11"):
    return ["--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", matrix, "--word-size", ws,
            *RM_SEARCH, "--dust", "no", "--complexity-adjust", "--outfmt", FIELDS, *extra]


@pytest.mark.parametrize("extra", [["--mask-level", "80"], ["--mask-level", "101"],
                                   ["--mask-level", "80", "--evalue", "1e-40"],
                                   ["--mask-level", "101", "--evalue", "1e-60"]])
def test_table_statistics_and_reap_match_oracle(extra):
    g, o = both(fixture_args(golden("comparison.matrix"), extra))
    assert g.stdout == o.stdout and g.stdout.count("\n") > 5
    assert "taken from the built-in table by matrix file name" in g.stderr
    evalues = [l.split("\t")[-2] for l in g.stdout.splitlines()]
    assert any("e-" in e for e in evalues)
    if "--evalue" in extra:  # the reap removed rows
        full = run([CLI] + fixture_args(golden("comparison.matrix"), extra[:2]))
        assert full.stdout.count("\n") > g.stdout.count("\n")


def test_ncbi_compat_prints_sentinels_and_reaps_nothing():
    g, o = both(fixture_args(golden("comparison.matrix"), ["--mask-level", "80", "--ncbi-compat"]))
    assert g.stdout == o.stdout and g.stdout.count("\n") > 5
    assert all(l.endswith("\t1.0\t 0.0") for l in g.stdout.splitlines())


@pytest.mark.parametrize("matrix_text,extra", [
    ("# KARLIN lambda 0.1276 k 0.0179 h 0.3562\n", []),
    ("", ["--matrix-lambda", "0.11", "--matrix-k", "0.05", "--matrix-alpha", "0.9", "--matrix-beta", "-20"]),
    ("", ["--evalue", "1e-50", "--matrix-lambda", "0.11", "--matrix-k", "0.05", "--matrix-alpha", "0.9"]),
])
def test_comment_and_overrides_match_oracle(tmp_path, matrix_text, extra):
    m = tmp_path / "custom.matrix"
    m.write_text(matrix_text + open(golden("comparison.matrix")).read())
    g, o = both(fixture_args(str(m), ["--mask-level", "80", *extra]))
    assert g.stdout == o.stdout and g.stdout.count("\n") > 5


def test_repeatmasker_matrix_name_at_its_gap_costs(tmp_path):
    m = tmp_path / "18p43g.matrix"
    m.write_text(open(golden("comparison.matrix")).read())
    args = ["--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", str(m), "--gapopen", "28",
            "--gapextend", "5", "--word-size", "7", "--min-raw-gapped-score", "250", "--dust", "no",
            "--mask-level", "80", "--complexity-adjust", "--evalue", "1e-20", "--outfmt", FIELDS]
    g, o = both(args)
    assert g.stdout == o.stdout and g.stdout.count("\n") > 5
