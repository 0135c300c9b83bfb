"""The C-ABI library loads and exports every entry point the headers under
include/ declare; host-side argument errors come back as error codes (no GPU
call)."""
import ctypes
import glob
import os
import re

import pytest

from conftest import CLI, ROOT

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared():
    names = set()
    for h in HEADERS:
        names |= set(re.findall(r"^int (rmb_\w+)\(", open(h).read(), re.M))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared()
    for n in ("rmb_ctx_create", "rmb_ctx_destroy", "rmb_library_load", "rmb_search", "rmb_results_size",
              "rmb_results_copy", "rmb_last_error", "rmb_matrix_load", "rmb_tap_ungapped",
              "rmb_tap_align_score_only", "rmb_tap_align_ex", "rmb_stats", "rmb_search_pairs",
              "rmb_triage_create", "rmb_triage_forward_backward", "rmb_triage_logits", "rmb_kmer_profiles"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import rmblast_amd
    lib = rmblast_amd.load_library()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing
    assert set(rmblast_amd.EXPORTED) <= set(declared())


def test_cli_was_built():
    assert os.access(CLI, os.X_OK)


def test_matrix_load_parses_and_rejects(tmp_path):
    import rmblast_amd
    m = rmblast_amd.load_matrix(os.path.join(ROOT, "tests", "golden", "comparison.matrix"))
    assert m.scores[0][0] == 9 and m.scores[15][0] == -(2 ** 30)
    assert abs(m.freqs[0] - 0.265) < 1e-12 and 0.1 < m.lambda_ < 0.2
    bad = tmp_path / "bad.matrix"
    bad.write_text("   A C\nA 1 -1\nC -1 1\n")       # no FREQS line: NCBI refuses it
    with pytest.raises(rmblast_amd.RmbError):
        rmblast_amd.load_matrix(str(bad))


def test_null_arguments_are_errors_not_crashes():
    import rmblast_amd
    lib = rmblast_amd.load_library()
    assert lib.rmb_ctx_create(0, None, None, None) != 0
    assert lib.rmb_search(None, 0, None, None) != 0
    assert lib.rmb_results_size(None, None, None) != 0
    assert lib.rmb_ctx_destroy(None) == 0


def test_missing_library_fails_loudly(tmp_path):
    import rmblast_amd
    with pytest.raises(rmblast_amd.RmbError):
        rmblast_amd.load_library(str(tmp_path / "nope.so"))


@pytest.mark.parametrize("fmt,needle", [
    ("6", "needs a field list"),
    ("7 score", "output format '7' is not supported"),
    ("10 score qseqid", "output format '10' is not supported"),
    ("6 score pident qseqid", "unknown -outfmt field 'pident'"),
    ("6 std", "unknown -outfmt field 'std'"),
    ("0 score", "takes no fields"),
])
def test_cli_rejects_unsupported_outfmt(fmt, needle):
    """The front end refuses these before it touches a device, naming the
    problem (rmblastn/tests/cli.rs:465-485)."""
    import subprocess
    g = os.path.join(ROOT, "tests", "golden")
    r = subprocess.run([CLI, "--db", os.path.join(g, "db.fa"), "--query", os.path.join(g, "query.fa"), "--matrix",
                        os.path.join(g, "comparison.matrix"), "--gapopen", "20", "--gapextend", "5", "--outfmt", fmt],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and needle in r.stderr and r.stdout == ""


def test_cli_requires_gap_costs():
    import subprocess
    g = os.path.join(ROOT, "tests", "golden")
    r = subprocess.run([CLI, "--db", os.path.join(g, "db.fa"), "--query", os.path.join(g, "query.fa"), "--matrix",
                        os.path.join(g, "comparison.matrix"), "--gapopen", "20"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--gapextend" in r.stderr
