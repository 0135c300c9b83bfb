"""Failure handling at the boundary (the reference's CLI turns errors into a
message and exit 1, rmblastn/src/cli.rs:375,493): a search that fails prints
no rows, on the CPU (no usable device) and on the GPU (a failure forced from
the host after each stage, RMB_FORCE_ERROR, while every kernel ends normally).
A later run in a fresh process gives the full rows again."""
import pytest

from conftest import CLI, RM_FIELDS, RM_SEARCH, golden, run

ARGS = ["--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"),
        "--word-size", "14", *RM_SEARCH, "--outfmt", RM_FIELDS, "--mask-level", "101", "--dust", "no"]


def test_no_such_device_exits_1_without_rows():
    r = run([CLI, *ARGS, "--device", "99"])
    assert r.returncode == 1
    assert r.stdout == ""
    assert r.stderr.startswith("error: device")


@pytest.mark.gpu
@pytest.mark.parametrize("stage", ["phase1", "phase2a", "phase2b"])
def test_forced_failure_exits_1_without_rows(stage):
    r = run([CLI, *ARGS], env={"RMB_FORCE_ERROR": stage})
    assert r.returncode == 1 and r.stdout == ""
    assert f"forced failure after {stage}" in r.stderr
    ok = run([CLI, *ARGS])
    assert ok.returncode == 0
    assert ok.stdout == open(golden("expected_nomask_nodust.ncbi-2.17.1.tsv")).read()
