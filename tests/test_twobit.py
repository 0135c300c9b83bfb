"""The .2bit subject database of the rmblastn front end (host side, no GPU):
the fixture writer against the reference's own alu.2bit, and malformed files
refused with an error before any device work (rmblast-lib/src/seq/twobit.rs:187-300)."""
import os
import subprocess

import pytest

from conftest import CLI, golden, write_2bit


def fasta_records(path):
    recs = []
    for line in open(path).read().splitlines():
        if line.startswith(">"):
            recs.append([line[1:].split()[0], ""])
        else:
            recs[-1][1] += line
    return recs


def test_writer_reproduces_the_reference_alu_2bit(tmp_path):
    write_2bit(tmp_path / "alu.2bit", fasta_records(golden("alu.fa")))
    assert (tmp_path / "alu.2bit").read_bytes() == open(golden("alu.2bit"), "rb").read()


@pytest.mark.parametrize("cut,needle", [(10, "truncated 2bit header"), (30, "truncated"), (100, "runs past the end")])
def test_cli_refuses_truncated_2bit(tmp_path, cut, needle):
    data = open(golden("alu.2bit"), "rb").read()
    bad = tmp_path / "bad.2bit"
    bad.write_bytes(data[:cut])
    r = subprocess.run([CLI, "--db", str(bad), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"),
                        "--gapopen", "20", "--gapextend", "5"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and needle in r.stderr and r.stdout == ""


def test_cli_refuses_overlapping_n_blocks(tmp_path):
    # two N runs written as overlapping blocks: the second block starts inside the first
    import struct
    write_2bit(tmp_path / "n.2bit", [("s", "ACGTNNNNACGTNNNNACGT")])
    data = bytearray((tmp_path / "n.2bit").read_bytes())
    off = 16 + 1 + 1 + 4
    size, nb = struct.unpack_from("<II", data, off)
    assert nb == 2
    struct.pack_into("<I", data, off + 12, 5)  # second block start 12 -> 5
    (tmp_path / "n.2bit").write_bytes(bytes(data))
    r = subprocess.run([CLI, "--db", str(tmp_path / "n.2bit"), "--query", golden("query.fa"), "--matrix",
                        golden("comparison.matrix"), "--gapopen", "20", "--gapextend", "5"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "overlap or are out of order" in r.stderr
