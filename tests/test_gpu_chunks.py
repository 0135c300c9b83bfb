"""Split queries (1,999,800 bases or more) on the HIP path: NCBI's chunked
search (rmblastn/src/main.rs:1150-1300) against the NCBI RMBlast 2.17.1
golden and the oracle, with copies placed across the chunk seams."""
import numpy as np
import pytest

from rmblast_amd import synth
from conftest import CHR22_3MB, CHUNKED_SEARCH, CLI, ORACLE_BIN, RM_FIELDS, golden, golden_unzipped, run

pytestmark = pytest.mark.gpu


def test_chr22_3mb_ws7_matches_ncbi_on_gpu(tmp_path):
    # three chunks: [0, 1,000,067), [999,967, 2,000,034), [1,999,934, 3,000,000)
    q = golden_unzipped(CHR22_3MB, tmp_path)
    r = run([CLI, "--db", golden("db.fa"), "--query", q, *CHUNKED_SEARCH], timeout=300)
    assert r.returncode == 0, r.stderr
    want = open(golden("expected_chr22_30M-33M_ws7.ncbi-2.17.1.tsv")).read()
    assert r.stdout.count("\n") == 4236
    assert r.stdout == want


def split_of(n):
    """NCBI's chunk size and step for a query of n bases (main.rs:1150-1193)."""
    count = n // 999_900
    size = (n + (count - 1) * 100) // count
    if count < size - 100:
        size += 1
    return count, size, size - 100


def seam_genome(seed, n):
    """A synthetic genome of n bases with repeat copies and microsatellite
    tracts; at the chunk seams of NCBI's split: a plus-strand copy across the
    first, a minus-strand copy across the second, and from the third on two
    copies that meet inside the 100-base overlap."""
    rng = np.random.default_rng(seed)
    fams = synth.make_families(rng, 12, [int(x) for x in rng.integers(3000, 6000, size=12)])
    g, _ = synth.insert_copies(rng, n, fams, 1200, (300, 3000), (0.05, 0.30), (0.0, 0.03))
    g = g.copy()
    for _ in range(60):
        unit = synth.random_bases(rng, int(rng.integers(1, 4)))
        at = int(rng.integers(0, n - 300))
        g[at:at + 300] = np.tile(unit, 300)[:300]
    count, size, step = split_of(n)
    for k in range(1, count):
        seam = k * step
        copy = synth.mutate(rng, fams[k % len(fams)][1], 0.06, 0.005)
        if k == 1:
            pieces = [(seam + 50 - len(copy) // 2, copy)]
        elif k == 2:
            pieces = [(seam + 50 - len(copy) // 2, synth.COMP[copy[::-1]])]
        else:
            head = copy[: len(copy) // 2]
            tail = synth.mutate(rng, fams[(k + 1) % len(fams)][1], 0.06, 0.005)[:2500]
            pieces = [(seam + 40 - len(head), head), (seam + 40, tail)]
        for at, piece in pieces:
            g[at:at + len(piece)] = piece
    fams.append(("sat", np.tile(np.array([1, 0], dtype=np.uint8), 400)))
    return g, fams


@pytest.mark.parametrize("dust", ["no", "yes"])
def test_split_query_with_seam_copies_matches_oracle(tmp_path, dust):
    n = 4_000_000
    g, fams = seam_genome(21, n)
    rng = np.random.default_rng(22)
    small = [("short%d" % i, g[int(a):int(a) + 40_000]) for i, a in enumerate(rng.integers(0, n - 40_000, size=3))]
    synth.to_fasta([("big", g)] + small, tmp_path / "q.fa")
    synth.to_fasta(fams, tmp_path / "db.fa")
    (tmp_path / "m.matrix").write_text(synth.MATRIX_TEXT)
    args = ["--db", str(tmp_path / "db.fa"), "--query", str(tmp_path / "q.fa"), "--matrix",
            str(tmp_path / "m.matrix"), "--word-size", "7", "--gapopen", "20", "--gapextend", "5",
            "--min-raw-gapped-score", "200", "--mask-level", "80", "--complexity-adjust", "--dust", dust,
            "--outfmt", RM_FIELDS]
    gp = run([CLI] + args, timeout=300)
    assert gp.returncode == 0, gp.stderr
    o = run([ORACLE_BIN] + args + ["--num-threads", "16"], timeout=600)
    assert o.returncode == 0, o.stderr
    assert gp.stdout == o.stdout
    rows = [l.split("\t") for l in gp.stdout.splitlines()]
    big = [r for r in rows if r[4] == "big"]
    assert len(big) > 500
    _, _, step = split_of(n)
    for seam, strand in ((step, "plus"), (2 * step, "minus")):   # one alignment across each of these seams
        assert any(int(r[5]) < seam - 1000 and int(r[6]) > seam + 1100 and r[8] == strand for r in big)
