"""HIP path parity with a RepeatMasker matrix of NCBI's ALP table: the
ungapped cutoff is then the Karlin-Altschul value, resolved per query record
from its whole length and the library's average length (rmblastn/src/main.rs:
1154-1170, rmblast-lib/src/search/ka_cutoff.rs:126-161).  The table is keyed
by the matrix file's name, so the scores are those of the fixtures' matrix
under the name 18p43g.matrix at its gap costs 28/5."""
import numpy as np
import pytest

from rmblast_amd import synth
from conftest import CLI, ORACLE_BIN, RM_FIELDS, golden, run

pytestmark = pytest.mark.gpu


def ka_args(tmp_path, db, query, ws, text=None, min_raw=300):
    m = tmp_path / "18p43g.matrix"
    m.write_text(text if text is not None else open(golden("comparison.matrix")).read())
    return ["--db", str(db), "--query", str(query), "--matrix", str(m), "--gapopen", "28", "--gapextend", "5",
            "--word-size", str(ws), "--min-raw-gapped-score", str(min_raw), "--dust", "no", "--mask-level", "80",
            "--complexity-adjust", "--outfmt", RM_FIELDS]


def gpu_and_oracle(args):
    g = run([CLI] + args)
    assert g.returncode == 0, g.stderr
    assert "Karlin-Altschul ungapped cutoff IN EFFECT (matrix=18p43g.matrix gapopen=28 gapextend=5)" in g.stderr
    o = run([ORACLE_BIN] + args)
    assert o.returncode == 0, o.stderr
    return g.stdout, o.stdout


@pytest.mark.parametrize("ws", [7, 11, 14])
def test_fixtures_with_a_table_matrix_match_oracle(tmp_path, ws):
    args = ka_args(tmp_path, golden("db.fa"), golden("query.fa"), ws)
    g, o = gpu_and_oracle(args)
    assert g == o and g.count("\n") > 20
    if ws == 7:  # the cutoff matters here: without mask level, the rows differ under another name
        nomask = list(args)
        nomask[nomask.index("--mask-level") + 1] = "101"
        g2, o2 = gpu_and_oracle(nomask)
        assert g2 == o2
        assert differs_under_another_name(tmp_path, nomask, g2)


def differs_under_another_name(tmp_path, args, rows):
    other = list(args)
    m = other.index("--matrix") + 1
    (tmp_path / "other.matrix").write_text(open(other[m]).read())
    other[m] = str(tmp_path / "other.matrix")
    r = run([CLI] + other)
    assert r.returncode == 0, r.stderr
    return r.stdout != rows


def test_records_of_mixed_lengths_get_their_own_cutoffs(tmp_path):
    """Records shorter than half the average family length search a smaller
    space and keep ungapped hits at lower scores; a batch mixing them with
    long records must give the oracle's rows record by record."""
    w = synth.parity_config(seed=21, genome_len=240_000, n_families=15, n_copies=500)
    genome = w.records[0][1]
    rng = np.random.default_rng(4)
    recs, at = [], 0
    while at < len(genome):
        n = int(rng.choice([150, 400, 900, 2500, 12_000, 40_000]))
        recs.append((f"r{len(recs)}", genome[at:at + n]))
        at += n
    synth.to_fasta(recs, tmp_path / "q.fa")
    synth.to_fasta(w.families, tmp_path / "db.fa")
    args = ka_args(tmp_path, tmp_path / "db.fa", tmp_path / "q.fa", 7, synth.MATRIX_TEXT, min_raw=600)
    g, o = gpu_and_oracle(args)
    assert g == o and g.count("\n") > 100
    assert differs_under_another_name(tmp_path, args, g)
