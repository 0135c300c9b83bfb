"""HIP path parity at the configurations' real shapes.

* The exact-path-at-scale workload (the bench's rows): 60 of its 1,667 60-kbp
  records, spread over the genome, against all 1,000 families, with the C1
  scoring line (mask level 80, complexity adjustment) and the RepeatMasker
  outfmt 6 fields, GPU front end against the oracle.
* The 1 Mbp x 50-family parity configuration at its full size.
* The divergent/gappy stress configuration as one 4.2 Mbp record, so the
  split-query path (five chunks, seams folded) runs at X-drop 120/400.
* The reference's chr22 ungapped-stage golden (content and order) on the two
  chunks of chr22 held by the committed 3 Mbp fixture, through the ungapped
  tap (rmblast-lib/tests/ungapped_regression.rs:74-88).
"""
import gzip
import os

import numpy as np
import pytest

import rmblast_amd
from rmblast_amd import synth
from conftest import CLI, CHR22_3MB, ORACLE_BIN, RM_FIELDS, golden, run
from test_oracle_chr22 import read_2bit

pytestmark = pytest.mark.gpu


def gpu_and_oracle(args, threads=16, tmo=900):
    g = run([CLI] + args, timeout=tmo)
    assert g.returncode == 0, g.stderr
    o = run([ORACLE_BIN] + args + ["--num-threads", str(threads)], timeout=tmo)
    assert o.returncode == 0, o.stderr
    return g.stdout, o.stdout


def synth_args(tmp_path, xdrop=(20, 30, 100)):
    (tmp_path / "m.matrix").write_text(synth.MATRIX_TEXT)
    return ["--db", str(tmp_path / "db.fa"), "--query", str(tmp_path / "q.fa"), "--matrix",
            str(tmp_path / "m.matrix"), "--word-size", "7", "--gapopen", "20", "--gapextend", "5",
            "--min-raw-gapped-score", "200", "--xdrop-ungap", str(xdrop[0]), "--xdrop-gap", str(xdrop[1]),
            "--xdrop-gap-final", str(xdrop[2]), "--mask-level", "80", "--complexity-adjust", "--dust", "no",
            "--outfmt", RM_FIELDS]


def test_scale_workload_records_match_oracle(tmp_path):
    w = synth.scale_config()
    recs = w.records[::28]
    assert len(recs) == 60 and len(w.families) == 1000
    synth.to_fasta(recs, tmp_path / "q.fa")
    synth.to_fasta(w.families, tmp_path / "db.fa")
    g, o = gpu_and_oracle(synth_args(tmp_path))
    assert g == o
    assert g.count("\n") > 3000 and len({l.split("\t")[4] for l in g.splitlines()}) == 60


def test_parity_config_full_size_matches_oracle(tmp_path):
    w = synth.parity_config()
    assert len(w.records[0][1]) == 1_000_000 and len(w.families) == 50
    synth.to_fasta(w.records, tmp_path / "q.fa")
    synth.to_fasta(w.families, tmp_path / "db.fa")
    g, o = gpu_and_oracle(synth_args(tmp_path))
    assert g == o and g.count("\n") > 1000


def test_parity_config_with_every_traceback_half_lane_run(tmp_path):
    """The 1 Mbp parity configuration with every traceback half through the
    lane-serial DP (tb_wave_halves -1; its default at this size runs them one
    per wave): the oracle's rows."""
    w = synth.parity_config(seed=7)
    synth.to_fasta(w.records, tmp_path / "q.fa")
    synth.to_fasta(w.families, tmp_path / "db.fa")
    args = synth_args(tmp_path)
    (tmp_path / "t.tsv").write_text("matrix\tword_size\txdrop_gap\txdrop_gap_final\ttb_wave_halves\n"
                                    "m.matrix\t7\t30\t100\t-1\n")
    o = run([ORACLE_BIN] + args + ["--num-threads", "16"], timeout=900)
    assert o.returncode == 0, o.stderr
    g = run([CLI] + args + ["--tuning-file", str(tmp_path / "t.tsv")], timeout=900)
    assert g.returncode == 0, g.stderr
    assert g.stdout == o.stdout and g.stdout.count("\n") > 1000


def test_gappy_config_split_query_matches_oracle(tmp_path):
    w = synth.gappy_config(genome_len=4_200_000, n_families=60, n_copies=1700)
    synth.to_fasta(w.records, tmp_path / "q.fa")
    synth.to_fasta(w.families, tmp_path / "db.fa")
    g, o = gpu_and_oracle(synth_args(tmp_path, (20, 120, 400)))
    assert g == o and g.count("\n") > 200


def test_chr22_ungapped_stage_golden_on_gpu():
    lines = gzip.open(golden("chr22_stage_chunks30_31.txt.gz"), "rt").read().splitlines()
    want = [l for l in lines if not l.startswith("#")]
    seq = "".join(l.strip() for l in gzip.open(golden(CHR22_3MB + ".gz"), "rt") if not l.startswith(">"))
    # chunks 30 and 31 of chr22's NCBI query chunks, in the fixture's
    # coordinates (it starts at chr22 position 30,000,000)
    chunk_bounds = [(491_040, 1_507_508), (1_507_408, 2_523_876)]
    m = rmblast_amd.load_matrix(golden("comparison.matrix"))
    s = rmblast_amd.Searcher(m, gap_open=20, gap_extend=5, word_size=14, xdrop_ungap=400, xdrop_gap=100,
                             xdrop_gap_final=200, min_raw_gapped_score=200, complexity_adjust=True)
    alu = read_2bit(golden("alu.2bit"))
    s.load_library([rmblast_amd.encode(x) for _, x in alu])
    got = []
    for a, b in chunk_bounds:
        for subj, (_, sseq) in enumerate(alu):
            for q_seed, s_seed, qs, qe, ss, se, score, minus in s.tap_ungapped(rmblast_amd.encode(seq[a:b]), subj):
                if minus:  # record frame (reverse-complement subject) -> forward subject
                    ss, se = len(sseq) - se, len(sseq) - ss
                got.append(f"UNGAP strand={'-' if minus else '+'} qs={qs} qe={qe} ss={ss} se={se} "
                           f"score={np.int32(score)}")
    assert len(want) == 20860
    first = next((i for i, (x, y) in enumerate(zip(got, want)) if x != y), None)
    assert first is None, (first, got[first], want[first])
    assert len(got) == len(want)


def test_uneven_batching_subset_matches_oracle_under_every_schedule(tmp_path):
    """300 of the uneven configuration's records (1-500 kbp) against 200 of
    its families (50-20,000 bp), with the default schedule and two others
    from the autotuner's space: the same rows as the oracle each time."""
    w = synth.uneven_config(n_records=300, n_families=200)
    synth.to_fasta(w.records, tmp_path / "q.fa")
    synth.to_fasta(w.families, tmp_path / "db.fa")
    args = synth_args(tmp_path)
    o = run([ORACLE_BIN] + args + ["--num-threads", "16"], timeout=900)
    assert o.returncode == 0, o.stderr
    assert o.stdout.count("\n") > 10000
    head = "matrix\tword_size\txdrop_gap\txdrop_gap_final\tbatches\tprelim_ring\tprelim_wg_per_cu\ttb_ring\ttb_wg_per_cu\n"
    for row in (None, "m.matrix\t7\t30\t100\t3\t64\t2\t256\t2\n", "m.matrix\t7\t30\t100\t12\t16\t4\t64\t4\n"):
        extra = []
        if row:
            (tmp_path / "t.tsv").write_text(head + row)
            extra = ["--tuning-file", str(tmp_path / "t.tsv")]
        g = run([CLI] + args + extra, timeout=900)
        assert g.returncode == 0, g.stderr
        assert g.stdout == o.stdout, row


def test_autotuner_keeps_results_and_writes_its_table(tmp_path):
    from rmblast_amd import autotune
    w = synth.uneven_config(n_records=24, n_families=150)
    (tmp_path / "m.matrix").write_text(synth.MATRIX_TEXT)
    s = rmblast_amd.Searcher(rmblast_amd.load_matrix(str(tmp_path / "m.matrix")), gap_open=20, gap_extend=5,
                             word_size=7, min_raw_gapped_score=200, complexity_adjust=True)
    s.load_library([f[1] for f in w.families])
    best, t, trials = autotune.tune(s, [r[1] for r in w.records], reps=1, log=lambda m: None)
    # a value is kept only when it beats the best so far by the margin, so no
    # trial is more than the margin faster than the schedule kept
    assert len(trials) >= 10 and min(x[1] for x in trials) >= t * (1.0 - autotune.MARGIN)
    assert any(x[1] == t for x in trials)
    path = tmp_path / "tuning.tsv"
    autotune.write_entry(str(path), ("m.matrix", 7, 30, 100), best, t * 1e3, "test")
    assert rmblast_amd.read_tuning(str(path))[("m.matrix", 7, 30, 100)] == {k: best[k] for k in best}
