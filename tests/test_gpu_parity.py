"""HIP path parity: the product (librmblast_hip.so through its C ABI, and the
rmblastn front end on top of it) against the NCBI golden and the CPU oracle.
Integer work throughout, so every comparison is byte/bit exact."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import rmblast_amd
from rmblast_amd import synth
from conftest import CLI, ORACLE_BIN, RM_FIELDS, RM_SEARCH, golden, ptr, run, seq_from_fasta, simple_matrix

pytestmark = pytest.mark.gpu


def both(args, tmo=300):
    g = run([CLI] + args, timeout=tmo)
    assert g.returncode == 0, g.stderr
    o = run([ORACLE_BIN] + args, timeout=tmo)
    assert o.returncode == 0, o.stderr
    return g.stdout, o.stdout


def test_nomask_nodust_golden_on_gpu():
    g = run([CLI, "--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"),
             "--word-size", "14", *RM_SEARCH, "--outfmt", RM_FIELDS, "--mask-level", "101", "--dust", "no"])
    assert g.returncode == 0, g.stderr
    assert g.stdout == open(golden("expected_nomask_nodust.ncbi-2.17.1.tsv")).read()


@pytest.mark.parametrize("ws", [7, 8, 11, 14, 19])
@pytest.mark.parametrize("extra", [["--mask-level", "80", "--complexity-adjust"], ["--mask-level", "101"]])
def test_fixture_runs_match_oracle(ws, extra):
    g, o = both(["--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"),
                 "--word-size", str(ws), *RM_SEARCH, "--outfmt", RM_FIELDS, "--dust", "no", *extra])
    assert g == o and g.count("\n") > 20


@pytest.mark.parametrize("db,query,matrix", [
    ("db_dash.fa", "query.fa", "comparison.matrix"),      # subject ambiguity: NcbiRandom + n_mask restore
    ("db.fa", "query_dash.fa", "comparison.matrix"),      # query N runs
    ("db.fa", "query.fa", "asym_freqs.matrix"),           # minus-strand complexity adjust
    ("db.fa", "query.fa", "identity.matrix"),
])
def test_ambiguity_and_matrices_match_oracle(db, query, matrix):
    g, o = both(["--db", golden(db), "--query", golden(query), "--matrix", golden(matrix), "--word-size", "11",
                 *RM_SEARCH, "--outfmt", RM_FIELDS, "--dust", "no", "--mask-level", "80", "--complexity-adjust"])
    assert g == o


@pytest.mark.parametrize("ws", [7, 14])
def test_chr22_excerpt_matches_oracle(tmp_path, ws):
    db = tmp_path / "db.fa"
    db.write_text(open(golden("alu.fa")).read() + open(golden("db.fa")).read())
    g, o = both(["--db", str(db), "--query", golden("chr22_30000000_30200000.fa"), "--matrix",
                 golden("comparison.matrix"), "--word-size", str(ws), "--gapopen", "20", "--gapextend", "5",
                 "--min-raw-gapped-score", "200", "--mask-level", "80", "--dust", "no",
                 "--outfmt", RM_FIELDS])
    assert g == o and g.count("\n") > 10


@pytest.mark.parametrize("seed", [11, 12])
def test_synthetic_parity_config_matches_oracle(tmp_path, seed):
    w = synth.parity_config(seed=seed, genome_len=300_000, n_families=20, n_copies=600)
    synth.to_fasta(w.records, tmp_path / "q.fa")
    synth.to_fasta(w.families, tmp_path / "db.fa")
    (tmp_path / "m.matrix").write_text(synth.MATRIX_TEXT)
    g, o = both(["--db", str(tmp_path / "db.fa"), "--query", str(tmp_path / "q.fa"), "--matrix",
                 str(tmp_path / "m.matrix"), "--word-size", "7", "--gapopen", "20", "--gapextend", "5",
                 "--min-raw-gapped-score", "200", "--mask-level", "80", "--complexity-adjust", "--dust", "no",
                 "--outfmt", RM_FIELDS])
    assert g == o and g.count("\n") > 100


@pytest.mark.parametrize("parts_log2", ["0", "1", "3", "4"])
@pytest.mark.parametrize("ws", [7, 11])
def test_items_split_over_bucket_ranges_match_oracle(tmp_path, parts_log2, ws):
    """A search of few (record, family) items splits each item over up to 16
    workgroups, each owning a range of the diagonal hash's 512 buckets; the
    rows are the oracle's whatever the split (RMB_P1_PARTS_LOG2 forces it)."""
    w = synth.parity_config(seed=13, genome_len=300_000, n_families=20, n_copies=600)
    synth.to_fasta(w.records, tmp_path / "q.fa")
    synth.to_fasta(w.families, tmp_path / "db.fa")
    (tmp_path / "m.matrix").write_text(synth.MATRIX_TEXT)
    args = ["--db", str(tmp_path / "db.fa"), "--query", str(tmp_path / "q.fa"), "--matrix",
            str(tmp_path / "m.matrix"), "--word-size", str(ws), "--gapopen", "20", "--gapextend", "5",
            "--min-raw-gapped-score", "200", "--mask-level", "80", "--complexity-adjust", "--dust", "no",
            "--outfmt", RM_FIELDS]
    g = run([CLI] + args, env={"RMB_P1_PARTS_LOG2": parts_log2})
    assert g.returncode == 0, g.stderr
    o = run([ORACLE_BIN] + args)
    assert o.returncode == 0, o.stderr
    assert g.stdout == o.stdout and g.stdout.count("\n") > 100


def test_gappy_divergent_and_microsatellites_match_oracle(tmp_path):
    rng = np.random.default_rng(5)
    fams = synth.make_families(rng, 6, [5000, 8000, 6000, 7000, 9000, 5500])
    g0, _ = synth.insert_copies(rng, 200_000, fams, 60, (500, 4000), (0.25, 0.40), (0.05, 0.08))
    g0 = g0.copy()
    for k in range(40):                              # 5% microsatellite tracts, period 2-6
        p = int(rng.integers(2, 7))
        unit = synth.random_bases(rng, p)
        L = 250
        at = int(rng.integers(0, len(g0) - L))
        g0[at:at + L] = np.tile(unit, L // p + 1)[:L]
    fams.append(("sat", np.tile(np.array([1, 0], dtype=np.uint8), 400)))
    synth.to_fasta([("q", g0)], tmp_path / "q.fa")
    synth.to_fasta(fams, tmp_path / "db.fa")
    (tmp_path / "m.matrix").write_text(synth.MATRIX_TEXT)
    args = ["--db", str(tmp_path / "db.fa"), "--query", str(tmp_path / "q.fa"), "--matrix",
            str(tmp_path / "m.matrix"), "--word-size", "7", "--gapopen", "20", "--gapextend", "5",
            "--min-raw-gapped-score", "200", "--xdrop-gap", "120", "--xdrop-gap-final", "400",
            "--complexity-adjust", "--dust", "no", "--outfmt", RM_FIELDS]
    g, o = both(args)
    assert g == o and g.count("\n") > 20
    # the preliminary hits past the LDS ring: one per wave by default, here
    # every one on the global-ring lane pass instead
    for env in ({"RMB_PRELIM_WAVE_HITS": "0", "RMB_STAGE_TIMES": "1"}, {"RMB_STAGE_TIMES": "1"}):
        r = run([CLI] + args, env=env)
        assert r.returncode == 0, r.stderr
        assert r.stdout == o
        wide = [l for l in r.stderr.splitlines() if "wider bands" in l]
        assert wide and any(" (0 hits)" not in l.split("wider bands")[1] for l in wide), wide


def test_traceback_two_pass_fallback_matches_oracle(tmp_path):
    """At X-drop 400 the traceback kernel takes its 256-cell LDS rings.
    Halves whose band outgrows a ring are rerun in the same launch with a
    retry slot; with the 64-cell rings and no retry slots they take the
    two-pass path (score pass, then a traceback pass into exactly sized
    storage).  All must give the oracle's rows."""
    rng = np.random.default_rng(11)
    fams = synth.make_families(rng, 3, [3000, 4000, 2500])
    g0, _ = synth.insert_copies(rng, 60_000, fams, 20, (400, 2500), (0.20, 0.35), (0.04, 0.08))
    g0 = g0.copy()
    for k in range(12):
        p = int(rng.integers(2, 5))
        unit = synth.random_bases(rng, p)
        at = int(rng.integers(0, len(g0) - 300))
        g0[at:at + 300] = np.tile(unit, 300 // p + 1)[:300]
    fams.append(("sat", np.tile(np.array([1, 0], dtype=np.uint8), 400)))
    synth.to_fasta([("q", g0)], tmp_path / "q.fa")
    synth.to_fasta(fams, tmp_path / "db.fa")
    (tmp_path / "m.matrix").write_text(synth.MATRIX_TEXT)
    args = ["--db", str(tmp_path / "db.fa"), "--query", str(tmp_path / "q.fa"), "--matrix",
            str(tmp_path / "m.matrix"), "--word-size", "7", "--gapopen", "20", "--gapextend", "5",
            "--min-raw-gapped-score", "200", "--xdrop-gap", "120", "--xdrop-gap-final", "400",
            "--complexity-adjust", "--dust", "no", "--outfmt", RM_FIELDS]
    o = run([ORACLE_BIN] + args)
    assert o.returncode == 0, o.stderr
    # wave "0": every half one per lane (the one-pass kernels and their retry
    # slots); "": the few halves here run one per wave (256-cell bands with the
    # wide rings, 64-cell ones without)
    for slots, wide, wave in (("32", "1", "0"), ("32", "0", "0"), ("0", "0", "0"), ("32", "1", ""), ("0", "0", "")):
        env = {"RMB_TB_RETRY_SLOTS": slots, "RMB_TB_WIDE": wide, "RMB_STAGE_TIMES": "1"}
        if wave:
            env["RMB_TB_WAVE_HALVES"] = wave
        g = run([CLI] + args, env=env)
        assert g.returncode == 0, g.stderr
        assert g.stdout == o.stdout and g.stdout.count("\n") > 10
        tb = [l for l in g.stderr.splitlines() if "two-pass fallback" in l]
        assert tb, g.stderr[-2000:]
        assert all((re.search(r"(\d+) wave-run", l).group(1) == "0") == (wave == "0") for l in tb), tb
        if slots == "0":  # the two-pass path ran
            assert any("(0 halves)" not in l.split("fallback")[1] for l in tb), tb


@pytest.mark.parametrize("xg,xf", [("300", "2500"), ("1200", "1500")])
def test_very_large_xdrops_take_the_wide_paths_with_the_oracles_rows(tmp_path, xg, xf):
    """X-drops past the 12-bit ring cells' range (X-drop + open + largest
    score < 2,000) or past 256-cell bands: the traceback halves go from the
    one-pass and wave-run kernels to the retry slots or the two-pass path,
    and the preliminary hits past the first pass's ring to the wave-run pass
    and (bands past 256 cells, X-drop 1,200) on to the global-ring pass; the
    rows stay the oracle's."""
    rng = np.random.default_rng(13)
    fams = synth.make_families(rng, 3, [2500, 3000, 2000])
    g0, _ = synth.insert_copies(rng, 40_000, fams, 16, (400, 2000), (0.10, 0.30), (0.02, 0.06))
    synth.to_fasta([("q", g0)], tmp_path / "q.fa")
    synth.to_fasta(fams, tmp_path / "db.fa")
    (tmp_path / "m.matrix").write_text(synth.MATRIX_TEXT)
    args = ["--db", str(tmp_path / "db.fa"), "--query", str(tmp_path / "q.fa"), "--matrix",
            str(tmp_path / "m.matrix"), "--word-size", "7", "--gapopen", "20", "--gapextend", "5",
            "--min-raw-gapped-score", "200", "--xdrop-gap", xg, "--xdrop-gap-final", xf,
            "--complexity-adjust", "--dust", "no", "--outfmt", RM_FIELDS]
    g, o = both(args)
    assert g == o and g.count("\n") > 5


def test_edge_cases_short_and_empty_records(tmp_path):
    q = tmp_path / "q.fa"
    q.write_text(">empty\n\n>short\nACGTA\n>nrun\n" + "N" * 500 + "\n>ok\n" +
                 seq_from_fasta(golden("query.fa"))[:3000] + "\n")
    db = tmp_path / "db.fa"
    db.write_text(">tiny\nACG\n" + open(golden("db.fa")).read())
    g, o = both(["--db", str(db), "--query", str(q), "--matrix", golden("comparison.matrix"), "--word-size", "7",
                 *RM_SEARCH, "--outfmt", RM_FIELDS, "--dust", "no", "--mask-level", "80"])
    assert g == o


def test_repeat_runs_are_deterministic(tmp_path):
    args = ["--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"),
            "--word-size", "7", *RM_SEARCH, "--outfmt", RM_FIELDS, "--dust", "no"]
    a = run([CLI] + args)
    b = run([CLI] + args)
    assert a.returncode == 0 and a.stdout == b.stdout


# ---- stage taps through the C ABI ---------------------------------------------------
def searcher_with(mat16, **kw):
    m = rmblast_amd.RmbMatrix()
    for i in range(16):
        for j in range(16):
            m.scores[i][j] = int(mat16[i, j])
    return rmblast_amd.Searcher(m, **kw)


def test_tap_word_finder_known_answer():
    # engine.rs:3431-3490 (NCBI testNuclWordFinder) on the GPU seeding kernel
    mat = simple_matrix(1, -3, other=0)
    s = searcher_with(mat, word_size=19, xdrop_ungap=11, min_raw_gapped_score=28)
    subj = rmblast_amd.encode(seq_from_fasta(golden("gi33383640_xylanase_subject.fa")))
    s.load_library([subj])
    q = rmblast_amd.encode(seq_from_fasta(golden("gi3090_xylanase_query.fa")))
    hits = s.tap_ungapped(q, 0)
    plus = sorted((int(h[2]), int(h[4]), int(h[3] - h[2]), int(h[6])) for h in hits if h[7] == 0)
    assert plus == sorted([(233, 0, 31, 31), (1037, 66, 945, 853), (263, 27, 685, 541), (1911, 811, 101, 73),
                           (1782, 940, 63, 51)])


@pytest.mark.parametrize("ws", [7, 11, 14])
def test_tap_ungapped_matches_oracle(oracle_lib, ws):
    w = synth.parity_config(seed=21, genome_len=120_000, n_families=4, n_copies=120)
    m = rmblast_amd.load_matrix(golden("comparison.matrix"))
    mat = np.ascontiguousarray(np.array([[m.scores[i][j] for j in range(16)] for i in range(16)], dtype=np.int32))
    s = rmblast_amd.Searcher(m, word_size=ws, xdrop_ungap=20, min_raw_gapped_score=60)
    s.load_library([f[1] for f in w.families])
    q = w.records[0][1]
    for f in range(len(w.families)):
        got = s.tap_ungapped(q, f)
        subj = w.families[f][1]
        out = np.zeros((200000, 8), dtype=np.uint32)
        n = oracle_lib.orc_ungapped_hits(ptr(q), len(q), ptr(subj), len(subj), ws, 20, 60,
                                         ptr(mat, ctypes.c_int32), ptr(out, ctypes.c_uint32), 200000)
        assert n > 0 and np.array_equal(got, out[:n]), (ws, f, len(got), n)


def random_pair(rng):
    kind = rng.integers(0, 4)
    la = int(rng.integers(1, 3000))
    a = rng.integers(0, 4, la).astype(np.uint8)
    if kind == 0:        # mutated copy with indels
        b = synth.mutate(rng, a, float(rng.uniform(0.0, 0.3)), float(rng.uniform(0, 0.08)))
    elif kind == 1:      # unrelated
        b = rng.integers(0, 4, int(rng.integers(1, 3000))).astype(np.uint8)
    elif kind == 2:      # tandem repeat against tandem repeat: wide bands
        unit = rng.integers(0, 4, int(rng.integers(1, 7))).astype(np.uint8)
        a = np.tile(unit, la // len(unit) + 1)[:la]
        b = synth.mutate(rng, np.tile(unit, 400)[:int(rng.integers(50, 2500))], 0.05, 0.02)
    else:                # ambiguity codes sprinkled in
        b = synth.mutate(rng, a, 0.1, 0.02)
        b[rng.random(len(b)) < 0.02] = 14
    if len(b) == 0:
        b = np.zeros(1, dtype=np.uint8)
    return a, b


def test_tap_dp_kernels_match_oracle(oracle_lib):
    m = rmblast_amd.load_matrix(golden("comparison.matrix"))
    mat = np.ascontiguousarray(np.array([[m.scores[i][j] for j in range(16)] for i in range(16)], dtype=np.int32))
    rng = np.random.default_rng(99)
    gaps = [(20, 5), (4, 4), (8, 2)]
    ctxs = {g: rmblast_amd.Searcher(m, gap_open=g[0], gap_extend=g[1]) for g in gaps}
    for trial in range(120):
        a, b = random_pair(rng)
        go, ge = gaps[trial % 3]
        s = ctxs[(go, ge)]
        x = [30, 100, 400, 10 ** 6][trial % 4]
        for rev in (0, 1):
            want = (ctypes.c_int64 * 3)()
            oracle_lib.orc_align_score_only(ptr(a), ptr(b), len(a), len(b), go, ge, x, ptr(mat, ctypes.c_int32), rev,
                                            want)
            assert s.tap_align_score_only(a, b, x, rev) == tuple(want), (trial, rev)
            o = (ctypes.c_int64 * 4)()
            ops = np.zeros(2 * 20000, dtype=np.uint32)
            n = oracle_lib.orc_align_ex(ptr(a), ptr(b), len(a), len(b), go, ge, x, ptr(mat, ctypes.c_int32), rev, o,
                                        ptr(ops, ctypes.c_uint32), 20000)
            sc, al, bl, gops = s.tap_align_ex(a, b, x, rev)
            assert (sc, al, bl) == tuple(o[:3]), (trial, rev)
            assert gops == [(int(ops[2 * k]), int(ops[2 * k + 1])) for k in range(n)], (trial, rev)


@pytest.mark.parametrize("wave_halves", [1 << 30, -1], ids=["wave_run", "lane_run"])
def test_traceback_dp_kernels_match_oracle(oracle_lib, wave_halves):
    """Every traceback half through the wave-cooperative DP (tb_wave_halves
    2^30), then through the lane-serial one (-1), with 64-cell rings: score,
    extents and ops identical to the oracle's, over
    mutated copies, unrelated pairs, tandem repeats (bands past 64 cells,
    which fall back to the two-pass path) and ambiguity codes, both
    directions, four X-drops and three gap costs; and single-base and
    empty-overlap edge cases."""
    m = rmblast_amd.load_matrix(golden("comparison.matrix"))
    mat = np.ascontiguousarray(np.array([[m.scores[i][j] for j in range(16)] for i in range(16)], dtype=np.int32))
    rng = np.random.default_rng(2024)
    gaps = [(20, 5), (4, 4), (8, 2)]
    ctxs = {}
    for g in gaps:
        ctxs[g] = rmblast_amd.Searcher(m, gap_open=g[0], gap_extend=g[1])
        ctxs[g].set_tuning({"tb_ring": 64, "tb_wave_halves": wave_halves})
    pairs = [random_pair(rng) for _ in range(150)]
    pairs += [(np.array([0], np.uint8), np.array([0], np.uint8)), (np.array([1], np.uint8), np.array([2], np.uint8)),
              (np.array([0, 1, 2, 3], np.uint8), np.array([3], np.uint8)),
              (rng.integers(0, 4, 5000).astype(np.uint8), rng.integers(0, 4, 3).astype(np.uint8))]
    for trial, (a, b) in enumerate(pairs):
        go, ge = gaps[trial % 3]
        s = ctxs[(go, ge)]
        x = [30, 100, 400, 10 ** 6][trial % 4]
        for rev in (0, 1):
            o = (ctypes.c_int64 * 4)()
            ops = np.zeros(2 * 20000, dtype=np.uint32)
            n = oracle_lib.orc_align_ex(ptr(a), ptr(b), len(a), len(b), go, ge, x, ptr(mat, ctypes.c_int32), rev, o,
                                        ptr(ops, ctypes.c_uint32), 20000)
            sc, al, bl, gops = s.tap_align_ex(a, b, x, rev)
            assert (sc, al, bl) == tuple(o[:3]), (trial, rev, len(a), len(b))
            assert gops == [(int(ops[2 * k]), int(ops[2 * k + 1])) for k in range(n)], (trial, rev)


# the NCBI goldens that run with DUST on (cli.rs:147-299), through the product CLI
GPU_DUST_GOLDENS = [
    ("rm_defaults", "query.fa", "db.fa", "comparison.matrix", ["--mask-level", "80", "--complexity-adjust"]),
    ("query_dash", "query_dash.fa", "db.fa", "comparison.matrix", ["--mask-level", "80", "--complexity-adjust"]),
    ("db_dash", "query.fa", "db_dash.fa", "comparison.matrix", ["--mask-level", "80", "--complexity-adjust"]),
    ("asym_freqs", "query.fa", "db.fa", "asym_freqs.matrix", ["--mask-level", "101", "--complexity-adjust"]),
]


@pytest.mark.parametrize("name,query,db,matrix,extra", GPU_DUST_GOLDENS)
def test_dust_goldens_on_gpu(name, query, db, matrix, extra):
    g = run([CLI, "--db", golden(db), "--query", golden(query), "--matrix", golden(matrix), "--word-size", "14",
             *RM_SEARCH, "--outfmt", RM_FIELDS, *extra, "--dust", "yes"])
    assert g.returncode == 0, g.stderr
    assert g.stdout == open(golden(f"expected_{name}.ncbi-2.17.1.tsv")).read()


def test_ncbi_defaults_golden_on_gpu():
    g = run([CLI, "--db", golden("db.fa"), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"),
             "--gapopen", "20", "--gapextend", "5", "--outfmt", RM_FIELDS])
    assert g.returncode == 0, g.stderr
    assert g.stdout == open(golden("expected_defaults.ncbi-2.17.1.tsv")).read()


def test_dust_on_synthetic_matches_oracle(tmp_path):
    # word size 7 with DUST on the parity-configuration generator, with
    # microsatellites so DUST has something to mask
    from rmblast_amd import synth
    w = synth.parity_config(seed=13, genome_len=150_000, n_families=15, n_copies=300)
    q = list(w.records[0][1])
    rng = np.random.default_rng(3)
    for k in range(40):  # (CA)n, (A)n and (TTG)n tracts
        at = int(rng.integers(0, len(q) - 200))
        unit = [[1, 0], [0], [3, 3, 2]][k % 3]
        for i in range(int(rng.integers(30, 150))):
            q[at + i] = unit[i % len(unit)]
    qp, dp, mp = tmp_path / "q.fa", tmp_path / "db.fa", tmp_path / "m.matrix"
    synth.to_fasta([(w.records[0][0], np.array(q, dtype=np.uint8))], str(qp))
    synth.to_fasta(w.families, str(dp))
    mp.write_text(synth.MATRIX_TEXT)
    args = ["--db", str(dp), "--query", str(qp), "--matrix", str(mp), "--gapopen", "20", "--gapextend", "5",
            "--word-size", "7", "--min-raw-gapped-score", "200", "--mask-level", "80", "--complexity-adjust",
            "--dust", "yes", "--outfmt", RM_FIELDS]
    g = run([CLI] + args)
    o = run([ORACLE_BIN, "--num-threads", "8"] + args, timeout=1200)
    assert g.returncode == 0 and o.returncode == 0, (g.stderr, o.stderr)
    assert g.stdout == o.stdout
    assert g.stdout.count("\n") > 50


# ---- pairwise output (outfmt 0, the default) ------------------------------------------
RM_PAIRWISE = ["--word-size", "14", "--mask-level", "80", "--complexity-adjust", "--min-raw-gapped-score", "200",
               "--xdrop-ungap", "400", "--xdrop-gap-final", "200", "--xdrop-gap", "100"]


def pairwise_run(cwd, db, extra=()):
    r = subprocess.run([CLI, "--db", db, "--query", "query.fa", "--matrix", "comparison.matrix", "--gapopen", "20",
                        "--gapextend", "5", "--num-threads", "2", *extra],
                       cwd=str(cwd), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("name,extra", [("defaults", []), ("rm", RM_PAIRWISE)])
def test_pairwise_goldens_on_gpu(tmp_path, name, extra):
    """NCBI 2.17.1 pairwise output on the fixtures, run from their directory
    so the database title is "db.fa"; the "Posted date" line (the database
    file's time) is dropped on both sides (rmblastn/tests/cli.rs:190-212)."""
    for f in ("db.fa", "query.fa", "comparison.matrix"):
        shutil.copy(golden(f), tmp_path / f)
    out = pairwise_run(tmp_path, "db.fa", extra)
    got = "".join(l + "\n" for l in out.splitlines() if not l.startswith("    Posted date:"))
    assert got == open(golden(f"expected_pairwise_{name}.ncbi-2.17.1.txt")).read()


def test_pairwise_titles_wrap_like_ncbi(tmp_path):
    """The known answers of rmblast-lib/src/output.rs:735-765: a database path
    wrapped at 68 columns after a hyphen, a subject title broken after the
    first space past each 60th character, a long hyphenated word never
    broken, an id-only defline printed with its trailing space."""
    shutil.copy(golden("query.fa"), tmp_path / "query.fa")
    shutil.copy(golden("comparison.matrix"), tmp_path / "comparison.matrix")
    recs = []
    for line in open(golden("db.fa")).read().splitlines():
        if line.startswith(">"):
            recs.append([line[1:], []])
        else:
            recs[-1][1].append(line)
    long5 = "id5 x averyveryveryveryveryveryveryveryveryveryveryveryverylongwordwithnospaces-and-hyphens-here-too"
    recs[0][0] = "id3 " + "a very long description with many words " * 4
    recs[1][0] = long5
    recs[2][0] = "onlyid"
    d = "databases/repeat-libraries/2026-11-release/curated-families-with-long-names"
    (tmp_path / d).mkdir(parents=True)
    (tmp_path / d / "db.fa").write_text("".join(">" + h + "\n" + "\n".join(s) + "\n" for h, s in recs))
    out = pairwise_run(tmp_path, d + "/db.fa")
    wrapped = "databases/repeat-libraries/2026-11-release/curated-families-with-\nlong-names/db.fa\n"
    assert out.startswith("RMBLASTN 2.17.1+\n")
    assert "\nDatabase: " + wrapped + "           17 sequences; 11,207 total letters\n" in out
    assert "\n  Database: " + wrapped in out
    assert ("\n>id3 a very long description with many words a very long description \n"
            "with many words a very long description with many words a \n"
            "very long description with many words \nLength=") in out
    assert "\n>" + long5 + "\nLength=" in out
    assert "\n>onlyid \nLength=" in out


def test_gilist_selects_subjects(tmp_path):
    """--gilist keeps the subjects it names (lines trimmed; a bare number also
    names "gi|<n>"); one that names nothing is a search with no hits and a
    warning (rmblastn/src/main.rs:578-600,1444-1470; rmblastn/tests/cli.rs:453-463)."""
    recs = []
    for line in open(golden("db.fa")).read().splitlines():
        if line.startswith(">"):
            recs.append([f"gi|{len(recs) + 1} " + line[1:], []])
        else:
            recs[-1][1].append(line)
    db = tmp_path / "db.fa"
    db.write_text("".join(">" + h + "\n" + "\n".join(s) + "\n" for h, s in recs))
    args = ["--db", str(db), "--query", golden("query.fa"), "--matrix", golden("comparison.matrix"), "--gapopen", "20",
            "--gapextend", "5", "--outfmt", "6 score qseqid qstart qend sstrand sseqid sstart send"]
    full = run([CLI] + args)
    assert full.returncode == 0, full.stderr
    for text, keep in (("3\n  gi|7 \n\n", {"gi|3", "gi|7"}), ("no-such-id\n", set())):
        gl = tmp_path / "gilist"
        gl.write_text(text)
        r = run([CLI] + args + ["--gilist", str(gl)])
        assert r.returncode == 0, r.stderr
        want = "".join(l + "\n" for l in full.stdout.splitlines() if l.split("\t")[5] in keep)
        assert r.stdout == want
        assert bool(keep) or "matches no sequence" in r.stderr
    assert sum(1 for l in full.stdout.splitlines() if l.split("\t")[5] in {"gi|3", "gi|7"}) > 5


def test_2bit_database_matches_fasta(tmp_path):
    """A .2bit library gives the rows of the same library in FASTA: the
    reference's alu.2bit against its FASTA on a chr22 excerpt, and families
    with N runs (N blocks, resolved like FASTA N: twobit.rs:346-384) and
    soft-masked stretches, in both byte orders (seq/twobit.rs:187-300)."""
    import random
    from conftest import write_2bit
    q = golden("chr22_30000000_30200000.fa")
    common = ["--matrix", golden("comparison.matrix"), "--word-size", "14", *RM_SEARCH, "--outfmt", RM_FIELDS,
              "--dust", "no"]
    a = run([CLI, "--db", golden("alu.2bit"), "--query", q] + common)
    b = run([CLI, "--db", golden("alu.fa"), "--query", q] + common)
    assert a.returncode == 0 and b.returncode == 0, (a.stderr, b.stderr)
    assert a.stdout == b.stdout and a.stdout.count("\n") > 5
    recs = []
    for line in open(golden("db.fa")).read().splitlines():
        if line.startswith(">"):
            recs.append([line[1:].split()[0], ""])
        else:
            recs[-1][1] += line
    rng = random.Random(5)
    for k in (0, 2, 6, 9):
        s = recs[k][1]
        for _ in range(3):
            at, z = rng.randrange(0, len(s) - 40), rng.randrange(1, 30)
            s = s[:at] + "N" * z + s[at + z:]
        at = rng.randrange(0, len(s) - 60)
        recs[k][1] = s[:at] + s[at:at + 50].lower() + s[at + 50:]
    (tmp_path / "db.fa").write_text("".join(f">{n}\n{s}\n" for n, s in recs))
    write_2bit(tmp_path / "db.2bit", recs)
    write_2bit(tmp_path / "db_swapped.2bit", recs, swapped=True)
    args = ["--query", golden("query.fa"), "--matrix", golden("comparison.matrix"), "--word-size", "11",
            "--gapopen", "20", "--gapextend", "5", "--outfmt", RM_FIELDS]
    outs = []
    for db in ("db.fa", "db.2bit", "db_swapped.2bit"):
        r = run([CLI, "--db", str(tmp_path / db)] + args)
        assert r.returncode == 0, r.stderr
        outs.append(r.stdout)
    assert outs[0] == outs[1] == outs[2] and outs[0].count("\n") > 100
    assert any("N" in l.split("\t")[-1] for l in outs[0].splitlines())  # N restored in the subject rows


def test_gappy_stress_config_matches_oracle(tmp_path):
    """The divergent/gappy stress configuration's parameter point (families of
    5-15 kbp, copies 25-40% diverged with 5-8% indels, 5% microsatellites,
    X-drop 120/400) on a 1.5 Mbp genome: identical rows to the oracle."""
    w = synth.gappy_config(genome_len=1_500_000, n_copies=600)
    synth.to_fasta(w.records, str(tmp_path / "q.fa"))
    synth.to_fasta(w.families, str(tmp_path / "db.fa"))
    (tmp_path / "m.matrix").write_text(synth.MATRIX_TEXT)
    args = ["--db", str(tmp_path / "db.fa"), "--query", str(tmp_path / "q.fa"), "--matrix", str(tmp_path / "m.matrix"),
            "--word-size", "7", "--gapopen", "20", "--gapextend", "5", "--min-raw-gapped-score", "200",
            "--xdrop-gap", "120", "--xdrop-gap-final", "400", "--complexity-adjust", "--mask-level", "80",
            "--dust", "no", "--outfmt", RM_FIELDS]
    g = run([CLI] + args)
    o = run([ORACLE_BIN, "--num-threads", "16"] + args, timeout=900)
    assert g.returncode == 0 and o.returncode == 0, (g.stderr, o.stderr)
    assert g.stdout == o.stdout and g.stdout.count("\n") > 100
