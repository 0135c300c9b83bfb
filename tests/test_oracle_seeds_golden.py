"""The oracle's post-dedup seeds on chunks 30 and 31 of chr22 (the committed
3 Mbp fixture) against the reference's dedup-seeds golden restricted to them
(rmblast-lib/tests/data/chr22_dedup_seeds.txt, tests/golden/
chr22_seeds_chunks30_31.txt.gz), as sorted lists per chunk
(dedup_seeds_regression.rs:121-139 compares sorted lists)."""
import gzip

from conftest import CHR22_3MB, ORACLE_BIN, golden, run


def test_oracle_seeds_on_chr22_chunks_match_golden(tmp_path):
    want = {}
    for line in gzip.open(golden("chr22_seeds_chunks30_31.txt.gz"), "rt"):
        if not line.startswith("#"):
            c, l = line.rstrip("\n").split("\t")
            want.setdefault(c, []).append(l)
    seq = "".join(l.strip() for l in gzip.open(golden(CHR22_3MB + ".gz"), "rt") if not l.startswith(">"))
    with open(tmp_path / "q.fa", "w") as f:
        for c, (a, b) in {"c30": (491_040, 1_507_508), "c31": (1_507_408, 2_523_876)}.items():
            f.write(f">{c}\n{seq[a:b]}\n")
    r = run([ORACLE_BIN, "--db", golden("alu.fa"), "--query", str(tmp_path / "q.fa"), "--matrix",
             golden("comparison.matrix"), "--gapopen", "20", "--gapextend", "5", "--word-size", "14",
             "--xdrop-ungap", "400", "--xdrop-gap", "100", "--xdrop-gap-final", "200",
             "--min-raw-gapped-score", "200", "--dust", "no", "--dump", "seeds"], timeout=600)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in r.stdout.splitlines():
        q, _, qo, so, strand = line.split("\t")
        got.setdefault(q, []).append(f"SEED strand={strand} q={qo} s={so}")
    assert {c: sorted(v) for c, v in got.items()} == want
    assert sum(len(v) for v in want.values()) == 20860
