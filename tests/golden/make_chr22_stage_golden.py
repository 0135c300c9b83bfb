"""Writes tests/golden/chr22_stage_chunks30_31.txt.gz: the reference's chr22
ungapped-stage golden (rmblast-lib/tests/data/chr22_ungapped_hits.txt,
content and order) restricted to query chunks 30 and 31 of chr22, the two
chunks inside the committed chr22:30,000,000-33,000,000 fixture; and
tests/golden/chr22_seeds_chunks30_31.txt.gz: the post-dedup seeds golden
(rmblast-lib/tests/data/chr22_dedup_seeds.txt, a multiset) restricted to the
same chunks, one "chunk<TAB>SEED ..." line per seed, sorted per chunk.

The seeds golden does not say which chunk a seed belongs to; the oracle's
seeds of the two chunks are checked to be a sub-multiset of it.

The golden lists every chunk's hits one chunk after another; the block of
chunks 30-31 is found by running the oracle on those two chunks and locating
its lines in the golden, which must hold them contiguously and identically
(the oracle's full-chromosome run matches the golden line for line:
tests/test_oracle_chr22.py).  Run from the repository root where the reference checkout is
mounted; the output is data (the golden's hit lines), not reference code.
"""
import collections
import gzip
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = "/root/reference"
CHR22 = os.path.join(REF, "tests", "data", "chr22.fa")
ALU = os.path.join(REF, "tests", "data", "alu.2bit")
UNG = os.path.join(REF, "rmblast-lib", "tests", "data", "chr22_ungapped_hits.txt")
SEEDS = os.path.join(REF, "rmblast-lib", "tests", "data", "chr22_dedup_seeds.txt")
sys.path.insert(0, os.path.join(ROOT, "tests"))
from test_oracle_chr22 import chunks, read_2bit  # noqa: E402

FIXTURE_START = 30_000_000   # tests/golden/chr22_30000000_33000000.fa.gz = chr22[30,000,000:33,000,000]
CHUNKS = (30, 31)


def main():
    seq = "".join(l.strip() for l in open(CHR22) if not l.startswith(">"))
    plan = chunks(len(seq))
    d = tempfile.mkdtemp()
    with open(os.path.join(d, "q.fa"), "w") as f:
        for k in CHUNKS:
            a, b = plan[k]
            assert FIXTURE_START <= a and b <= FIXTURE_START + 3_000_000
            f.write(f">c{k}\n{seq[a:b]}\n")
    alu = read_2bit(ALU)
    with open(os.path.join(d, "db.fa"), "w") as f:
        for name, s in alu:
            f.write(f">{name}\n{s}\n")
    lens = {name: len(s) for name, s in alu}
    out = {}
    for kind in ("ungapped", "seeds"):
        r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "rmblastn_oracle"), "--db", os.path.join(d, "db.fa"),
                            "--query", os.path.join(d, "q.fa"), "--matrix",
                            os.path.join(ROOT, "tests", "golden", "comparison.matrix"), "--gapopen", "20",
                            "--gapextend", "5", "--word-size", "14", "--xdrop-ungap", "400", "--xdrop-gap", "100",
                            "--xdrop-gap-final", "200", "--min-raw-gapped-score", "200", "--dust", "no",
                            "--dump", kind], capture_output=True, text=True, check=True)
        out[kind] = r.stdout.splitlines()
    ung = []
    for line in out["ungapped"]:
        q, s, strand, qs, qe, ss, se, sc, _, _ = line.split("\t")
        ss, se = int(ss), int(se)
        if strand == "-":
            ss, se = lens[s] - se, lens[s] - ss
        ung.append(f"UNGAP strand={strand} qs={qs} qe={qe} ss={ss} se={se} score={sc}")
    gold = open(UNG).read().splitlines()
    at = gold.index(ung[0])
    block = gold[at:at + len(ung)]
    assert block == ung, "the oracle's hits of these chunks are not a contiguous block of the golden"
    path = os.path.join(ROOT, "tests", "golden", "chr22_stage_chunks30_31.txt.gz")
    with gzip.open(path, "wt") as f:
        f.write(f"# chunks {CHUNKS}: {[plan[k] for k in CHUNKS]} (chr22 coordinates), golden lines "
                f"{at}..{at + len(block)}\n")
        for l in block:
            f.write(l + "\n")
    print(f"{len(block)} ungapped hits (golden lines {at}..{at + len(block)}) -> {path}")
    seeds = sorted((q, f"SEED strand={strand} q={qo} s={so}") for q, _, qo, so, strand in
                   (l.split("\t") for l in out["seeds"]))
    have = collections.Counter(open(SEEDS).read().splitlines())
    need = collections.Counter(l for _, l in seeds)
    assert not need - have, "the oracle's seeds of these chunks are not all in the golden"
    path = os.path.join(ROOT, "tests", "golden", "chr22_seeds_chunks30_31.txt.gz")
    with gzip.open(path, "wt") as f:
        f.write(f"# chunks {CHUNKS}: {[plan[k] for k in CHUNKS]} (chr22 coordinates); query offsets chunk-relative, "
                f"as in the golden; every line is in the golden\n")
        for q, l in seeds:
            f.write(f"{q}\t{l}\n")
    print(f"{len(seeds)} seeds -> {path}")


if __name__ == "__main__":
    main()
