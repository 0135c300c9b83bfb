"""The oracle's seeding and ungapped stages against the reference's chr22
regressions (rmblast-lib/tests/dedup_seeds_regression.rs and
ungapped_regression.rs): chr22 split into NCBI query chunks, each searched
against the Alu of tests/data/alu.2bit with word size 14, X_ungapped 400,
min_raw_gapped_score 200, dust off.  247,629 post-dedup seeds (as a set) and
247,629 ungapped hits (content and order), both produced by instrumented NCBI
2.17.0.  The inputs (51 MB) live only in the reference checkout, so this test
runs where it is mounted and is skipped elsewhere."""
import os
import struct
import subprocess

import pytest

from conftest import ORACLE_BIN, REFERENCE, golden, run

CHR22 = os.path.join(REFERENCE, "tests", "data", "chr22.fa")
ALU = os.path.join(REFERENCE, "tests", "data", "alu.2bit")
UNG = os.path.join(REFERENCE, "rmblast-lib", "tests", "data", "chr22_ungapped_hits.txt")
SEEDS = os.path.join(REFERENCE, "rmblast-lib", "tests", "data", "chr22_dedup_seeds.txt")

pytestmark = pytest.mark.skipif(not all(os.path.exists(p) for p in (CHR22, ALU, UNG, SEEDS)),
                                reason="reference chr22 regression data not mounted")


def read_2bit(path):
    """UCSC .2bit (T=0 C=1 A=2 G=3; N blocks -> N), as seq/twobit.rs reads it."""
    data = open(path, "rb").read()
    sig, ver, count, _ = struct.unpack_from("<IIII", data, 0)
    assert sig == 0x1A412743
    off = 16
    out = []
    for _ in range(count):
        n = data[off]
        name = data[off + 1:off + 1 + n].decode()
        (pos,) = struct.unpack_from("<I", data, off + 1 + n)
        off += 1 + n + 4
        (size, nb) = struct.unpack_from("<II", data, pos)
        p = pos + 8
        nstarts = struct.unpack_from(f"<{nb}I", data, p); p += 4 * nb
        nsizes = struct.unpack_from(f"<{nb}I", data, p); p += 4 * nb
        (mb,) = struct.unpack_from("<I", data, p); p += 4 + 8 * mb + 4
        seq = []
        for i in range(size):
            b = (data[p + i // 4] >> (6 - 2 * (i % 4))) & 3
            seq.append("TCAG"[b])
        for st, sz in zip(nstarts, nsizes):
            for i in range(st, st + sz):
                seq[i] = "N"
        out.append((name, "".join(seq)))
    return out


def chunks(n):  # rmblastn/src/main.rs:1149-1193, as the regression tests compute them
    step0 = 1_000_000 - 100
    k = n // step0
    if k <= 1:
        return [(0, n)]
    cs = (n + (k - 1) * 100) // k
    if k < cs - 100:
        cs += 1
    step = cs - 100
    return [(i * step, min(i * step + cs, n)) for i in range(k) if i * step < n]


@pytest.fixture(scope="module")
def dumps(tmp_path_factory):
    d = tmp_path_factory.mktemp("chr22")
    seq = "".join(l.strip() for l in open(CHR22) if not l.startswith(">"))
    with open(d / "q.fa", "w") as f:
        for i, (a, b) in enumerate(chunks(len(seq))):
            f.write(f">c{i}\n{seq[a:b]}\n")
    alu = read_2bit(ALU)
    with open(d / "db.fa", "w") as f:
        for name, s in alu:
            f.write(f">{name}\n{s}\n")
    out = {}
    for kind in ("ungapped", "seeds"):
        r = run([ORACLE_BIN, "--db", str(d / "db.fa"), "--query", str(d / "q.fa"), "--matrix",
                 golden("comparison.matrix"), "--gapopen", "20", "--gapextend", "5", "--word-size", "14",
                 "--xdrop-ungap", "400", "--xdrop-gap", "100", "--xdrop-gap-final", "200",
                 "--min-raw-gapped-score", "200", "--dust", "no", "--dump", kind], timeout=1200)
        assert r.returncode == 0, r.stderr
        out[kind] = r.stdout
    out["s_len"] = {name: len(s) for name, s in alu}
    return out


def test_chr22_ungapped_hits_content_and_order(dumps):
    got = []
    for line in dumps["ungapped"].splitlines():
        q, s, strand, qs, qe, ss, se, sc, _, _ = line.split("\t")
        qs, qe, ss, se = int(qs), int(qe), int(ss), int(se)
        if strand == "-":  # record frame -> forward subject (engine.rs:1918-1921)
            L = dumps["s_len"][s]
            ss, se = L - se, L - ss
        got.append(f"UNGAP strand={strand} qs={qs} qe={qe} ss={ss} se={se} score={sc}")
    want = open(UNG).read().splitlines()
    assert len(got) == len(want) == 247629
    first = next((i for i, (a, b) in enumerate(zip(got, want)) if a != b), None)
    assert first is None, (first, got[first], want[first])


def test_chr22_dedup_seeds_multiset(dumps):
    # dedup_seeds_regression.rs:121-139 compares the sorted lists
    got = []
    for line in dumps["seeds"].splitlines():
        q, s, qo, so, strand = line.split("\t")
        got.append(f"SEED strand={strand} q={qo} s={so}")
    got.sort()
    want = open(SEEDS).read().splitlines()
    assert len(got) == len(want) == 247629
    assert got == want
