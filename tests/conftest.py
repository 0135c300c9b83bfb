import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_BIN = os.path.join(ROOT, "oracle", "_build", "rmblastn_oracle")
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
CLI = os.path.join(ROOT, "rmblast_amd", "_build", "rmblastn")
REFERENCE = "/root/reference"

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

RM_FIELDS = ("6 score perc_sub perc_query_gap perc_db_gap qseqid qstart qend qlen sstrand sseqid sstart send "
             "slen kdiv cpg_kdiv transi transv cpg_sites qseq sseq")
# The parameters every reference fixture run shares (rmblastn/tests/cli.rs:107-117).
RM_SEARCH = ["--gapopen", "20", "--gapextend", "5", "--xdrop-ungap", "400", "--xdrop-gap-final", "200",
             "--xdrop-gap", "100", "--min-raw-gapped-score", "200"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def golden(name):
    return os.path.join(GOLDEN, name)


def golden_unzipped(name, dest):
    """Decompress tests/golden/<name>.gz into directory `dest`; returns the path."""
    import gzip
    out = os.path.join(str(dest), name)
    with gzip.open(golden(name + ".gz"), "rb") as f, open(out, "wb") as g:
        g.write(f.read())
    return out


# chr22:30,000,000-33,000,000 of the reference's tests/data/chr22.fa, the query
# of rmblastn/tests/cli.rs:252-274 (three chunks of the split query path)
CHR22_3MB = "chr22_30000000_33000000.fa"
CHUNKED_SEARCH = ["--matrix", golden("comparison.matrix"), "--gapopen", "20", "--gapextend", "5",
                  "--word-size", "7", "--min-raw-gapped-score", "200", "--mask-level", "80",
                  "--outfmt", "6 score qseqid qstart qend sstrand sseqid sstart send"]


def run(cmd, timeout=600, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=e)
    return r


@pytest.fixture(scope="session")
def oracle_lib():
    if not os.path.exists(ORACLE_LIB):
        pytest.fail("oracle not built: run `python -m rmblast_amd.build` (or make -C oracle)")
    L = ctypes.CDLL(ORACLE_LIB)
    P = ctypes.POINTER
    L.orc_extend_ungapped.argtypes = [P(ctypes.c_uint8), ctypes.c_size_t, P(ctypes.c_uint8), ctypes.c_size_t,
                                      ctypes.c_uint32, ctypes.c_uint32, P(ctypes.c_int32), ctypes.c_int,
                                      P(ctypes.c_int32)]
    L.orc_align_score_only.argtypes = [P(ctypes.c_uint8), P(ctypes.c_uint8), ctypes.c_size_t, ctypes.c_size_t,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int32), ctypes.c_int,
                                       P(ctypes.c_int64)]
    L.orc_align_ex.argtypes = [P(ctypes.c_uint8), P(ctypes.c_uint8), ctypes.c_size_t, ctypes.c_size_t,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int32), ctypes.c_int,
                               P(ctypes.c_int64), P(ctypes.c_uint32), ctypes.c_int]
    L.orc_diag_hash_run.argtypes = [P(ctypes.c_int32), ctypes.c_int, ctypes.c_int, P(ctypes.c_uint32)]
    L.orc_ncbi_random.argtypes = [ctypes.c_uint32, P(ctypes.c_uint8), ctypes.c_int, P(ctypes.c_uint8)]
    L.orc_ungapped_hits.argtypes = [P(ctypes.c_uint8), ctypes.c_size_t, P(ctypes.c_uint8), ctypes.c_size_t,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int32),
                                    P(ctypes.c_uint32), ctypes.c_int]
    L.orc_adjusted_score.argtypes = [ctypes.c_char_p, ctypes.c_int32, P(ctypes.c_uint8), ctypes.c_int,
                                     P(ctypes.c_double)]
    L.orc_mask_level.argtypes = [P(ctypes.c_uint32), ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                 P(ctypes.c_uint8)]
    L.orc_purge_prelims.argtypes = [P(ctypes.c_uint32), ctypes.c_int, ctypes.c_uint32, P(ctypes.c_uint8)]
    L.orc_gapped_bidir.argtypes = [P(ctypes.c_uint8), ctypes.c_size_t, P(ctypes.c_uint8), ctypes.c_size_t,
                                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   P(ctypes.c_int32), P(ctypes.c_int64)]
    L.orc_improve_seed.argtypes = [P(ctypes.c_uint8), ctypes.c_size_t, P(ctypes.c_uint8), ctypes.c_size_t,
                                   P(ctypes.c_uint32), ctypes.c_int, P(ctypes.c_uint32)]
    L.orc_chunk_plan.argtypes = [ctypes.c_uint64, P(ctypes.c_uint64)]
    L.orc_merge_chunk_prelims.argtypes = [P(ctypes.c_uint32), ctypes.c_int, P(ctypes.c_uint32), ctypes.c_int,
                                          ctypes.c_uint32, P(ctypes.c_uint32)]
    L.orc_merge_chunk_prelims.restype = ctypes.c_int
    L.orc_lookup.argtypes = [P(ctypes.c_uint8), ctypes.c_size_t, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_uint32),
                             ctypes.c_size_t, P(ctypes.c_uint32), ctypes.c_size_t]
    L.orc_lookup.restype = ctypes.c_int64
    L.orc_seeds.argtypes = [P(ctypes.c_uint8), ctypes.c_size_t, P(ctypes.c_uint8), ctypes.c_size_t, ctypes.c_int,
                            ctypes.c_int, ctypes.c_int, P(ctypes.c_int32), P(ctypes.c_uint32), ctypes.c_int]
    L.orc_prelims.argtypes = [P(ctypes.c_uint8), ctypes.c_size_t, P(ctypes.c_uint8), ctypes.c_size_t, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int32),
                              P(ctypes.c_uint32), ctypes.c_int]
    return L


def ptr(a, ct=ctypes.c_uint8):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def simple_matrix(match, mismatch, other=0, sentinel=None):
    """16x16 BLASTNA matrix: match/mismatch on ACGT, `other` elsewhere."""
    m = np.full((16, 16), other, dtype=np.int32)
    for i in range(4):
        for j in range(4):
            m[i, j] = match if i == j else mismatch
    if sentinel is not None:
        m[15, :] = sentinel
        m[:, 15] = sentinel
    return np.ascontiguousarray(m)


def seq_from_fasta(path):
    with open(path) as f:
        return "".join(l.strip() for l in f if not l.startswith(">"))


def write_2bit(path, records, swapped=False):
    """UCSC .2bit file of (name, bases) records: bases as text, N runs stored
    as N blocks, lower case as mask blocks, the rest packed T C A G = 0 1 2 3
    from the high bits (the layout rmblast-lib/src/seq/twobit.rs:1-12 reads)."""
    import re
    import struct
    e = ">" if swapped else "<"
    code = {"T": 0, "C": 1, "A": 2, "G": 3, "N": 0}
    names = [n.encode() for n, _ in records]
    off = 16 + sum(1 + len(n) + 4 for n in names)
    index, body = b"", b""
    for (name, seq), nb in zip(records, names):
        up = seq.upper()
        nblk = [(m.start(), m.end() - m.start()) for m in re.finditer("N+", up)]
        mblk = [(m.start(), m.end() - m.start()) for m in re.finditer("[a-z]+", seq)]
        rec = struct.pack(e + "II", len(seq), len(nblk))
        rec += b"".join(struct.pack(e + "I", s) for s, _ in nblk) + b"".join(struct.pack(e + "I", z) for _, z in nblk)
        rec += struct.pack(e + "I", len(mblk))
        rec += b"".join(struct.pack(e + "I", s) for s, _ in mblk) + b"".join(struct.pack(e + "I", z) for _, z in mblk)
        rec += struct.pack(e + "I", 0)
        packed = bytearray((len(seq) + 3) // 4)
        for i, ch in enumerate(up):
            packed[i // 4] |= code[ch] << (6 - 2 * (i % 4))
        index += struct.pack("B", len(nb)) + nb + struct.pack(e + "I", off + len(body))
        body += rec + bytes(packed)
    with open(path, "wb") as f:
        f.write(struct.pack(e + "IIII", 0x1A412743, 0, len(records), 0) + index + body)
