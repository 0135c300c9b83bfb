"""Symmetric DUST (host code of the library, rmb_dust_intervals) against the
known answers in the reference's tests (rmblast-lib/src/filter/dust.rs:
671-846): NCBI dustmasker intervals for two chr22 regions (fixtures copied to
tests/golden), the AluY poly-A tail, a chr22 region cut from tests/data/chr22.fa
of the reference checkout where it is mounted, and the trivial cases."""
import os

import numpy as np
import pytest

from conftest import REFERENCE, golden

CODE = {ord(c): v for c, v in zip("ACGTacgt", [0, 1, 2, 3, 0, 1, 2, 3])}


def encode(text):
    return np.array([CODE.get(b, 14) for b in text.encode()], dtype=np.uint8)


def fasta_bases(path):
    return encode("".join(l.strip() for l in open(path) if not l.startswith(">")))


def dust(bases):
    import rmblast_amd
    return rmblast_amd.dust_intervals(bases)


def test_chr22_regions_match_ncbi_dustmasker():
    # dust.rs:761-799
    assert dust(fasta_bases(golden("chr22_39030802-39031701.fa"))) == [(366, 387), (612, 618), (816, 889)]
    assert dust(fasta_bases(golden("chr22_40060802-40061701.fa"))) == [(580, 592), (800, 891)]


def test_aluy_poly_a_tail_is_masked():
    # dust.rs:745-758: NCBI dustmasker reports [281, 310]
    aluy = ("GGCCGGGCGCGGTGGCTCACGCCTGTAATCCCAGCACTTTGGGAGGCCGAGGCGGGCGGATCACGAGGTCAGGAGATCGAGACCATCCTGGCTAAC"
            "ACGGTGAAACCCCGTCTCTACTAAAAATACAAAAAATTAGCCGGGCGTGGTGGCGGGCGCCTGTAGTCCCAGCTACTCGGGAGGCTGAGGCAGGAGA"
            "ATGGCGTGAACCCGGGAGGCGGAGCTTGCAGTGAGCCGAGATCGCGCCACTGCACTCCAGCCTGGGCGACAGAGCGAGACTCCGTCTCAAAAAAAAA"
            "AAAAAAAAAAAAAAAAAAAAA")
    iv = dust(encode(aluy))
    assert any(s <= 281 and e >= 310 for s, e in iv), iv


def test_trivial_cases():
    assert dust(np.zeros(100, dtype=np.uint8))          # pure A repeat is masked
    diverse = [0, 1, 2, 3, 0, 2, 1, 3, 0, 3, 1, 2, 1, 0, 3, 2, 1, 3, 0, 2, 2, 0, 3, 1, 2, 1, 0, 3, 3, 0, 2, 1,
               3, 1, 0, 2, 3, 2, 0, 1, 0, 0, 1, 1, 2, 2, 3, 3, 0, 1, 3, 2, 1, 2, 0, 3, 2, 3, 1, 0, 3, 3, 2, 1]
    assert dust(np.array(diverse, dtype=np.uint8)) == []
    assert dust(np.zeros(2, dtype=np.uint8)) == []


CHR22 = os.path.join(REFERENCE, "tests", "data", "chr22.fa")


@pytest.mark.skipif(not os.path.exists(CHR22), reason="reference chr22.fa not mounted")
def test_chr22_region_10510100_matches_ncbi():
    # dust.rs:673-697: NCBI dustmasker 10510356-10510364 and 10510518-10510527
    want_len, start = 10510600, 10510100
    parts, n = [], 0
    with open(CHR22) as f:
        for line in f:
            if line.startswith(">"):
                continue
            s = line.strip()
            parts.append(s)
            n += len(s)
            if n >= want_len:
                break
    seq = encode("".join(parts)[start:want_len])
    iv = dust(seq)
    assert (10510356 - start, 10510364 - start) in iv and (10510518 - start, 10510527 - start) in iv
