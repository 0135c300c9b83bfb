"""Seeded synthetic genome/library generator for the benchmark configurations.

Sequences are numpy uint8 arrays of BLASTNA codes (A0 C1 G2 T3).  Background
composition is A .295 C .205 G .205 T .295 (41% GC).  A copy of a family
fragment carries substitutions at a per-copy rate (transition:transversion
2:1, the C of a CpG mutating to T, and the G to A, at 5x the base rate) and
indels at a per-copy rate with geometric lengths of mean 3, and is inserted on
a random strand over the background at a random position (later copies may
cover parts of earlier ones, so the genome keeps its stated length).
"""
from __future__ import annotations

import dataclasses

import numpy as np

FREQ = np.array([0.295, 0.205, 0.205, 0.295])
COMP = np.array([3, 2, 1, 0], dtype=np.uint8)
TRANSITION = np.array([2, 3, 0, 1], dtype=np.uint8)        # A<->G, C<->T
TRANSVERSIONS = np.array([[1, 3], [0, 2], [1, 3], [0, 2]], dtype=np.uint8)

# The int 4x4 matrix of the configurations: +9 match, -6 transition,
# -15 transversion, N -1 (everything against N), with the background FREQS.
MATRIX_TEXT = """# FREQS A 0.295 C 0.205 G 0.205 T 0.295
   A   C   G   T   N
A  9 -15  -6 -15  -1
C -15  9 -15  -6  -1
G -6 -15   9 -15  -1
T -15 -6 -15   9  -1
N -1  -1  -1  -1  -1
"""


@dataclasses.dataclass
class Workload:
    records: list          # list of (id, uint8 array)
    families: list         # list of (id, uint8 array)
    inserts: list          # (record index, start, end, family index, strand, divergence)


def random_bases(rng, n):
    return rng.choice(4, size=n, p=FREQ).astype(np.uint8)


def mutate(rng, seq, sub_rate, indel_rate):
    """Substitutions (Ts:Tv 2:1, CpG 5x) then indels of geometric length (mean 3)."""
    s = seq.copy()
    n = len(s)
    if n == 0:
        return s
    cpg = np.zeros(n, dtype=bool)
    is_c = s[:-1] == 1
    is_g = s[1:] == 2
    both = is_c & is_g
    cpg[:-1] |= both
    cpg[1:] |= both
    w = np.where(cpg, 5.0, 1.0)
    p = np.minimum(sub_rate * w / w.mean(), 0.95)
    hit = rng.random(n) < p
    idx = np.nonzero(hit)[0]
    if len(idx):
        cur = s[idx]
        ts = rng.random(len(idx)) < (2.0 / 3.0)
        ts |= cpg[idx]                     # CpG decay is C->T / G->A, both transitions
        tv_pick = TRANSVERSIONS[cur, rng.integers(0, 2, len(idx))]
        s[idx] = np.where(ts, TRANSITION[cur], tv_pick)
    n_indel = rng.binomial(n, indel_rate)
    if n_indel == 0:
        return s
    pos = np.sort(rng.integers(0, n, n_indel))
    lens = rng.geometric(1.0 / 3.0, n_indel)
    ins = rng.random(n_indel) < 0.5
    out = []
    last = 0
    for p_, l_, i_ in zip(pos, lens, ins):
        if p_ < last:
            continue
        out.append(s[last:p_])
        if i_:
            out.append(random_bases(rng, int(l_)))
            last = p_
        else:
            last = min(n, p_ + int(l_))
    out.append(s[last:])
    return np.concatenate(out) if out else s


def make_families(rng, n, lengths):
    return [(f"fam{i:05d}", random_bases(rng, int(L))) for i, L in enumerate(lengths)]


def insert_copies(rng, genome_len, families, n_copies, frag_range, div_range, indel_range):
    """Returns the genome and insert records.  Each copy overwrites the
    background at a uniformly random position, so a later copy can cover part
    of an earlier one and the genome keeps exactly genome_len bases."""
    g = random_bases(rng, genome_len)
    inserts = []
    for _ in range(n_copies):
        f = int(rng.integers(0, len(families)))
        L = len(families[f][1])
        lo, hi = frag_range
        fl = int(rng.integers(min(lo, L), min(hi, L) + 1))
        off = int(rng.integers(0, L - fl + 1))
        div = float(rng.uniform(*div_range))
        ind = float(rng.uniform(*indel_range))
        strand = int(rng.integers(0, 2))
        piece = mutate(rng, families[f][1][off:off + fl], div, ind)
        if strand:
            piece = COMP[piece[::-1]]
        piece = piece[:genome_len]
        pos = int(rng.integers(0, genome_len - len(piece) + 1))
        g[pos:pos + len(piece)] = piece
        inserts.append((pos, pos + len(piece), f, strand, div))
    return g, inserts


def parity_config(seed=1, genome_len=1_000_000, n_families=50, n_copies=2000):
    """The 1 Mbp x 50-family parity run (families 300-6,000 bp; fragments
    100-3,000 bp; 5-25% substitutions, 0.5-3% indels)."""
    rng = np.random.default_rng(seed)
    fams = make_families(rng, n_families, rng.integers(300, 6001, n_families))
    g, ins = insert_copies(rng, genome_len, fams, n_copies, (100, 3000), (0.05, 0.25), (0.005, 0.03))
    return Workload([("chr1", g)], fams, [(0,) + x for x in ins])


def scale_config(seed=2, n_records=1667, record_len=60_000, n_families=1000, n_copies=200_000,
                 frag_range=(50, 1000)):
    """The 100 Mbp (1,667 x 60 kbp) x 1,000-family configuration: family lengths
    log-normal (median 1,225 bp, sigma 1.0) truncated to 100-15,000 bp; copies at
    2-35% divergence with 0.5-3% indels; fragment lengths uniform in frag_range
    (about one genome's worth of copies in total)."""
    rng = np.random.default_rng(seed)
    L = np.exp(rng.normal(np.log(1225.0), 1.0, n_families))
    L = np.clip(L, 100, 15000).astype(int)
    fams = make_families(rng, n_families, L)
    g, ins = insert_copies(rng, n_records * record_len, fams, n_copies, frag_range, (0.02, 0.35), (0.005, 0.03))
    recs = [(f"chunk{i:05d}", g[i * record_len:(i + 1) * record_len]) for i in range(n_records)]
    inserts = [(s // record_len, s, e, f, st, d) for (s, e, f, st, d) in ins]
    return Workload(recs, fams, inserts)


def gappy_config(seed=4, genome_len=10_000_000, n_families=300, n_copies=4000, frag_range=(300, 5000),
                 tract_frac=0.05):
    """The divergent/gappy stress configuration: one 10 Mbp record (the split
    query path) against 300 families of 5-15 kbp; copies 25-40% diverged with
    5-8% indels, fragments uniform in frag_range; then microsatellite tracts
    (period 2-6, 50-500 bp) overwrite tract_frac of the genome."""
    rng = np.random.default_rng(seed)
    fams = make_families(rng, n_families, rng.integers(5000, 15001, n_families))
    g, ins = insert_copies(rng, genome_len, fams, n_copies, frag_range, (0.25, 0.40), (0.05, 0.08))
    covered = 0
    while covered < tract_frac * genome_len:
        n = int(rng.integers(50, 501))
        unit = random_bases(rng, int(rng.integers(2, 7)))
        at = int(rng.integers(0, genome_len - n))
        g[at:at + n] = np.tile(unit, n // len(unit) + 1)[:n]
        covered += n
    return Workload([("chr1", g)], fams, [(0,) + x for x in ins])


def uneven_config(seed=3, n_records=20_000, n_families=5_000, rec_range=(1_000, 500_000),
                  fam_range=(50, 20_000), copies_per_kbp=2.0, frag_range=(50, 1000)):
    """The uneven-batching configuration: records of log-uniform lengths in
    rec_range (1 kbp - 500 kbp) against families of log-uniform lengths in
    fam_range (50 - 20,000 bp); each record carries about copies_per_kbp
    copies per kbp (the density of the exact-path-at-scale genome), at
    2-35% divergence with 0.5-3% indels.  Family i and record i each come from
    their own generator, so the first k records (or families) of a smaller
    configuration are those of the full one: a subset is generated without the
    whole (20,000 records are about 1.6 Gbp)."""
    def log_uniform(rng, lo, hi):
        return int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
    fams = []
    for i in range(n_families):
        rng = np.random.default_rng([seed, 1, i])
        fams.append((f"fam{i:05d}", random_bases(rng, log_uniform(rng, *fam_range))))
    recs, inserts = [], []
    for r in range(n_records):
        rng = np.random.default_rng([seed, 2, r])
        n = log_uniform(rng, *rec_range)
        g, ins = insert_copies(rng, n, fams, int(rng.poisson(copies_per_kbp * n / 1000.0)), frag_range,
                               (0.02, 0.35), (0.005, 0.03))
        recs.append((f"rec{r:05d}", g))
        inserts += [(r,) + x for x in ins]
    return Workload(recs, fams, inserts)


def subset(w: Workload, n_records: int, n_families: int | None = None) -> Workload:
    fams = w.families if n_families is None else w.families[:n_families]
    return Workload(w.records[:n_records], fams, [x for x in w.inserts if x[0] < n_records])


def to_fasta(items, path, width=60):
    letters = np.frombuffer(b"ACGTRYMKWSBDHVN-", dtype=np.uint8)
    with open(path, "w") as f:
        for name, seq in items:
            f.write(f">{name}\n")
            s = letters[seq].tobytes().decode()
            for i in range(0, len(s), width):
                f.write(s[i:i + width] + "\n")
