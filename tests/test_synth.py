"""The synthetic workload generator: seeded, 41% GC, sizes as configured."""
import numpy as np

from rmblast_amd import synth


def test_parity_config_is_seeded_and_sized():
    a = synth.parity_config(seed=3, genome_len=200_000, n_families=10, n_copies=100)
    b = synth.parity_config(seed=3, genome_len=200_000, n_families=10, n_copies=100)
    assert np.array_equal(a.records[0][1], b.records[0][1])
    g = a.records[0][1]
    assert len(g) == 200_000 and g.max() < 4
    gc = np.isin(g, [1, 2]).mean()
    assert 0.38 < gc < 0.44
    assert all(300 <= len(f[1]) <= 6000 for f in a.families)


def test_mutation_rates_are_close_to_requested():
    rng = np.random.default_rng(0)
    s = synth.random_bases(rng, 200_000)
    m = synth.mutate(rng, s, 0.15, 0.0)
    assert len(m) == len(s)
    d = (m != s).mean()
    assert 0.13 < d < 0.17
    sub = s != m
    ts = (synth.TRANSITION[s[sub]] == m[sub]).mean()
    assert 0.62 < ts < 0.80          # 2:1, raised by the CpG transitions at 5x
    m2 = synth.mutate(rng, s, 0.0, 0.02)
    assert abs(len(m2) - len(s)) < 0.02 * len(s)


def test_scale_config_shape_small():
    w = synth.scale_config(seed=5, n_records=4, record_len=10_000, n_families=20, n_copies=50)
    assert len(w.records) == 4 and all(len(r[1]) == 10_000 for r in w.records)
    assert all(100 <= len(f[1]) <= 15000 for f in w.families)
