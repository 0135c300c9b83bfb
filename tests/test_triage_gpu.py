"""Learned triage on the GPU, through the C ABI: the BF16 MFMA GEMM against
numpy for every operand layout, the k-mer profiles, the model's logits, loss
and gradients against the CPU oracle (oracle/triage_ref.py, same rounding
points), training reduces the loss, and the pair search used by the pruned
path returns exactly the full search's HSPs for the pairs it is given."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import triage_ref as T  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model():
    from rmblast_amd.triage import Triage
    m = Triage(seed=3)
    yield m
    m.close()


def _bf16_rand(rng, shape, scale=1.0):
    return T.bf16_bits(rng.standard_normal(shape).astype(np.float32) * scale)


@pytest.mark.parametrize("ta,tb", [(False, True), (True, False), (False, False), (True, True)])
@pytest.mark.parametrize("M,N,K", [(70, 45, 100), (64, 64, 32), (5, 130, 2080), (129, 7, 3)])
def test_gemm_tap_matches_numpy(model, ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = _bf16_rand(rng, (K, M) if ta else (M, K))
    B = _bf16_rand(rng, (N, K) if tb else (K, N))
    C = model.tap_gemm(ta, tb, M, N, K, A, A.shape[1], B, B.shape[1])
    a = T.from_bf16_bits(A).astype(np.float64)
    b = T.from_bf16_bits(B).astype(np.float64)
    ref = (a.T if ta else a) @ (b.T if tb else b)
    np.testing.assert_allclose(C, ref, rtol=1e-4, atol=1e-4 * np.sqrt(K))


def test_kmer_profiles_match_oracle(model):
    rng = np.random.default_rng(11)
    seq = rng.integers(0, 4, 9000).astype(np.uint8)
    seq[rng.choice(9000, 40, replace=False)] = 14
    starts = np.array([0, 1000, 2500, 7000, 8995, 4000], dtype=np.uint64)
    lens = np.array([2000, 2000, 3000, 2000, 5, 0], dtype=np.uint32)
    got = model.profiles(seq, starts, lens)
    for k in range(len(starts)):
        ref = T.bf16_bits(T.kmer_profile(seq, int(starts[k]), int(lens[k])))
        diff = np.abs(got[k].astype(np.int32) - ref.astype(np.int32))
        assert diff.max() <= 1, k  # log1pf and numpy's log1p may differ by one BF16 step


def test_forward_backward_matches_oracle(model):
    rng = np.random.default_rng(2)
    nw, nf = 37, 5
    xw = T.bf16_bits(np.log1p(rng.poisson(1.0, (nw, T.KMER_DIM))).astype(np.float32))
    xf = T.bf16_bits(np.log1p(rng.poisson(1.5, (nf, T.KMER_DIM))).astype(np.float32))
    lab = (rng.random(nw * nf) < 0.2).astype(np.uint8)
    lab[::7] = 2  # ignored pairs
    flat = T.init_params(4) * 0.5
    flat[T.layout()[0]["a"][0]] = 0.05
    flat[T.layout()[0]["c"][0]] = 0.7
    model.set_params(flat)
    loss, grads = model.forward_backward(xw, xf, lab, pos_weight=4.0, want_grads=True)
    logits = model.logits(xw, xf).reshape(-1)
    z, loss_ref, g_ref = T.forward_backward(flat, T.from_bf16_bits(xw), T.from_bf16_bits(xf), lab, pos_weight=4.0)
    np.testing.assert_allclose(logits, z, rtol=2e-3, atol=2e-3)
    assert abs(loss - loss_ref) <= 1e-3 * max(1.0, abs(loss_ref))
    off, _ = T.layout()
    for k, (o, shape) in off.items():
        n = int(np.prod(shape)) if shape else 1
        a, b = grads[o:o + n].astype(np.float64), g_ref[o:o + n].astype(np.float64)
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 2e-2, (k, rel)


def test_training_reduces_loss(model):
    rng = np.random.default_rng(9)
    nw, nf = 64, 8
    # positives: the family profile is mixed into the window profile
    fam = rng.poisson(2.0, (nf, T.KMER_DIM))
    win = rng.poisson(1.0, (nw, T.KMER_DIM))
    lab = np.zeros((nw, nf), dtype=np.uint8)
    for i in range(nw):
        j = rng.integers(0, nf)
        if rng.random() < 0.5:
            win[i] += fam[j]
            lab[i, j] = 1
    xw = T.bf16_bits(np.log1p(win).astype(np.float32))
    xf = T.bf16_bits(np.log1p(fam).astype(np.float32))
    model.set_params(T.init_params(6) * 0.3)
    first = None
    for step in range(60):
        loss, _ = model.forward_backward(xw, xf, lab, pos_weight=4.0)
        first = loss if first is None else first
        model.adam_step(1e-3)
    assert loss < 0.5 * first, (first, loss)


@pytest.fixture(scope="module")
def small_instance():
    import rmblast_amd
    from rmblast_amd import synth
    import tempfile
    w = synth.parity_config(seed=21, genome_len=120_000, n_families=12, n_copies=240)
    d = tempfile.mkdtemp()
    mp = os.path.join(d, "synth.matrix")
    open(mp, "w").write(synth.MATRIX_TEXT)
    m = rmblast_amd.load_matrix(mp)
    s = rmblast_amd.Searcher(m, gap_open=20, gap_extend=5, word_size=7, min_raw_gapped_score=200,
                             complexity_adjust=True)
    s.load_library([f[1] for f in w.families])
    return s, w


def test_search_pairs_equals_full_search_on_those_pairs(small_instance):
    from rmblast_amd.triage import hsp_key
    s, w = small_instance
    q = w.records[0][1]
    half = len(q) // 2
    queries = [q[:half], q[half:]]
    full, fops = s.search(queries)
    full_keys = {(int(h["query"]),) + hsp_key(h, fops) for h in full}
    nf = len(w.families)
    pq = [0, 0, 1, 1, 1, 0]
    ps = [0, 3, 1, 5, nf - 1, 3]  # duplicates are searched once
    got, gops = s.search_pairs(queries, pq, ps)
    got_keys = {(int(h["query"]),) + hsp_key(h, gops) for h in got}
    want = {k for k in full_keys if (k[0], k[1]) in set(zip(pq, ps))}
    assert got_keys == want
    allp, aops = s.search_pairs(queries, [i for i in range(2) for _ in range(nf)], list(range(nf)) * 2)
    assert {(int(h["query"]),) + hsp_key(h, aops) for h in allp} == full_keys


def test_pruned_search_rows_are_plain_search_rows(small_instance):
    """Pruning whole (record, family) pairs: with the true pair labels the
    rows are the plain search's, and with any subset of pairs they are
    exactly the plain search's rows of those pairs (nothing added)."""
    from rmblast_amd.triage import make_instance, pruned_search, row_keys
    s, w = small_instance
    inst = make_instance(s, None, 0, record=4000, genome=w)
    assert len(inst.recs) == 30 and len(inst.hsps) > 50
    full = row_keys(inst.hsps, inst.ops)
    keys, n, _ = pruned_search(s, inst, inst.pair_labels > 0)
    assert keys == full and n == int(inst.pair_labels.sum())
    rng = np.random.default_rng(4)
    for frac in (0.5, 0.1):
        passed = rng.random(inst.pair_labels.shape) < frac
        keys, _, _ = pruned_search(s, inst, passed)
        assert keys == {k for k in full if passed[k[0], k[1]]}
    keys, n, _ = pruned_search(s, inst, np.zeros_like(passed))
    assert keys == set() and n == 0


@pytest.fixture(scope="module")
def trained():
    """The network trained on 8 parity-configuration genomes (4 kbp records)
    and warm-started from the cosine logistic; 2 validation genomes and the
    held-out genome (seed 100000, 1 Mbp x 50 families x 2,000 copies)."""
    from rmblast_amd import triage as tr
    searcher = tr.make_searcher()
    model = tr.Triage(seed=5)
    train_set = [tr.make_instance(searcher, model, 500 + k) for k in range(8)]
    val_set = [tr.make_instance(searcher, model, 600 + k) for k in range(2)]
    held = tr.make_instance(searcher, model, 100_000)
    th = tr.fit_cosine_logistic(train_set, 4.0)
    tr.warm_start_from_cosine(model, th)
    tr.train(model, train_set, epochs=10, lr=1e-3, pos_weight=4.0, val=val_set)
    yield searcher, model, train_set, val_set, held, th
    model.close()


def test_trained_network_gains_over_cosine_only(trained):
    """The method without the network (a logistic fit on the cosine of the
    profiles) against the trained network warm-started from it: on held-out
    genomes the network's pair loss is lower."""
    from rmblast_amd import triage as tr
    _, model, _, val_set, _, th = trained
    cos_val = np.mean([tr._weighted_logistic_loss(th[0] * tr.cosine_scores(v) + th[1], v.train_labels, 4.0)
                       for v in val_set])
    net_val = np.mean([tr._weighted_logistic_loss(model.logits(v.xw, v.xf).astype(np.float64), v.train_labels, 4.0)
                       for v in val_set])
    assert net_val < cos_val, (net_val, cos_val)


def test_triage_rows_on_held_out_genome_are_exact_with_recall(trained):
    """The held-out genome through the whole triage path (profiles, logits,
    pair scores, pruned search) at a threshold set on the validation genomes:
    every row is a row of the plain search (precision 1) and at least 99.5%
    of the plain search's rows are kept."""
    from rmblast_amd import triage as tr
    searcher, model, _, val_set, held, _ = trained
    val_scores = [tr.hsp_pair_scores(v, model.logits(v.xw, v.xf)) for v in val_set]
    thr = tr.threshold_for_recall(val_scores, 0.999)
    xw = model.profiles(held.seq, held.starts, held.lens)
    assert np.array_equal(xw, held.xw)
    passed = tr.pair_scores(model.logits(xw, held.xf), held.win_rec, len(held.recs)) >= thr
    ev = tr.evaluate(searcher, held, passed, repeats=1)
    assert ev["rows_not_in_plain_search"] == 0 and ev["precision"] == 1.0, ev
    assert ev["hsp_recall"] >= 0.995, ev
