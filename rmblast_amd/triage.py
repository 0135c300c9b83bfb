"""Learned triage: a ctypes binding of include/rmblast_triage.h plus the data
plumbing around it (windows, labels from the exact search, training loop,
pruned search and its evaluation against the full exact search).

The genome is searched as fixed records (RECORD bp); the triage scores every
window of a record against every family and prunes whole (record, family)
pairs, a pair passing when one of its windows does.  The pairs it passes are
searched by the exact path (Searcher.search_pairs), whose search of a pair is
the plain search's for that pair, so every row it reports is a row of the
plain search of those records (precision 1) and pruning can only lose rows.
Windows are 2 kbp at a 1 kbp stride inside each record; features are log1p
canonical 6-mer counts.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import lib as _lib_loader

KMER_DIM = 2080
WINDOW, STRIDE = 2000, 1000
TRAIN_MIN_OVERLAP = 300  # bp of an HSP a window must hold to be a training positive


def _bind(lib):
    P = ctypes.POINTER
    if getattr(lib, "_triage_bound", False):
        return lib
    vp, u8p, u16p, f32p = ctypes.c_void_p, P(ctypes.c_uint8), P(ctypes.c_uint16), P(ctypes.c_float)
    lib.rmb_triage_create.argtypes = [ctypes.c_int, ctypes.c_uint64, P(vp)]
    lib.rmb_triage_destroy.argtypes = [vp]
    lib.rmb_triage_last_error.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    lib.rmb_triage_n_params.argtypes = [vp, P(ctypes.c_uint64)]
    lib.rmb_triage_get_params.argtypes = [vp, f32p]
    lib.rmb_triage_set_params.argtypes = [vp, f32p]
    lib.rmb_kmer_profiles.argtypes = [vp, u8p, ctypes.c_uint64, P(ctypes.c_uint64), P(ctypes.c_uint32),
                                      ctypes.c_uint32, u16p]
    lib.rmb_triage_forward_backward.argtypes = [vp, u16p, ctypes.c_uint32, u16p, ctypes.c_uint32, u8p,
                                                ctypes.c_float, f32p, f32p]
    lib.rmb_triage_adam_step.argtypes = [vp, ctypes.c_float]
    lib.rmb_triage_logits.argtypes = [vp, u16p, ctypes.c_uint32, u16p, ctypes.c_uint32, f32p]
    lib.rmb_triage_tap_gemm.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        u16p, ctypes.c_int, u16p, ctypes.c_int, f32p]
    lib._triage_bound = True
    return lib


def _p(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class Triage:
    """One triage model on one GPU."""

    def __init__(self, seed: int = 1, device: int = 0):
        self._lib = _bind(_lib_loader())
        h = ctypes.c_void_p()
        rc = self._lib.rmb_triage_create(device, seed, ctypes.byref(h))
        if rc:
            raise RuntimeError(f"rmb_triage_create failed ({rc})")
        self._h = h
        n = ctypes.c_uint64()
        self._lib.rmb_triage_n_params(self._h, ctypes.byref(n))
        self.n_params = n.value

    def _check(self, rc, what):
        if rc:
            buf = ctypes.create_string_buffer(512)
            self._lib.rmb_triage_last_error(self._h, buf, 512)
            raise RuntimeError(f"{what} failed ({rc}): {buf.value.decode()}")

    def close(self):
        if self._h:
            self._lib.rmb_triage_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        self._check(self._lib.rmb_triage_get_params(self._h, _p(out, ctypes.c_float)), "get_params")
        return out

    def set_params(self, flat: np.ndarray):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == self.n_params
        self._check(self._lib.rmb_triage_set_params(self._h, _p(flat, ctypes.c_float)), "set_params")

    def profiles(self, seq: np.ndarray, starts, lens) -> np.ndarray:
        """n x 2080 BF16 bit patterns (uint16)."""
        seq = np.ascontiguousarray(seq, dtype=np.uint8)
        st = np.ascontiguousarray(starts, dtype=np.uint64)
        ln = np.ascontiguousarray(lens, dtype=np.uint32)
        out = np.zeros((len(st), KMER_DIM), dtype=np.uint16)
        self._check(self._lib.rmb_kmer_profiles(self._h, _p(seq, ctypes.c_uint8), len(seq), _p(st, ctypes.c_uint64),
                                                _p(ln, ctypes.c_uint32), len(st), _p(out, ctypes.c_uint16)),
                    "rmb_kmer_profiles")
        return out

    def forward_backward(self, xw, xf, labels, pos_weight=1.0, want_grads=False):
        xw = np.ascontiguousarray(xw, dtype=np.uint16)
        xf = np.ascontiguousarray(xf, dtype=np.uint16)
        lab = np.ascontiguousarray(labels, dtype=np.uint8).reshape(-1)
        assert lab.size == len(xw) * len(xf)
        loss = np.zeros(1, dtype=np.float32)
        grads = np.zeros(self.n_params, dtype=np.float32) if want_grads else None
        self._check(self._lib.rmb_triage_forward_backward(
            self._h, _p(xw, ctypes.c_uint16), len(xw), _p(xf, ctypes.c_uint16), len(xf), _p(lab, ctypes.c_uint8),
            pos_weight, _p(loss, ctypes.c_float), _p(grads, ctypes.c_float) if want_grads else None),
            "rmb_triage_forward_backward")
        return float(loss[0]), grads

    def adam_step(self, lr: float):
        self._check(self._lib.rmb_triage_adam_step(self._h, lr), "rmb_triage_adam_step")

    def logits(self, xw, xf) -> np.ndarray:
        xw = np.ascontiguousarray(xw, dtype=np.uint16)
        xf = np.ascontiguousarray(xf, dtype=np.uint16)
        out = np.zeros(len(xw) * len(xf), dtype=np.float32)
        self._check(self._lib.rmb_triage_logits(self._h, _p(xw, ctypes.c_uint16), len(xw), _p(xf, ctypes.c_uint16),
                                                len(xf), _p(out, ctypes.c_float)), "rmb_triage_logits")
        return out.reshape(len(xw), len(xf))

    def tap_gemm(self, ta, tb, M, N, K, A, lda, B, ldb) -> np.ndarray:
        A = np.ascontiguousarray(A, dtype=np.uint16)
        B = np.ascontiguousarray(B, dtype=np.uint16)
        C = np.zeros((M, N), dtype=np.float32)
        self._check(self._lib.rmb_triage_tap_gemm(self._h, int(ta), int(tb), M, N, K, _p(A, ctypes.c_uint16), lda,
                                                  _p(B, ctypes.c_uint16), ldb, _p(C, ctypes.c_float)), "tap_gemm")
        return C


# --------------------------------------------------------------- data ------
def windows(length: int, window: int = WINDOW, stride: int = STRIDE):
    """(starts, lens) of the query windows: `window` bp every `stride` bp; the
    last window ends at the record end."""
    if length <= window:
        return np.array([0], dtype=np.uint64), np.array([length], dtype=np.uint32)
    starts = list(range(0, length - window + 1, stride))
    if starts[-1] + window < length:
        starts.append(length - window)
    return np.array(starts, dtype=np.uint64), np.full(len(starts), window, dtype=np.uint32)


def labels_from_hsps(hsps, starts, lens, n_families, query=0, min_overlap=0):
    """labels[i, j] = 1 if an HSP of family j on record `query` overlaps
    window i.  With min_overlap > 0, a window overlapping an HSP by less than
    min(min_overlap, half the HSP) is labelled 2 (ignored in training) unless
    another HSP makes it positive."""
    lab = np.zeros((len(starts), n_families), dtype=np.uint8)
    s = starts.astype(np.int64)
    e = s + lens.astype(np.int64)
    for h in hsps:
        if h["query"] != query:
            continue
        qs, qe, f = int(h["q_start"]), int(h["q_end"]), int(h["subject"])
        ov = np.minimum(e, qe) - np.maximum(s, qs)
        hit = ov > 0
        if min_overlap > 0:
            need = min(min_overlap, max(1, (qe - qs) // 2))
            strong = ov >= need
            lab[strong, f] = 1
            weak = hit & ~strong & (lab[:, f] == 0)
            lab[weak, f] = 2
        else:
            lab[hit, f] = 1
    return lab


RECORD = 4_000  # the genome is searched as records of this many bp


def records_of(length: int, size: int = RECORD):
    """[(begin, end)] of the records a genome of `length` bp is searched as:
    consecutive, non-overlapping, `size` bp each (the last may be shorter)."""
    return [(b, min(length, b + size)) for b in range(0, max(length, 1), size)]


def record_windows(recs):
    """Windows of every record (record-local windows shifted to genome
    coordinates, so no window crosses a record end): (starts, lens, record of
    each window).  Windows are grouped by record, in record order."""
    st, ln, wr = [], [], []
    for r, (b, e) in enumerate(recs):
        s, l = windows(e - b)
        st.append(s + np.uint64(b))
        ln.append(l)
        wr.append(np.full(len(s), r, dtype=np.uint32))
    return np.concatenate(st), np.concatenate(ln), np.concatenate(wr)


@dataclass
class Instance:
    seq: np.ndarray            # one C1-style genome
    families: list             # family sequences
    recs: list                 # [(begin, end)]: the records the genome is searched as
    starts: np.ndarray         # windows (genome coordinates), grouped by record
    lens: np.ndarray
    win_rec: np.ndarray        # record of each window
    hsps: np.ndarray           # the plain exact search of every (record, family) pair
    ops: np.ndarray
    labels: np.ndarray         # windows x families: 1 if the window overlaps an HSP of the family
    pair_labels: np.ndarray    # records x families: 1 if the pair has an HSP
    xw: np.ndarray = None      # BF16 profiles
    xf: np.ndarray = None
    train_labels: np.ndarray = None  # as labels, marginal overlaps marked 2 (ignored)

    def record_seqs(self):
        return [self.seq[b:e] for b, e in self.recs]


def window_labels(hsps, recs, starts, lens, win_rec, n_families, min_overlap=0):
    """labels_from_hsps over records: an HSP of record r (record-local
    coordinates) labels the windows of record r it overlaps."""
    lab = np.zeros((len(starts), n_families), dtype=np.uint8)
    first = np.searchsorted(win_rec, np.arange(len(recs) + 1))
    if not len(hsps):
        return lab
    order = np.argsort(hsps["query"], kind="stable")
    hq = hsps["query"][order]
    for r in np.unique(hq):
        a, b = int(first[r]), int(first[r + 1])
        mine = hsps[order[np.searchsorted(hq, r):np.searchsorted(hq, r, side="right")]]
        lab[a:b] = labels_from_hsps(mine, starts[a:b] - np.uint64(recs[r][0]), lens[a:b], n_families, query=r,
                                    min_overlap=min_overlap)
    return lab


def pair_scores(scores: np.ndarray, win_rec: np.ndarray, n_records: int) -> np.ndarray:
    """records x families: a pair's score is the best score among its
    record's windows (windows are grouped by record)."""
    first = np.searchsorted(win_rec, np.arange(n_records))
    return np.maximum.reduceat(scores, first, axis=0)


def hsp_key(h, ops, q_shift=0):
    o = ops[int(h["ops_offset"]):int(h["ops_offset"]) + int(h["ops_count"])]
    return (int(h["subject"]), int(h["score"]), int(h["q_start"]) + q_shift, int(h["q_end"]) + q_shift,
            int(h["s_start"]), int(h["s_end"]), int(h["strand"]), o.tobytes())


def row_keys(hsps, ops):
    """One key per HSP: (record,) + hsp_key."""
    return {(int(h["query"]),) + hsp_key(h, ops) for h in hsps}


def pruned_search(searcher, inst: Instance, passed: np.ndarray):
    """Exact search of the passed (record, family) pairs through
    Searcher.search_pairs.  A pair's search is the plain search's for that
    pair (HSPs of one record against one family depend on nothing else), so
    every row is a row of the plain search; a pruned pair loses its rows.
    Returns (set of row keys, pairs searched, seconds)."""
    t0 = time.perf_counter()
    pq, ps = np.nonzero(passed)
    keys = set()
    if len(pq):
        hsps, ops = searcher.search_pairs(inst.record_seqs(), pq.astype(np.uint32), ps.astype(np.uint32))
        keys = row_keys(hsps, ops)
    return keys, int(len(pq)), time.perf_counter() - t0


# ------------------------------------------------- training and evaluation --
C1_PARAMS = dict(gap_open=20, gap_extend=5, word_size=7, xdrop_ungap=20, xdrop_gap=30, xdrop_gap_final=100,
                 min_raw_gapped_score=200, complexity_adjust=True)


def make_searcher():
    import os
    import tempfile
    import rmblast_amd
    from . import synth
    d = tempfile.mkdtemp(prefix="rmb_triage_")
    mp = os.path.join(d, "synth.matrix")
    with open(mp, "w") as f:
        f.write(synth.MATRIX_TEXT)
    return rmblast_amd.Searcher(rmblast_amd.load_matrix(mp), **C1_PARAMS)


def family_profiles(model, families):
    cat = np.concatenate([np.asarray(f, dtype=np.uint8) for f in families])
    lens = np.array([len(f) for f in families], dtype=np.uint32)
    starts = np.concatenate([[0], np.cumsum(lens[:-1])]).astype(np.uint64)
    return model.profiles(cat, starts, lens)


def make_instance(searcher, model, seed: int, record: int = RECORD, genome=None) -> Instance:
    """One parity-configuration genome (1 Mbp, 50 families, 2,000 copies) as
    `record`-bp records: the plain exact search of every (record, family)
    pair, windows, labels and profiles."""
    from . import synth
    w = genome if genome is not None else synth.parity_config(seed=seed)
    q = w.records[0][1]
    fams = [f[1] for f in w.families]
    searcher.load_library(fams)
    recs = records_of(len(q), record)
    st, ln, wr = record_windows(recs)
    hsps, ops = searcher.search([q[b:e] for b, e in recs])
    lab = window_labels(hsps, recs, st, ln, wr, len(fams))
    pl = np.zeros((len(recs), len(fams)), dtype=np.uint8)
    if len(hsps):
        pl[hsps["query"].astype(np.int64), hsps["subject"].astype(np.int64)] = 1
    return Instance(seq=q, families=fams, recs=recs, starts=st, lens=ln, win_rec=wr, hsps=hsps, ops=ops,
                    labels=lab, pair_labels=pl, xw=model.profiles(q, st, ln) if model else None,
                    xf=family_profiles(model, fams) if model else None,
                    train_labels=window_labels(hsps, recs, st, ln, wr, len(fams), min_overlap=TRAIN_MIN_OVERLAP))


def hsp_pair_scores(inst: Instance, scores: np.ndarray) -> np.ndarray:
    """Per plain-search HSP, the score of its (record, family) pair: the HSP
    is reported iff that score passes the threshold."""
    ps = pair_scores(scores, inst.win_rec, len(inst.recs))
    if not len(inst.hsps):
        return np.empty(0)
    return ps[inst.hsps["query"].astype(np.int64), inst.hsps["subject"].astype(np.int64)].astype(np.float64)


def _from_bf16(bits):
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


def cosine_scores(inst: Instance) -> np.ndarray:
    """The method without the network: cosine similarity of the same log1p
    canonical 6-mer profiles, window against family."""
    a = _from_bf16(inst.xw).astype(np.float64)
    b = _from_bf16(inst.xf).astype(np.float64)
    a /= np.maximum(np.linalg.norm(a, axis=1, keepdims=True), 1e-12)
    b /= np.maximum(np.linalg.norm(b, axis=1, keepdims=True), 1e-12)
    return a @ b.T


def _weighted_logistic_loss(z, lab, pos_weight):
    y = (lab == 1).astype(np.float64)
    wt = np.where(lab == 1, pos_weight, np.where(lab == 0, 1.0, 0.0))
    return float((wt * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))).sum() / max(wt.sum(), 1.0))


def fit_cosine_logistic(instances, pos_weight, iters=200):
    """The method without the network as a classifier: logit = c cos + b,
    fitted by Newton's method on the same weighted loss."""
    x = np.concatenate([cosine_scores(i).reshape(-1) for i in instances])
    lab = np.concatenate([(i.train_labels if i.train_labels is not None else i.labels).reshape(-1)
                          for i in instances])
    y = (lab == 1).astype(np.float64)
    wt = np.where(lab == 1, pos_weight, np.where(lab == 0, 1.0, 0.0))
    X = np.stack([x, np.ones_like(x)], axis=1)
    th = np.zeros(2)
    for _ in range(iters):
        z = X @ th
        p = 1 / (1 + np.exp(-z))
        g = X.T @ (wt * (p - y))
        H = (X * (wt * p * (1 - p))[:, None]).T @ X + 1e-9 * np.eye(2)
        step = np.linalg.solve(H, g)
        th -= step
        if np.abs(step).max() < 1e-10:
            break
    return th


HEAD_HIDDEN = 64


def head_offsets(n_params):
    """Offsets of wh2 (64), bh2 and c in the flat parameter vector (the head's
    scalars close it: bh1[64], wh2[64], bh2, a, c)."""
    hp = n_params - (2 * HEAD_HIDDEN + 3)
    return {"wh2": hp + HEAD_HIDDEN, "bh2": hp + 2 * HEAD_HIDDEN, "a": hp + 2 * HEAD_HIDDEN + 1,
            "c": hp + 2 * HEAD_HIDDEN + 2}


def warm_start_from_cosine(model, th):
    """Start the network at the fitted cosine logistic: c and bh2 from the
    fit, wh2 = 0 and a = 0, so the towers and the head add corrections to it."""
    p = model.get_params()
    o = head_offsets(model.n_params)
    p[o["wh2"]:o["wh2"] + HEAD_HIDDEN] = 0.0
    p[o["a"]] = 0.0
    p[o["c"]] = th[0]
    p[o["bh2"]] = th[1]
    model.set_params(p)


def threshold_for_recall(hsp_scores, target):
    """Largest threshold passing at least `target` of the HSPs."""
    s = np.sort(np.concatenate(hsp_scores))
    if not len(s):
        return -np.inf
    k = int(np.floor((1.0 - target) * len(s)))
    return float(s[min(k, len(s) - 1)])


def _loss(model, inst, pos_weight):
    lab = inst.train_labels if inst.train_labels is not None else inst.labels
    return model.forward_backward(inst.xw, inst.xf, lab, pos_weight)[0]


def train(model, instances, epochs, lr, pos_weight, seed=0, log=None, val=None):
    """Adam over the training genomes, one genome (all its window x family
    pairs) per step.  With validation genomes, the parameters of the epoch
    with the lowest validation loss are kept (early stopping)."""
    rng = np.random.default_rng(seed)
    losses, val_losses = [], []
    best = (np.inf, None)
    if val:  # the starting point is a candidate too
        best = (float(np.mean([_loss(model, v, pos_weight) for v in val])), model.get_params())
    for ep in range(epochs):
        tot = 0.0
        for k in rng.permutation(len(instances)):
            tot += _loss(model, instances[k], pos_weight)
            model.adam_step(lr)
        losses.append(tot / len(instances))
        msg = f"epoch {ep + 1}/{epochs}: mean loss {losses[-1]:.4f}"
        if val:
            vl = float(np.mean([_loss(model, v, pos_weight) for v in val]))
            val_losses.append(vl)
            msg += f", validation loss {vl:.4f}"
            if vl < best[0]:
                best = (vl, model.get_params())
        if log:
            log(msg)
    if best[1] is not None:
        model.set_params(best[1])
    return losses, val_losses


def evaluate(searcher, inst: Instance, passed: np.ndarray, repeats: int = 9):
    """Pruned search of `inst` with the given passed (record, family) pairs,
    against the plain exact search of every pair: row recall, score-weighted
    recall, precision (rows the plain search does not report lower it),
    pairs searched and the timed search seconds (best of `repeats`)."""
    searcher.load_library(inst.families)
    full_keys = row_keys(inst.hsps, inst.ops)
    best, keys, n_pairs = None, set(), 0
    for _ in range(repeats):
        keys, n_pairs, secs = pruned_search(searcher, inst, passed)
        best = secs if best is None else min(best, secs)
    hit = full_keys & keys
    score = {k: k[2] for k in full_keys}
    return {
        "hsp_recall": len(hit) / max(1, len(full_keys)),
        "score_weighted_recall": sum(score[k] for k in hit) / max(1, sum(score.values())),
        "precision": len(hit) / max(1, len(keys)),
        "rows_not_in_plain_search": len(keys - full_keys),
        "passed_pair_fraction": float(passed.mean()),
        "pairs_searched": n_pairs,
        "search_seconds": best,
    }


def plain_search_seconds(searcher, inst: Instance, repeats: int = 9, whole: bool = False):
    """Best-of-`repeats` seconds of the plain exact search of every pair
    (with `whole`, of the genome as one record instead)."""
    searcher.load_library(inst.families)
    qs = [inst.seq] if whole else inst.record_seqs()
    best = None
    for _ in range(repeats):
        t = time.perf_counter()
        searcher.search(qs)
        secs = time.perf_counter() - t
        best = secs if best is None else min(best, secs)
    return best


def record_sweep(searcher, seed: int, sizes=(2_000, 4_000, 8_000, 16_000, 60_000), log=print):
    """Per record size, on one genome: the plain search's time and rows, and
    a perfect triage's (the true pair labels) pair fraction and time, i.e.
    the speedup bound of pruning at (record, family) granularity."""
    from . import synth
    g = synth.parity_config(seed=seed)
    out = {}
    for size in sizes:
        inst = make_instance(searcher, None, seed, size, genome=g)
        t_plain = plain_search_seconds(searcher, inst)
        ev = evaluate(searcher, inst, inst.pair_labels > 0)
        out[str(size)] = {"records": len(inst.recs), "rows": int(len(inst.hsps)), "plain_search_seconds": t_plain,
                          "true_pair_fraction": float(inst.pair_labels.mean()),
                          "perfect_triage_seconds": ev["search_seconds"], "perfect_recall": ev["hsp_recall"],
                          "perfect_precision": ev["precision"], "bound": t_plain / ev["search_seconds"]}
        log(f"record {size} bp: {out[str(size)]}")
    return out


def run_experiment(n_train=40, n_val=4, eval_seed=100_000, epochs=30, lr=1e-3, pos_weight=4.0, target=0.995,
                   record=RECORD, model_seed=1, sweep=True, log=print):
    """Train on parity-configuration genomes of seeds 1..n_train (labels from
    the plain exact search of `record`-bp records), pick the threshold on
    n_val validation genomes so that `target` of their rows pass, and
    evaluate on the held-out genome `eval_seed`: the learned scores and, for
    comparison, the same procedure on the cosine similarity of the profiles
    (the method without the network)."""
    searcher = make_searcher()
    model = Triage(seed=model_seed)
    report = {"record_bp": record}
    if sweep:
        report["record_sweep"] = record_sweep(searcher, eval_seed, log=log)
    t0 = time.perf_counter()
    train_set = [make_instance(searcher, model, 1 + k, record) for k in range(n_train)]
    val_set = [make_instance(searcher, model, 10_000 + k, record) for k in range(n_val)]
    held = make_instance(searcher, model, eval_seed, record)
    log(f"instances built in {time.perf_counter() - t0:.1f}s")
    n_pairs = sum(i.labels.size for i in train_set)
    pos_rate = float(np.mean([i.labels.mean() for i in train_set]))
    th = fit_cosine_logistic(train_set, pos_weight)
    warm_start_from_cosine(model, th)
    t0 = time.perf_counter()
    losses, val_losses = train(model, train_set, epochs, lr, pos_weight, log=log, val=val_set)
    train_s = time.perf_counter() - t0

    def learned(inst):
        return model.logits(inst.xw, inst.xf)

    # pair classification on the validation genomes: the network against the
    # fitted cosine-only logistic (the method without the network)
    val_lab = [v.train_labels for v in val_set]
    cos_val = float(np.mean([_weighted_logistic_loss(th[0] * cosine_scores(v) + th[1], l, pos_weight)
                             for v, l in zip(val_set, val_lab)]))
    net_val = float(np.mean([_weighted_logistic_loss(learned(v).astype(np.float64), l, pos_weight)
                             for v, l in zip(val_set, val_lab)]))
    log(f"validation pair loss: network {net_val:.4f}, cosine-only logistic {cos_val:.4f}")

    report.update({"train_window_pairs": n_pairs, "train_positive_rate": pos_rate, "epochs": epochs,
                   "train_seconds": train_s, "loss_first_epoch": losses[0], "loss_last_epoch": losses[-1],
                   "validation_loss_by_epoch": val_losses,
                   "validation_pair_loss": {"network": net_val, "cosine_only_logistic": cos_val},
                   "best_epoch": int(np.argmin(val_losses)) + 1 if val_losses else epochs,
                   "target_hsp_recall": target, "held_out_seed": eval_seed, "held_out_rows": int(len(held.hsps)),
                   "held_out_records": len(held.recs)})
    t_plain = plain_search_seconds(searcher, held)
    report["plain_search_seconds"] = t_plain
    report["one_record_search_seconds"] = plain_search_seconds(searcher, held, whole=True)
    # true pair labels (a perfect triage): the bound of pruning at this granularity
    perfect = evaluate(searcher, held, held.pair_labels > 0)
    perfect["speedup_vs_plain_search"] = t_plain / perfect["search_seconds"]
    report["true_labels"] = perfect
    log(f"plain search {t_plain * 1e3:.1f} ms; true labels: {perfect}")
    targets = sorted(set([0.9, 0.95, 0.99, target]))
    for name, fn in (("learned", learned), ("cosine_profile", cosine_scores)):
        val_scores = [hsp_pair_scores(v, fn(v)) for v in val_set]
        # the whole triage path on the held-out genome: profiles, scores, pruned search
        t = time.perf_counter()
        xw = model.profiles(held.seq, held.starts, held.lens)
        xf = family_profiles(model, held.families)
        t_feat = time.perf_counter() - t
        t = time.perf_counter()
        probe = Instance(held.seq, held.families, held.recs, held.starts, held.lens, held.win_rec, held.hsps,
                         held.ops, held.labels, held.pair_labels, xw, xf)
        scores = pair_scores(fn(probe), held.win_rec, len(held.recs))
        t_score = time.perf_counter() - t
        curve = {}
        for tg in targets:
            thr = threshold_for_recall(val_scores, tg)
            passed = scores >= thr
            ev = evaluate(searcher, held, passed)
            total = t_feat + t_score + ev["search_seconds"]
            ev.update({"threshold": thr, "profile_seconds": t_feat, "score_seconds": t_score,
                       "triage_path_seconds": total, "speedup_vs_plain_search": t_plain / total,
                       "pair_recall": float((passed & (held.pair_labels > 0)).sum()
                                            / max(1, held.pair_labels.sum()))})
            curve[str(tg)] = ev
            log(f"{name} @ validation HSP recall {tg}: {ev}")
        report[name] = curve
    model.close()
    return report


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train and evaluate the learned triage on the parity configuration")
    ap.add_argument("--train", type=int, default=40)
    ap.add_argument("--val", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--pos-weight", type=float, default=4.0)
    ap.add_argument("--target", type=float, default=0.995)
    ap.add_argument("--record", type=int, default=RECORD)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    import sys
    rep = run_experiment(a.train, a.val, epochs=a.epochs, lr=a.lr, pos_weight=a.pos_weight, target=a.target,
                         record=a.record, sweep=not a.no_sweep, log=lambda s: print("[triage]", s, file=sys.stderr, flush=True))
    line = json.dumps(rep)
    print(line)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
