"""Learned triage: a ctypes binding of include/rmblast_triage.h plus the data
plumbing around it (windows, labels from the exact search, training loop,
pruned search and its evaluation against the full exact search).

The triage prunes (query window, family) pairs before seeding; the pairs it
passes are searched by the exact path (Searcher.search_pairs), so every HSP
it reports is computed by the same kernels as the full search.  Windows are
2 kbp at a 1 kbp stride; features are log1p canonical 6-mer counts.
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


@dataclass
class Instance:
    seq: np.ndarray            # the query record (one C1-style genome)
    families: list             # family sequences
    starts: np.ndarray
    lens: np.ndarray
    hsps: np.ndarray           # full exact-search HSPs
    ops: np.ndarray
    labels: np.ndarray         # windows x families: 1 if the window overlaps an HSP of the family
    xw: np.ndarray = None      # BF16 profiles
    xf: np.ndarray = None
    train_labels: np.ndarray = None  # as labels, marginal overlaps marked 2 (ignored)


def segments_for(passed_windows, starts, lens, q_len, margin):
    """Merge the passed windows of one family into query segments, each
    widened by `margin` on both sides; returns [(begin, end)]."""
    iv = sorted((max(0, int(starts[i]) - margin), min(q_len, int(starts[i]) + int(lens[i]) + margin))
                for i in passed_windows)
    out = []
    for b, e in iv:
        if out and b <= out[-1][1]:
            out[-1] = (out[-1][0], max(out[-1][1], e))
        else:
            out.append((b, e))
    return out


def hsp_key(h, ops, q_shift=0):
    o = ops[int(h["ops_offset"]):int(h["ops_offset"]) + int(h["ops_count"])]
    return (int(h["subject"]), int(h["score"]), int(h["q_start"]) + q_shift, int(h["q_end"]) + q_shift,
            int(h["s_start"]), int(h["s_end"]), int(h["strand"]), o.tobytes())


CHUNK, CHUNK_OVERLAP = 16_000, 4_000  # segment pieces searched as separate records


def chunks_of(b, e, chunk=CHUNK, overlap=CHUNK_OVERLAP):
    """Cut [b, e) into pieces of at most `chunk` bp overlapping by `overlap`
    bp, so every HSP shorter than the overlap lies whole in some piece."""
    if e - b <= chunk:
        return [(b, e)]
    out, s = [], b
    while True:
        out.append((s, min(e, s + chunk)))
        if s + chunk >= e:
            return out
        s += chunk - overlap


def pruned_search(searcher, inst: Instance, passed: np.ndarray, margin: int):
    """Exact search of the segments the passed (window, family) pairs select,
    each cut into overlapping pieces searched as separate records; HSPs are
    mapped back to record coordinates and deduplicated.  Returns (set of HSP
    keys, number of piece-family pairs searched, seconds)."""
    t0 = time.perf_counter()
    q_len = len(inst.seq)
    seg_seqs, seg_off, pair_q, pair_s = [], [], [], []
    seg_index = {}
    for j in range(passed.shape[1]):
        idx = np.nonzero(passed[:, j])[0]
        if not len(idx):
            continue
        for sb, se in segments_for(idx, inst.starts, inst.lens, q_len, margin):
            for b, e in chunks_of(sb, se):
                if (b, e) not in seg_index:
                    seg_index[(b, e)] = len(seg_seqs)
                    seg_seqs.append(inst.seq[b:e])
                    seg_off.append(b)
                pair_q.append(seg_index[(b, e)])
                pair_s.append(j)
    keys = set()
    if pair_q:
        hsps, ops = searcher.search_pairs(seg_seqs, pair_q, pair_s)
        for h in hsps:
            k = int(h["query"])
            b, n = seg_off[k], len(seg_seqs[k])
            # an HSP reaching a cut end of its piece is truncated there; its
            # whole copy lies inside the overlapping neighbour piece
            if (b > 0 and int(h["q_start"]) == 0) or (b + n < q_len and int(h["q_end"]) == n):
                continue
            keys.add(hsp_key(h, ops, b))
    return keys, len(pair_q), time.perf_counter() - t0


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


def make_instance(searcher, model, seed: int) -> Instance:
    """One parity-configuration genome (1 Mbp, 50 families, 2,000 copies):
    full exact search, windows, labels and profiles."""
    from . import synth
    w = synth.parity_config(seed=seed)
    q = w.records[0][1]
    fams = [f[1] for f in w.families]
    searcher.load_library(fams)
    hsps, ops = searcher.search([q])
    st, ln = windows(len(q))
    lab = labels_from_hsps(hsps, st, ln, len(fams))
    return Instance(seq=q, families=fams, starts=st, lens=ln, hsps=hsps, ops=ops, labels=lab,
                    xw=model.profiles(q, st, ln), xf=family_profiles(model, fams),
                    train_labels=labels_from_hsps(hsps, st, ln, len(fams), min_overlap=TRAIN_MIN_OVERLAP))


def hsp_window_scores(inst: Instance, scores: np.ndarray) -> np.ndarray:
    """Per full-search HSP, the best score among the windows it overlaps (for
    its family): the HSP is searched iff that score passes the threshold."""
    s = inst.starts.astype(np.int64)
    e = s + inst.lens.astype(np.int64)
    out = np.empty(len(inst.hsps), dtype=np.float64)
    for k, h in enumerate(inst.hsps):
        hit = (s < int(h["q_end"])) & (e > int(h["q_start"]))
        out[k] = scores[hit, int(h["subject"])].max() if hit.any() else -np.inf
    return out


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


def evaluate(searcher, inst: Instance, passed: np.ndarray, margin: int, repeats: int = 3):
    """Pruned search of `inst` with the given passed (window, family) pairs,
    against the full exact search: HSP recall, score-weighted recall,
    precision, searched pairs, and the timed search seconds (best of
    `repeats`)."""
    searcher.load_library(inst.families)
    full_keys = {hsp_key(h, inst.ops) for h in inst.hsps}
    best = None
    for _ in range(repeats):
        keys, n_pairs, secs = pruned_search(searcher, inst, passed, margin)
        best = secs if best is None else min(best, secs)
    hit = full_keys & keys
    score = {k: k[1] for k in full_keys}
    return {
        "hsp_recall": len(hit) / max(1, len(full_keys)),
        "score_weighted_recall": sum(score[k] for k in hit) / max(1, sum(score.values())),
        "precision": len(hit) / max(1, len(keys)),
        "passed_pair_fraction": float(passed.mean()),
        "segment_pairs_searched": n_pairs,
        "search_seconds": best,
    }


def full_search_seconds(searcher, inst: Instance, repeats: int = 3):
    searcher.load_library(inst.families)
    best = None
    for _ in range(repeats):
        t = time.perf_counter()
        searcher.search([inst.seq])
        secs = time.perf_counter() - t
        best = secs if best is None else min(best, secs)
    return best


def run_experiment(n_train=40, n_val=4, eval_seed=100_000, epochs=30, lr=1e-3, pos_weight=4.0, target=0.995,
                   margin=2000, model_seed=1, log=print):
    """Train on parity-configuration genomes of seeds 1..n_train (labels from
    the full exact search), pick the threshold on n_val validation genomes so
    that `target` of their HSPs pass, and evaluate on the held-out genome
    `eval_seed`: the learned scores and, for comparison, the same procedure on
    the cosine similarity of the profiles (the method without the network)."""
    searcher = make_searcher()
    model = Triage(seed=model_seed)
    t0 = time.perf_counter()
    train_set = [make_instance(searcher, model, 1 + k) for k in range(n_train)]
    val_set = [make_instance(searcher, model, 10_000 + k) for k in range(n_val)]
    held = make_instance(searcher, model, eval_seed)
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

    report = {"train_pairs": n_pairs, "train_positive_rate": pos_rate, "epochs": epochs, "train_seconds": train_s,
              "loss_first_epoch": losses[0], "loss_last_epoch": losses[-1],
              "validation_loss_by_epoch": val_losses,
              "validation_pair_loss": {"network": net_val, "cosine_only_logistic": cos_val},
              "best_epoch": int(np.argmin(val_losses)) + 1 if val_losses else epochs, "target_hsp_recall": target,
              "margin_bp": margin, "held_out_seed": eval_seed, "held_out_hsps": int(len(held.hsps))}
    t_full = full_search_seconds(searcher, held)
    report["full_search_seconds"] = t_full
    # the same chunked exact search with every pair passed: what chunking alone gives
    every = evaluate(searcher, held, np.ones_like(held.labels, dtype=bool), margin)
    every["speedup_vs_full_search"] = t_full / every["search_seconds"]
    report["all_pairs_chunked"] = every
    log(f"all pairs, chunked: {every}")
    # true labels (a perfect triage) as the upper bound of pruning
    perfect = evaluate(searcher, held, held.labels > 0, margin)
    perfect["speedup_vs_full_search"] = t_full / perfect["search_seconds"]
    report["true_labels"] = perfect
    log(f"true labels: {perfect}")
    targets = sorted(set([0.9, 0.95, 0.99, target]))
    for name, fn in (("learned", learned), ("cosine_profile", cosine_scores)):
        val_scores = [hsp_window_scores(v, fn(v)) for v in val_set]
        # the whole triage path on the held-out genome: profiles, scores, pruned search
        t = time.perf_counter()
        xw = model.profiles(held.seq, held.starts, held.lens)
        xf = family_profiles(model, held.families)
        t_feat = time.perf_counter() - t
        t = time.perf_counter()
        scores = fn(Instance(held.seq, held.families, held.starts, held.lens, held.hsps, held.ops, held.labels, xw,
                             xf))
        t_score = time.perf_counter() - t
        curve = {}
        for tg in targets:
            thr = threshold_for_recall(val_scores, tg)
            passed = scores >= thr
            ev = evaluate(searcher, held, passed, margin)
            total = t_feat + t_score + ev["search_seconds"]
            ev.update({"threshold": thr, "profile_seconds": t_feat, "score_seconds": t_score,
                       "triage_path_seconds": total, "speedup_vs_full_search": t_full / total,
                       "speedup_vs_all_pairs_chunked": every["search_seconds"] / total,
                       "pair_recall": float((passed & (held.labels > 0)).sum() / max(1, held.labels.sum()))})
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
    ap.add_argument("--margin", type=int, default=2000)
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    import sys
    rep = run_experiment(a.train, a.val, epochs=a.epochs, lr=a.lr, pos_weight=a.pos_weight, target=a.target,
                         margin=a.margin, log=lambda s: print("[triage]", s, file=sys.stderr, flush=True))
    line = json.dumps(rep)
    print(line)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
