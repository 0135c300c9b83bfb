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


def labels_from_hsps(hsps, starts, lens, n_families, query=0):
    """labels[i, j] = 1 if an HSP of family j on record `query` overlaps window i."""
    lab = np.zeros((len(starts), n_families), dtype=np.uint8)
    s = starts.astype(np.int64)
    e = s + lens.astype(np.int64)
    for h in hsps:
        if h["query"] != query:
            continue
        hit = (s < int(h["q_end"])) & (e > int(h["q_start"]))
        lab[hit, int(h["subject"])] = 1
    return lab


@dataclass
class Instance:
    seq: np.ndarray            # the query record (one C1-style genome)
    families: list             # family sequences
    starts: np.ndarray
    lens: np.ndarray
    hsps: np.ndarray           # full exact-search HSPs
    ops: np.ndarray
    labels: np.ndarray         # windows x families
    xw: np.ndarray = None      # BF16 profiles
    xf: np.ndarray = None


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


def pruned_search(searcher, inst: Instance, passed: np.ndarray, margin: int):
    """Exact search of the segments the passed (window, family) pairs select.
    Returns (set of HSP keys in record coordinates, number of segment-family
    pairs searched, seconds)."""
    t0 = time.perf_counter()
    q_len = len(inst.seq)
    seg_seqs, seg_off, pair_q, pair_s = [], [], [], []
    seg_index = {}
    for j in range(passed.shape[1]):
        idx = np.nonzero(passed[:, j])[0]
        if not len(idx):
            continue
        for b, e in segments_for(idx, inst.starts, inst.lens, q_len, margin):
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
            keys.add(hsp_key(h, ops, seg_off[int(h["query"])]))
    return keys, len(pair_q), time.perf_counter() - t0
