"""CPU oracle of the learned triage (test infrastructure only).

A numpy restatement of the model the HIP kernels run (rmblast_amd/csrc/
triage_kernels.hip, triage.cpp): canonical 6-mer profiles, the two towers,
the dot-plus-two-layer head, the weighted logistic loss and the full
backward pass.  The triage has no counterpart in the reference; this file is
pinned by the mathematics instead: canonical classes counted in closed form,
and the backward pass checked against central finite differences in float64
(tests/test_triage_oracle.py).  `bf16=True` rounds at the points where the GPU
path rounds (GEMM operands), so GPU results can be compared to it closely.
"""
from __future__ import annotations

import numpy as np

KMER_DIM = 2080
H1, EMB, HEAD = 512, 256, 64


# ----------------------------------------------------------------- features --
def canonical_index() -> np.ndarray:
    """4096 6-mer codes (first base in the high bits, A0 C1 G2 T3) -> rank of
    min(code, reverse complement) among the canonical codes."""
    codes = np.arange(4096)
    rc = np.zeros(4096, dtype=np.int64)
    for i in range(6):
        rc = (rc << 2) | (3 - ((codes >> (2 * i)) & 3))
    canon = np.minimum(codes, rc)
    classes = np.unique(canon)
    return np.searchsorted(classes, canon)


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


def kmer_profile(seq: np.ndarray, start: int, length: int) -> np.ndarray:
    """log1p of canonical 6-mer counts of seq[start:start+length] (words with
    a code >= 4 skipped), float32."""
    s = np.asarray(seq[start:start + length], dtype=np.int64)
    out = np.zeros(KMER_DIM, dtype=np.float32)
    if len(s) < 6:
        return out
    idx = canonical_index()
    win = np.lib.stride_tricks.sliding_window_view(s, 6)
    ok = (win < 4).all(axis=1)
    code = np.zeros(len(win), dtype=np.int64)
    for k in range(6):
        code = (code << 2) | (win[:, k] & 3)
    np.add.at(out, idx[code[ok]], 1.0)
    return np.log1p(out).astype(np.float32)


# -------------------------------------------------------------- parameters --
def layout():
    """Offsets of the flat parameter vector (triage.cpp Layout)."""
    W1, B1, W2, B2 = H1 * KMER_DIM, H1, EMB * H1, EMB
    tower = W1 + B1 + W2 + B2
    off = {}
    for t in range(2):
        b = t * tower
        off[f"W1_{t}"] = (b, (H1, KMER_DIM))
        off[f"b1_{t}"] = (b + W1, (H1,))
        off[f"W2_{t}"] = (b + W1 + B1, (EMB, H1))
        off[f"b2_{t}"] = (b + W1 + B1 + W2, (EMB,))
    h = 2 * tower
    off["Wh1"] = (h, (HEAD, EMB))
    off["bh1"] = (h + HEAD * EMB, (HEAD,))
    off["wh2"] = (h + HEAD * EMB + HEAD, (HEAD,))
    off["bh2"] = (h + HEAD * EMB + 2 * HEAD, ())
    off["a"] = (h + HEAD * EMB + 2 * HEAD + 1, ())
    off["c"] = (h + HEAD * EMB + 2 * HEAD + 2, ())
    total = h + HEAD * EMB + 2 * HEAD + 3
    return off, total


def unpack(flat: np.ndarray) -> dict:
    off, _ = layout()
    out = {}
    for k, (o, shape) in off.items():
        n = int(np.prod(shape)) if shape else 1
        out[k] = flat[o:o + n].reshape(shape) if shape else flat[o]
    return out


def pack(d: dict, dtype=np.float32) -> np.ndarray:
    off, total = layout()
    flat = np.zeros(total, dtype=dtype)
    for k, (o, shape) in off.items():
        n = int(np.prod(shape)) if shape else 1
        flat[o:o + n] = np.asarray(d[k], dtype=dtype).reshape(n)
    return flat


# ------------------------------------------------------- forward / backward --
def forward_backward(flat, xw, xf, labels=None, pos_weight=1.0, bf16=True, dtype=np.float32):
    """Logits of all nw x nf pairs (pair p = i*nf + j); with labels (0
    negative, 1 positive, 2 ignored) also the weighted logistic loss and the
    gradient as a flat vector."""
    r = bf16_round if bf16 else (lambda x: x)
    P_ = unpack(np.asarray(flat, dtype=dtype))
    xw = np.asarray(xw, dtype=dtype)
    xf = np.asarray(xf, dtype=dtype)

    def tower(t, X):
        h1 = r(X) @ r(P_[f"W1_{t}"]).T + P_[f"b1_{t}"]
        h1 = np.maximum(h1, 0)
        e = r(h1) @ r(P_[f"W2_{t}"]).T + P_[f"b2_{t}"]
        return h1, e

    h1w, ew = tower(0, xw)
    h1f, ef = tower(1, xf)
    nw, nf = len(xw), len(xf)
    prod = ew[:, None, :] * ef[None, :, :]
    P = r(prod).reshape(nw * nf, EMB)
    dot = prod.sum(-1).reshape(-1)
    xr_w, xr_f = r(xw), r(xf)
    nrm = np.linalg.norm(xr_w, axis=1)[:, None] * np.linalg.norm(xr_f, axis=1)[None, :]
    cos = np.where(nrm > 0, (xr_w @ xr_f.T) / np.where(nrm > 0, nrm, 1), 0).reshape(-1)
    hh = P @ r(P_["Wh1"]).T + P_["bh1"]
    act = np.maximum(hh, 0)
    z = P_["bh2"] + P_["a"] * dot + P_["c"] * cos + act @ P_["wh2"]
    if labels is None:
        return z
    lab = np.asarray(labels).reshape(-1)
    y = (lab == 1).astype(dtype)
    wt = np.where(lab == 1, pos_weight, np.where(lab == 0, 1.0, 0.0)).astype(dtype)
    wsum = max(wt.sum(), 1.0)
    loss = (wt * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))).sum() / wsum
    g = wt * (1 / (1 + np.exp(-z)) - y) / wsum
    G = {}
    G["bh2"] = g.sum()
    G["a"] = (g * dot).sum()
    G["c"] = (g * cos).sum()
    G["wh2"] = act.T @ g
    dh = (hh > 0) * g[:, None] * P_["wh2"][None, :]
    G["bh1"] = dh.sum(0)
    G["Wh1"] = r(dh).T @ P
    dP = (r(dh) @ r(P_["Wh1"])).reshape(nw, nf, EMB)
    gg = g.reshape(nw, nf)
    dEw = ((dP + P_["a"] * gg[:, :, None]) * ef[None, :, :]).sum(1)
    dEf = ((dP + P_["a"] * gg[:, :, None]) * ew[:, None, :]).sum(0)
    for t, X, h1, dE in ((0, xw, h1w, dEw), (1, xf, h1f, dEf)):
        G[f"b2_{t}"] = dE.sum(0)
        G[f"W2_{t}"] = r(dE).T @ r(h1)
        dh1 = (r(dE) @ r(P_[f"W2_{t}"])) * (h1 > 0)
        G[f"b1_{t}"] = dh1.sum(0)
        G[f"W1_{t}"] = r(dh1).T @ r(X)
    return z, loss, pack(G, dtype)


def init_params(seed: int) -> np.ndarray:
    """A small random parameter vector for tests (not the library's init)."""
    rng = np.random.default_rng(seed)
    off, total = layout()
    flat = np.zeros(total, dtype=np.float32)
    for k, (o, shape) in off.items():
        n = int(np.prod(shape)) if shape else 1
        fan = shape[-1] if shape else 1
        flat[o:o + n] = rng.uniform(-1, 1, n) * np.sqrt(3.0 / fan)
    return flat
