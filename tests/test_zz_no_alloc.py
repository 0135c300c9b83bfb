"""rmb_align_score_only_noalloc: score-only X-drop DPs on device-resident
halves with the library's fixed scratch area, against the existing entry
point rmb_tap_align_score_only, on two calls in a row (the second smaller),
with halves whose bands outgrow the 64-cell LDS ring (the rerun path) and the
size limit.  The test's device buffers come from the HIP runtime the library
itself is linked against."""
import ctypes

import numpy as np
import pytest

import rmblast_amd
from rmblast_amd import synth
from conftest import golden

pytestmark = pytest.mark.gpu

HALF_DTYPE = np.dtype([("a", "<u8"), ("b", "<u8"), ("m", "<u4"), ("n", "<u4"), ("xdrop", "<i4"),
                       ("reverse", "<u4")])
H2D, D2H = 1, 2


def hip_runtime():
    """The libamdhip64 that librmblast_hip.so loaded (same process, same runtime)."""
    rmblast_amd.lib()
    path = next(l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64.so" in l)
    h = ctypes.CDLL(path)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(nbytes, 1)) == 0

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data, arr.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = ctypes.c_void_p()


def make_halves(rng, n):
    """Mutated copies, unrelated pairs, tandem repeats (inner sequence at most
    1,000 bases, so no band can outgrow 1,024 cells) and ambiguity codes, at
    X-drop 30, 100 and 400 (bands past 64 cells take the rerun path)."""
    pairs = []
    for i in range(n):
        kind = int(rng.integers(0, 4))
        la = int(rng.integers(1, 1500))
        a = rng.integers(0, 4, la).astype(np.uint8)
        if kind == 0:
            b = synth.mutate(rng, a, float(rng.uniform(0.0, 0.3)), float(rng.uniform(0, 0.08)))
        elif kind == 1:
            b = rng.integers(0, 4, int(rng.integers(1, 1500))).astype(np.uint8)
        elif kind == 2:
            unit = rng.integers(0, 4, int(rng.integers(1, 7))).astype(np.uint8)
            a = np.tile(unit, la // len(unit) + 1)[:la]
            b = synth.mutate(rng, np.tile(unit, 200)[:int(rng.integers(50, 1000))], 0.05, 0.02)[:1000]
        else:
            b = synth.mutate(rng, a, 0.1, 0.02)
            b[rng.random(len(b)) < 0.02] = 14
        if len(b) == 0:
            b = np.zeros(1, dtype=np.uint8)
        pairs.append((a, b, [30, 100, 400][i % 3], i % 2))
    return pairs


def run_noalloc(s, hip, pairs):
    # 64 bytes of tail: the DP kernels read the sequences in 16-byte blocks
    seqs = np.concatenate([np.concatenate([a, b]) for a, b, _, _ in pairs] + [np.zeros(64, dtype=np.uint8)])
    dseq = DevBuf(hip, seqs.nbytes)
    dseq.put(seqs)
    halves = np.zeros(len(pairs), dtype=HALF_DTYPE)
    at = 0
    for i, (a, b, x, rev) in enumerate(pairs):
        halves[i] = (dseq.p.value + at, dseq.p.value + at + len(a), len(a), len(b), x, rev)
        at += len(a) + len(b)
    dh = DevBuf(hip, halves.nbytes)
    dh.put(halves)
    dout = DevBuf(hip, 12 * len(pairs))
    try:
        rc = s._lib.rmb_align_score_only_noalloc(s._ctx, dh.p, len(pairs), dout.p)
        s._check(rc, "rmb_align_score_only_noalloc")
        return dout.get(np.int32, 3 * len(pairs)).reshape(-1, 3)
    finally:
        for d in (dseq, dh, dout):
            d.free()


def test_noalloc_score_dps_match_the_tap():
    hip = hip_runtime()
    m = rmblast_amd.load_matrix(golden("comparison.matrix"))
    s = rmblast_amd.Searcher(m, gap_open=20, gap_extend=5)
    s._lib.rmb_align_score_only_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                    ctypes.c_void_p]
    rng = np.random.default_rng(7)
    for n in (3000, 700):    # two calls in a row, the second smaller: the scratch keeps the first's contents
        pairs = make_halves(rng, n)
        got = run_noalloc(s, hip, pairs)
        want = np.array([s.tap_align_score_only(a, b, x, rev) for a, b, x, rev in pairs], dtype=np.int64)
        assert np.array_equal(got.astype(np.int64), want), n
        assert (want[:, 0] > 100).sum() > n // 10
    # the size limit is refused before anything runs
    assert s._lib.rmb_align_score_only_noalloc(s._ctx, ctypes.c_void_p(8), 65537, ctypes.c_void_p(8)) == 1
    assert s._lib.rmb_align_score_only_noalloc(s._ctx, None, 5, None) == 1
    assert s._lib.rmb_align_score_only_noalloc(s._ctx, None, 0, None) == 0
