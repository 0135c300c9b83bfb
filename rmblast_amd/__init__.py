"""rmblast_amd: MI355X (gfx950) search path of rmblastn behind a C ABI.

This package is a thin ctypes binding of librmblast_hip.so
(include/rmblast_hip.h), mirroring the reference's rmblast-lib search API
(rmblast-lib/src/search/mod.rs:9-21): a context holds the parameters and the
matrix, the library (TE families) is loaded once, and `search` runs a batch of
query records against it on the GPU.  There is no CPU fallback: importing the
package fails if the HIP library has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "librmblast_hip.so")
CLI_PATH = os.path.join(_HERE, "_build", "rmblastn")

EXPORTED = [
    "rmb_matrix_load", "rmb_ctx_create", "rmb_ctx_destroy", "rmb_last_error", "rmb_library_load",
    "rmb_search", "rmb_search_pairs", "rmb_dust_intervals", "rmb_results_size", "rmb_results_copy", "rmb_stats", "rmb_tap_ungapped",
    "rmb_tap_align_score_only", "rmb_tap_align_ex", "rmb_ka_lookup", "rmb_ungapped_cutoff", "rmb_mask_level",
    "rmb_report_order", "rmb_report_tabular", "rmb_report_copy", "rmb_report_text", "rmb_align_stats_of",
    "rmb_ka_stats_resolve", "rmb_ka_search_space", "rmb_ka_evalue", "rmb_ctx_set_tuning",
    "rmb_tap_lookup", "rmb_tap_seeds", "rmb_tap_prelims", "rmb_tap_subject",
]


class RmbParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "gap_open", "gap_extend", "word_size", "xdrop_ungap", "xdrop_gap", "xdrop_gap_final",
        "min_raw_gapped_score", "ungapped_cutoff", "complexity_adjust", "dust")] + [
        ("ka_lambda", ctypes.c_double), ("ka_k", ctypes.c_double)]


class RmbTuning(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("batches", "prelim_ring", "prelim_wg_per_cu", "tb_ring",
                                              "tb_wg_per_cu", "tb_wave_halves")]


TUNING_FIELDS = ("batches", "prelim_ring", "prelim_wg_per_cu", "tb_ring", "tb_wg_per_cu", "tb_wave_halves")
TUNING_PATH = os.path.join(_HERE, "tuning.tsv")


def tuning_key(matrix_name: str, word_size: int, xdrop_gap: int, xdrop_gap_final: int):
    return (os.path.basename(matrix_name), int(word_size), int(xdrop_gap), int(xdrop_gap_final))


def read_tuning(path: str = TUNING_PATH) -> dict:
    """The autotuner's choices: {(matrix, ws, xdrop_gap, xdrop_gap_final): {field: value}}."""
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        head = None
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            cols = line.rstrip("\n").split("\t")
            if head is None:
                head = cols
                continue
            row = dict(zip(head, cols))
            key = (row["matrix"], int(row["word_size"]), int(row["xdrop_gap"]), int(row["xdrop_gap_final"]))
            out[key] = {k: int(row.get(k, 0)) for k in TUNING_FIELDS}
    return out


class RmbKaStatsOptions(ctypes.Structure):
    _fields_ = [("matrix_lambda", ctypes.c_double), ("matrix_k", ctypes.c_double), ("matrix_alpha", ctypes.c_double),
                ("matrix_beta", ctypes.c_double), ("ncbi_compat", ctypes.c_int32), ("pad", ctypes.c_int32)]


class RmbKaStats(ctypes.Structure):
    _fields_ = [("source", ctypes.c_int32), ("pad", ctypes.c_int32), ("lambda_", ctypes.c_double),
                ("k", ctypes.c_double), ("log_k", ctypes.c_double), ("h", ctypes.c_double),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double)]


KA_SOURCES = ("sentinel", "table", "cli", "comment")


class RmbReportSpec(ctypes.Structure):
    _fields_ = [("mask_level", ctypes.c_int32), ("pad", ctypes.c_int32), ("fields", ctypes.c_char_p),
                ("query_ids", ctypes.POINTER(ctypes.c_char_p)), ("subject_ids", ctypes.POINTER(ctypes.c_char_p)),
                ("ka", ctypes.POINTER(RmbKaStats)), ("evalue_max", ctypes.c_double)]


class RmbKaCutoffInfo(ctypes.Structure):
    _fields_ = [("found", ctypes.c_int32), ("es", ctypes.c_int32), ("lambda_", ctypes.c_double),
                ("k", ctypes.c_double), ("searchsp", ctypes.c_uint64)]


class RmbMatrix(ctypes.Structure):
    _fields_ = [("scores", (ctypes.c_int32 * 16) * 16), ("freqs", ctypes.c_double * 16),
                ("lambda_", ctypes.c_double)]


HSP_DTYPE = np.dtype([("query", "<u4"), ("subject", "<u4"), ("score", "<i4"), ("q_start", "<u4"),
                      ("q_end", "<u4"), ("s_start", "<u4"), ("s_end", "<u4"), ("strand", "<u4"),
                      ("ops_offset", "<u4"), ("ops_count", "<u4")])

STAT_NAMES = ["seed_evals", "ungapped_hits", "prelim_dps", "traceback_dps", "dp_cells", "tb_cells",
              "hsps", "ungapped_steps", "t_lookup_us", "t_seed_us", "t_prelim_us", "t_traceback_us",
              "t_total_us", "t_phase1_kernel_ns", "t_gapped_kernel_ns", "phase1_launches", "seeds_staged",
              "p1_ticks_scan", "p1_ticks_reserve", "p1_ticks_extend", "p1_ticks_bucket", "p1_ticks_replay",
              "p1_ticks_emit", "t_prelim_kernel_ns", "t_report_tail_us"]


class RmbError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise RmbError(f"{path} is missing: build it with `python -m rmblast_amd.build`")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    u8p = P(ctypes.c_uint8)
    lib.rmb_matrix_load.argtypes = [ctypes.c_char_p, P(RmbMatrix), ctypes.c_char_p, ctypes.c_size_t]
    lib.rmb_ctx_create.argtypes = [ctypes.c_int, P(RmbParams), P(RmbMatrix), P(ctypes.c_void_p)]
    lib.rmb_ctx_destroy.argtypes = [ctypes.c_void_p]
    lib.rmb_ctx_set_tuning.argtypes = [ctypes.c_void_p, P(RmbTuning)]
    lib.rmb_last_error.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.rmb_library_load.argtypes = [ctypes.c_void_p, ctypes.c_uint32, P(u8p), P(ctypes.c_uint32)]
    lib.rmb_search.argtypes = [ctypes.c_void_p, ctypes.c_uint32, P(u8p), P(ctypes.c_uint32)]
    lib.rmb_search_pairs.argtypes = [ctypes.c_void_p, ctypes.c_uint32, P(u8p), P(ctypes.c_uint32), ctypes.c_uint64,
                                     P(ctypes.c_uint32), P(ctypes.c_uint32)]
    lib.rmb_results_size.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), P(ctypes.c_uint64)]
    lib.rmb_results_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.rmb_stats.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), ctypes.c_int]
    lib.rmb_ka_lookup.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, P(ctypes.c_double),
                                  P(ctypes.c_int32)]
    lib.rmb_ungapped_cutoff.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_uint64, ctypes.c_uint64, P(ctypes.c_int32), P(RmbKaCutoffInfo)]
    lib.rmb_mask_level.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32,
                                   P(ctypes.c_uint8)]
    lib.rmb_report_order.argtypes = [ctypes.c_void_p, P(RmbReportSpec), P(ctypes.c_uint64), P(ctypes.c_uint64)]
    lib.rmb_ka_stats_resolve.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, P(RmbKaStatsOptions),
                                         P(RmbKaStats), ctypes.c_char_p, ctypes.c_size_t]
    lib.rmb_ka_search_space.argtypes = [P(RmbKaStats), ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                        P(ctypes.c_int64), P(ctypes.c_int32)]
    lib.rmb_ka_evalue.argtypes = [P(RmbKaStats), ctypes.c_int64, ctypes.c_int32, P(ctypes.c_double),
                                  P(ctypes.c_double)]
    lib.rmb_report_tabular.argtypes = [ctypes.c_void_p, P(RmbReportSpec), P(ctypes.c_uint64)]
    lib.rmb_search_report.argtypes = [ctypes.c_void_p, ctypes.c_uint32, P(u8p), P(ctypes.c_uint32), P(RmbReportSpec),
                                      P(ctypes.c_uint64)]
    lib.rmb_report_copy.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.rmb_report_text.argtypes = [ctypes.c_void_p, P(ctypes.c_void_p), P(ctypes.c_uint64)]
    lib.rmb_dust_intervals.argtypes = [u8p, ctypes.c_uint64, P(ctypes.c_uint64), P(ctypes.c_uint64)]
    lib.rmb_tap_ungapped.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint32, ctypes.c_uint32,
                                     P(ctypes.c_uint32), P(ctypes.c_uint64)]
    lib.rmb_tap_subject.argtypes = [ctypes.c_void_p, ctypes.c_uint32, u8p, u8p, P(ctypes.c_uint32)]
    lib.rmb_tap_seeds.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint32, ctypes.c_uint32, P(ctypes.c_uint32),
                                  P(ctypes.c_uint64)]
    lib.rmb_tap_prelims.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint32, ctypes.c_uint32, P(ctypes.c_uint32),
                                    P(ctypes.c_uint64)]
    lib.rmb_tap_lookup.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint32, P(ctypes.c_int32), P(ctypes.c_uint32),
                                   P(ctypes.c_uint64), P(ctypes.c_uint32), P(ctypes.c_uint64)]
    lib.rmb_tap_align_score_only.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint32, u8p, ctypes.c_uint32,
                                             ctypes.c_int32, ctypes.c_int32, P(ctypes.c_int64)]
    lib.rmb_tap_align_ex.argtypes = [ctypes.c_void_p, u8p, ctypes.c_uint32, u8p, ctypes.c_uint32,
                                     ctypes.c_int32, ctypes.c_int32, P(ctypes.c_int64), P(ctypes.c_uint32),
                                     ctypes.c_uint32]
    return lib


_LIB = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


def load_matrix(path: str) -> RmbMatrix:
    m = RmbMatrix()
    err = ctypes.create_string_buffer(256)
    if lib().rmb_matrix_load(path.encode(), ctypes.byref(m), err, 256):
        raise RmbError(f"matrix {path}: {err.value.decode()}")
    return m


def dust_intervals(bases) -> list:
    """Symmetric DUST intervals [(first, last), ...] (0-based, inclusive) of
    BLASTNA bases without sentinels (host code of the library)."""
    b = np.ascontiguousarray(bases, dtype=np.uint8)
    n = ctypes.c_uint64(0)
    ptr = b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    if lib().rmb_dust_intervals(ptr, len(b), None, ctypes.byref(n)):
        raise RmbError("rmb_dust_intervals failed")
    out = np.zeros(2 * max(n.value, 1), dtype=np.uint64)
    lib().rmb_dust_intervals(ptr, len(b), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(n))
    return [(int(out[2 * k]), int(out[2 * k + 1])) for k in range(n.value)]


def ka_lookup(matrix_name: str, gap_open: int, gap_extend: int):
    """(lambda, K) of a RepeatMasker matrix at its gap costs from NCBI's ALP
    table (ka_cutoff.rs:105-114), or None."""
    out = (ctypes.c_double * 2)()
    found = ctypes.c_int32(0)
    if lib().rmb_ka_lookup(matrix_name.encode(), gap_open, gap_extend, out, ctypes.byref(found)):
        raise RmbError("rmb_ka_lookup failed")
    return (out[0], out[1]) if found.value else None


def ungapped_cutoff(matrix_name: str, gap_open: int, gap_extend: int, min_raw_gapped_score: int,
                    query_length: int, avg_subj_length: int):
    """(cutoff, info or None) as ka_cutoff.rs:126-161 computes them; info is
    a dict {lambda, k, searchsp, es} when the KA branch applied."""
    c = ctypes.c_int32(0)
    info = RmbKaCutoffInfo()
    if lib().rmb_ungapped_cutoff(matrix_name.encode(), gap_open, gap_extend, min_raw_gapped_score, query_length,
                                 avg_subj_length, ctypes.byref(c), ctypes.byref(info)):
        raise RmbError("rmb_ungapped_cutoff failed")
    if not info.found:
        return c.value, None
    return c.value, {"lambda": info.lambda_, "k": info.k, "searchsp": info.searchsp, "es": info.es}


def ka_stats(matrix_path: str, gap_open: int, gap_extend: int, matrix_lambda=0.0, matrix_k=0.0,
             matrix_alpha=0.0, matrix_beta=0.0, ncbi_compat=False) -> RmbKaStats:
    """E-value statistics of a matrix file at its gap costs (ka_stats.rs:153-300)."""
    opt = RmbKaStatsOptions(matrix_lambda, matrix_k, matrix_alpha, matrix_beta, int(bool(ncbi_compat)), 0)
    out = RmbKaStats()
    err = ctypes.create_string_buffer(256)
    if lib().rmb_ka_stats_resolve(matrix_path.encode(), gap_open, gap_extend, ctypes.byref(opt), ctypes.byref(out),
                                  err, 256):
        raise RmbError(err.value.decode())
    return out


def ka_search_space(ka: RmbKaStats, query_length: int, db_length: int, db_num_seqs: int):
    """(effective search space, length adjustment) of one query."""
    sp, la = ctypes.c_int64(0), ctypes.c_int32(0)
    lib().rmb_ka_search_space(ctypes.byref(ka), query_length, db_length, db_num_seqs, ctypes.byref(sp),
                              ctypes.byref(la))
    return sp.value, la.value


def ka_evalue(ka: RmbKaStats, eff_searchsp: int, score: int):
    """(E-value, bit score) of a raw score."""
    e, b = ctypes.c_double(0), ctypes.c_double(0)
    lib().rmb_ka_evalue(ctypes.byref(ka), eff_searchsp, score, ctypes.byref(e), ctypes.byref(b))
    return e.value, b.value


def mask_level(hsps: np.ndarray, q_len: int, level: int) -> np.ndarray:
    """apply_mask_level (engine.rs:863-974, cross-subject form) over one
    query's HSPs (an HSP_DTYPE array): a bool array of the survivors."""
    h = np.ascontiguousarray(hsps, dtype=HSP_DTYPE)
    keep = np.zeros(len(h), dtype=np.uint8)
    if lib().rmb_mask_level(h.ctypes.data_as(ctypes.c_void_p), len(h), q_len, level,
                            keep.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))):
        raise RmbError("rmb_mask_level failed")
    return keep.astype(bool)


def _u8ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


class Searcher:
    """One GPU context: parameters + matrix + loaded library.

    Mirrors build_query_lookup + search_with_query_lookup (engine.rs:606-729)
    over every subject, as rmblastn's search_db_parallel drives them
    (rmblastn/src/main.rs:1133-1234)."""

    def __init__(self, matrix: RmbMatrix, *, gap_open=20, gap_extend=5, word_size=11, xdrop_ungap=20,
                 xdrop_gap=30, xdrop_gap_final=100, min_raw_gapped_score=0, ungapped_cutoff=-1,
                 complexity_adjust=False, device=0, matrix_name=None, tuning=None):
        """matrix_name: the -matrix path; when it names a matrix of NCBI's
        ALP table at these gap costs, the ungapped cutoff is resolved per
        query record as rmblastn does (ungapped_cutoff = -1 only).
        tuning: a dict of rmb_tuning fields, or "auto" for the autotuner's
        entry for (matrix_name, word_size, X-drops) in tuning.tsv if any."""
        self._lib = lib()
        ka = ka_lookup(matrix_name, gap_open, gap_extend) if matrix_name else None
        p = RmbParams(gap_open, gap_extend, word_size, xdrop_ungap, xdrop_gap, xdrop_gap_final,
                      min_raw_gapped_score, ungapped_cutoff, int(bool(complexity_adjust)), 0,
                      ka[0] if ka else 0.0, ka[1] if ka else 0.0)
        self.params = p
        self.matrix = matrix
        ctx = ctypes.c_void_p()
        rc = self._lib.rmb_ctx_create(device, ctypes.byref(p), ctypes.byref(matrix), ctypes.byref(ctx))
        self._ctx = ctx
        self._check(rc, "rmb_ctx_create")
        self._keep = []
        self.tuning = None
        if tuning == "auto":
            tuning = read_tuning().get(tuning_key(matrix_name or "", word_size, xdrop_gap, xdrop_gap_final))
        if tuning:
            self.set_tuning(tuning)

    def set_tuning(self, tuning: dict):
        t = RmbTuning(*[int(tuning.get(k, 0)) for k in TUNING_FIELDS])
        self._check(self._lib.rmb_ctx_set_tuning(self._ctx, ctypes.byref(t)), "rmb_ctx_set_tuning")
        self.tuning = dict(tuning)

    def _check(self, rc, what):
        if rc:
            buf = ctypes.create_string_buffer(512)
            if self._ctx:
                self._lib.rmb_last_error(self._ctx, buf, 512)
            raise RmbError(f"{what} failed ({rc}): {buf.value.decode()}")

    def close(self):
        if self._ctx:
            self._lib.rmb_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_library(self, seqs):
        arrs = [np.ascontiguousarray(s, dtype=np.uint8) for s in seqs]
        ptrs = (ctypes.POINTER(ctypes.c_uint8) * len(arrs))(*[_u8ptr(a) for a in arrs])
        lens = (ctypes.c_uint32 * len(arrs))(*[len(a) for a in arrs])
        self._check(self._lib.rmb_library_load(self._ctx, len(arrs), ptrs, lens), "rmb_library_load")
        self.subjects = arrs

    def search(self, queries):
        arrs = [np.ascontiguousarray(q, dtype=np.uint8) for q in queries]
        ptrs = (ctypes.POINTER(ctypes.c_uint8) * len(arrs))(*[_u8ptr(a) for a in arrs])
        lens = (ctypes.c_uint32 * len(arrs))(*[len(a) for a in arrs])
        self._check(self._lib.rmb_search(self._ctx, len(arrs), ptrs, lens), "rmb_search")
        return self._results()

    def search_pairs(self, queries, pair_query, pair_subject):
        """Search only the listed (query index, subject index) pairs."""
        arrs = [np.ascontiguousarray(q, dtype=np.uint8) for q in queries]
        ptrs = (ctypes.POINTER(ctypes.c_uint8) * len(arrs))(*[_u8ptr(a) for a in arrs])
        lens = (ctypes.c_uint32 * len(arrs))(*[len(a) for a in arrs])
        pq = np.ascontiguousarray(pair_query, dtype=np.uint32)
        ps = np.ascontiguousarray(pair_subject, dtype=np.uint32)
        assert pq.shape == ps.shape
        self._check(self._lib.rmb_search_pairs(self._ctx, len(arrs), ptrs, lens, len(pq),
                                               pq.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                               ps.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
                    "rmb_search_pairs")
        return self._results()

    def _results(self):
        nh, no = ctypes.c_uint64(), ctypes.c_uint64()
        self._lib.rmb_results_size(self._ctx, ctypes.byref(nh), ctypes.byref(no))
        hsps = np.zeros(nh.value, dtype=HSP_DTYPE)
        ops = np.zeros(no.value, dtype=np.uint32)
        self._lib.rmb_results_copy(self._ctx, hsps.ctypes.data_as(ctypes.c_void_p),
                                   ops.ctypes.data_as(ctypes.c_void_p))
        return hsps, ops

    def report_order(self, mask_level: int, ka: RmbKaStats = None, evalue_max: float = 0.0) -> np.ndarray:
        """Indices of the last search's HSPs that rmblastn reports, in its order."""
        spec = RmbReportSpec(mask_level, 0, None, None, None, ctypes.pointer(ka) if ka is not None else None,
                             evalue_max)
        n = ctypes.c_uint64(0)
        self._check(self._lib.rmb_report_order(self._ctx, ctypes.byref(spec), None, ctypes.byref(n)),
                    "rmb_report_order")
        rows = np.zeros(n.value, dtype=np.uint64)
        self._check(self._lib.rmb_report_order(self._ctx, ctypes.byref(spec),
                                               rows.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                               ctypes.byref(n)), "rmb_report_order")
        return rows

    def search_report(self, queries, fields: str, query_ids, subject_ids, mask_level: int = 101,
                      copy: bool = True, ka: RmbKaStats = None, evalue_max: float = 0.0):
        """rmb_search_report: search `queries` and write their outfmt 6 text
        while the search runs (the bytes report_tabular gives after search).
        Returns the text as report_tabular does; results() gives the HSPs."""
        arrs = [np.ascontiguousarray(q, dtype=np.uint8) for q in queries]
        ptrs = (ctypes.POINTER(ctypes.c_uint8) * len(arrs))(*[_u8ptr(a) for a in arrs])
        lens = (ctypes.c_uint32 * len(arrs))(*[len(a) for a in arrs])
        qids = (ctypes.c_char_p * len(query_ids))(*[q.encode() for q in query_ids])
        sids = (ctypes.c_char_p * len(subject_ids))(*[x.encode() for x in subject_ids])
        spec = RmbReportSpec(mask_level, 0, fields.encode(), qids, sids,
                             ctypes.pointer(ka) if ka is not None else None, evalue_max)
        n = ctypes.c_uint64(0)
        self._check(self._lib.rmb_search_report(self._ctx, len(arrs), ptrs, lens, ctypes.byref(spec),
                                                ctypes.byref(n)), "rmb_search_report")
        return self._report_text(copy)

    def results(self):
        """The last search's HSPs and ops (HSP_DTYPE array, uint32 ops)."""
        return self._results()

    def _report_text(self, copy: bool):
        ptr, size = ctypes.c_void_p(), ctypes.c_uint64()
        self._check(self._lib.rmb_report_text(self._ctx, ctypes.byref(ptr), ctypes.byref(size)), "rmb_report_text")
        if size.value == 0:
            return b"" if copy else memoryview(b"")
        if copy:
            return ctypes.string_at(ptr.value, size.value)
        return memoryview((ctypes.c_char * size.value).from_address(ptr.value)).toreadonly()

    def report_tabular(self, fields: str, query_ids, subject_ids, mask_level: int = 101, copy: bool = True,
                       ka: RmbKaStats = None, evalue_max: float = 0.0):
        """outfmt 6 text of the last search (mask level, order, statistics and
        rows on the library's host threads): bytes, or with copy=False a
        read-only memoryview of the library's buffer, valid until the next
        search or report."""
        qids = (ctypes.c_char_p * len(query_ids))(*[q.encode() for q in query_ids])
        sids = (ctypes.c_char_p * len(subject_ids))(*[x.encode() for x in subject_ids])
        spec = RmbReportSpec(mask_level, 0, fields.encode(), qids, sids,
                             ctypes.pointer(ka) if ka is not None else None, evalue_max)
        n = ctypes.c_uint64(0)
        self._check(self._lib.rmb_report_tabular(self._ctx, ctypes.byref(spec), ctypes.byref(n)),
                    "rmb_report_tabular")
        return self._report_text(copy)

    def stats(self):
        out = (ctypes.c_uint64 * len(STAT_NAMES))()
        self._lib.rmb_stats(self._ctx, out, len(STAT_NAMES))
        return dict(zip(STAT_NAMES, list(out)))

    def tap_ungapped(self, query, subject: int):
        q = np.ascontiguousarray(query, dtype=np.uint8)
        n = ctypes.c_uint64(0)
        self._check(self._lib.rmb_tap_ungapped(self._ctx, _u8ptr(q), len(q), subject, None, ctypes.byref(n)),
                    "rmb_tap_ungapped")
        out = np.zeros((max(n.value, 1), 8), dtype=np.uint32)
        self._check(self._lib.rmb_tap_ungapped(self._ctx, _u8ptr(q), len(q), subject,
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                               ctypes.byref(n)), "rmb_tap_ungapped")
        return out[:n.value]

    def _tap_rows(self, fn, query, subject, width, what):
        q = np.ascontiguousarray(query, dtype=np.uint8)
        n = ctypes.c_uint64(0)
        self._check(fn(self._ctx, _u8ptr(q), len(q), subject, None, ctypes.byref(n)), what)
        out = np.zeros((max(n.value, 1), width), dtype=np.uint32)
        self._check(fn(self._ctx, _u8ptr(q), len(q), subject, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                       ctypes.byref(n)), what)
        return out[:n.value]

    def tap_subject(self, subject: int):
        """Subject `subject` as the search scans it: (resolved plus strand, its
        reverse complement), from the device (rmb_tap_subject)."""
        n = ctypes.c_uint32(0)
        self._check(self._lib.rmb_tap_subject(self._ctx, subject, None, None, ctypes.byref(n)), "rmb_tap_subject")
        plus = np.zeros(max(n.value, 1), dtype=np.uint8)
        rc = np.zeros(max(n.value, 1), dtype=np.uint8)
        self._check(self._lib.rmb_tap_subject(self._ctx, subject, _u8ptr(plus), _u8ptr(rc), ctypes.byref(n)),
                    "rmb_tap_subject")
        return plus[:n.value], rc[:n.value]

    def tap_seeds(self, query, subject: int):
        """Post-dedup seeds {q_off, s_off, strand} in scan order (rmb_tap_seeds)."""
        return self._tap_rows(self._lib.rmb_tap_seeds, query, subject, 3, "rmb_tap_seeds")

    def tap_prelims(self, query, subject: int):
        """Phase 2a's preliminary HSPs {q_seed, s_seed, q_start, q_end, s_start,
        s_end, score, strand} (rmb_tap_prelims)."""
        return self._tap_rows(self._lib.rmb_tap_prelims, query, subject, 8, "rmb_tap_prelims")

    def tap_lookup(self, query):
        """The query's lookup table: (lut_width, cells[4^lw + 1], entries)."""
        q = np.ascontiguousarray(query, dtype=np.uint8)
        lw = ctypes.c_int32(0)
        nc, ne = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._check(self._lib.rmb_tap_lookup(self._ctx, _u8ptr(q), len(q), ctypes.byref(lw), None, ctypes.byref(nc),
                                             None, ctypes.byref(ne)), "rmb_tap_lookup")
        cells = np.zeros(nc.value, dtype=np.uint32)
        entries = np.zeros(max(ne.value, 1), dtype=np.uint32)
        u32 = ctypes.POINTER(ctypes.c_uint32)
        self._check(self._lib.rmb_tap_lookup(self._ctx, _u8ptr(q), len(q), ctypes.byref(lw), cells.ctypes.data_as(u32),
                                             ctypes.byref(nc), entries.ctypes.data_as(u32), ctypes.byref(ne)),
                    "rmb_tap_lookup")
        return lw.value, cells, entries[:ne.value]

    def tap_align_score_only(self, a, b, xdrop, reverse):
        a = np.ascontiguousarray(a, dtype=np.uint8)
        b = np.ascontiguousarray(b, dtype=np.uint8)
        out = (ctypes.c_int64 * 3)()
        self._check(self._lib.rmb_tap_align_score_only(self._ctx, _u8ptr(a), len(a), _u8ptr(b), len(b), xdrop,
                                                       int(reverse), out), "rmb_tap_align_score_only")
        return tuple(out)

    def tap_align_ex(self, a, b, xdrop, reverse, max_ops=100000):
        a = np.ascontiguousarray(a, dtype=np.uint8)
        b = np.ascontiguousarray(b, dtype=np.uint8)
        out = (ctypes.c_int64 * 4)()
        ops = np.zeros(max_ops, dtype=np.uint32)
        self._check(self._lib.rmb_tap_align_ex(self._ctx, _u8ptr(a), len(a), _u8ptr(b), len(b), xdrop,
                                               int(reverse), out,
                                               ops.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), max_ops),
                    "rmb_tap_align_ex")
        score, alen, blen, nops = tuple(out)
        return score, alen, blen, [(int(o & 3), int(o >> 2)) for o in ops[:nops]]


_CODES = np.full(256, 14, dtype=np.uint8)
for _i, _c in enumerate("ACGTRYMKWSBDHVN"):
    _CODES[ord(_c)] = _i
    _CODES[ord(_c.lower())] = _i
_CODES[ord("U")] = _CODES[ord("u")] = 3


def encode(text: str) -> np.ndarray:
    """IUPAC text -> BLASTNA codes (unknown letters, '-' and 'X' read as N)."""
    return _CODES[np.frombuffer(text.encode(), dtype=np.uint8)]


def read_fasta(path: str):
    recs, name, buf = [], None, []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\r\n")
            if line.startswith(">"):
                if name is not None:
                    recs.append((name, encode("".join(buf))))
                name = line[1:].split()[0]
                buf = []
            elif line.strip():
                buf.append(line.strip())
    if name is not None:
        recs.append((name, encode("".join(buf))))
    return recs
