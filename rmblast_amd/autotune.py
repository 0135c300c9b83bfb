"""Autotuner of the search's kernel schedule, per matrix and word size.

Sweeps the rmb_tuning choices (pipeline batches, the preliminary and traceback
DPs' LDS band rings and workgroups per CU, lane- or wave-run traceback) on a
sample workload, by
coordinate descent from the library's defaults: each parameter in turn takes
each of its values with the others held at the best so far, and a value is
kept when its step is faster by more than MARGIN (measurement noise; the
defaults stay otherwise).  A trial is rmblastn's whole step, the search
with its outfmt 6 rows written as it runs (rmb_search_report, the
RepeatMasker field list, mask level 80), timed as the median of `reps` runs:
the batch count trades the pipeline's overlap windows against its tail, and
a run that happens to keep every window should not decide it.  Every configuration must return the same
HSPs as the defaults (the schedule changes only the time); a difference is an
error.  The best configuration is written to rmblast_amd/tuning.tsv under
(matrix file name, word size, X-drop gapped, X-drop final), where
Searcher(..., tuning="auto") and the rmblastn front end's --tuning-file find it.

    python -m rmblast_amd.autotune --workload uneven --records 120 --families 1500
"""
import argparse
import hashlib
import os
import sys
import tempfile
import time

import numpy as np

import rmblast_amd
from rmblast_amd import synth

MARGIN = 0.005  # a value is kept when it is this much faster than the best so far

SPACE = {  # candidate values per field; 0 is the library's default
    "batches": [0, 4, 12],
    "prelim_ring": [0, 16, 64],
    "prelim_wg_per_cu": [0, 2, 4, 7, 8, 9],
    "tb_ring": [0, 64, 256],
    "tb_wg_per_cu": [0, 2, 4, 7],
    "tb_wave_halves": [0, -1, 1 << 30],  # default regime, always lane-run, always wave-run
}


def result_digest(hsps, ops):
    h = hashlib.sha256()
    order = np.lexsort((hsps["s_start"], hsps["q_start"], hsps["strand"], hsps["subject"], hsps["query"]))
    for f in ("query", "subject", "score", "q_start", "q_end", "s_start", "s_end", "strand", "ops_count"):
        h.update(np.ascontiguousarray(hsps[f][order]).tobytes())
    return h.hexdigest(), len(hsps)


RM_FIELDS = ("score perc_sub perc_query_gap perc_db_gap qseqid qstart qend qlen sstrand sseqid sstart send slen "
             "kdiv cpg_kdiv transi transv cpg_sites qseq sseq")


def timed(s, queries, reps):
    qids = [f"q{i}" for i in range(len(queries))]
    sids = [f"s{i}" for i in range(len(s.subjects))]
    ts, dig = [], None
    for _ in range(reps):
        t = time.perf_counter()
        s.search_report(queries, RM_FIELDS, qids, sids, 80, copy=False)
        ts.append(time.perf_counter() - t)
        dig = result_digest(*s.results())
    return float(np.median(ts)), dig


def tune(s, queries, reps=3, log=print):
    """Coordinate descent over SPACE; returns (best config, its time, trials)."""
    s.set_tuning({})
    timed(s, queries, 1)  # buffers sized by a first search
    t0, ref = timed(s, queries, reps)
    log(f"defaults: {t0 * 1e3:.1f} ms, {ref[1]} HSPs")
    best, best_t, trials = {k: 0 for k in SPACE}, t0, [({k: 0 for k in SPACE}, t0)]
    for field, values in SPACE.items():
        for v in values:
            if v == best[field]:
                continue
            cfg = dict(best, **{field: v})
            s.set_tuning(cfg)
            t, dig = timed(s, queries, reps)
            if dig != ref:
                raise RuntimeError(f"tuning {cfg} changed the results: {dig} vs {ref}")
            trials.append((cfg, t))
            better = t < best_t * (1.0 - MARGIN)
            log(f"{field}={v}: {t * 1e3:.1f} ms" + ("  (best)" if better else ""))
            if better:
                best, best_t = cfg, t
    s.set_tuning(best)
    return best, best_t, trials


def write_entry(path, key, cfg, ms, workload):
    rows = {}
    head = ["matrix", "word_size", "xdrop_gap", "xdrop_gap_final", *rmblast_amd.TUNING_FIELDS, "ms", "workload"]
    if os.path.exists(path):
        with open(path) as f:
            lines = [l.rstrip("\n").split("\t") for l in f if l.strip() and not l.startswith("#")]
        for cols in lines[1:]:
            rows[tuple(cols[:4])] = cols
    rows[tuple(str(x) for x in key)] = [str(x) for x in key] + [str(cfg.get(k, 0)) for k in rmblast_amd.TUNING_FIELDS] + \
        [f"{ms:.1f}", workload]
    with open(path, "w") as f:
        f.write("# written by rmblast_amd/autotune.py: the fastest rmb_tuning per (matrix, word size, X-drops)\n")
        f.write("\t".join(head) + "\n")
        for k in sorted(rows):
            f.write("\t".join(rows[k]) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("uneven", "scale", "gappy"), default="uneven")
    ap.add_argument("--records", type=int, default=120)
    ap.add_argument("--families", type=int, default=1500)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=rmblast_amd.TUNING_PATH)
    a = ap.parse_args(argv)
    if a.workload == "uneven":
        w = synth.uneven_config(n_records=a.records, n_families=a.families)
        xdrop = (20, 30, 100)
    elif a.workload == "scale":
        w = synth.subset(synth.scale_config(), a.records, a.families)
        xdrop = (20, 30, 100)
    else:
        w = synth.gappy_config()
        xdrop = (20, 120, 400)
    tmp = tempfile.mkdtemp(prefix="rmb_tune_")
    mpath = os.path.join(tmp, "synth.matrix")
    with open(mpath, "w") as f:
        f.write(synth.MATRIX_TEXT)
    s = rmblast_amd.Searcher(rmblast_amd.load_matrix(mpath), gap_open=20, gap_extend=5, word_size=7,
                             xdrop_ungap=xdrop[0], xdrop_gap=xdrop[1], xdrop_gap_final=xdrop[2],
                             min_raw_gapped_score=200, complexity_adjust=True)
    s.load_library([f[1] for f in w.families])
    queries = [r[1] for r in w.records]
    mbp = sum(len(q) for q in queries) / 1e6
    print(f"[autotune] {a.workload}: {len(queries)} records ({mbp:.1f} Mbp) x {len(w.families)} families",
          flush=True)
    best, t, trials = tune(s, queries, a.reps, log=lambda m: print("[autotune] " + m, flush=True))
    key = rmblast_amd.tuning_key(mpath, 7, xdrop[1], xdrop[2])
    write_entry(a.out, key, best, t * 1e3, f"{a.workload} {len(queries)} records x {len(w.families)} families")
    print(f"[autotune] best {best}: {t * 1e3:.1f} ms -> {a.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
