"""Benchmark: the exact search path at scale on one MI355X.

Workload (the exact-path-at-scale configuration): a 100 Mbp synthetic genome
split into 1,667 records of 60 kbp, searched on both strands against 1,000
synthetic families (log-normal lengths, 100-15,000 bp), word size 7, gap
costs 20/5, X-drop 20/30/100, min_raw_gapped_score 200, complexity adjustment,
mask level 80, no DUST.  One step = one rmb_search_report over all records:
the search (lookup build, seeding, ungapped, preliminary and traceback DP,
ordered replay) and the host tail into final outfmt-6 rows in memory (mask
level, output order, alignment statistics, the RepeatMasker field list), the
rows of each batch of queries written on the host while the next batches
run, with the library already resident on the GPU.  Metric: query Mbp x
families per second.

--workload gappy measures the divergent/gappy stress configuration instead:
one 10 Mbp record (the split query path) against 300 families of 5-15 kbp,
copies 25-40% diverged with 5-8% indels, 5% microsatellite tracts, X-drop
gapped/final 120/400; its CPU baseline sample is the record's first 1.5 Mbp.
--workload parity measures the parity configuration: one 1 Mbp record
against 50 families of 300-6,000 bp (2,000 copies, 5-25% divergence), a
latency-bound search (the traceback DPs of its long copies set the time);
its CPU baseline sample is the whole record on all cores and its first
200 kbp on one.
--workload uneven measures the uneven-batching configuration on its first 400
records (log-uniform 1-500 kbp) against all 5,000 families (log-uniform
50-20,000 bp).  The kernel schedule is the autotuner's entry for the matrix,
word size and X-drops when rmblast_amd/tuning.tsv has one (--tuning default
ignores it).

The CPU baseline is the oracle (a C++ restatement of the reference's search
path) built for this host (-march=native) at run time, on all the host cores
this process may use and on one core, over a bounded sample of the same
records; the GPU's rows for those records are compared with the oracle's.

Prints one JSON line.  --gpus N with N > 1 is refused: this build runs on one GPU.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INT32_VALU_PEAK = 256 * 64 * 2.4e9   # 256 CUs x 64 lanes x 2.4 GHz, int32 ops/s
OPS_PER_UNGAPPED_STEP = 6
OPS_PER_DP_CELL = 12
RM_FIELDS = ("score perc_sub perc_query_gap perc_db_gap qseqid qstart qend qlen sstrand sseqid sstart send slen "
             "kdiv cpg_kdiv transi transv cpg_sites qseq sseq")
MASK_LEVEL = 80


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("scale", "gappy", "uneven", "parity"), default="scale")
    ap.add_argument("--records", type=int, default=None,
                    help="records to search (scale: of the 1,667; uneven: of the 20,000; default 1,667 / 400)")
    ap.add_argument("--families", type=int, default=None, help="families (default 1,000 / 5,000)")
    ap.add_argument("--cpu-records", type=int, default=None,
                    help="records in the all-cores CPU baseline sample (about 15 s of 16-core work; default 300 / 12)")
    ap.add_argument("--cpu1-records", type=int, default=None,
                    help="records in the one-core CPU baseline sample (default 20 / 1)")
    ap.add_argument("--tuning", default="auto", help="'auto' (the autotuner's table entry, if any) or 'default'")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    import torch
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda")
    src.fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[bench] copied {n} bytes x 3", file=sys.stderr)
    return 0


def host_cores():
    """Cores this process may use: its CPU affinity, within the job's CPU share
    where the environment states one (OMP_NUM_THREADS)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def kernel_source_sha():
    return hashlib.sha256(open(os.path.join(ROOT, "rmblast_amd", "csrc", "rmb_kernels.hip"), "rb").read()).hexdigest()[:16]


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported", file=sys.stderr)
        return 2
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)
    import rmblast_amd
    from rmblast_amd import synth

    t0 = time.time()
    gappy = a.workload == "gappy"
    uneven = a.workload == "uneven"
    parity_wl = a.workload == "parity"
    if gappy:
        w = synth.gappy_config()
    elif parity_wl:
        w = synth.parity_config()
    elif uneven:
        w = synth.uneven_config(n_records=a.records or 400, n_families=a.families or 5000)
    else:
        w = synth.scale_config()
        w = synth.subset(w, a.records or 1667, a.families or 1000)
    cpu_records = a.cpu_records or (12 if uneven else 300)
    cpu1_records = a.cpu1_records or (1 if uneven else 20)
    gen_s = time.time() - t0
    print(f"[bench] workload generated in {gen_s:.1f}s: {len(w.records)} records x {len(w.families)} families",
          file=sys.stderr, flush=True)
    tmp = tempfile.mkdtemp(prefix="rmb_bench_")
    oracle_build = None
    if not a.no_cpu:  # the CPU baseline's oracle, built for this host while the GPU runs
        oracle_build = subprocess.Popen(
            ["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8", "MARCH=native", "OUT=" + tmp + "/oracle",
             tmp + "/oracle/rmblastn_oracle"], stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
    mpath = os.path.join(tmp, "synth.matrix")
    with open(mpath, "w") as f:
        f.write(synth.MATRIX_TEXT)
    m = rmblast_amd.load_matrix(mpath)
    xdrop = [20, 120, 400] if gappy else [20, 30, 100]
    params = dict(gap_open=20, gap_extend=5, word_size=7, xdrop_ungap=xdrop[0], xdrop_gap=xdrop[1],
                  xdrop_gap_final=xdrop[2], min_raw_gapped_score=200, complexity_adjust=True)
    s = rmblast_amd.Searcher(m, **params, matrix_name=mpath, tuning="auto" if a.tuning == "auto" else None)
    if s.tuning:
        print(f"[bench] tuning from the autotuner's table: {s.tuning}", file=sys.stderr, flush=True)
    s.load_library([f[1] for f in w.families])
    queries = [r[1] for r in w.records]
    qids = [r[0] for r in w.records]
    sids = [f[0] for f in w.families]
    mbp = sum(len(q) for q in queries) / 1e6
    nfam = len(w.families)

    def step():
        # rmb_search_report: the search, with each batch's rows (mask level,
        # order, statistics, outfmt-6 text) written on the host while the
        # next batches run; the rows end in memory
        t = time.perf_counter()
        rows = s.search_report(queries, RM_FIELDS, qids, sids, MASK_LEVEL, copy=False)
        dt = time.perf_counter() - t
        st = s.stats()
        return dt, (st["t_total_us"] - st["t_report_tail_us"]) / 1e6, st["hsps"], rows

    for i in range(a.warmup):
        step()
        print(f"[bench] warmup {i + 1}/{a.warmup} done", file=sys.stderr, flush=True)
    times, search_times, stats = [], [], []
    rows = b""
    for i in range(a.steps):
        dt, ds, nh, rows = step()
        times.append(dt)
        search_times.append(ds)
        stats.append(s.stats())
        nrows = int(np.count_nonzero(np.frombuffer(rows, dtype=np.uint8) == 10))
        print(f"[bench] step {i + 1}/{a.steps}: {dt:.3f}s ({ds:.3f}s search), {nh} HSPs, {nrows} rows",
              file=sys.stderr, flush=True)
    rows = bytes(rows)  # the last step's rows, copied out of the library after the timed steps
    step_s = sum(times) / a.steps
    search_s = sum(search_times) / a.steps
    st = stats[-1]
    value = mbp * nfam / step_s

    # Roofline of the dominant kernel, by measured time: the seeding kernel
    # (6 int32 ops per ungapped extension step) or the preliminary DPs (12
    # int32 ops per DP cell), against the int32 VALU peak; each kernel's time
    # is its HIP-event time on the stream it is launched on, summed over the
    # search's launches.
    p1_ns = max(st["t_phase1_kernel_ns"], 1)
    pre_ns = max(st["t_prelim_kernel_ns"], 1)
    prelim_cells = st["dp_cells"] - st["tb_cells"]
    p1_ops = OPS_PER_UNGAPPED_STEP * st["ungapped_steps"]
    pre_ops = OPS_PER_DP_CELL * prelim_cells
    launches = max(st["phase1_launches"], 1)
    fam_len = np.array([len(f[1]) for f in w.families], dtype=np.float64)
    scan_bytes = float(len(queries)) * float(np.ceil(fam_len / 4).sum())
    hbm = {"kernel": "k_phase1", "algorithmic_bytes": scan_bytes,
           "achieved": round(scan_bytes / (p1_ns * 1e-9) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    hbm["frac"] = round(hbm["achieved"] / HBM_PEAK_GBS, 6)
    if p1_ns >= pre_ns:
        kern, ops, ns = "k_phase1", p1_ops, p1_ns
        per = f"{OPS_PER_UNGAPPED_STEP} int32 ops per ungapped extension step x {st['ungapped_steps']} steps"
    else:
        kern, ops, ns = "k_prelim_fused", pre_ops, pre_ns
        per = f"{OPS_PER_DP_CELL} int32 ops per DP cell x {prelim_cells} preliminary DP cells"
    achieved = ops / (ns * 1e-9) / 1e12
    # HBM bytes per launch of the seeding kernel, measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE on this workload and scaled by a plain copy's
    # readings; used only while the kernel source is the one measured.
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "phase1_traffic.json")
    if kern == "k_phase1" and os.path.exists(tpath) and len(queries) == 1667 and nfam == 1000 and not gappy:
        tj = json.load(open(tpath))
        if tj.get("kernel_source_sha16") == kernel_source_sha():
            traffic = tj.get("hbm_bytes_per_launch_measured")
    roofline = {"bound": "int32_valu", "achieved": round(achieved, 4), "peak": round(INT32_VALU_PEAK / 1e12, 2),
                "unit": "Tops/s", "frac": round(achieved / (INT32_VALU_PEAK / 1e12), 5), "traffic": traffic,
                "kernel": kern, "kernel_ms": round(ns / 1e6, 3), "launches": launches if kern == "k_phase1" else None,
                "algorithmic_ops": ops, "ops_definition": per, "hbm": hbm}
    # gapped DP throughput: DP cells of the preliminary and traceback DPs (the
    # reference's TOTAL_DP_CELLS unit) over the time inside the gapped kernels
    g_ns = st["t_gapped_kernel_ns"]
    gcups = st["dp_cells"] / g_ns if g_ns > 0 else 0.0
    ticks = [st[k] for k in ("p1_ticks_scan", "p1_ticks_reserve", "p1_ticks_extend", "p1_ticks_bucket",
                             "p1_ticks_replay", "p1_ticks_emit")]
    tick_sum = max(sum(ticks), 1)
    p1_stages = {n: round(t / tick_sum, 4)
                 for n, t in zip(("scan", "reserve", "extend", "bucket", "replay", "emit"), ticks)}

    cpu = None
    if not a.no_cpu:
        err = oracle_build.communicate()[1]
        exe = os.path.join(tmp, "oracle", "rmblastn_oracle")
        if oracle_build.returncode != 0 or not os.path.exists(exe):
            print(f"[bench] native oracle build failed: {err[-500:]}", file=sys.stderr)
        else:
            if gappy:
                sample = [(w.records[0][0], w.records[0][1][:1_500_000])]
                cpu = cpu_baseline(exe, sample, sample, w.families, tmp, mpath, xdrop, "the record's first 1.5 Mbp")
            elif parity_wl:
                one = [(w.records[0][0], w.records[0][1][:200_000])]
                cpu = cpu_baseline(exe, w.records, one, w.families, tmp, mpath, xdrop,
                                   "the whole record (one core: its first 200 kbp)", gpu_rows=rows,
                                   sample_ids=set(qids))
            else:
                cpu = cpu_baseline(exe, w.records[:cpu_records], w.records[:cpu1_records], w.families, tmp,
                                   mpath, xdrop, f"first {cpu_records} records",
                                   gpu_rows=rows, sample_ids=set(qids[:cpu_records]))
    if gappy:
        workload = ("divergent/gappy stress: 1 x 10 Mbp record x 300 families of 5-15 kbp, 25-40% divergence, "
                    "5-8% indels, X-drop 4x, ws 7")
    elif parity_wl:
        workload = ("parity configuration: 1 x 1 Mbp record x 50 families of 300-6,000 bp, 2,000 copies, ws 7, "
                    "search + host tail to outfmt-6 rows")
    elif uneven:
        workload = (f"uneven batching: the first {len(queries)} of 20,000 records of log-uniform 1-500 kbp "
                    f"x {nfam} families of log-uniform 50-20,000 bp, ws 7, search + host tail to outfmt-6 rows")
    else:
        workload = ("exact path at scale: 1,667 x 60 kbp records x 1,000 families, ws 7, search + host tail to "
                    "outfmt-6 rows")
    line = {
        "metric": "query Mbp x families per second (exact search path, final outfmt-6 rows in memory)",
        "value": round(value, 2),
        "unit": "Mbp*families/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(step_s * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generator, rmblast_amd/synth.py)",
        "config": {"workload": workload,
                   "records": len(queries), "families": nfam, "query_mbp": round(mbp, 3), "word_size": 7,
                   "min_raw_gapped_score": 200, "xdrop": xdrop, "complexity_adjust": True, "dust": False,
                   "mask_level": MASK_LEVEL, "outfmt": "6 " + RM_FIELDS},
        "search_only": {"ms_per_step": round(search_s * 1e3, 2), "value": round(mbp * nfam / search_s, 2),
                        "report_tail_ms": round(st["t_report_tail_us"] / 1e3, 2)},
        "roofline": roofline,
        "gapped_gcups": round(gcups, 3),
        "gapped_kernel_ms": round(g_ns / 1e6, 3),
        "prelim_kernel_ms": round(st["t_prelim_kernel_ns"] / 1e6, 3),
        "stage_ms": {k: round(st[k] / 1e3, 2) for k in ("t_seed_us", "t_prelim_us", "t_traceback_us", "t_total_us")},
        "phase1_stage_frac": p1_stages,
        "tuning": s.tuning,
        "counters": {k: st[k] for k in ("seeds_staged", "seed_evals", "ungapped_hits", "prelim_dps",
                                         "traceback_dps", "dp_cells", "tb_cells", "hsps", "ungapped_steps")},
        "rows": rows.count(b"\n"),
        "cpu_baseline": cpu,
    }
    if cpu and "parity" in cpu:
        line["parity"] = cpu["parity"]
    print(json.dumps(line))
    return 0


def run_oracle(exe, records, families, tmp, mpath, xdrop, threads, tag):
    from rmblast_amd import synth
    q = os.path.join(tmp, f"cpu_q_{tag}.fa")
    db = os.path.join(tmp, "cpu_db.fa")
    synth.to_fasta(records, q)
    if not os.path.exists(db):
        synth.to_fasta(families, db)
    args = [exe, "--db", db, "--query", q, "--matrix", mpath, "--gapopen", "20", "--gapextend", "5", "--word-size",
            "7", "--min-raw-gapped-score", "200", "--xdrop-ungap", str(xdrop[0]), "--xdrop-gap", str(xdrop[1]),
            "--xdrop-gap-final", str(xdrop[2]), "--complexity-adjust", "--dust", "no", "--mask-level",
            str(MASK_LEVEL), "--num-threads", str(threads), "--outfmt", "6 " + RM_FIELDS]
    r = subprocess.run(args, capture_output=True, env=dict(os.environ, RMB_ORACLE_TIME="1"), timeout=900)
    if r.returncode != 0:
        print(f"[bench] oracle failed: {r.stderr[-500:]}", file=sys.stderr)
        return None, None
    secs = float([l for l in r.stderr.decode().splitlines() if l.startswith("oracle_search_seconds")][0].split()[1])
    return secs, r.stdout


def cpu_baseline(exe, records, records1, families, tmp, mpath, xdrop, what, gpu_rows=None, sample_ids=None):
    """The CPU oracle (a restatement of the reference's search path, built
    -march=native on this host) on `records` x all families with threads over
    subjects per record like rmblastn's default thread mode, on every core this
    process may use, and on `records1` with one thread.  Its timed region runs
    from encoded inputs to the final outfmt-6 rows in memory."""
    cores = host_cores()
    secs, out = run_oracle(exe, records, families, tmp, mpath, xdrop, cores, "n")
    if secs is None:
        return None
    print(f"[bench] CPU baseline ({cores} threads): {secs:.2f}s", file=sys.stderr, flush=True)
    mbp = sum(len(x[1]) for x in records) / 1e6
    res = {"value": round(mbp * len(families) / secs, 2), "unit": "Mbp*families/s", "cores": cores,
           "kind": "port", "seconds": round(secs, 3),
           "sample": f"{what} ({mbp:.2f} Mbp) x {len(families)} families, same parameters and fields, "
                     f"oracle built -march=native on this host"}
    secs1, _ = run_oracle(exe, records1, families, tmp, mpath, xdrop, 1, "1")
    if secs1 is not None:
        mbp1 = sum(len(x[1]) for x in records1) / 1e6
        res["single_core"] = {"value": round(mbp1 * len(families) / secs1, 2), "seconds": round(secs1, 3),
                              "sample": f"first {len(records1)} records ({mbp1:.2f} Mbp)"}
        print(f"[bench] CPU baseline (1 thread): {secs1:.2f}s", file=sys.stderr, flush=True)
    if gpu_rows is not None:
        # the GPU's rows of the sample's records, in its order, against the oracle's
        g = [l for l in gpu_rows.split(b"\n") if l and l.split(b"\t")[4].decode() in sample_ids]
        o = [l for l in out.split(b"\n") if l]
        same = g == o
        res["parity"] = {"records": len(records), "gpu_rows": len(g), "oracle_rows": len(o),
                         "identical": same}
        print(f"[bench] parity on the CPU sample: {'identical' if same else 'DIFFERENT'} "
              f"({len(g)} GPU rows, {len(o)} oracle rows)", file=sys.stderr, flush=True)
    return res


if __name__ == "__main__":
    sys.exit(main())
