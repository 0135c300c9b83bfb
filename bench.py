"""Benchmark: the exact search path at scale on one MI355X.

Workload (the exact-path-at-scale configuration): a 100 Mbp synthetic genome
split into 1,667 records of 60 kbp, searched on both strands against 1,000
synthetic families (log-normal lengths, 100-15,000 bp), word size 7, gap
costs 20/5, X-drop 20/30/100, min_raw_gapped_score 200, complexity adjustment,
no DUST.  One step = one rmb_search over all records (lookup build, seeding,
ungapped, preliminary and traceback DP, ordered replay), with the library
already resident on the GPU.  Metric: query Mbp x families per second.

--workload gappy measures the divergent/gappy stress configuration instead:
one 10 Mbp record (the split query path) against 300 families of 5-15 kbp,
copies 25-40% diverged with 5-8% indels, 5% microsatellite tracts, X-drop
gapped/final 120/400; its CPU baseline sample is the record's first 1.5 Mbp.

Prints one JSON line.  --gpus N with N > 1 is refused: this build runs on one GPU.
"""
import argparse
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("scale", "gappy"), default="scale")
    ap.add_argument("--records", type=int, default=1667, help="records of the 100 Mbp genome to search")
    ap.add_argument("--families", type=int, default=1000)
    ap.add_argument("--cpu-records", type=int, default=300,
                    help="records in the CPU baseline sample (about 15 s of 16-core work)")
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
    if gappy:
        w = synth.gappy_config()
    else:
        w = synth.scale_config()
        w = synth.subset(w, a.records, a.families)
    gen_s = time.time() - t0
    print(f"[bench] workload generated in {gen_s:.1f}s: {len(w.records)} records x {len(w.families)} families",
          file=sys.stderr, flush=True)
    tmp = tempfile.mkdtemp(prefix="rmb_bench_")
    mpath = os.path.join(tmp, "synth.matrix")
    with open(mpath, "w") as f:
        f.write(synth.MATRIX_TEXT)
    m = rmblast_amd.load_matrix(mpath)
    xdrop = [20, 120, 400] if gappy else [20, 30, 100]
    params = dict(gap_open=20, gap_extend=5, word_size=7, xdrop_ungap=xdrop[0], xdrop_gap=xdrop[1],
                  xdrop_gap_final=xdrop[2], min_raw_gapped_score=200, complexity_adjust=True)
    s = rmblast_amd.Searcher(m, **params)
    s.load_library([f[1] for f in w.families])
    queries = [r[1] for r in w.records]
    mbp = sum(len(q) for q in queries) / 1e6
    nfam = len(w.families)

    for i in range(a.warmup):
        s.search(queries)
        print(f"[bench] warmup {i + 1}/{a.warmup} done", file=sys.stderr, flush=True)
    times, stats = [], []
    for i in range(a.steps):
        t = time.perf_counter()
        hsps, ops = s.search(queries)
        times.append(time.perf_counter() - t)
        stats.append(s.stats())
        print(f"[bench] step {i + 1}/{a.steps}: {times[-1]:.3f}s, {len(hsps)} HSPs", file=sys.stderr, flush=True)
    total = sum(times)
    step_s = total / a.steps
    st = stats[-1]
    value = mbp * nfam / step_s

    # roofline of the dominant kernel (the seeding kernel): algorithmic scan
    # bytes = ceil(L/4) of the streamed sequence per (streamed sequence, lookup
    # table) pair; here every family is streamed against each record's table,
    # so records x sum_f ceil(L_f/4); achieved = bytes / HIP-event kernel time.
    fam_len = np.array([len(f[1]) for f in w.families], dtype=np.float64)
    algo_bytes = float(len(queries)) * float(np.ceil(fam_len / 4).sum())
    k_ns = max(st["t_phase1_kernel_ns"], 1)
    achieved = algo_bytes / (k_ns * 1e-9) / 1e9
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on the
    # full workload, scaled by a plain copy's readings (profiles/)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "phase1_traffic.json")
    if os.path.exists(tpath) and len(queries) == 1667 and nfam == 1000:
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch_measured")
    valu_ops = OPS_PER_UNGAPPED_STEP * st["ungapped_steps"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": "k_phase1", "kernel_ms": round(k_ns / 1e6, 3),
                "int32_valu": {"ops": valu_ops, "achieved_tops": round(valu_ops / (k_ns * 1e-9) / 1e12, 3),
                               "peak_tops": round(INT32_VALU_PEAK / 1e12, 1),
                               "frac": round(valu_ops / (k_ns * 1e-9) / INT32_VALU_PEAK, 5)}}
    # gapped DP throughput: score-pass cells of the preliminary and traceback
    # DPs over the time inside the gapped kernels (HIP events on their stream)
    g_ns = st["t_gapped_kernel_ns"]
    gcups = st["dp_cells"] / g_ns if g_ns > 0 else 0.0
    ticks = [st[k] for k in ("p1_ticks_scan", "p1_ticks_reserve", "p1_ticks_extend", "p1_ticks_bucket",
                             "p1_ticks_replay", "p1_ticks_emit")]
    tick_sum = max(sum(ticks), 1)
    p1_stages = {n: round(t / tick_sum, 4)
                 for n, t in zip(("scan", "reserve", "extend", "bucket", "replay", "emit"), ticks)}

    cpu = None
    if not a.no_cpu:
        if gappy:
            sample = [(w.records[0][0], w.records[0][1][:1_500_000])]
            cpu = cpu_baseline(sample, w.families, tmp, mpath, xdrop, "the record's first 1.5 Mbp")
        else:
            cpu = cpu_baseline(w.records[:a.cpu_records], w.families, tmp, mpath, xdrop,
                               f"first {a.cpu_records} records")
    workload = ("divergent/gappy stress: 1 x 10 Mbp record x 300 families of 5-15 kbp, 25-40% divergence, "
                "5-8% indels, X-drop 4x, ws 7" if gappy else
                "exact path at scale: 1,667 x 60 kbp records x 1,000 families, ws 7")
    line = {
        "metric": "query Mbp x families per second (exact search path)",
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
                   "min_raw_gapped_score": 200, "xdrop": xdrop, "complexity_adjust": True, "dust": False},
        "roofline": roofline,
        "gapped_gcups": round(gcups, 3),
        "gapped_kernel_ms": round(g_ns / 1e6, 3),
        "prelim_kernel_ms": round(st["t_prelim_kernel_ns"] / 1e6, 3),
        "stage_ms": {k: round(st[k] / 1e3, 2) for k in ("t_seed_us", "t_prelim_us", "t_traceback_us", "t_total_us")},
        "phase1_stage_frac": p1_stages,
        "counters": {k: st[k] for k in ("seeds_staged", "seed_evals", "ungapped_hits", "prelim_dps",
                                         "traceback_dps", "dp_cells", "tb_cells", "hsps", "ungapped_steps")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


def cpu_baseline(records, families, tmp, mpath, xdrop, what):
    """The CPU oracle (a restatement of the reference's search path) on the
    sample `records` x all families, threads over subjects per record like
    rmblastn's default thread mode."""
    from rmblast_amd import synth
    exe = os.path.join(ROOT, "oracle", "_build", "rmblastn_oracle")
    if not os.path.exists(exe):
        return None
    cores = min(16, os.cpu_count() or 1)
    q = os.path.join(tmp, "cpu_q.fa")
    db = os.path.join(tmp, "cpu_db.fa")
    synth.to_fasta(records, q)
    synth.to_fasta(families, db)
    args = [exe, "--db", db, "--query", q, "--matrix", mpath, "--gapopen", "20", "--gapextend", "5", "--word-size",
            "7", "--min-raw-gapped-score", "200", "--xdrop-ungap", str(xdrop[0]), "--xdrop-gap", str(xdrop[1]),
            "--xdrop-gap-final", str(xdrop[2]), "--complexity-adjust", "--dust", "no", "--num-threads", str(cores),
            "--outfmt", "6 score qseqid qstart qend sstrand sseqid sstart send"]
    r = subprocess.run(args, capture_output=True, text=True, env=dict(os.environ, RMB_ORACLE_TIME="1"),
                       timeout=900)
    if r.returncode != 0:
        return None
    secs = float([l for l in r.stderr.splitlines() if l.startswith("oracle_search_seconds")][0].split()[1])
    mbp = sum(len(x[1]) for x in records) / 1e6
    return {"value": round(mbp * len(families) / secs, 2), "unit": "Mbp*families/s", "cores": cores,
            "kind": "port", "seconds": round(secs, 3),
            "sample": f"{what} ({mbp:.2f} Mbp) x {len(families)} families, same parameters"}


if __name__ == "__main__":
    sys.exit(main())
