#!/usr/bin/env python3
"""Benchmark: distinct k-mers/s over phases P1..P4 on one MI355X.

Workload (--workload): genome300m by default (24 records, 300 Mbp,
Zipf-like lengths, GC 41% with drift, 5% interspersed repeat families,
poly-N runs; seed 3), k=31; genome5m and transcriptome are the other
configurations. The autotuner picks the bin count and segment target first
(untimed). A step is one pass of P1..P4 over the records, starting from
them as ASCII in host memory: P1 packs the 2-bit stream on host threads and
copies it to HBM in chunks, overlapped with the scan (cf1g_run_p1_p4_stream);
P2..P4 replay their phase launches from hipGraphs. value = D /
seconds-per-step (D = distinct canonical k-mers). The same step over
records already resident in HBM is reported beside it ("resident").

Also reported on the same JSON line:
  roofline     : the dominant kernel's algorithmic bytes per launch over its
                 mean launch time (HIP events on the context stream in a
                 second, kernel-by-kernel pass of the same steps), against
                 8 TB/s HBM, with its measured HBM traffic from profiles/;
                 "issue": the same kernel's VALU instructions (profiles/)
                 per second against the 614.4 G/s issue ceiling.
  cpu_baseline : the multithreaded C++ CPU oracle (the restatement of the
                 reference's P1..P4) on the same records from host memory,
                 on every core this process may use: per-phase medians of 3
                 runs after a warm-up, with nproc and the CPU model.
  end_to_end   : FASTA file -> host parse and packing -> P1..P4, and the
                 same records from host memory as ASCII.
  p5           : unitig output (GPU-marked unitigs, host formatting).
  p4_cas       : P4 in the reference's form (per-occurrence lookup + CAS).
  autotune     : every configuration the autotuner measured.
  lookup       : learned rank index vs BBHash lookups/s over 4 x 2^28
                 uniform lookups of present keys (every learned rank
                 verified), the index's training summary, and the
                 classification pass through each index.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cf1-rs_amd"))

METRIC = "distinct k-mers/s phases 1-4 at k=31 vs CPU; learned-index vs BBHash lookups/s"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_G = 614.4  # wave64 VALU instructions per second (1,024 SIMDs x 2.4 GHz / 4 cycles), in G

# workload name -> (generator(synth, scale, k) -> (ascii, rec_off, names), description(bases, records))
WORKLOADS = {
    "genome300m": (lambda sy, sc, k: sy.genome_300m(scale=sc),
                   lambda nb, nr: f"genome-like {nb / 1e6:.0f} Mbp, {nr} records (Zipf lengths, GC 41% with drift, "
                                  f"5% repeat families, poly-N runs), seed 3"),
    "genome5m": (lambda sy, sc, k: sy.genome_5m(scale=sc),
                 lambda nb, nr: f"one {nb / 1e6:.0f} Mbp uniform record with 200 planted 2 kbp repeats (half reverse-"
                                f"complemented), seed 1"),
    "transcriptome": (lambda sy, sc, k: sy.transcriptome(scale=sc, k=k),
                      lambda nb, nr: f"transcriptome-like {nr} records, {nb / 1e6:.0f} Mbp (lognormal lengths from "
                                     f"400k shared 150 bp exons, 2% shorter than k), seed 2"),
}

PHASE_OF = {
    "pack": "P1", "p1_scan": "P1", "p1_hist": "P1", "p1_reduce": "P1",
    "p2_partition": "P2", "p2_route": "P2",
    "p3_gate": "P3", "p3_rcount": "P3", "p3_rscatter": "P3", "p3_expand": "P3", "p3_dedup": "P3", "p3_dedup_list": "P3", "p3_bigsort": "P3",
    "p3_compact": "P3", "p1_excl": "P1", "scan_total": "scan", "zero": "clear",
    "scan_tile": "scan", "scan_add": "scan", "p3b_bbh_big": "P3", "p3b_bbh_huge": "P3",
    "p4_states": "P4", "p4_classify": "P4",
}


def mphf_level_stats(mphf_words):
    """Keys placed per level (summed over segments) of the exported MPHF and
    the mean number of levels a lookup of a present key probes."""
    import numpy as np

    ns = int(mphf_words[5])
    D = int(mphf_words[6])
    base = 16 + 65536 // 4
    key_off = mphf_words[base:base + ns + 1].astype(np.int64)
    blk_off = mphf_words[base + ns + 1:base + 2 * (ns + 1)].astype(np.int64)
    blocks = mphf_words[base + 2 * (ns + 1):].reshape(-1, 8)
    words = blocks[:, 1:].reshape(-1)
    pc = np.concatenate([[0], np.cumsum(np.bitwise_count(words).astype(np.int64))])
    t = np.diff(key_off)
    start = blk_off[:-1] * 7
    off = np.zeros(ns, np.int64)
    placed = []
    left = int(t.sum())
    while left > 0 and len(placed) < 64:
        n = ((np.maximum(2 * t, 1) + 63) // 64) * 64
        live = t > 0
        a = start + off // 64
        b = start + (off + n) // 64
        got = int((pc[np.where(live, b, 0)] - pc[np.where(live, a, 0)]).sum())
        placed.append(got)
        left -= got
        off = off + n
        t = (9 * t + 19) // 20
    e_probe = sum((l + 1) * p for l, p in enumerate(placed)) / max(D, 1)
    return placed, e_probe


def algorithmic_bytes(info, kb, rw):
    """Algorithmic HBM bytes per kernel over one step: what each kernel's
    function must read and write at least once (inputs, outputs, no
    intermediate re-reads). NG = stream positions, W / S / D = k-mer
    occurrences / super k-mers / distinct k-mers, B = BBHash bytes, rec = S
    fixed-width records (16 or 32 bytes).
      pack       : ASCII in, 2-bit bases + placeholder bits out (the resident
                   step; the streaming step packs on the host)
      p1_scan    : packed stream in, one 8-byte entry and its 4-byte
                   minimizer hash per super k-mer out
      p1_hist    : the hashes in (the 512 KiB histogram is noise)
      p2_route   : entries and hashes in, the super k-mers' bases in (one
                   record's worth each), records out with a 2-byte segment
                   and a 4-byte destination per super k-mer
      p3_rcount  : the segments and one word per record in
      p3_rscatter: records and segments in, records (segment order) and a
                   4-byte destination per record out
      p3_dedup   : records in; D distinct keys and D masks (local-rank
                   order) and the BBHash blocks out
      p3_compact : the BBHash blocks B in and out (dense; the masks stay in
                   their segments' runs, and the dense key array is built only
                   when a caller asks for the keys)
      p4_states  : D masks (segment runs) in, D state bytes (dense) out
      p4_classify: D state bytes in, ceil(D/12) state words out
    """
    NG = info["stream_len"]
    S, D = info["n_superkmers"], info["n_distinct"]
    B = info["mphf_words"] * 8
    seq = NG / 4 + NG / 8
    rec = S * rw * 8
    return {
        "pack": NG + seq,
        "p1_scan": seq + 12 * S,
        "p1_hist": 4 * S,
        "p2_route": 12 * S + 2 * rec + 6 * S,
        "p3_rcount": 10 * S,
        "p3_rscatter": 2 * rec + 6 * S,
        "p3_dedup": rec + D * (kb + 2) + B,
        "p3_compact": 2 * B,
        "p4_states": 2 * D + D,
        "p4_classify": D + 8 * ((D + 11) // 12),
    }


def profiled_traffic(kernel, workload, k, bins):
    """HBM bytes per dispatch of `kernel` (FETCH_SIZE + WRITE_SIZE, each
    scaled by the counter's reading for a 1 GiB device copy) from the
    committed counter summary of this workload, k and bin count; when the
    autotuner chose a bin count without a summary, the summary of the same
    workload and k at another bin count (named in the returned source), or
    None."""
    import glob

    exact = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}_k{k}_b{bins}.md")
    paths = [exact] if os.path.exists(exact) else sorted(glob.glob(os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}_k{k}_b*.md")))
    for path in paths:
        for line in open(path):
            if line.startswith("traffic_bytes_per_dispatch:"):
                for tok in line.split(":", 1)[1].split():
                    name, val = tok.split("=")
                    if name == kernel:
                        return float(val), os.path.relpath(path, ROOT)
    return None, None


def profiled_valu(kernel, workload, k, bins):
    """VALU wave-instructions per dispatch of `kernel` from the committed
    instruction-counter summary of this workload and k (the bin count's own
    summary when there is one), or None."""
    import glob

    exact = os.path.join(ROOT, "profiles", f"sq_counters_{workload}_k{k}_b{bins}.md")
    paths = [exact] if os.path.exists(exact) else sorted(glob.glob(os.path.join(ROOT, "profiles", f"sq_counters_{workload}_k{k}_b*.md")))
    for path in paths:
        for line in open(path):
            if line.startswith("valu_insts_per_dispatch:"):
                for tok in line.split(":", 1)[1].split():
                    name, val = tok.split("=")
                    if name == kernel:
                        return float(val), os.path.relpath(path, ROOT)
    return None, None


def cpu_model():
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return None


def cpu_baseline(ascii_, rec_off, k, num_bins, sample_bases, threads, runs):
    """The CPU oracle (C++ restatement of the reference's P1..P4) on the same
    records (or the first `sample_bases` bases; 0 = all), from records in
    host memory to the state vector: one warm-up, then the median of `runs`
    timed runs, per phase and in total."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from oracle import RefPipeline

    if sample_bases and sample_bases < int(rec_off[-1]):
        seqs, tot = [], 0
        for i in range(len(rec_off) - 1):
            a, b = int(rec_off[i]), int(rec_off[i + 1])
            take = min(b - a, sample_bases - tot)
            if take <= 0:
                break
            seqs.append(ascii_[a:a + take].tobytes())
            tot += take
        p = RefPipeline(k, num_bins=num_bins, threads=threads)
        p.load(seqs)
        sample = f"first {tot / 1e6:.0f} Mbp of the same records ({len(seqs)} records)"
    else:
        p = RefPipeline(k, num_bins=num_bins, threads=threads)
        p.load_arrays(ascii_, rec_off)
        sample = f"all {int(rec_off[-1]) / 1e6:.0f} Mbp of the same records ({len(rec_off) - 1} records)"
    rows = []
    for r in range(runs + 1):
        t = time.perf_counter()
        d = p.run_p1_p4()
        dt = time.perf_counter() - t
        print(f"[bench] CPU run {r + 1}/{runs + 1}{' (warm-up)' if r == 0 else ''}: {dt:.2f} s", file=sys.stderr, flush=True)
        if r:
            rows.append((dt, p.times()))
    med = sorted(x[0] for x in rows)[len(rows) // 2]
    phases = {}
    for ph in ("p1", "p2", "p3a", "p3b", "p4"):
        v = sorted(x[1][ph] for x in rows)
        phases[ph] = round(v[len(v) // 2], 3)
    return {"value": d / med, "unit": "distinct k-mers/s", "cores": threads, "kind": "port",
            "sample": f"{sample}, D={d}, {num_bins} bins; median of {runs} runs after a warm-up: {med:.2f} s for P1-P4",
            "nproc": os.cpu_count(), "cpu_model": cpu_model(), "runs_s": [round(x[0], 3) for x in rows],
            "phases_s_median": phases,
            "label": "C++ restatement of cf1-rs v0.3.1 (multithreaded, same phases and constants), not the Rust binary"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="genome300m", choices=sorted(WORKLOADS),
                    help="genome300m (the headline: 24 records, 300 Mbp), genome5m (one 5 Mbp record with 200 planted "
                         "2 kbp repeats), transcriptome (100k records built from shared exons)")
    ap.add_argument("--k", type=int, default=31)
    ap.add_argument("--num-bins", type=int, default=128, help="bins when the autotuner is off")
    ap.add_argument("--p3-target-seg", type=int, default=0, help="P3 segment target when the autotuner is off (0: 1536)")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--cpu-sample-mbp", type=float, default=0.0, help="CPU baseline on the first N Mbp (0: all records)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: every core this process may use)")
    ap.add_argument("--cpu-runs", type=int, default=3, help="timed CPU baseline runs after one warm-up (median reported)")
    ap.add_argument("--host-threads", type=int, default=0, help="host packing threads of the streaming P1 (0: every core)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--lookups", type=int, default=1 << 28, help="lookups per batch (0: skip the lookup benchmark)")
    ap.add_argument("--lookup-batches", type=int, default=4)
    ap.add_argument("--lidx-leaf-bits", type=int, default=0, help="learned index: log2 leaves (0: the library default, 17)")
    ap.add_argument("--lidx-line-bytes", type=int, default=0, help="learned index: 64 or 128 (0: 64)")
    ap.add_argument("--lidx-line-keys", type=int, default=0, help="learned index: keys per line (0: the default for the size)")
    ap.add_argument("--end-to-end", type=int, default=1, help="also time FASTA file -> P1..P4 (0: skip)")
    ap.add_argument("--ingest-threads", type=int, default=16)
    ap.add_argument("--p5", type=int, default=1, help="also time P5 output (0: skip)")
    ap.add_argument("--no-autotune", action="store_true", help="keep --num-bins instead of the autotuner's choice")
    ap.add_argument("--p4-cas-runs", type=int, default=3, help="timed runs of the per-occurrence CAS classification (0: skip)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run device-to-device copies of this size (MiB) to scale HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every run uses one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import cf1g
    from cf1g import synth

    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb << 20)
    t = time.perf_counter()
    gen, wdesc = WORKLOADS[args.workload]
    ascii_, rec_off, names = gen(synth, args.scale, args.k)
    print(f"[bench] generated {len(ascii_) / 1e6:.0f} Mbp in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    g = cf1g.Cf1g(k=args.k, num_bins=args.num_bins, p3_target_seg=args.p3_target_seg)
    g.load_records(ascii=ascii_, rec_off=rec_off)
    tune = None
    if not args.no_autotune:
        t = time.perf_counter()
        tune = g.autotune(reps=2)
        print(f"[bench] autotuner: {tune['num_bins']} bins, segment target {tune['p3_target_seg']}, P4 chunk "
              f"{tune['p4_chunk']} x {tune['p4_wgs_per_cu']} WG/CU ({time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
    host_threads = args.host_threads or len(os.sched_getaffinity(0))
    for i in range(args.warmup):
        g.run_p1_p4_stream(ascii_, rec_off, threads=host_threads)
        g.run_p1_p4()
        print(f"[bench] warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    # per-phase wall times through the phase API (untimed; the second of two
    # passes, so one-time runtime costs such as the first pageable copy of
    # the bucket->bin table are not counted)
    for _ in range(2):
        ph = {}
        t = time.perf_counter(); g.count_minimizer_histogram(want=False); ph["P1"] = time.perf_counter() - t
        t = time.perf_counter(); g.partition_minimizers_and_route(num_bins=tune["num_bins"] if tune else args.num_bins); ph["P2"] = time.perf_counter() - t
        t = time.perf_counter(); g.build_mphf(); ph["P3"] = time.perf_counter() - t
        t = time.perf_counter(); g.classify_vertices(); ph["P4"] = time.perf_counter() - t

    # timed steps: from the records as ASCII in host memory (pageable) to the
    # state vector; P1 streams the records to HBM, P2..P4 replay their
    # hipGraphs; no per-kernel events in the stream
    g.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        D = g.run_p1_p4_stream(ascii_, rec_off, threads=host_threads)
    g.sync()
    t1 = time.perf_counter()
    print(f"[bench] {args.steps} timed steps done", file=sys.stderr, flush=True)
    # the same steps over records already resident in HBM
    g.sync()
    tr0 = time.perf_counter()
    for i in range(args.steps):
        g.run_p1_p4()
    g.sync()
    tr1 = time.perf_counter()
    ms_resident = (tr1 - tr0) * 1000.0 / args.steps
    # the same steps launched kernel by kernel with HIP events around each
    # launch: per-kernel times (roofline) and the graph-free step time
    g.set_profiling(True)
    g.kernel_times(reset=True)
    g.sync()
    t2 = time.perf_counter()
    for i in range(args.steps):
        g.run_p1_p4()
    g.sync()
    t3 = time.perf_counter()
    kt = g.kernel_times(reset=True)
    g.set_profiling(False)
    ms_eager = (t3 - t2) * 1000.0 / args.steps
    ms = (t1 - t0) * 1000.0 / args.steps
    info = g.info()
    kb = 16 if args.k > 32 else 8
    rw = 4 if args.k > 32 else 2
    placed, e_probe = mphf_level_stats(g.mphf_export())
    ab = algorithmic_bytes(info, kb, rw)
    kern = {}
    for name, (cnt, tot) in kt.items():
        per_step = cnt / args.steps
        kern[name] = {"launches_per_step": per_step, "ms_per_step": tot / args.steps, "mean_launch_ms": tot / cnt}
        if name in ab:
            byt = ab[name] / per_step  # per launch
            kern[name]["gbs"] = byt / (tot / cnt * 1e-3) / 1e9
    dom = max((n for n in kern if n in ab), key=lambda n: kern[n]["ms_per_step"])
    dk = kern[dom]
    achieved = dk["gbs"]
    traffic, tsrc = profiled_traffic(dom, args.workload, args.k, info["num_bins"]) if args.scale == 1.0 else (None, None)
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
            "bytes_per_launch": ab[dom] / dk["launches_per_step"], "mean_launch_ms": round(dk["mean_launch_ms"], 4)}
    # the same kernel against the instruction-issue ceiling: 1,024 SIMDs at
    # 2.4 GHz issue one wave64 VALU instruction per 4 cycles each
    valu, vsrc = profiled_valu(dom, args.workload, args.k, info["num_bins"]) if args.scale == 1.0 else (None, None)
    if valu:
        ach = valu / (dk["mean_launch_ms"] * 1e-3) / 1e9
        roof["issue"] = {"bound": "valu", "achieved": round(ach, 1), "peak": VALU_PEAK_G, "unit": "G wave-instructions/s",
                         "frac": round(ach / VALU_PEAK_G, 4), "valu_per_launch": valu, "source": vsrc}
    copy_gbs = measured_copy_peak()
    roof["measured_copy_gbs"] = round(copy_gbs, 1)
    roof["frac_of_measured_copy"] = round(achieved / copy_gbs, 4)
    tot_bytes = sum(ab.values())
    roof["pipeline"] = {"algorithmic_bytes_per_step": tot_bytes, "achieved": round(tot_bytes / (ms_resident * 1e-3) / 1e9, 1)}
    phases = {}
    for name, v in kern.items():
        phases.setdefault(PHASE_OF.get(name, name), 0.0)
        phases[PHASE_OF.get(name, name)] += v["ms_per_step"]

    # end to end from a FASTA file: host parse + host packing + copy of the
    # packed stream, then P1..P4 (the timed step starts from records in HBM)
    e2e = end_to_end(g, ascii_, rec_off, names, args.ingest_threads) if args.end_to_end else None

    # P5 (host, per the reference's phase split) over the GPU-marked unitigs
    p5 = p5_bench(g, names, args.ingest_threads) if args.p5 else None

    # P4 in the reference's form (per-occurrence lookup + CAS join) beside
    # the mask-join P4 the step uses; both must give the same state words
    p4cas = p4_cas_bench(g, info, e_probe) if args.p4_cas_runs > 0 else None

    # learned index vs BBHash lookups over device-resident present keys
    lookup = None
    if args.lookups > 0 and D > 0 and kb == 8:
        lidx_args = {"leaf_bits": args.lidx_leaf_bits, "line_bytes": args.lidx_line_bytes, "line_keys": args.lidx_line_keys}
        lookup = lookup_bench(g, D, args.lookups, args.lookup_batches,
                              far=profiled_far_requests(args.workload, args.k, info["num_bins"]) if args.scale == 1.0 else None,
                              lidx_args=lidx_args)

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        print(f"[bench] CPU baseline on {threads} threads ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(ascii_, rec_off, args.k, info["num_bins"], int(args.cpu_sample_mbp * 1e6), threads, args.cpu_runs)

    out = {
        "metric": METRIC,
        "value": D / (ms / 1000.0),
        "unit": "distinct k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"name": args.workload, "workload": f"{wdesc(len(ascii_), len(rec_off) - 1)}, k={args.k}, "
                               f"{info['num_bins']} bins{' (autotuned)' if tune else ''}",
                   "k": args.k, "bases": int(len(ascii_)), "records": int(len(rec_off) - 1),
                   "stream_len": info["stream_len"], "mphf_words": info["mphf_words"],
                   "kmer_occurrences": info["n_kmers"], "superkmers": info["n_superkmers"], "distinct_kmers": D,
                   "num_bins": info["num_bins"], "seg_bits": info["seg_bits"], "mphf_big_segments": info["mphf_big_segments"],
                   "mphf_mean_levels_probed": round(e_probe, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (D / (ms / 1000.0)) / cpu["value"] if cpu else None,
        "timed_region": "records as ASCII in pageable host memory -> state vector complete (P1..P4, the host-to-device "
                        f"transfer of the records inside P1; {host_threads} host packing threads)",
        "resident": {"ms_per_step": round(ms_resident, 4), "value": D / (ms_resident / 1000.0),
                     "note": "the same P1..P4 step over records already resident in HBM"},
        "ms_per_step_kernel_by_kernel": round(ms_eager, 4),
        "graphs": {"captures": info["graph_captures"], "replays": info["graph_replays"]},
        "phase_wall_ms": {k2: round(v * 1000, 3) for k2, v in ph.items()},
        "phase_kernel_ms": {k2: round(v, 3) for k2, v in phases.items()},
        "kernels": {k2: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()} for k2, v in kern.items()},
        "end_to_end": e2e,
        "p5": p5,
        "autotune": tune,
        "p4_cas": p4cas,
        "lookup": lookup,
    }
    print(json.dumps(out), flush=True)
    g.close()


def end_to_end(g, ascii_, rec_off, names, threads, runs=3):
    """FASTA file on disk (page cache) -> cf1g_load_fasta (parse, host
    packing, copy of the packed stream) -> P1..P4, and the same records
    given as ASCII in host memory (cf1g_load_records: 300 MB copied, packed
    on the device). Median of `runs`."""
    import tempfile

    from cf1g import synth

    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "records.fa")
        synth.write_fasta(path, ascii_, rec_off, names)
        fbytes = os.path.getsize(path)
        rows = []
        for _ in range(runs):
            t0 = time.perf_counter()
            st = g.load_fasta(path, threads=threads)
            t1 = time.perf_counter()
            g.run_p1_p4()
            t2 = time.perf_counter()
            g.load_records(ascii=ascii_, rec_off=rec_off)
            t3 = time.perf_counter()
            g.run_p1_p4()
            t4 = time.perf_counter()
            rows.append((t1 - t0, t2 - t1, t3 - t2, t4 - t3, st))
    rows.sort(key=lambda r: r[0] + r[1])
    m = rows[len(rows) // 2]
    st = m[4]
    print(f"[bench] end to end: FASTA load {m[0] * 1e3:.1f} ms + P1-P4 {m[1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    return {"fasta_bytes": fbytes, "threads": threads,
            "fasta_ms": round((m[0] + m[1]) * 1e3, 3), "fasta_load_ms": round(m[0] * 1e3, 3),
            "parse_ms": round(st["parse_ms"], 3), "pack_ms": round(st["pack_ms"], 3), "copy_ms": round(st["h2d_ms"], 3),
            "packed_bytes_copied": st["stream_bytes"], "p1_p4_ms": round(m[1] * 1e3, 3),
            "ascii_in_host_memory_ms": round((m[2] + m[3]) * 1e3, 3), "ascii_copy_ms": round(m[2] * 1e3, 3)}


def p5_bench(g, names, threads):
    """cf1g_p5_write into a temporary directory: unitig spans on the GPU
    (annotation, cuts, ids, claims), then the host formats and writes
    cf_seg / cf_seq / JSON."""
    import tempfile

    with tempfile.TemporaryDirectory() as td:
        pre = os.path.join(td, "out")
        g.set_profiling(True)
        g.kernel_times(reset=True)
        t = time.perf_counter()
        meta = g.p5_write(pre, names=list(names), input_desc="records.fa", threads=threads)
        dt = time.perf_counter() - t
        kt = g.kernel_times(reset=True)
        g.set_profiling(False)
        sizes = {ext: os.path.getsize(pre + ext) for ext in (".cf_seg", ".cf_seq")}
    gpu = {k2: round(v[1], 3) for k2, v in kt.items() if k2.startswith("p5_") or k2 in ("bbh_lookup", "lidx_lookup")}
    print(f"[bench] P5: {dt * 1e3:.1f} ms, {meta['unitigs']} unitigs", file=sys.stderr, flush=True)
    return {"ms": round(dt * 1e3, 3), "threads": threads, "unitigs": meta["unitigs"], "vertices": meta["vertices"],
            "gpu_kernel_ms": gpu, "cf_seg_bytes": sizes[".cf_seg"], "cf_seq_bytes": sizes[".cf_seq"]}


def p4_cas_bench(g, info, e_probe, runs=3):
    """cf1g_p4_classify_cas on the same records: one BBHash lookup and one
    64-bit CAS join per k-mer occurrence (persistent work-stealing launch).
    Algorithmic bytes: the packed stream once plus one 64-byte granule per
    dependent random access of each occurrence (the BBHash levels probed,
    the segment tables, the state word)."""
    import numpy as np

    want = g.states()
    g.classify_vertices_cas()  # warm-up
    st = [g.classify_vertices_cas() for _ in range(runs)]
    same = bool(np.array_equal(g.states(), want))
    ms = sorted(x["kernel_ms"] for x in st)[len(st) // 2]
    W, NG = info["n_kmers"], info["stream_len"]
    ab = NG / 4 + NG / 8 + 64.0 * W * (e_probe + 2)
    s0 = st[0]
    print(f"[bench] CAS classification: {ms:.3f} ms, states equal the mask join: {same}", file=sys.stderr, flush=True)
    return {"kernel_ms": round(ms, 4), "occurrences": s0["occurrences"], "occurrences_per_s": W / (ms * 1e-3),
            "cas_issued": s0["cas_issued"], "cas_failed": s0["cas_failed"], "cas_per_s": s0["cas_issued"] / (ms * 1e-3),
            "chunk": s0["chunk"], "wgs_per_cu": s0["wgs_per_cu"], "grid": s0["grid"], "n_chunks": s0["n_chunks"],
            "algorithmic_bytes": ab, "achieved_gbs": round(ab / (ms * 1e-3) / 1e9, 1),
            "states_equal_mask_join": same}


def copy_calibration(nbytes, reps=3):
    """hipMemcpy device-to-device of nbytes (a plain copy kernel of known
    size: reads and writes nbytes each) for scaling FETCH_SIZE/WRITE_SIZE."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy calibration: {reps} x {nbytes} bytes device-to-device", file=sys.stderr, flush=True)


def measured_copy_peak(nbytes=2 << 30, reps=5):
    """Device-to-device hipMemcpyAsync of nbytes, timed with HIP events on
    the null stream: the copy rate the box reaches (GB/s, reads plus
    writes), the ceiling beside the published 8 TB/s. Best of `reps`."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    best = None
    for i in range(reps + 1):
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None) == 0
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        if i and (best is None or ms.value < best):  # the first copy is a warm-up
            best = ms.value
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    gbs = 2.0 * nbytes / (best * 1e-3) / 1e9
    print(f"[bench] device copy: {gbs:.0f} GB/s ({nbytes >> 20} MiB, best of {reps})", file=sys.stderr, flush=True)
    return gbs


def lookup_bench(g, D, batch, n_batches, far=None, seed=5, lidx_args=None):
    """The learned index against the BBHash on the same n_batches x batch
    lookups of present keys (uniform over the distinct keys, seeded), keys
    and results resident in HBM. Each index gets one untimed batch of other
    keys first (its structure into the caches), then its timed batches back
    to back. Lookups/s from the lookup kernels' HIP-event times. Every
    learned rank is checked against the drawn index (ranks are positions in
    the sorted keys); the BBHash values are checked against the BBHash of the
    sorted keys. Then P4 in the reference's form through each index, and the
    same lookups through 2^20 leaf segments."""
    import ctypes as C

    import numpy as np

    t = time.perf_counter()
    info = g.lidx_train(**(lidx_args or {}))
    train_wall = time.perf_counter() - t
    print(f"[bench] learned index trained in {train_wall:.2f} s", file=sys.stderr, flush=True)
    keys = g.lidx_export()["keys"]
    hip = C.CDLL("libamdhip64.so")
    L = g.L
    dq, dout, dall = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dq), C.c_size_t(batch * 8)) == 0
    assert hip.hipMalloc(C.byref(dout), C.c_size_t(batch * 8)) == 0
    assert hip.hipMalloc(C.byref(dall), C.c_size_t(max(D, 1) * 8)) == 0
    # BBHash values of the sorted keys (for checking the BBHash batches)
    assert hip.hipMemcpy(dall, keys.ctypes.data_as(C.c_void_p), C.c_size_t(D * 8), 1) == 0
    vall = np.empty(D, np.uint64)
    g._check(L.cf1g_mphf_lookup_device(g.h, dall, D, dall))
    assert hip.hipMemcpy(vall.ctypes.data_as(C.c_void_p), dall, C.c_size_t(D * 8), 2) == 0
    rng = np.random.default_rng(seed)
    out = np.empty(batch, np.uint64)
    errs = {"learned": 0, "bbhash": 0}
    batches = [rng.integers(0, D, batch).astype(np.uint64) for _ in range(n_batches + 1)]  # the last: warm-up
    kt = {}
    # each index: one untimed batch (its structure into the caches), then
    # the timed batches back to back; the other index's batches after
    for name, fn, kname in (("learned", L.cf1g_lidx_lookup_device, "lidx_lookup"), ("bbhash", L.cf1g_mphf_lookup_device, "bbh_lookup")):
        for bi in range(-1, n_batches):
            idx = batches[bi]
            assert hip.hipMemcpy(dq, keys[idx].ctypes.data_as(C.c_void_p), C.c_size_t(batch * 8), 1) == 0
            if bi == 0:
                g.set_profiling(True)
                g.kernel_times(reset=True)
            g._check(fn(g.h, dq, batch, dout))
            if bi < 0:
                continue
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), dout, C.c_size_t(batch * 8), 2) == 0
            errs[name] += int(np.count_nonzero(out != (idx if name == "learned" else vall[idx])))
            print(f"[bench] {name} lookup batch {bi + 1}/{n_batches} done", file=sys.stderr, flush=True)
        kts = g.kernel_times(reset=True)
        # the learned index's call is its lookup kernel and the kernel that
        # resolves the searches the lookup deferred
        kt[kname] = tuple(sum(kts.get(kn, (0, 0.0))[j] for kn in KERNELS_OF[kname]) for j in (0, 1))
        g.set_profiling(False)
    # P4 in the reference's form through each index: one lookup per k-mer
    # occurrence in sequence order plus the CAS join
    via = {"bbhash": g.classify_via(0), "learned": g.classify_via(1)}
    n = batch * n_batches
    lms = kt.get("lidx_lookup", (0, 0.0))[1]
    bms = kt.get("bbh_lookup", (0, 0.0))[1]
    lrate = n / (lms * 1e-3) if lms else None
    brate = n / (bms * 1e-3) if bms else None
    keep = ("n_leaves", "n_pieces", "fallback_leaves", "fallback_keys", "n_lines", "line_bytes", "line_keys",
            "overflow_lines", "search_keys", "max_err_lo", "max_err_hi", "mean_abs_err", "line_bits_per_key", "bytes_total",
            "root_loss", "leaf_loss", "sort_ms", "train_ms")
    summary = {k2: (round(info[k2], 4) if isinstance(info[k2], float) else info[k2]) for k2 in keep}
    # the same lookups through 2^20 leaf segments (the leaf table then leaves L2)
    shape20 = None
    if D >= (1 << 21):
        info20 = g.lidx_train(leaf_bits=20, line_bytes=(lidx_args or {}).get("line_bytes", 0),
                              line_keys=(lidx_args or {}).get("line_keys", 0))
        assert hip.hipMemcpy(dq, keys[batches[-1]].ctypes.data_as(C.c_void_p), C.c_size_t(batch * 8), 1) == 0
        g._check(L.cf1g_lidx_lookup_device(g.h, dq, batch, dout))  # warm-up
        idx = batches[0]
        assert hip.hipMemcpy(dq, keys[idx].ctypes.data_as(C.c_void_p), C.c_size_t(batch * 8), 1) == 0
        g.set_profiling(True)
        g.kernel_times(reset=True)
        g._check(L.cf1g_lidx_lookup_device(g.h, dq, batch, dout))
        k20 = g.kernel_times(reset=True)
        g.set_profiling(False)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), dout, C.c_size_t(batch * 8), 2) == 0
        ms20 = sum(k20.get(kn, (0, 0.0))[1] for kn in KERNELS_OF["lidx_lookup"])
        shape20 = {"n_leaves": info20["n_leaves"], "lookups": batch, "lookups_per_s": batch / (ms20 * 1e-3) if ms20 else None,
                   "rank_errors": int(np.count_nonzero(out != idx)), "mean_abs_err": round(info20["mean_abs_err"], 3),
                   "max_err": [info20["max_err_lo"], info20["max_err_hi"]]}
        g.lidx_train(**(lidx_args or {}))  # back to the measured shape (P5 through the index uses it)
        print(f"[bench] 2^20 leaves: {shape20['lookups_per_s'] / 1e9:.2f} G lookups/s", file=sys.stderr, flush=True)
    for p in (dq, dout, dall):
        hip.hipFree(p)
    return {
        "n": n, "batch": batch, "keys": "uniform over the distinct keys, seed %d" % seed,
        "bbhash_lookups_per_s": brate, "learned_lookups_per_s": lrate,
        "learned_vs_bbhash": (lrate / brate) if (lrate and brate) else None,
        "learned_rank_errors": errs["learned"], "bbhash_value_errors": errs["bbhash"],
        "lookup_kernel_ms": {"learned": round(lms, 3), "bbhash": round(bms, 3)},
        "p4_cas_via": {k2: {"kernel_ms": round(v["kernel_ms"], 3), "occurrences": v["occurrences"],
                            "occurrences_per_s": v["occurrences"] / (v["kernel_ms"] * 1e-3) if v["kernel_ms"] else None,
                            "cas_issued": v["cas_issued"], "cas_failed": v["cas_failed"]} for k2, v in via.items()},
        "learned_index": summary,
        "leaf_segments_2p20": shape20,
        "train_wall_s": round(train_wall, 3),
        "roofline": lookup_roofline(far, {"bbh_lookup": brate, "lidx_lookup": lrate}),
    }


# kernels each index's lookup call launches
KERNELS_OF = {"lidx_lookup": ("lidx_lookup", "lidx_search"), "bbh_lookup": ("bbh_lookup",)}


def profiled_far_requests(workload, k, bins):
    """L2 read requests to the fabric per lookup (TCC_EA0_RDREQ / lookups) of
    each lookup kernel, from the committed lookup counter summary, and its
    path."""
    import glob

    exact = os.path.join(ROOT, "profiles", f"lookup_counters_{workload}_k{k}_b{bins}.md")
    paths = [exact] if os.path.exists(exact) else sorted(glob.glob(os.path.join(ROOT, "profiles", f"lookup_counters_{workload}_k{k}_b*.md")))
    for path in paths:
        for line in open(path):
            if line.startswith("far_requests_per_lookup:"):
                return {t.split("=")[0]: float(t.split("=")[1]) for t in line.split(":", 1)[1].split()}, os.path.relpath(path, ROOT)
    return None, None


def lookup_roofline(far, rates):
    """Per lookup kernel: far requests per lookup (profiled), bytes per
    lookup at 64 bytes per request, and the rate those bytes move at against
    8 TB/s."""
    if not far or not far[0]:
        return None
    req, src = far
    out = {"source": src, "bytes_per_request": 64}
    for name, rate in rates.items():
        if name in req and rate:
            b = req[name] * 64
            out[name] = {"far_requests_per_lookup": req[name], "bytes_per_lookup": round(b, 1),
                         "achieved_gbs": round(b * rate / 1e9, 1), "frac_of_hbm_peak": round(b * rate / 1e9 / HBM_PEAK_GBS, 4)}
    return out


if __name__ == "__main__":
    main()
