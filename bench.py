"""Benchmark: aligned pairs/s of the MI355X extension engines.

One step = one launch of the hot path over one batch already resident in HBM
(phx_plan_run).  Default workload: 10,000 reads of 150 bp against 170 bp
windows, gap-affine WFA 4/6/2 with score + CIGAR (see phraya_amd/synth.py).

Prints ONE JSON line (rank 0, one GPU).  `roofline` uses the algorithmic bytes
of one launch (packed inputs + 12 B of results + 4 B per CIGAR run) over the
alignment kernel's mean duration from HIP events on its launch stream.
`cpu_baseline` times the CPU restatement of the upstream engines (oracle/) on
the host's cores over a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "aligned pairs/s + p99 latency, 150bp & mixed-length WFA, bit-exact vs CPU"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (synth workload, engine, costs, description)
    "short150": ("short150", "affine", (4, 6, 2),
                 "10k x (150 bp read vs 170 bp window), affine WFA 4/6/2, score+CIGAR"),
    "myers150": ("myers150", "myers_dist", (0, 0, 0),
                 "1M x (150 bp read vs 180 bp window), Myers edit distance + end column"),
    "mixed": ("mixed", "affine", (4, 6, 2),
              "200k reads log-uniform 100 bp-20 kbp vs 1.1x windows, affine WFA 4/6/2, score+CIGAR"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, engine_name, costs, seconds, sample_n=None):
    """Time the oracle (the C++ restatement of phraya-align's CPU engines) on the
    host's cores over repeated passes of a bounded, seeded random sample of the
    workload; also a one-thread figure on a smaller sample."""
    from oracle import oracle as O
    # phraya's myers_edit_distance computes the CIGAR too (wfa_simd.rs:3657-3673),
    # so the Myers baseline is the CIGAR-producing engine; forward-only is extra.
    eng = {"affine": O.AFFINE, "linear": O.LINEAR, "myers_dist": O.MYERS_GLOBAL,
           "myers": O.MYERS}[engine_name]
    threads = len(os.sched_getaffinity(0))
    sub, note = w, "the whole batch"
    if sample_n is not None and w.n > sample_n:
        idx = np.sort(np.random.default_rng(12345).choice(w.n, sample_n, replace=False))
        sub = w.subset(idx)
        note = f"a seeded uniform random sample of {sample_n} of the {w.n} pairs"

    def timed(e, t, budget, subset):
        passes, pairs, t0 = 0, 0, time.perf_counter()
        while True:
            O.batch(e, subset.q_seq, subset.q_off, subset.t_seq, subset.t_off, costs=costs,
                    threads=t, want_cigars=True)
            passes += 1
            pairs += subset.n
            el = time.perf_counter() - t0
            if el >= budget:
                return pairs / el, passes, el

    rate, passes, el = timed(eng, threads, seconds, sub)
    one_sub = sub.subset(np.arange(min(sub.n, max(200, sub.n // 16))))
    rate1, _, _ = timed(eng, 1, max(2.0, seconds / 5), one_sub)
    out = {"value": rate, "unit": "pairs/s", "cores": threads, "kind": "port",
           "sample": f"{passes} passes over {note} ({sub.n} pairs), {el:.1f} s wall, "
                     f"CIGAR strings formatted; oracle/ C++ restatement of phraya-align "
                     f"(-O3, SSE2 match-extend), one pair per task over {threads} threads",
           "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "one_thread_pairs_per_s": rate1}
    if engine_name == "myers_dist":
        fwd, _, _ = timed(O.MYERS_FWD, threads, max(2.0, seconds / 5), sub)
        out["forward_only_pairs_per_s"] = fwd
    return out


def calibrate_copy(nbytes: int) -> None:
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE
    counters of the alignment kernels (run under rocprofv3 --pmc)."""
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="short150", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=None, help="override pair count")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", type=int, default=2000, help="pairs checked against the oracle")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(pairs shard trivially; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import phraya_amd as px
    from phraya_amd import synth

    wname, engine, costs, desc = WORKLOADS[args.workload]
    log(f"generating {wname}")
    w = synth.workload(wname, n=args.n)
    n = w.n
    eng_id = {"affine": px.ENGINE_AFFINE, "linear": px.ENGINE_LINEAR,
              "myers_dist": px.ENGINE_MYERS_DIST, "myers": px.ENGINE_MYERS}[engine]
    pb = px.PackedBatch(w.q_seq, w.q_off, w.t_seq, w.t_off, np.zeros(n, np.uint64),
                        np.zeros(n, np.uint64))
    log(f"uploading {n} pairs")
    plan = px.Plan(pb, eng_id, px.AffineCosts(*costs) if engine == "affine" else px.AffineCosts())
    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        plan.run()
    log(f"timing {args.steps} steps")
    step_ms, kern_ms, launches = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run()  # returns after the device finished (event sync)
        step_ms.append(plan.last_ms())
        kern_ms.append(plan.last_kernel_ms())
        launches.append(plan.last_launches())
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    value = n * args.steps / wall
    log(f"{value:.4g} pairs/s, {ms_per_step:.3f} ms/step")

    res = plan.results()
    st = res.aln["status"]
    alg_bytes = plan.algorithmic_bytes()
    qn = np.diff(w.q_off).astype(np.float64)
    tn = np.diff(w.t_off).astype(np.float64)
    cells = float((qn * tn).sum())
    kmean = float(np.mean(kern_ms))
    achieved = alg_bytes / (kmean / 1000.0) / 1e9

    # Parity on a sample of this very run.
    from oracle import oracle as O
    chk = min(args.check, n)
    sub = w.subset(np.arange(chk))
    orc_eng = {"affine": O.AFFINE, "linear": O.LINEAR, "myers_dist": O.MYERS_FWD,
               "myers": O.MYERS}[engine]
    ores, ocig = O.batch(orc_eng, sub.q_seq, sub.q_off, sub.t_seq, sub.t_off, costs=costs,
                         threads=len(os.sched_getaffinity(0)), want_cigars=engine != "myers_dist")
    mism = 0
    for i in range(chk):
        if int(st[i]) != int(ores["status"][i]) or \
           int(res.aln["edit_distance"][i]) != int(ores["edit_distance"][i]) or \
           int(res.aln["target_end"][i]) != int(ores["target_end"][i]) or \
           (engine != "myers_dist" and res.cigar(i) != ocig[i]) or \
           (engine == "affine" and int(res.aln["score"][i]) != int(ores["score"][i])):
            mism += 1
    log(f"parity: {chk} pairs checked, {mism} mismatches")

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath) and args.n is None:
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline ~{args.cpu_seconds:.0f} s")
        cpu = cpu_baseline(w, engine, costs, args.cpu_seconds,
                           sample_n={"short150": None, "myers150": 20_000, "mixed": 500}[args.workload])

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "pairs_per_step": n,
                   "engine": engine, "costs": list(costs) if engine == "affine" else None,
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                       "p99": float(np.percentile(step_ms, 99)),
                       "what": "device time of one whole batch (events around phx_plan_run)"},
        "gcups_equiv": cells / (ms_per_step / 1000.0) / 1e9,
        "kernel_ms_mean": kmean,
        "launches_per_step": float(np.mean(launches)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "parity": {"checked": chk, "mismatches": mism},
        "status_counts": {str(k): int(v) for k, v in zip(*np.unique(st, return_counts=True))},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
