#!/usr/bin/env python3
"""Benchmark of the banded signal-to-reference pair-HMM forward-backward on one MI355X.

Headline (`value`): batched FP32 GCUPS on the batched configuration -- 4096 reads,
log-normal reference lengths (median 8 kb, sigma 0.5, clipped to 1-50 kb) over the
ACEGT alphabet (E = 5mC at 30% of CpG sites), 15625-k-mer 6-mer Gaussian table,
band expansion 50, one length-sorted launch sequence.  A step is one pass of the
whole path over the batch with inputs resident in HBM (sa_batch_run), timed with
HIP events on the library's stream from the first prep kernel to the last pair
written to HBM.  GCUPS = band cells / time, each cell counted once per
forward-backward-posterior pass.

Also reported: the plain per-read GPU drop-in (one 50 kb read, e=100, FP64 and FP32,
and a loop of single-read calls over the batched reads), the FP64 CPU oracle timed
on this host's cores over a bounded sample of the batched reads, and the roofline
of the dominant kernel.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "banded FB GCUPS batched vs plain GPU vs CPU; p50/p99 read latency @ rate"
# Algorithmic work per band cell (Gaussian emissions; FMA = 2 FLOP): per direction
# 2 emissions x 13 + 8 transitions x 11 = 114; the backward pass adds the posterior (4),
# the amortised total refresh (~11) and the window-overlap recompute (~5).
FLOP_FWD = 114.0
FLOP_BWD = 134.0
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_native_reads(native, reads):
    return [native.Read(r["ref"], r["events"], r["anchors"], r["scale"], r["shift"], r["var"]) for r in reads]


def cpu_baseline(model, reads, params, target_s, threads):
    """The FP64 oracle (the reference path restated in C) on this host's cores: one
    read per thread, longest first, over a bounded sample of the batched reads."""
    import oracle
    from tests import _parity as P
    oracle.build()
    t0 = time.time()
    r0 = reads[0]
    _, st0 = P.run_oracle(model, r0, params)
    rate1 = st0["cells"] / max(time.time() - t0, 1e-6)  # one core, cells/s
    budget = rate1 * threads * target_s
    sample, tot = [], 0
    for r in reads[1:]:
        n = (len(r["ref"]) + len(r["events"])) * 80
        if tot + n > budget and sample:
            break
        sample.append(r)
        tot += n
    sample.sort(key=lambda r: -(len(r["ref"]) + len(r["events"])))
    cells = [0] * len(sample)
    idx = [0]
    lock = threading.Lock()

    def work():
        while True:
            with lock:
                i = idx[0]
                idx[0] += 1
            if i >= len(sample):
                return
            _, st = P.run_oracle(model, sample[i], params)
            cells[i] = st["cells"]

    t0 = time.time()
    ths = [threading.Thread(target=work) for _ in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.time() - t0
    return {"value": sum(cells) / wall / 1e9, "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"{len(sample)} of the batched reads ({sum(cells)} cells), FP64 oracle, one read per "
                      f"thread, {wall:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reads", type=int, default=4096)
    ap.add_argument("--median", type=int, default=8000)
    ap.add_argument("--expansion", type=int, default=50)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--plain-reads", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (reads shard across GPUs with no "
              "collective; the multi-GPU schedule is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    from signalalign_amd import native, synth

    prec = native.SA_FP32 if args.precision == "fp32" else native.SA_FP64
    dtype = "f32" if args.precision == "fp32" else "f64"
    t0 = time.time()
    model = synth.gaussian_model("ACEGT", 6, seed=1)
    reads = synth.batch_reads(model, args.reads, args.median, seed=args.seed, methyl_cpg=0.3)
    log(f"generated {len(reads)} reads in {time.time() - t0:.1f}s: "
        f"{sum(len(r['ref']) for r in reads)} nt, {sum(len(r['events']) for r in reads)} events")

    ctx = native.Context(0)
    dm = ctx.model(model)
    params = native.Params(diagonal_expansion=args.expansion)
    batch = native.Batch(ctx, dm, make_native_reads(native, reads), params)
    log("batch uploaded")

    for i in range(args.warmup):
        _, tm = batch.run(prec, fetch=False)
        log(f"warmup {i}: device {tm['total_ms']:.1f} ms")
    dev_ms, wall_ms, fwd_ms, bwd_ms = [], [], [], []
    cells = 0
    for i in range(args.steps):
        tw = time.perf_counter()
        res, tm = batch.run(prec, fetch=True)
        wall_ms.append((time.perf_counter() - tw) * 1e3)
        dev_ms.append(tm["total_ms"])
        fwd_ms.append(tm["forward_ms"])
        bwd_ms.append(tm["backward_ms"])
        cells = tm["cells"]
        log(f"step {i}: device {tm['total_ms']:.1f} ms (prep {tm['prep_ms']:.1f} fwd {tm['forward_ms']:.1f} "
            f"bwd {tm['backward_ms']:.1f} collect {tm['collect_ms']:.1f}), wall {wall_ms[-1]:.1f} ms, "
            f"{cells / tm['total_ms'] / 1e6:.1f} GCUPS")
    n_pairs = sum(len(r.pairs) for r in res)
    ms = float(np.mean(dev_ms))
    gcups = cells / (ms * 1e-3) / 1e9
    f_ms, b_ms = float(np.mean(fwd_ms)), float(np.mean(bwd_ms))
    if b_ms >= f_ms:
        dom, k_ms, flop = "backward_kernel", b_ms, FLOP_BWD
    else:
        dom, k_ms, flop = "forward_kernel", f_ms, FLOP_FWD
    achieved = flop * cells / (k_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[dtype]
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None, "kernel": dom,
                "kernel_ms": round(k_ms, 3), "flop_per_cell": flop,
                "other_kernel_ms": round(f_ms if dom == "backward_kernel" else b_ms, 3)}
    batch.close()

    plain = {}
    if not args.no_plain:
        # plain drop-in: one 50 kb read, e=100 (FP64 and FP32)
        r50 = synth.single_read(synth.gaussian_model("ACGT", 6, seed=1), 50000, seed=args.seed + 50)
        dm6 = ctx.model(synth.gaussian_model("ACGT", 6, seed=1))
        for name, pr in (("fp64", native.SA_FP64), ("fp32", native.SA_FP32)):
            b1 = native.Batch(ctx, dm6, make_native_reads(native, [r50]), native.Params(diagonal_expansion=100))
            b1.run(pr, fetch=False)
            _, tm1 = b1.run(pr, fetch=False)
            plain[f"read50kb_{name}_gcups"] = round(tm1["cells"] / tm1["total_ms"] / 1e6, 3)
            plain[f"read50kb_{name}_ms"] = round(tm1["total_ms"], 2)
            b1.close()
        plain["read50kb_cells"] = tm1["cells"]
        plain["read50kb_events"] = len(r50["events"])
        # plain drop-in called once per read over the first batched reads
        sub = reads[: args.plain_reads]
        tot_ms, tot_cells = 0.0, 0
        for r in sub:
            b1 = native.Batch(ctx, dm, make_native_reads(native, [r]), params)
            _, tm1 = b1.run(prec, fetch=True)
            tot_ms += tm1["total_ms"]
            tot_cells += tm1["cells"]
            b1.close()
        plain["per_read_loop_gcups"] = round(tot_cells / tot_ms / 1e6, 3)
        plain["per_read_loop_reads"] = len(sub)
        log(f"plain: {plain}")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        from tests import _parity  # noqa: F401  (oracle helper)
        cpu = cpu_baseline(model, reads, params, args.cpu_seconds, threads)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC, "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": f"batched: {args.reads} reads, log-normal lengths median {args.median} nt "
                               "(sigma 0.5, 1-50 kb), ACEGT 6-mer Gaussian table (15625 k-mers), band expansion "
                               f"{args.expansion}, {args.precision} recurrence with FP64 totals",
                   "reads": args.reads, "cells": cells, "pairs": n_pairs, "precision": args.precision},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "wall_ms_per_step_pcie_inclusive": round(float(np.mean(wall_ms)), 3),
    }
    if cpu:
        line["batched_vs_cpu"] = round(gcups / cpu["value"], 1)
    if plain.get("per_read_loop_gcups"):
        line["batched_vs_plain_loop"] = round(gcups / plain["per_read_loop_gcups"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
