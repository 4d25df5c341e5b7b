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


def copy_calibration(mib):
    """One device-to-device copy of a known size: the HBM byte counters of this
    dispatch (FETCH_SIZE, WRITE_SIZE) calibrate those of the alignment kernels."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    log(f"copy calibration: {mib} MiB read + {mib} MiB written")
    del x, y


def traffic_for(kname, cells, args):
    """HBM bytes per launch of `kname`: the calibrated rocprofv3 FETCH_SIZE +
    WRITE_SIZE bytes per band cell kept with the profiles, times the launch's cells
    (None if that kernel is not listed or the workload is not the profiled one)."""
    if args.reads != 4096 or args.median != 8000 or args.expansion != 50:
        return None
    try:
        with open(args.traffic_file) as f:
            for line in f:
                parts = line.strip().split("\t")
                if len(parts) >= 2 and parts[0] == kname:
                    return round(float(parts[1]) * cells)
    except OSError:
        pass
    return None


def fp32_error(r32, r64, threshold):
    """Per-read error of the FP32 recurrence against the FP64 one (which matches the
    FP64 oracle to 1e-4 in posteriors): max |dp| over pairs in both lists, pairs
    present in only one list (and how many of those lie more than 1e-3 from the
    threshold), and |dlogL|/|logL|."""
    dps, lls, only, only_far = [], [], 0, 0
    for a, b in zip(r32, r64):
        ka = a.pairs[:, 1] * (1 << 24) + a.pairs[:, 2]
        kb = b.pairs[:, 1] * (1 << 24) + b.pairs[:, 2]
        _, ia, ib = np.intersect1d(ka, kb, assume_unique=True, return_indices=True)
        dps.append(float(np.max(np.abs(a.pairs[ia, 0] - b.pairs[ib, 0]))) / 1e7 if len(ia) else 0.0)
        xa = np.setdiff1d(np.arange(len(ka)), ia)
        xb = np.setdiff1d(np.arange(len(kb)), ib)
        only += len(xa) + len(xb)
        pa = np.concatenate([a.pairs[xa, 0], b.pairs[xb, 0]]) / 1e7
        only_far += int(np.sum(np.abs(pa - threshold) > 1e-3))
        lls.append(abs(a.log_likelihood - b.log_likelihood) / max(abs(b.log_likelihood), 1e-300))
    dps, lls = np.array(dps), np.array(lls)
    return {"reads": len(dps), "max_abs_dp": float(dps.max()), "p99_abs_dp": float(np.percentile(dps, 99)),
            "median_abs_dp": float(np.median(dps)), "max_ll_rel": float(lls.max()),
            "p99_ll_rel": float(np.percentile(lls, 99)), "pairs_in_one_list": only,
            "pairs_in_one_list_beyond_1e-3_of_threshold": only_far}


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
    ap.add_argument("--no-fp64-check", action="store_true", help="skip the FP32-vs-FP64 per-read error report")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run a device copy of this many MiB (HBM byte-counter calibration under rocprofv3)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic_batched_4096reads.txt"))
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

    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
    ctx = native.Context(0)
    dm = ctx.model(model)
    params = native.Params(diagonal_expansion=args.expansion)
    batch = native.Batch(ctx, dm, make_native_reads(native, reads), params)
    log("batch uploaded")

    for i in range(args.warmup):
        _, tm = batch.run(prec, fetch=False)
        log(f"warmup {i}: device {tm['total_ms']:.1f} ms")
    dev_ms, wall_ms, fwd_ms, bwd_ms, tms = [], [], [], [], []
    cells = 0
    for i in range(args.steps):
        tw = time.perf_counter()
        res, tm = batch.run(prec, fetch=True)
        wall_ms.append((time.perf_counter() - tw) * 1e3)
        dev_ms.append(tm["total_ms"])
        fwd_ms.append(tm["forward_ms"])
        bwd_ms.append(tm["backward_ms"])
        tms.append(tm)
        cells = tm["cells"]
        log(f"step {i}: device {tm['total_ms']:.1f} ms (prep {tm['prep_ms']:.1f} fwd {tm['forward_ms']:.1f} "
            f"bwd {tm['backward_ms']:.1f} collect {tm['collect_ms']:.1f}), wall {wall_ms[-1]:.1f} ms, "
            f"{cells / tm['total_ms'] / 1e6:.1f} GCUPS")
    n_pairs = sum(len(r.pairs) for r in res)
    ms = float(np.mean(dev_ms))
    gcups = cells / (ms * 1e-3) / 1e9
    # roofline of the dominant kernel: the kernel (one instantiation) with the most
    # launch time; its average launch duration (HIP events on the stream each launch
    # runs on) and average cells per launch, as rocprofv3 averages per kernel name
    tname = "float" if dtype == "f32" else "double"
    ek = tms[-1]["emission_path"]
    agg = {}
    for t in tms:
        for L in t["launches"]:
            if L["direction"] == 0:
                name = (f"sa::forward_reg_kernel<{tname}, {L['cpt']}, {ek}, {L['waves']}>" if L["cpt"]
                        else f"sa::forward_kernel<{tname}, ...>")
            else:
                name = (f"sa::backward_reg_kernel<{tname}, {L['cpt']}, {ek}>" if L["cpt"]
                        else f"sa::backward_kernel<{tname}, ...>")
            a_ = agg.setdefault(name, {"n": 0, "ms": 0.0, "cells": 0, "flop": FLOP_FWD if L["direction"] == 0
                                       else FLOP_BWD})
            a_["n"] += 1
            a_["ms"] += L["ms"]
            a_["cells"] += L["cells"]
    kname = max(agg, key=lambda k: agg[k]["ms"])
    k = agg[kname]
    k_ms, k_cells, flop = k["ms"] / k["n"], k["cells"] / k["n"], k["flop"]
    achieved = flop * k_cells / (k_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[dtype]
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": traffic_for(kname, k_cells, args), "kernel": kname,
                "kernel_avg_ms": round(k_ms, 3), "kernel_cells_per_launch": int(k_cells),
                "flop_per_cell": flop,
                "launches": {n: {"calls": v["n"], "avg_ms": round(v["ms"] / v["n"], 3),
                                 "cells_per_launch": int(v["cells"] / v["n"])} for n, v in agg.items()},
                "forward_ms": round(float(np.mean(fwd_ms)), 3), "backward_ms": round(float(np.mean(bwd_ms)), 3),
                "sweeps_ms": round(float(np.mean([t["fb_ms"] for t in tms])), 3)}
    fp_err = None
    if not args.no_fp64_check and prec == native.SA_FP32:
        res64, tm64 = batch.run(native.SA_FP64, fetch=True)
        fp_err = fp32_error(res, res64, params.threshold)
        fp_err["fp64_gcups"] = round(tm64["cells"] / tm64["total_ms"] / 1e6, 3)
        log(f"fp32 vs fp64: {fp_err}")
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
        "fp32_error_vs_fp64": fp_err,
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
