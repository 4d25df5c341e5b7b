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
# HDP emissions: each emission is a descale, a grid index, four gathers, the Hermite
# cubic, the clamp and scale (~24 FLOP) plus one log: forward 2 x 24 + 88 = 136, backward
# 136 + posterior, refresh and overlap (20) = 156.
FLOP_FWD_HDP = 136.0
FLOP_BWD_HDP = 156.0
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_native_reads(native, reads):
    return [native.Read(r["ref"], r["events"], r["anchors"], r["scale"], r["shift"], r["var"]) for r in reads]


def cpu_info():
    """Host CPU model and core counts (the CPU rows run on this host's cores)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": model, "logical_cpus": os.cpu_count() or 1, "allowed_cpus": allowed}


def run_parallel(fn, items, threads):
    """fn(item) over items, longest first on `threads` threads (the oracle releases the
    GIL inside its C call); returns (results in item order, wall seconds)."""
    out = [None] * len(items)
    idx = [0]
    lock = threading.Lock()

    def work():
        while True:
            with lock:
                i = idx[0]
                idx[0] += 1
            if i >= len(items):
                return
            out[i] = fn(items[i])

    t0 = time.time()
    ths = [threading.Thread(target=work) for _ in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    return out, time.time() - t0


def cpu_baseline(model, reads, params, target_s, threads):
    """The FP64 oracle (the reference path restated in C) on this host's cores: one
    read per thread, longest first, over a bounded sample of the batched reads.
    Returns the baseline record and the oracle's results of the sampled reads
    (read index -> (pairs, log-likelihood)) for the parity check of the GPU runs."""
    import oracle
    from tests import _parity as P
    oracle.build()
    t0 = time.time()
    r0 = reads[0]
    _, st0 = P.run_oracle(model, r0, params)
    rate1 = st0["cells"] / max(time.time() - t0, 1e-6)  # one core, cells/s
    budget = rate1 * threads * target_s
    sample, tot = [], 0
    for i, r in enumerate(reads[1:], start=1):
        n = (len(r["ref"]) + len(r["events"])) * 80
        if tot + n > budget and sample:
            break
        sample.append(i)
        tot += n
    sample.sort(key=lambda i: -(len(reads[i]["ref"]) + len(reads[i]["events"])))
    res, wall = run_parallel(lambda i: P.run_oracle(model, reads[i], params), sample, threads)
    cells = sum(st["cells"] for _, st in res)
    rec = {"value": cells / wall / 1e9, "unit": "GCUPS", "cores": threads, "kind": "port",
           "sample": f"{len(sample)} of the batched reads ({cells} cells), FP64 oracle, one read per thread on "
                     f"{threads} threads, {wall:.1f} s wall", "cpu": cpu_info()}
    return rec, {i: (pairs, st["log_likelihood"]) for i, (pairs, st) in zip(sample, res)}


def parity_vs_oracle(gpu, ref, threshold):
    """Per-pair error of GPU results against the FP64 oracle over the sampled reads:
    max and p99 |dp| over pairs in both lists, pairs beyond 1e-4, pairs in one list
    only (and how many of them lie more than 1e-3 from the threshold), max |dlogL|/|logL|."""
    dps, lls, only, only_far, n_pairs = [], [], 0, 0, 0
    for i, (rp, rll) in ref.items():
        g = gpu[i]
        ka = g.pairs[:, 1].astype(np.int64) * (1 << 24) + g.pairs[:, 2]
        kb = rp[:, 1].astype(np.int64) * (1 << 24) + rp[:, 2]
        _, ia, ib = np.intersect1d(ka, kb, assume_unique=True, return_indices=True)
        dps.append(np.abs(g.pairs[ia, 0].astype(np.int64) - rp[ib, 0]) / 1e7)
        n_pairs += len(kb)
        xa = np.setdiff1d(np.arange(len(ka)), ia)
        xb = np.setdiff1d(np.arange(len(kb)), ib)
        only += len(xa) + len(xb)
        pa = np.concatenate([g.pairs[xa, 0], rp[xb, 0]]) / 1e7
        only_far += int(np.sum(np.abs(pa - threshold) > 1e-3))
        lls.append(abs(g.log_likelihood - rll) / max(abs(rll), 1e-300))
    d = np.concatenate(dps) if dps else np.zeros(1)
    return {"reads": len(ref), "oracle_pairs": n_pairs, "max_abs_dp": float(d.max()),
            "p99_abs_dp": float(np.percentile(d, 99)), "pairs_beyond_1e-4": int(np.sum(d > 1e-4)),
            "pairs_in_one_list": only, "pairs_in_one_list_beyond_1e-3_of_threshold": only_far,
            "max_ll_rel": float(max(lls)) if lls else 0.0}


def cpu_rows(native, synth, ctx, args, threads):
    """The CPU rows of the plain and HDP configurations on this host: the FP64 oracle
    on one core for one 2 kb read (ACGT 6-mer table, e=20; median of three runs) and
    one 50 kb read (e=100), and on every allowed core over a bounded sample of the HDP
    configuration's reads (both strands); the plain GPU drop-in timed beside each."""
    import oracle
    from tests import _parity as P
    oracle.build()
    out = {}
    g6 = synth.gaussian_model("ACGT", 6, seed=1)
    dm6 = ctx.model(g6)
    for name, n, e, runs in (("read2kb_e20", 2000, 20, 3), ("read50kb_e100", 50000, 100, 1)):
        r = synth.single_read(g6, n, seed=args.seed + (50 if n == 50000 else 2))
        params = native.Params(diagonal_expansion=e)
        ts = []
        for _ in range(runs):
            t0 = time.time()
            _, st = P.run_oracle(g6, r, params)
            ts.append(time.time() - t0)
        t = float(np.median(ts))
        b1 = native.Batch(ctx, dm6, make_native_reads(native, [r]), params)
        b1.run(native.SA_FP64, fetch=False)
        _, tm = b1.run(native.SA_FP64, fetch=False)
        b1.close()
        out[name] = {"events": len(r["events"]), "cells": st["cells"], "cpu_1core_s": round(t, 4),
                     "cpu_1core_gcups": round(st["cells"] / t / 1e9, 5),
                     "gpu_plain_fp64_ms": round(tm["total_ms"], 3),
                     "gpu_plain_fp64_gcups": round(tm["cells"] / tm["total_ms"] / 1e6, 3)}
    dm6.close()
    # HDP configuration: every allowed core over a sample of the strand-reads
    mt = synth.hdp_model("ACEGT", 5, seed=3)
    mc = synth.hdp_model("ACEGT", 5, seed=4)
    reads = synth.batch_reads(mt, max(64, threads * 4), 5000, seed=args.seed + 1, methyl_cpg=0.3, strands=2,
                              complement_model=mc)
    params = native.Params(diagonal_expansion=50)
    items = [(mt if r["strand"] == 0 else mc, r) for r in reads]
    items.sort(key=lambda it: -(len(it[1]["ref"]) + len(it[1]["events"])))
    res, wall = run_parallel(lambda it: P.run_oracle(it[0], it[1], params), items, threads)
    cells = sum(st["cells"] for _, st in res)
    out["hdp_all_core"] = {"strand_reads": len(items), "cells": cells, "threads": threads, "wall_s": round(wall, 2),
                           "cpu_gcups": round(cells / wall / 1e9, 5)}
    log(f"cpu rows: {out}")
    return out


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


def kernel_roofline(tms, dtype, flop_fwd, flop_bwd):
    """Roofline of the dominant kernel of a run: the kernel instantiation that carries
    the most of the job's work (band cells x arithmetic per cell; launch time would
    pick the latency-bound sweeps of the few longest reads, which overlap everything
    else), its average launch duration (HIP events on the stream each launch runs on)
    and average cells per launch, as rocprofv3 averages per kernel name."""
    tname = "float" if dtype == "f32" else "double"
    ek = tms[-1]["emission_path"]
    agg = {}
    for t in tms:
        for L in t["launches"]:
            if L["direction"] == 0 and L.get("persistent", 0):
                name = f"sa::fb_persistent_kernel<{tname}, {ek}>"
            elif L["direction"] == 0:
                name = (f"sa::forward_sw_kernel<{tname}, {L['cpt']}, {ek}>" if L["cpt"] and L["waves"] == 1
                        else f"sa::forward_halo_kernel<{tname}, {ek}, {L['waves']}>" if L.get("sync_every", 0) > 1
                        else f"sa::forward_reg_kernel<{tname}, {L['cpt']}, {ek}, {L['waves']}>" if L["cpt"]
                        else f"sa::forward_kernel<{tname}, ...>")
            else:
                name = (f"sa::backward_reg_kernel<{tname}, {L['cpt']}, {ek}>" if L["cpt"]
                        else f"sa::backward_kernel<{tname}, ...>")
            # the persistent launch runs the forward sweeps and the windows of its items
            flop = (flop_fwd + flop_bwd if L.get("persistent", 0) else flop_fwd if L["direction"] == 0
                    else flop_bwd)
            a_ = agg.setdefault(name, {"n": 0, "ms": 0.0, "cells": 0, "flop": flop})
            a_["n"] += 1
            a_["ms"] += L["ms"]
            a_["cells"] += L["cells"]
    kname = max(agg, key=lambda k: (agg[k]["cells"] * agg[k]["flop"], agg[k]["ms"]))
    k = agg[kname]
    k_ms, k_cells, flop = k["ms"] / k["n"], k["cells"] / k["n"], k["flop"]
    achieved = flop * k_cells / (k_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[dtype]
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": None, "kernel": kname,
            "kernel_avg_ms": round(k_ms, 3), "kernel_cells_per_launch": int(k_cells), "flop_per_cell": flop,
            "launches": {n: {"calls": v["n"], "avg_ms": round(v["ms"] / v["n"], 3),
                             "cells_per_launch": int(v["cells"] / v["n"])} for n, v in agg.items()}}


def hdp_leg(native, synth, ctx, args):
    """The HDP configuration: 5-mer ACEGT template and complement HDP models (3125
    k-mers, mixture densities on a 2000-point grid), 2048 reads (median 5 kb) on
    both strands (the complement read aligns to the reverse-complement target with
    the complement model), band expansion 50, FP32.  Two schedules of the same work,
    both timed: "per strand", each strand's reads a batch under its own model, one
    after the other; "one batch", all strand-reads in one batch under a table of both
    models (models.merge_strand_models; complement references lower-cased), whose
    long sweeps and emission plane overlap across the strands.  GCUPS = cells / device
    time of one step (both strands); the leg reports the faster schedule."""
    from signalalign_amd import models
    t0 = time.time()
    mt = synth.hdp_model("ACEGT", 5, seed=3)
    mc = synth.hdp_model("ACEGT", 5, seed=4)
    reads = synth.batch_reads(mt, args.hdp_reads, 5000, seed=args.seed + 1, methyl_cpg=0.3, strands=2,
                              complement_model=mc)
    log(f"hdp: generated {len(reads)} strand-reads in {time.time() - t0:.1f}s")
    params = native.Params(diagonal_expansion=50)

    def timed(label, batches):
        ms, cells, tms = [], 0, []
        for step in range(args.warmup + args.steps + 1):
            tot, c = 0.0, 0
            for si, (_, b) in enumerate(batches):
                _, tm = b.run(native.SA_FP32, fetch=False)
                tot += tm["total_ms"]
                c += tm["cells"]
                log(f"hdp {label} step {step} batch {si}: device {tm['total_ms']:.1f} ms (prep {tm['prep_ms']:.1f} "
                    f"emission plane {tm['emission_ms']:.1f} fwd {tm['forward_ms']:.1f} bwd {tm['backward_ms']:.1f}), "
                    f"emission path {tm['emission_path']}")
                if step > args.warmup:
                    tms.append(tm)
            if step > args.warmup:  # the first runs plan and record each batch's graph
                ms.append(tot)
                cells = c
        for L in tms[-1]["launches"]:
            log(f"  hdp {label} launch stream {L['stream']} {'fwd' if L['direction'] == 0 else 'bwd'} cpt "
                f"{L['cpt']} waves {L['waves']} items {L['items']} cells {L['cells']} start {L['start_ms']:.1f} ms "
                f"dur {L['ms']:.1f} ms")
        for dm, b in batches:
            b.close()
            dm.close()
        return {"gcups": round(cells / float(np.mean(ms)) / 1e6, 3), "ms_per_step": round(float(np.mean(ms)), 3),
                "cells": cells, "roofline": kernel_roofline(tms, "f32", FLOP_FWD_HDP, FLOP_BWD_HDP)}

    sched = {}
    if args.hdp_schedule in ("both", "per-strand"):
        per_strand = []
        for strand, m in ((0, mt), (1, mc)):
            dm = ctx.model(m)
            rs = [r for r in reads if r["strand"] == strand]
            per_strand.append((dm, native.Batch(ctx, dm, make_native_reads(native, rs), params)))
        sched["per strand"] = timed("per-strand", per_strand)
    if args.hdp_schedule in ("both", "one-batch"):
        dmm = ctx.model(models.merge_strand_models(mt, mc))
        merged = [native.Read(r["ref"] if r["strand"] == 0 else r["ref"].lower(), r["events"], r["anchors"],
                              r["scale"], r["shift"], r["var"]) for r in reads]
        sched["one batch"] = timed("one-batch", [(dmm, native.Batch(ctx, dmm, merged, params))])
    best = max(sched, key=lambda k: sched[k]["gcups"])
    if best == "one batch" and args.hdp_reads == 2048:  # the profiled schedule
        r = sched[best]["roofline"]
        r["traffic"] = traffic_for(r["kernel"], r["kernel_cells_per_launch"], args.hdp_traffic_file)
    out = {"workload": f"HDP: {args.hdp_reads} reads x 2 strands, log-normal lengths median 5000 nt, ACEGT 5-mer "
                       "template/complement HDP grids (3125 k-mers x 2000 points), band expansion 50, fp32",
           "schedule": best, **sched[best],
           "schedules": {k: {"gcups": v["gcups"], "ms_per_step": v["ms_per_step"]} for k, v in sched.items()}}
    log(f"hdp: {out}")
    return out


def expectations_leg(native, ctx, dm, model, reads, params, args):
    """EM expectations (sa_expect_batch, FP64): the first --expect-reads reads of the
    batched workload in one run, transition and k-mer expectations without records;
    GCUPS = band cells / device time.  Beside it the FP64 oracle's expectations on a
    bounded sample (one core)."""
    sub = make_native_reads(native, reads[: args.expect_reads])
    tm = {}
    ctx.expect_batch(dm, sub, params, record_threshold=1.0, timing=tm)  # warm-up
    runs = []
    for _ in range(2):
        tm = {}
        ctx.expect_batch(dm, sub, params, record_threshold=1.0, timing=tm)
        runs.append(tm)
    ms = float(np.mean([t["total_ms"] for t in runs]))
    cells = runs[-1]["cells"]
    out = {"workload": f"EM expectations: the first {len(sub)} reads of the batched workload, FP64, band expansion "
                       f"{args.expansion}, sums only", "reads": len(sub), "cells": cells,
           "gcups": round(cells / ms / 1e6, 3), "ms_per_step": round(ms, 3)}
    if not args.no_cpu:
        import oracle
        from tests import _parity as P
        oracle.build()
        t0, ccells, n = time.time(), 0, 0
        for r in reads:
            if time.time() - t0 > 5.0:
                break
            om = P.oracle_model(model, r)
            oracle.get_expectations(om, r["ref"], r["events"], r["anchors"],
                                    oracle.Params(diagonal_expansion=args.expansion))
            lo, hi = oracle.band(r["anchors"], len(r["ref"]) - (model.kmer_length - 1), len(r["events"]),
                                 args.expansion)
            ccells += int(np.sum((hi - lo) // 2 + 1))
            n += 1
        out["cpu_oracle_gcups_1core"] = round(ccells / (time.time() - t0) / 1e9, 4)
        out["cpu_oracle_sample_reads"] = n
    log(f"expectations: {out}")
    return out


def degenerate_leg(native, ctx, dm, model, reads, params, args):
    """Degenerate references (multi-path cells, FP64): the first --degenerate-reads
    reads of the batched workload with the cytosine of every CpG (C or E) made 'X'
    over the options "CE", in one batch.  GCUPS = band cells / device time, beside
    the same reads without 'X' (FP64) and the FP64 oracle on a bounded sample (one
    core)."""
    import re

    def xref(ref):
        return re.sub(b"[CE](?=G)", b"X", ref if isinstance(ref, bytes) else ref.encode())

    sub = reads[: args.degenerate_reads]
    xreads = [native.Read(xref(r["ref"]), r["events"], r["anchors"], r["scale"], r["shift"], r["var"],
                          degenerate_options="CE") for r in sub]
    k = model.kmer_length
    sites = sum(xref(r["ref"]).count(b"X") for r in sub)
    paths, cols = 0, 0
    for r in sub:
        x = np.frombuffer(xref(r["ref"]), dtype=np.uint8) == ord("X")
        c = np.convolve(x.astype(np.int64), np.ones(k, dtype=np.int64), "valid")
        paths += int(np.sum(2 ** c))
        cols += len(c)
    b = native.Batch(ctx, dm, xreads, params)
    b.run(native.SA_FP64, fetch=False)  # warm-up
    tms = [b.run(native.SA_FP64, fetch=False)[1] for _ in range(2)]
    b.close()
    ms = float(np.mean([t["total_ms"] for t in tms]))
    cells = tms[-1]["cells"]
    bp = native.Batch(ctx, dm, make_native_reads(native, sub), params)
    bp.run(native.SA_FP64, fetch=False)
    tmp = [bp.run(native.SA_FP64, fetch=False)[1] for _ in range(2)]
    bp.close()
    pms = float(np.mean([t["total_ms"] for t in tmp]))
    out = {"workload": f"degenerate references: the first {len(sub)} reads of the batched workload, every CpG "
                       f"cytosine 'X' over \"CE\", FP64, band expansion {args.expansion}",
           "reads": len(sub), "cells": cells, "x_sites": sites, "mean_paths_per_column": round(paths / cols, 3),
           "gcups": round(cells / ms / 1e6, 3), "ms_per_step": round(ms, 3),
           "forward_ms": round(float(np.mean([t["forward_ms"] for t in tms])), 3),
           "backward_ms": round(float(np.mean([t["backward_ms"] for t in tms])), 3),
           "backward_launches": int(tms[-1]["backward_launches"]),
           "same_reads_without_x_fp64_gcups": round(tmp[-1]["cells"] / pms / 1e6, 3)}
    if not args.no_cpu:
        import oracle
        from tests import _parity as P
        oracle.build()
        t0, ccells, n = time.time(), 0, 0
        for r in sub:
            if time.time() - t0 > 5.0:
                break
            om = P.oracle_model(model, r)
            oracle.get_aligned_pairs(om, xref(r["ref"]), r["events"], r["anchors"],
                                     oracle.Params(diagonal_expansion=args.expansion), options="CE")
            lo, hi = oracle.band(r["anchors"], len(r["ref"]) - (k - 1), len(r["events"]), args.expansion)
            ccells += int(np.sum((hi - lo) // 2 + 1))
            n += 1
        out["cpu_oracle_gcups_1core"] = round(ccells / (time.time() - t0) / 1e9, 4)
        out["cpu_oracle_sample_reads"] = n
    log(f"degenerate: {out}")
    return out


def serving_leg(native, synth, ctx, gmodel, args):
    """The serving configuration: Poisson arrivals of single-read requests at a fixed
    rate (a fraction --serve-load of the capacity a closed burst measures first),
    lengths from the batched configuration's distribution, a 50/50 mix of the ACEGT
    6-mer Gaussian model and an ACEGT 5-mer HDP model (both resident), FP32, reads
    admitted by band-matrix footprint into waves (continuous batching).  Requests
    are drawn from a pool of synthetic reads per model.  Latency = result ready on
    the host - submission; GCUPS = cells / (last result - first submission)."""
    import queue
    t0 = time.time()
    hmodel = synth.hdp_model("ACEGT", 5, seed=3)
    pools = []
    for m, seed in ((gmodel, args.seed + 11), (hmodel, args.seed + 12)):
        rs = synth.batch_reads(m, args.serve_pool, args.median, seed=seed, methyl_cpg=0.3)
        objs = make_native_reads(native, rs)
        pools.append((ctx.model(m), objs, [o._struct() for o in objs]))
    log(f"serving: pools of {args.serve_pool} reads per model in {time.time() - t0:.1f}s")
    ps = native.Params(diagonal_expansion=args.expansion)._struct()
    srv = native.Server(0, hbm_budget_bytes=96 << 30, lanes=args.serve_lanes, max_wave_reads=args.serve_max_wave,
                        length_ratio=args.serve_length_ratio, batch_delay_ms=args.serve_batch_delay)
    rng = np.random.default_rng(args.seed + 13)

    def pick(n):
        mi = rng.integers(0, 2, n)
        ri = rng.integers(0, args.serve_pool, n)
        return list(zip(mi.tolist(), ri.tolist()))

    def run(reqs, rate):
        """Submit `reqs` (at `rate` per second, or all at once), collect every result."""
        q = queue.Queue()
        lag = [0.0]

        sub_s = [0.0]

        def submitter():
            start = time.perf_counter()
            if not rate:
                # a closed burst: each model's reads in one sa_server_submit_batch call
                for mi in (0, 1):
                    dm, _, structs = pools[mi]
                    for t in srv.submit_structs(dm, [structs[ri] for m_, ri in reqs if m_ == mi], ps,
                                                native.SA_FP32):
                        q.put(t)
                sub_s[0] = time.perf_counter() - start
                q.put(None)
                return
            arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
            for (mi, ri), at in zip(reqs, arrivals):
                now = time.perf_counter() - start
                if at > now:
                    time.sleep(at - now)
                else:
                    lag[0] = max(lag[0], now - at)
                dm, _, structs = pools[mi]
                q.put(srv.submit_struct(dm, structs[ri], ps, native.SA_FP32))
            sub_s[0] = time.perf_counter() - start
            q.put(None)

        th = threading.Thread(target=submitter)
        th.start()
        lat, cells, t_first, t_last, waves = [], 0, None, 0.0, set()
        while True:
            t = q.get()
            if t is None:
                break
            res, tm = srv.wait(t, timeout=300.0, fetch=False)
            srv.release(t)
            lat.append(tm["done_s"] - tm["submit_s"])
            cells += res.cells
            t_first = tm["submit_s"] if t_first is None else min(t_first, tm["submit_s"])
            t_last = max(t_last, tm["done_s"])
            waves.add(tm["wave"])
        th.join()
        log(f"serving: {len(reqs)} submissions took {sub_s[0] * 1e3:.0f} ms")
        return np.array(lat), cells, t_last - t_first, len(waves), lag[0]

    # a server runs for long: one unmeasured burst first grows every lane's device and
    # pinned host buffers to their working sizes (a growth of pinned memory waits for the
    # device), as hours of traffic would
    run(pick(args.serve_burst), 0.0)
    burst = pick(args.serve_burst)
    st0 = srv.stats()
    lat, cells, span, nw, _ = run(burst, 0.0)
    capacity = len(burst) / span
    st1 = srv.stats()
    log(f"serving burst: lanes' fill {st1['fill_ms'] - st0['fill_ms']:.0f} ms, run {st1['run_ms'] - st0['run_ms']:.0f} ms, "
        f"device {st1['device_ms'] - st0['device_ms']:.0f} ms over {st1['waves'] - st0['waves']} waves")
    log(f"serving: burst of {len(burst)} reads in {span:.2f}s ({capacity:.0f} reads/s, "
        f"{cells / span / 1e9:.1f} GCUPS, {nw} waves)")
    rate = args.serve_load * capacity
    reqs = pick(args.serve_requests)
    lat, cells, span, nw, lag = run(reqs, rate)
    st = srv.stats()
    srv.close()
    for dm, _, _ in pools:
        dm.close()
    out = {"workload": f"serving: {len(reqs)} single-read requests, Poisson arrivals at {rate:.0f} reads/s "
                       f"({args.serve_load:.0%} of the {capacity:.0f} reads/s a {len(burst)}-read burst sustains), "
                       f"lengths log-normal median {args.median} nt, 50/50 ACEGT 6-mer Gaussian / ACEGT 5-mer HDP "
                       f"models resident, band expansion {args.expansion}, fp32, drawn from pools of "
                       f"{args.serve_pool} synthetic reads per model",
           "rate_reads_per_s": round(rate, 1), "capacity_reads_per_s": round(capacity, 1),
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
           "max_ms": round(float(lat.max()) * 1e3, 3), "gcups": round(cells / span / 1e9, 3),
           "achieved_reads_per_s": round(len(reqs) / span, 1), "waves": nw,
           "mean_wave_reads": round(len(reqs) / max(nw, 1), 1), "submitter_max_lag_ms": round(lag * 1e3, 3),
           "max_wave_reads": st["max_wave_reads"], "lanes": args.serve_lanes,
           "wave_cap_reads": args.serve_max_wave, "length_ratio": args.serve_length_ratio,
           "batch_delay_ms": args.serve_batch_delay}
    log(f"serving: {out}")
    return out


def traffic_for(kname, cells, path):
    """HBM bytes per launch of `kname`: the calibrated rocprofv3 FETCH_SIZE +
    WRITE_SIZE bytes per band cell kept with the profiles (the file `path` of the
    profiled workload), times the launch's cells (None if that kernel is not listed)."""
    try:
        with open(path) as f:
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
    ap.add_argument("--no-cpu-rows", action="store_true", help="skip the single-core 2 kb / 50 kb and all-core HDP "
                                                                 "CPU rows")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true", help="skip the FP32-vs-FP64 per-read error report")
    ap.add_argument("--no-hdp", action="store_true", help="skip the HDP-emission configuration")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch tuning")
    ap.add_argument("--tuning", default="", help="fixed launch tuning instead of the autotuner, e.g. "
                                                 "long_factor=1.5,stream_split=0.65,waves_per_simd=0")
    ap.add_argument("--no-graph", action="store_true", help="timed steps without graph replay")
    ap.add_argument("--hdp-reads", type=int, default=2048)
    ap.add_argument("--hdp-only", action="store_true", help="run only the HDP configuration (for profiling it)")
    ap.add_argument("--hdp-schedule", choices=["both", "per-strand", "one-batch"], default="both",
                    help="the HDP leg's schedules to time (one of them: for profiling it)")
    ap.add_argument("--hdp-traffic-file", default=os.path.join(ROOT, "profiles", "traffic_hdp_2048x2reads.txt"))
    ap.add_argument("--no-serve", action="store_true", help="skip the serving configuration")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving configuration")
    ap.add_argument("--no-expect", action="store_true", help="skip the EM expectations configuration")
    ap.add_argument("--expect-reads", type=int, default=512)
    ap.add_argument("--no-degenerate", action="store_true", help="skip the degenerate-reference configuration")
    ap.add_argument("--degenerate-reads", type=int, default=256)
    ap.add_argument("--serve-requests", type=int, default=100000)
    ap.add_argument("--serve-pool", type=int, default=4096)
    ap.add_argument("--serve-burst", type=int, default=8192)
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--serve-lanes", type=int, default=4)
    ap.add_argument("--serve-batch-delay", type=float, default=30.0,
                    help="ms a free lane waits for requests to gather while other waves run (negative: none)")
    ap.add_argument("--serve-length-ratio", type=float, default=2.0,
                    help="a wave's reads are within this factor of the oldest request's length (negative: any)")
    ap.add_argument("--serve-max-wave", type=int, default=1024)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run a device copy of this many MiB (HBM byte-counter calibration under rocprofv3)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic_batched_4096reads.txt"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (reads shard across GPUs with no "
              "collective: each GPU would take whole reads, balanced by band cells; that schedule is not built)", file=sys.stderr)
        sys.exit(2)

    from signalalign_amd import native, synth

    if args.serve_only:
        ctx = native.Context(0)
        print(json.dumps({"serving": serving_leg(native, synth, ctx, synth.gaussian_model("ACEGT", 6, seed=1), args)}),
              flush=True)
        return
    if args.hdp_only:
        if args.copy_calibration_mib:
            copy_calibration(args.copy_calibration_mib)
        ctx = native.Context(0)
        print(json.dumps({"hdp": hdp_leg(native, synth, ctx, args)}), flush=True)
        return
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
    tuning = None
    if args.tuning:
        kv = dict(item.split("=") for item in args.tuning.split(","))
        batch.set_tuning(**{k: (float(v) if k in ("long_factor", "stream_split") else int(v)) for k, v in kv.items()})
        tuning = batch.tuning()
        log(f"tuning: {tuning}")
    elif not args.no_autotune:
        tuning, tuned_ms = batch.autotune(prec)
        log(f"autotune: {tuning} ({tuned_ms:.1f} ms)")
    if args.no_graph:
        batch.set_tuning(use_graph=0)
    for i in range(2):  # the second run records the graph the timed steps replay
        _, tm = batch.run(prec, fetch=False)
        log(f"prime {i}: device {tm['total_ms']:.1f} ms")
    dev_ms, wall_ms, fwd_ms, bwd_ms, tms = [], [], [], [], []
    cells = 0
    for i in range(args.steps):
        # wall: the library call, inputs resident in HBM to the pairs in host memory
        # (sa_batch_run), without the Python harness's conversion of the pairs
        tw = time.perf_counter()
        res, tm = batch.run(prec, fetch=False)
        wall_ms.append((time.perf_counter() - tw) * 1e3)
        dev_ms.append(tm["total_ms"])
        fwd_ms.append(tm["forward_ms"])
        bwd_ms.append(tm["backward_ms"])
        tms.append(tm)
        cells = tm["cells"]
        log(f"step {i}: device {tm['total_ms']:.1f} ms{' (graph replay)' if tm['replayed'] else ''} "
            f"(prep {tm['prep_ms']:.1f} fwd {tm['forward_ms']:.1f} "
            f"bwd {tm['backward_ms']:.1f} collect {tm['collect_ms']:.1f}), wall {wall_ms[-1]:.1f} ms, "
            f"{cells / tm['total_ms'] / 1e6:.1f} GCUPS")
    for L in tms[-1]["launches"]:
        log(f"  launch stream {L['stream']} {'fwd' if L['direction'] == 0 else 'bwd'} cpt {L['cpt']} waves "
            f"{L['waves']} sync {L['sync_every']} persistent {L['persistent']} items {L['items']} cells {L['cells']} "
            f"start {L['start_ms']:.1f} ms "
            f"dur {L['ms']:.1f} ms")
    res, _ = batch.run(prec, fetch=True)  # the pairs, for the error reports
    n_pairs = sum(len(r.pairs) for r in res)
    ms = float(np.mean(dev_ms))
    gcups = cells / (ms * 1e-3) / 1e9
    # roofline of the dominant kernel: the kernel (one instantiation) that carries the
    # most of the job's work (band cells x arithmetic per cell; launch time would pick
    # the latency-bound sweeps of the few longest reads, which overlap everything else);
    # its average launch duration (HIP events on the stream each launch runs on) and
    # average cells per launch, as rocprofv3 averages per kernel name
    roofline = kernel_roofline(tms, dtype, FLOP_FWD, FLOP_BWD)
    if args.reads == 4096 and args.median == 8000 and args.expansion == 50:  # the profiled workload
        roofline["traffic"] = traffic_for(roofline["kernel"], roofline["kernel_cells_per_launch"], args.traffic_file)
    roofline["forward_ms"] = round(float(np.mean(fwd_ms)), 3)
    roofline["backward_ms"] = round(float(np.mean(bwd_ms)), 3)
    roofline["sweeps_ms"] = round(float(np.mean([t["fb_ms"] for t in tms])), 3)
    fp_err, res64, fp64_gcups = None, None, None
    if not args.no_fp64_check and prec == native.SA_FP32:
        batch.run(native.SA_FP64, fetch=False)
        batch.run(native.SA_FP64, fetch=False)  # records the FP64 graph
        t64 = [batch.run(native.SA_FP64, fetch=False)[1]["total_ms"] for _ in range(max(1, args.steps))]
        res64, tm64 = batch.run(native.SA_FP64, fetch=True)
        fp64_gcups = round(tm64["cells"] / float(np.mean(t64)) / 1e6, 3)
        fp_err = fp32_error(res, res64, params.threshold)
        fp_err["fp64_gcups"] = fp64_gcups
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

    expect = None if args.no_expect else expectations_leg(native, ctx, dm, model, reads, params, args)
    degenerate = None if args.no_degenerate else degenerate_leg(native, ctx, dm, model, reads, params, args)
    hdp = None if args.no_hdp else hdp_leg(native, synth, ctx, args)
    serving = None if args.no_serve else serving_leg(native, synth, ctx, model, args)

    cpu, parity, rows = None, None, None
    if not args.no_cpu:
        threads = args.cpu_threads or cpu_info()["allowed_cpus"]
        from tests import _parity  # noqa: F401  (oracle helper)
        cpu, ref = cpu_baseline(model, reads, params, args.cpu_seconds, threads)
        log(f"cpu baseline: {cpu}")
        parity = {args.precision: parity_vs_oracle(res, ref, params.threshold)}
        if res64 is not None:
            parity["fp64"] = parity_vs_oracle(res64, ref, params.threshold)
        log(f"parity vs the oracle on the sampled reads: {parity}")
        if not args.no_cpu_rows:
            rows = cpu_rows(native, synth, ctx, args, threads)

    line = {
        "metric": METRIC, "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": f"batched: {args.reads} reads, log-normal lengths median {args.median} nt "
                               "(sigma 0.5, 1-50 kb), ACEGT 6-mer Gaussian table (15625 k-mers), band expansion "
                               f"{args.expansion}, {args.precision} recurrence with FP64 totals",
                   "reads": args.reads, "cells": cells, "pairs": n_pairs, "precision": args.precision,
                   "launch_tuning": tuning, "graph_replayed_steps": int(sum(t["replayed"] for t in tms))},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp32_error_vs_fp64": fp_err,
        "parity_vs_oracle_on_cpu_sample": parity,
        "oracle_matching": None if fp64_gcups is None else {
            "precision": "fp64", "gcups": fp64_gcups,
            "note": "the FP64 recurrence matches the oracle to the bar on every pair; the FP32 recurrence moves a "
                    "posterior by up to 4.5e-4 where a rounding crosses a knot of the reference's lookup logAdd"},
        "cpu_rows": rows,
        "plain_gpu": plain,
        "hdp": hdp,
        "serving": serving,
        "expectations": expect,
        "degenerate": degenerate,
        "wall_ms_per_step_pcie_inclusive": round(float(np.mean(wall_ms)), 3),
        "wall_gcups_pcie_inclusive": round(cells / float(np.mean(wall_ms)) / 1e6, 3),
    }
    if cpu:
        line["batched_vs_cpu"] = round(gcups / cpu["value"], 1)
        if fp64_gcups:
            line["batched_fp64_vs_cpu"] = round(fp64_gcups / cpu["value"], 1)
    if plain.get("per_read_loop_gcups"):
        line["batched_vs_plain_loop"] = round(gcups / plain["per_read_loop_gcups"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
