"""Benchmark: histories/s of the k-eigenvalue solve on one MI355X.

Workload (the configuration the headline metric is quoted on): the 9-region
4-group slab (5 fuel regions of 1.0 cm, 4 water regions of 0.5 cm, 20 cells
per region, reflective left / vacuum right), 10^7 particles per generation,
20 inactive + 100 active generations, seeded synthetic cross sections.

A step is one full k-eigenvalue solve of that workload (120 generations).
`value` is the FP32 event-kernel histories/s over the K timed steps
(region-by-region surface tracking; `--method woodcock` selects Woodcock
delta tracking); the plain FP64 port (one thread per history, surface tracking) runs
the same workload once beside it, and the FP64 CPU oracle is timed on a
bounded sample on the host cores.  The serving form runs a 10 s Poisson
request stream (50 req/s) beside them for the p50/p99 solve latency.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--method region|woodcock]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_STATE_FP32, B_STATE_FP64 = 24, 32
# HBM bytes per event-kernel launch of 10^7 histories on this workload, from
# two rocprofv3 counter passes (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000,
# factors calibrated on a 1 GiB device copy): region tracking 97.6 KiB x 2
# read + 105869.3 KiB written (profiles/k_event_region_9region_4group_hbm_traffic.md),
# Woodcock 15514.7 KiB x 2 + 159322.7 KiB (profiles/k_woodcock_9region_4group_hbm_traffic.md).
TRAFFIC_BYTES_PER_LAUNCH = {"region": (97.5625 * 2 + 105869.25) * 1024,
                            "woodcock": (15514.7 * 2 + 159322.7) * 1024}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 port leg")
    ap.add_argument("--cpu-histories", type=int, default=10_000_000)
    ap.add_argument("--cpu-generations", type=int, default=24)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving latency leg")
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--no-diffusion", action="store_true", help="skip the diffusion leg")
    ap.add_argument("--method", choices=["region", "woodcock"], default="region",
                    help="tracking of the FP32 event kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np

    import nraps_amd as nr
    from nraps_amd import _abi

    st, info = nr.device_info(0)
    if st != 0:
        raise SystemExit("no HIP device: " + info)
    with open(os.path.join(ROOT, "tests", "golden", "workloads.json")) as f:
        fbar = json.load(f)["workloads"]["slab9_g4_1e7"]["mean_flights"]

    prob, o = nr.generate(1)           # 10^7 x (20 + 100), FP32 event kernel
    o.method = _abi.EVENT_REGION if args.method == "region" else _abi.EVENT_WOODCOCK
    o.use_graph = 0                    # per-generation HIP events around the transport kernel
    N, gens = int(o.histories), int(o.generations)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    for i in range(args.warmup):
        nr.solve(prob, o)
        log(f"warmup {i + 1}/{args.warmup} done")
    tr_ms, tr_launch, dev_ms, tot_ms, ks = 0.0, 0, 0.0, 0.0, []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = nr.solve(prob, o)
        tr_ms += r.ms_transport
        tr_launch += r.transport_launches
        dev_ms += r.ms_device
        tot_ms += r.ms_total
        ks.append((r.k_mean, r.k_sigma))
        log(f"step {i + 1}/{args.steps}: k = {r.k_mean:.6f} +- {r.k_sigma:.6f}")
    t1 = time.perf_counter()
    wall = t1 - t0
    value = args.steps * N * gens / wall
    k32, s32 = ks[-1]

    out = {
        "metric": "histories/s on 1 GPU (FP32 vs plain FP64) at k within 3σ; p50/p99 latency",
        "value": value, "unit": "histories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "9-region 4-group slab: 1e7 particles/generation, "
                               "20 inactive + 100 active generations, FP32 event kernel "
                               + ("with region-by-region surface tracking" if args.method == "region"
                                  else "with Woodcock delta tracking") + ", FP64 k",
                   "histories_per_generation": N,
                   "generations": gens, "cells": prob.n_cells, "groups": prob.n_groups},
        "k_fp32": k32, "k_fp32_sigma": s32,
        # the solve's own clocks: generation loop on the device (HIP events)
        # and upload-to-download inside the library
        "ms_device_per_step": dev_ms / args.steps, "ms_solve_per_step": tot_ms / args.steps,
        "device_histories_per_s": args.steps * N * gens / (dev_ms * 1e-3),
    }
    # roofline of the dominant kernel (the transport launch)
    b_hist = B_STATE_FP32 * (1 + 2 * fbar)
    avg_ms = tr_ms / max(tr_launch, 1)
    achieved = b_hist * N / (avg_ms * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS,
                       "traffic": TRAFFIC_BYTES_PER_LAUNCH[args.method],
                       "kernel": f"k_event<float,lds,{args.method}>", "avg_launch_ms": avg_ms,
                       "bytes_per_history": b_hist, "mean_flights": fbar}

    if not args.no_fp64:
        o64 = nr.generate(1)[1]
        o64.method, o64.precision, o64.use_graph = _abi.HISTORY_SURFACE, _abi.FP64, 0
        t = time.perf_counter()
        r64 = nr.solve(prob, o64)
        w64 = time.perf_counter() - t
        v64 = N * gens / w64
        comb = math.hypot(s32, r64.k_sigma)
        out["fp64_plain_histories_per_s"] = v64
        out["speedup_vs_fp64_plain"] = value / v64
        out["k_fp64_plain"] = r64.k_mean
        out["k_fp64_plain_sigma"] = r64.k_sigma
        out["k_delta_over_sigma"] = abs(k32 - r64.k_mean) / comb if comb > 0 else None
        log(f"plain FP64 port: {v64:.4g} histories/s, k = {r64.k_mean:.6f} +- {r64.k_sigma:.6f}")

    if not args.no_serve:
        # serving form: Poisson requests of uneven slab problems (fission bank,
        # stop at 50 pcm relative k standard error), continuous batching
        from nraps_amd.serve_bench import run_serving
        sv = run_serving(args.serve_rate, args.serve_seconds, drain_s=300.0, log=log)
        out["latency"] = {k: sv[k] for k in ("rate_per_s", "seconds", "submitted", "completed",
                                             "p50_ms", "p99_ms", "histories_per_s",
                                             "sigma_k_target_pcm")}
        log(f"serving at {args.serve_rate:g} req/s: p50 {sv['p50_ms']:.1f} ms, "
            f"p99 {sv['p99_ms']:.1f} ms")

    if not args.no_diffusion:
        # diffusion solver (src/discrete.rs): the 201-region 8-group slab
        # (2000 cells) alone, and the 256 uneven slabs in one launch
        pd = nr.generate(2)[0]
        nr.diffusion_solve(pd)
        d1 = [nr.diffusion_solve(pd) for _ in range(5)]
        dms = sorted(r.ms_device for r in d1)[2]
        batch = [nr.generate(3, i)[0] for i in range(256)]
        nr.diffusion_solve_batch(batch[:4])
        db = [nr.diffusion_solve_batch(batch) for _ in range(3)]
        bms = sorted(r[0].ms_device for r in db)[1]
        its = d1[0].iterations
        out["diffusion"] = {
            "workload": "201-region 8-group slab, 2000 cells: power iteration to the "
                        "reference's tolerances", "iterations": its, "k": d1[0].k,
            "ms_device": dms, "us_per_group_solve": dms * 1e3 / (its * pd.n_groups),
            "bound": "latency (8 sequential group solves per iteration, 2 block scans each)",
            "batch_problems": 256, "batch_ms_device": bms,
            "batch_problems_per_s": 256 / (bms * 1e-3)}
        log(f"diffusion: {its} iterations in {dms:.2f} ms; 256-problem batch {bms:.2f} ms")

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        threads = max((os.cpu_count() or 2) - 1, 1)
        threads = min(threads, 15)
        oc = nr.run_opts(histories=args.cpu_histories, generations=args.cpu_generations, skip=2)
        t = time.perf_counter()
        rc = oracle_lib.solve(prob, oc, threads=threads)
        wc = time.perf_counter() - t
        cpu_v = args.cpu_histories * args.cpu_generations / wc
        out["cpu_baseline"] = {"value": cpu_v, "unit": "histories/s", "cores": threads,
                               "kind": "port",
                               "sample": f"FP64 CPU oracle (reference semantics, surface tracking) "
                                         f"on the same 9-region 4-group slab: {args.cpu_histories} histories x "
                                         f"{args.cpu_generations} generations, {wc:.1f} s"}
        out["speedup_vs_cpu"] = value / cpu_v
        log(f"CPU oracle: {cpu_v:.4g} histories/s on {threads} threads")
        # the same sample on one more thread (all of the 16-core share: the
        # reference uses available_parallelism() - 1, src/mc_code.rs:302)
        t = time.perf_counter()
        oracle_lib.solve(prob, oc, threads=threads + 1)
        wa = time.perf_counter() - t
        out["cpu_baseline"]["value_one_more_thread"] = args.cpu_histories * args.cpu_generations / wa
        out["cpu_baseline"]["cores_one_more_thread"] = threads + 1
        log(f"CPU oracle: {out['cpu_baseline']['value_one_more_thread']:.4g} histories/s "
            f"on {threads + 1} threads")
        if "diffusion" in out:
            # the FP64 CPU oracle of the diffusion solver (one core, O(n) solves)
            t = time.perf_counter()
            oracle_lib.diffusion(pd)
            c1 = (time.perf_counter() - t) * 1e3
            t = time.perf_counter()
            for q in batch:
                oracle_lib.diffusion(q)
            cb = (time.perf_counter() - t) * 1e3
            out["diffusion"]["cpu_baseline"] = {
                "ms": c1, "batch_ms": cb, "cores": 1, "kind": "port",
                "sample": "FP64 CPU oracle (tridiagonal solves): the 2000-cell slab once, "
                          "the 256 batch problems one after another"}
            out["diffusion"]["speedup_vs_cpu"] = c1 / out["diffusion"]["ms_device"]
            out["diffusion"]["batch_speedup_vs_cpu"] = cb / out["diffusion"]["batch_ms_device"]
            log(f"CPU diffusion oracle: {c1:.1f} ms; batch {cb:.0f} ms")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
