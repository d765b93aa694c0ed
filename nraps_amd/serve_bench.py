"""Serving and batched-form measurements on one GPU.

Serving: a Poisson stream of solve requests (exponential gaps drawn from
PCG32, seed 42) is submitted to one nraps_server for a fixed window.  Each
request is an uneven slab problem drawn like the batched workload (3..201
regions of 10 cells, 2/4/8 groups, 10^4..10^6 histories per cycle, fission
bank, 20 inactive cycles) and stops once its relative k standard error is at
most 50 pcm (or after 120 cycles).  Each cycle of the resident mix is
replayed from a hipGraph captured whenever the membership changes.
Reported: requests completed,
p50/p99 submit-to-result latency, and histories/s over the run.

Batched: the 256 uneven problems solved by one nraps_mc_solve_batch call
(one hipGraph per cycle, every cycle of every problem in the same launches).

    python -m nraps_amd.serve_bench --rates 50,200,800 --seconds 60
    python -m nraps_amd.serve_bench --batch 256
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import nraps_amd as nr  # noqa: E402
from nraps_amd import _abi  # noqa: E402

MASK64 = (1 << 64) - 1


class Pcg32:
    """PCG32 (XSH-RR) seeded like src/rand.rs:49-71 with an explicit seed."""

    def __init__(self, seed, stream):
        self.inc = ((stream << 1) | 1) & MASK64
        self.state = 0
        self.next()
        self.state = (self.state + seed) & MASK64
        self.next()

    def next(self):
        s = self.state
        self.state = (s * 6364136223846793005 + self.inc) & MASK64
        x = (((s >> 18) ^ s) >> 27) & 0xFFFFFFFF
        r = s >> 59
        return ((x >> r) | (x << ((32 - r) & 31))) & 0xFFFFFFFF


def arrival_times(rate, seconds, seed=42, stream=0xA7):
    """Poisson arrivals on [0, seconds): exponential gaps -ln(U)/rate."""
    g = Pcg32(seed, stream)
    t, out = 0.0, []
    while True:
        u = (g.next() + 1.0) / 4294967296.0  # (0, 1]
        t += -math.log(u) / rate
        if t >= seconds:
            return out
        out.append(t)


def percentile(xs, q):
    if not xs:
        return None
    s = sorted(xs)
    k = (len(s) - 1) * q
    lo, hi = int(math.floor(k)), int(math.ceil(k))
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def run_serving(rate, seconds, *, first_index=0, sigma_pcm=50.0, min_active=10,
                budget_bytes=200 * 10**9, max_slots=1024, drain_s=600.0, precision=_abi.FP32,
                use_graph=1, log=None, method=_abi.EVENT_REGION):
    times = arrival_times(rate, seconds)
    reqs = [nr.generate(4, first_index + i) for i in range(len(times))]
    lat, qlat, hist, cycles = [], [], 0, []
    with nr.Server(precision=precision, budget_bytes=budget_bytes, max_slots=max_slots,
                   min_active=min_active, sigma_k_target=sigma_pcm * 1e-5,
                   use_graph=use_graph, method=method) as srv:
        tickets = []
        t0 = time.perf_counter()
        for (p, o), ta in zip(reqs, times):
            d = t0 + ta - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(srv.submit(p, o))
        t_sub = time.perf_counter() - t0
        if log:
            log(f"rate {rate}/s: {len(tickets)} requests submitted in {t_sub:.1f} s")
        deadline = time.perf_counter() + drain_s
        done, last = 0, time.perf_counter()
        for t in tickets:
            while True:
                if log and time.perf_counter() - last > 15.0:
                    last = time.perf_counter()
                    log(f"rate {rate}/s: {done} of {len(tickets)} requests done")
                left_ms = int(max(0.0, deadline - time.perf_counter()) * 1e3)
                r = srv.wait(t, timeout_ms=min(left_ms, 5000))
                if r.status_code != _abi.NRAPS_PENDING or left_ms <= 5000:
                    break
            if r.status_code == _abi.NRAPS_PENDING:
                break
            st, q, tot, cyc = srv.latency(t)
            if r.status_code == 0 and st == 0:
                done += 1
                lat.append(tot)
                qlat.append(q)
                hist += r.histories_run
                cycles.append(cyc)
            srv.forget(t)
        wall = time.perf_counter() - t0
    return {
        "rate_per_s": rate, "seconds": seconds, "submitted": len(tickets), "completed": done,
        "p50_ms": percentile(lat, 0.50), "p99_ms": percentile(lat, 0.99),
        "queue_p50_ms": percentile(qlat, 0.50), "queue_p99_ms": percentile(qlat, 0.99),
        "histories": hist, "wall_s": wall, "histories_per_s": hist / wall if wall > 0 else None,
        "mean_cycles": sum(cycles) / len(cycles) if cycles else None,
        "sigma_k_target_pcm": sigma_pcm, "graph": bool(use_graph),
    }


def run_batch(n, *, use_graph=1, precision=_abi.FP32, generations=None, method=_abi.EVENT_REGION):
    probs, opts = [], []
    for i in range(n):
        p, o = nr.generate(3, i)
        o.use_graph = use_graph
        o.precision = precision
        o.method = method
        if generations:
            o.generations = generations
        probs.append(p)
        opts.append(o)
    # warm-up: a short batch of the first problems loads the kernels and
    # grows the device memory pool, as bench.py's warm-up steps do
    wo = []
    for o in opts[:8]:
        w = type(o).from_buffer_copy(o)
        w.generations, w.skip = 3, 2
        wo.append(w)
    nr.solve_batch(probs[:8], wo)
    t = time.perf_counter()
    res = nr.solve_batch(probs, opts)
    wall = time.perf_counter() - t
    hist = sum(r.histories_run for r in res)
    return {"problems": n, "histories": hist, "wall_s": wall, "histories_per_s": hist / wall,
            "generations": int(opts[0].generations),
            "library_s": res[0].ms_total * 1e-3, "device_loop_s": res[0].ms_device * 1e-3,
            "k_min": min(r.k_mean for r in res), "k_max": max(r.k_mean for r in res)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="50,200,800")
    ap.add_argument("--seconds", type=float, default=60.0)
    ap.add_argument("--drain", type=float, default=600.0)
    ap.add_argument("--batch", type=int, default=0, help="run the batched form with n problems")
    ap.add_argument("--generations", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true",
                    help="serving: launch each cycle directly instead of replaying a hipGraph "
                         "captured while membership is unchanged")
    ap.add_argument("--method", choices=["region", "woodcock"], default="region",
                    help="tracking of the event kernel")
    a = ap.parse_args()
    method = _abi.EVENT_REGION if a.method == "region" else _abi.EVENT_WOODCOCK

    def log(m):
        print(m, file=sys.stderr, flush=True)

    if a.batch:
        out = run_batch(a.batch, generations=a.generations or None, method=method)
        print(json.dumps({"form": "batched", **out}), flush=True)
        return
    idx = 0
    for rate in [float(x) for x in a.rates.split(",") if x]:
        out = run_serving(rate, a.seconds, first_index=idx, drain_s=a.drain,
                          use_graph=0 if a.no_graph else 1, log=log, method=method)
        idx += out["submitted"]
        print(json.dumps({"form": "serving", **out}), flush=True)


if __name__ == "__main__":
    main()
