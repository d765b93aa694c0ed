"""Precision-form comparison on the 201-region 8-group fission-bank slab.

Runs the workload twice on the GPU with the same seed: FP32 tracking and
tallies (FP64 k, bank and variance sums) and the FP64 event path. Both use
the workload's event tracking (region-by-region surface tracking) and the
fission-site bank, and the histories follow the same PCG32 streams. It reports the k of each (mean ± standard error over the
active generations), their difference in combined standard errors, the
relative differences of the eight group-total fluxes, and per-bin agreement
of the active-mean flux (fraction of bins within 3 combined standard errors,
largest relative difference among bins holding at least 1e-3 of their
group's peak).

    python -m nraps_amd.precision_run [--histories 100000000] [--generations 120]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import nraps_amd as nr  # noqa: E402
from nraps_amd import _abi  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--histories", type=int, default=0)
    ap.add_argument("--generations", type=int, default=0)
    ap.add_argument("--skip", type=int, default=0)
    ap.add_argument("--seed", type=int, default=42)
    a = ap.parse_args()
    p, o = nr.generate(2)
    if a.histories:
        o.histories = a.histories
    if a.generations:
        o.generations = a.generations
    if a.skip:
        o.skip = a.skip
    o.seed = a.seed
    o.use_graph = 1
    out = {"workload": "201-region 8-group upscatter slab, fission bank", "histories": int(o.histories),
           "generations": int(o.generations), "skip": int(o.skip), "seed": a.seed}
    res = {}
    for name, prec in (("fp32", _abi.FP32), ("fp64", _abi.FP64)):
        o.precision = prec
        t = time.perf_counter()
        r = nr.solve(p, o)
        wall = time.perf_counter() - t
        res[name] = r
        out[name] = {"k": r.k_mean, "k_sigma": r.k_sigma, "wall_s": wall,
                     "histories_per_s": int(o.histories) * int(o.generations) / wall}
        print(f"{name}: k = {r.k_mean:.7f} +- {r.k_sigma:.7f}, {wall:.1f} s", file=sys.stderr,
              flush=True)
    a32, a64 = res["fp32"], res["fp64"]
    G = p.n_groups
    out["k_delta_over_sigma"] = abs(a32.k_mean - a64.k_mean) / math.hypot(a32.k_sigma, a64.k_sigma)
    f32, f64 = a32.flux_mean.reshape(G, -1), a64.flux_mean.reshape(G, -1)
    s32, s64 = a32.flux_sigma.reshape(G, -1), a64.flux_sigma.reshape(G, -1)
    g32, g64 = f32.sum(1), f64.sum(1)
    out["group_flux_rel_diff"] = [float(x) for x in (g32 / g64 - 1)]
    sig = np.hypot(s32, s64)
    ok = np.abs(f32 - f64) <= 3 * sig
    out["bins_within_3sigma"] = float(ok.mean())
    big = f64 >= 1e-3 * f64.max(axis=1, keepdims=True)
    out["max_bin_rel_diff"] = float(np.max(np.abs(f32[big] / f64[big] - 1)))
    out["median_bin_rel_sigma"] = float(np.median(s64[big] / f64[big]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
