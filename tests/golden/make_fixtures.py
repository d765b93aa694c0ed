"""Write tests/golden/workloads.json: the seeded synthetic workloads, pinned.

For each workload the file holds the generated arrays (so a change of the
generator shows up as a diff) and, for the 9-region and 201-region slabs,
the mean number of flights per history F (real collisions + leaks)
measured by the FP64 CPU oracle at 10^6 histories with reference semantics.
F enters the algorithmic-bytes model of the roofline:
    B_hist = B_state * (1 + 2 F) [+ 2 B_site for the fission bank].

    python tests/golden/make_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import nraps_amd as nr  # noqa: E402
import oracle_lib  # noqa: E402
from nraps_amd import _abi  # noqa: E402


def pin(p, o):
    return {
        "n_cells": p.n_cells, "n_groups": p.n_groups, "n_mats": p.n_mats,
        "boundl": p.boundl, "boundr": p.boundr,
        "histories": int(o.histories), "generations": int(o.generations), "skip": int(o.skip),
        "method": int(o.method), "precision": int(o.precision), "source": int(o.source),
        "problem_id": int(o.problem_id),
        "matid": p.matid.tolist(), "mesh_left": p.mesh_left.tolist(),
        "mesh_right": p.mesh_right.tolist(), "delta_x": p.delta_x.tolist(),
        **{k: getattr(p, k).tolist() for k in SlabXS},
        "scat_matrix": p.scat_matrix.tolist(),
    }


SlabXS = ("sigt", "sigs", "mu", "siga", "sigf", "nut", "chit")


def mean_flights(p, histories=10**6, source=_abi.SOURCE_UNIFORM_FUEL):
    gens = 4 if source == _abi.SOURCE_UNIFORM_FUEL else 6
    o = nr.run_opts(histories=histories // (gens if source == _abi.SOURCE_UNIFORM_FUEL else 1),
                    generations=gens, skip=2, source=source)
    r = oracle_lib.solve(p, o)
    assert r.status_code == 0
    return r.mean_flights, r.histories_run


def main():
    out = {"seed": 42, "workloads": {}}
    p0, o0 = nr.generate(0)
    f0, n0 = mean_flights(p0)
    out["workloads"]["slab9_g4"] = pin(p0, o0)
    _, o1 = nr.generate(1)
    out["workloads"]["slab9_g4_1e7"] = {"same_problem_as": "slab9_g4",
                                        "histories": int(o1.histories),
                                        "generations": int(o1.generations), "skip": int(o1.skip),
                                        "mean_flights": f0, "mean_flights_histories": n0}
    p2, o2 = nr.generate(2)
    w2 = pin(p2, o2)
    f2, n2 = mean_flights(p2, histories=10**5, source=_abi.SOURCE_FISSION_BANK)
    w2["mean_flights"] = f2
    w2["mean_flights_histories"] = n2
    out["workloads"]["slab201_g8"] = w2
    for i in range(4):
        p3, o3 = nr.generate(3, i)
        out["workloads"][f"batch_{i}"] = pin(p3, o3)
    path = os.path.join(HERE, "workloads.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(path, f"F(9-region slab) = {f0:.6f}", f"F(201-region slab) = {f2:.6f}")


if __name__ == "__main__":
    main()
