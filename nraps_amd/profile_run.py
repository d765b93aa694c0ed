"""Profiling driver: run one workload for a few generations, then a
device-to-device copy of known size (a blit kernel) whose FETCH_SIZE /
WRITE_SIZE counters calibrate the HBM-byte counters of the run.

    rocprofv3 --kernel-trace --stats -d OUT -o run -f csv -- \\
        python3 -m nraps_amd.profile_run --config 1 --generations 4
    rocprofv3 --pmc FETCH_SIZE --kernel-trace -d OUT -o fetch -f csv -- \\
        python3 -m nraps_amd.profile_run --config 1 --generations 2 --copy-mib 1024
"""
import argparse
import ctypes as C
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import nraps_amd as nr  # noqa: E402
from nraps_amd import _abi  # noqa: E402


def copy_probe(mib):
    """hipMemcpy device-to-device of `mib` MiB (runs as a blit kernel)."""
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--index", type=int, default=0)
    ap.add_argument("--generations", type=int, default=4)
    ap.add_argument("--histories", type=int, default=0)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default=None)
    ap.add_argument("--method", choices=["event", "region", "plain"], default=None,
                    help="event: Woodcock event kernel; region: region surface-tracking "
                         "event kernel; plain: FP64 one-thread-per-history port")
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0, help="histories per wave claim (0: automatic)")
    ap.add_argument("--bank", action="store_true", help="fission-bank source (skip >= 2)")
    ap.add_argument("--diffusion", type=int, default=0,
                    help="instead: diffusion solve of this many problems of the config "
                         "(the uneven batch for config 3/4), --generations times")
    a = ap.parse_args()
    if a.diffusion:
        probs = [nr.generate(a.config, a.index + i)[0] for i in range(a.diffusion)] \
            if a.config >= 3 else [nr.generate(a.config, a.index)[0]] * a.diffusion
        for _ in range(max(a.generations, 1)):
            rs = nr.diffusion_solve_batch(probs)
        its = [r.iterations for r in rs]
        print(f"config {a.config}: diffusion of {len(probs)} problem(s), iterations "
              f"{min(its)}..{max(its)}, device {rs[0].ms_device:.3f} ms, k[0] {rs[0].k:.8f}",
              flush=True)
        return
    p, o = nr.generate(a.config, a.index)
    if a.bank:
        o.source = _abi.SOURCE_FISSION_BANK
    fission = o.source == _abi.SOURCE_FISSION_BANK
    o.generations = max(a.generations, 3 if fission else 2)
    o.skip = 2 if fission else 1
    o.use_graph = 0
    if a.histories:
        o.histories = a.histories
    if a.chunk:
        o.chunk = a.chunk
    if a.precision:
        o.precision = _abi.FP32 if a.precision == "fp32" else _abi.FP64
    if a.method:
        o.method = {"event": _abi.EVENT_WOODCOCK, "region": _abi.EVENT_REGION,
                    "plain": _abi.HISTORY_SURFACE}[a.method]
        if o.method == _abi.HISTORY_SURFACE:
            o.precision = _abi.FP64
    r = nr.solve(p, o)
    print(f"config {a.config}: {o.histories} x {o.generations} histories, "
          f"transport {r.ms_transport / max(r.transport_launches, 1):.3f} ms/launch, "
          f"device {r.ms_device / o.generations:.3f} ms/generation, "
          f"k {r.k_mean:.6f}", flush=True)
    if a.copy_mib:
        copy_probe(a.copy_mib)
        print(f"copy probe: {a.copy_mib} MiB device-to-device", flush=True)


if __name__ == "__main__":
    main()
