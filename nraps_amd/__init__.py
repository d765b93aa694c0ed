"""MI355X-native Monte Carlo k-eigenvalue solver for 1D multigroup slabs.

Python access to the C ABI of include/nraps.h (libnraps.so, built in-tree by
``python -m nraps_amd.build``).  The library is the product; this module only
marshals numpy arrays through ctypes.  There is no CPU fallback: if the
library or a GPU is missing, calls fail with an error.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import *  # noqa: F401,F403  (constants)
from .problem import DiffusionBuffers, ResultBuffers, SlabProblem, diffusion_opts, run_opts

__all__ = ["lib", "solve", "solve_batch", "generate", "load_deck", "write_outputs",
           "pcg32_fill", "device_info", "validate", "Server", "SlabProblem", "run_opts",
           "ResultBuffers", "NrapsError", "diffusion_solve", "diffusion_solve_batch",
           "diffusion_opts", "DiffusionBuffers", "write_diffusion_outputs"]

HERE = os.path.dirname(os.path.abspath(__file__))
# NRAPS_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("NRAPS_LIB") or os.path.join(HERE, "libnraps.so")


class NrapsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"nraps status {status}: {msg}")
        self.status = status


_lib = None


def lib():
    """The loaded libnraps.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m nraps_amd.build`")
        L = C.CDLL(LIB_PATH)
        P, O, R = C.POINTER(_abi.Problem), C.POINTER(_abi.RunOpts), C.POINTER(_abi.Result)
        DO, DR = C.POINTER(_abi.DiffusionOpts), C.POINTER(_abi.DiffusionResult)
        sig = {
            "nraps_run_opts_default": (None, [O]),
            "nraps_mc_solve": (C.c_int, [P, O, R]),
            "nraps_mc_solve_batch": (C.c_int, [C.c_int32, C.POINTER(P), O, C.POINTER(R)]),
            "nraps_last_error": (C.c_char_p, []),
            "nraps_device_info": (C.c_int, [C.c_int32, C.c_char_p, C.c_int32]),
            "nraps_pcg32_fill": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, C.c_int32,
                                           C.POINTER(C.c_uint32)]),
            "nraps_deck_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p),
                                          C.POINTER(_abi.DeckInfo)]),
            "nraps_problem_generate": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64,
                                                 C.POINTER(C.c_void_p), O]),
            "nraps_problem_view": (P, [C.c_void_p]),
            "nraps_problem_free": (None, [C.c_void_p]),
            "nraps_validate": (C.c_int, [P, O]),
            "nraps_write_outputs": (C.c_int, [C.c_char_p, P, O, R]),
            "nraps_server_create": (C.c_int, [C.POINTER(_abi.ServerOpts), C.POINTER(C.c_void_p)]),
            "nraps_server_submit": (C.c_int, [C.c_void_p, P, O, C.POINTER(C.c_uint64)]),
            "nraps_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, R]),
            "nraps_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, R, C.c_int64]),
            "nraps_server_cancel": (C.c_int, [C.c_void_p, C.c_uint64]),
            "nraps_server_latency": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
            "nraps_server_forget": (C.c_int, [C.c_void_p, C.c_uint64]),
            "nraps_server_destroy": (None, [C.c_void_p]),
            "nraps_write_diffusion_outputs": (C.c_int, [C.c_char_p, P, C.c_int64, DR]),
            "nraps_diffusion_opts_default": (None, [DO]),
            "nraps_diffusion_solve": (C.c_int, [P, DO, DR]),
            "nraps_diffusion_solve_batch": (C.c_int, [C.c_int32, C.POINTER(P), DO,
                                                      C.POINTER(DR)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().nraps_last_error().decode()


def _check(status, ok=(0,)):
    if status not in ok:
        raise NrapsError(status, last_error())
    return status


def validate(problem, opts):
    return lib().nraps_validate(C.byref(problem.view()), C.byref(opts))


def solve(problem, opts, *, trace=False, check=True):
    """Run nraps_mc_solve; returns the ResultBuffers (numpy arrays + scalars)."""
    rb = ResultBuffers(problem, opts, trace=trace)
    st = lib().nraps_mc_solve(C.byref(problem.view()), C.byref(opts), C.byref(rb.c))
    rb.status_code = st
    if check:
        _check(st)
    return rb


def solve_batch(problems, opts_list, *, check=True):
    n = len(problems)
    views = [p.view() for p in problems]
    rbs = [ResultBuffers(p, o) for p, o in zip(problems, opts_list)]
    PA = C.POINTER(_abi.Problem) * n
    RA = C.POINTER(_abi.Result) * n
    OA = _abi.RunOpts * n
    pa = PA(*[C.pointer(v) for v in views])
    ra = RA(*[C.pointer(r.c) for r in rbs])
    oa = OA(*opts_list)
    st = lib().nraps_mc_solve_batch(n, pa, oa, ra)
    for rb in rbs:  # each problem's own status, as solve() reports it
        rb.status_code = rb.c.status
    if check:
        _check(st)
    return rbs


def diffusion_solve(problem, opts=None, *, check=True):
    """Run nraps_diffusion_solve; returns the DiffusionBuffers."""
    opts = opts if opts is not None else diffusion_opts()
    rb = DiffusionBuffers(problem)
    st = lib().nraps_diffusion_solve(C.byref(problem.view()), C.byref(opts), C.byref(rb.c))
    rb.status_code = st
    if check:
        _check(st)
    return rb


def diffusion_solve_batch(problems, opts=None, *, check=True):
    opts = opts if opts is not None else diffusion_opts()
    n = len(problems)
    views = [p.view() for p in problems]
    rbs = [DiffusionBuffers(p) for p in problems]
    pa = (C.POINTER(_abi.Problem) * n)(*[C.pointer(v) for v in views])
    ra = (C.POINTER(_abi.DiffusionResult) * n)(*[C.pointer(r.c) for r in rbs])
    st = lib().nraps_diffusion_solve_batch(n, pa, C.byref(opts), ra)
    if check:
        _check(st)
    return rbs


def generate(config, index=0, seed=42):
    """Synthetic workload `config` (0..4); returns (SlabProblem, RunOpts)."""
    h = C.c_void_p()
    o = _abi.RunOpts()
    _check(lib().nraps_problem_generate(config, index, seed, C.byref(h), C.byref(o)))
    try:
        p = SlabProblem.from_view(lib().nraps_problem_view(h).contents)
    finally:
        lib().nraps_problem_free(h)
    return p, o


def load_deck(path):
    """Reference input deck -> (SlabProblem, DeckInfo)."""
    h = C.c_void_p()
    info = _abi.DeckInfo()
    _check(lib().nraps_deck_load(os.fsencode(path), C.byref(h), C.byref(info)))
    try:
        p = SlabProblem.from_view(lib().nraps_problem_view(h).contents)
    finally:
        lib().nraps_problem_free(h)
    return p, info


def write_outputs(directory, problem, opts, rb):
    _check(lib().nraps_write_outputs(os.fsencode(directory), C.byref(problem.view()),
                                     C.byref(opts), C.byref(rb.c)))


def write_diffusion_outputs(directory, problem, generations, db):
    _check(lib().nraps_write_diffusion_outputs(os.fsencode(directory), C.byref(problem.view()),
                                               int(generations), C.byref(db.c)))


def pcg32_fill(seed, stream0, n_streams, draws):
    out = np.zeros(n_streams * draws, dtype=np.uint32)
    _check(lib().nraps_pcg32_fill(seed, stream0, n_streams, draws,
                                  out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out.reshape(n_streams, draws)


def device_info(device=0):
    buf = C.create_string_buffer(256)
    st = lib().nraps_device_info(device, buf, 256)
    return st, buf.value.decode()


class Server:
    """Continuous-batching request loop (nraps_server_*)."""

    def __init__(self, *, precision=_abi.FP32, budget_bytes=200 * 10**9, max_slots=256,
                 min_active=10, sigma_k_target=0.0, use_graph=0, device=0,
                 method=_abi.EVENT_REGION):
        o = _abi.ServerOpts()
        o.device = device
        o.precision = precision
        o.budget_bytes = int(budget_bytes)
        o.max_slots = max_slots
        o.min_active = min_active
        o.sigma_k_target = sigma_k_target
        o.use_graph = use_graph
        o.method = method
        self._h = C.c_void_p()
        _check(lib().nraps_server_create(C.byref(o), C.byref(self._h)))
        self._keep = {}

    def submit(self, problem, opts):
        t = C.c_uint64()
        _check(lib().nraps_server_submit(self._h, C.byref(problem.view()), C.byref(opts),
                                         C.byref(t)))
        self._keep[t.value] = (problem, opts)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        problem, opts = self._keep[ticket]
        rb = ResultBuffers(problem, opts)
        st = lib().nraps_server_wait(self._h, ticket, C.byref(rb.c), timeout_ms)
        rb.status_code = st
        return rb

    def poll(self, ticket):
        problem, opts = self._keep[ticket]
        rb = ResultBuffers(problem, opts)
        st = lib().nraps_server_poll(self._h, ticket, C.byref(rb.c))
        rb.status_code = st
        return rb

    def cancel(self, ticket):
        return lib().nraps_server_cancel(self._h, ticket)

    def latency(self, ticket):
        q, t, c = C.c_double(), C.c_double(), C.c_int32()
        st = lib().nraps_server_latency(self._h, ticket, C.byref(q), C.byref(t), C.byref(c))
        return st, q.value, t.value, c.value

    def forget(self, ticket):
        """Drop a finished request's record (and the arrays kept alive for it)."""
        st = lib().nraps_server_forget(self._h, ticket)
        if st == 0:
            self._keep.pop(ticket, None)
        return st

    def close(self):
        if self._h:
            lib().nraps_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
