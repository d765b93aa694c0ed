"""ctypes mirror of include/nraps.h (structs, constants, function signatures).

The C ABI in include/nraps.h is the product boundary; this module only
describes its memory layout so Python callers (tests, bench) can pass plain
pointers through it.
"""
import ctypes as C

NRAPS_OK = 0
NRAPS_PENDING = 1
NRAPS_E_INVALID = -1
NRAPS_E_NOMEM = -2
NRAPS_E_DEVICE = -3
NRAPS_E_OVER_BUDGET = -4
NRAPS_E_CANCELLED = -5
NRAPS_E_EVENT_CAP = -6
NRAPS_E_NO_FISSION = -7
NRAPS_E_BANK_OVERFLOW = -8
NRAPS_E_NOT_FOUND = -9
NRAPS_E_IO = -10
NRAPS_E_NOT_CONVERGED = -11

FP64, FP32 = 0, 1
HISTORY_SURFACE, EVENT_WOODCOCK, EVENT_REGION = 0, 1, 2
SOURCE_UNIFORM_FUEL, SOURCE_FISSION_BANK = 0, 1
XS_ON_MATERIAL_CHANGE, XS_EVERY_COLLISION = 0, 1
BANK_INIT_UNIFORM_FUEL, BANK_INIT_DIFFUSION = 0, 1

EV_BIRTH, EV_ABSORB, EV_SCATTER, EV_LEAK, EV_CAP = 0, 1, 2, 3, 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)


class Problem(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int32), ("n_groups", C.c_int32), ("n_mats", C.c_int32),
        ("numass", C.c_int32), ("boundl", C.c_double), ("boundr", C.c_double),
        ("matid", _ip), ("delta_x", _dp), ("mesh_left", _dp), ("mesh_right", _dp),
        ("n_fuel", C.c_int32), ("fuel_indices", _ip),
        ("sigt", _dp), ("sigs", _dp), ("mu", _dp), ("siga", _dp), ("sigf", _dp),
        ("nut", _dp), ("chit", _dp), ("scat_matrix", _dp), ("inv_sigtr", _dp),
    ]


class RunOpts(C.Structure):
    _fields_ = [
        ("histories", C.c_int64), ("generations", C.c_int32), ("skip", C.c_int32),
        ("seed", C.c_uint64), ("problem_id", C.c_uint64),
        ("precision", C.c_int32), ("method", C.c_int32), ("source", C.c_int32),
        ("xs_refresh", C.c_int32), ("k_initial", C.c_double), ("max_events", C.c_int64),
        ("device", C.c_int32), ("use_graph", C.c_int32), ("chunk", C.c_int64),
        ("trace_histories", C.c_int64), ("trace_events", C.c_int32),
        ("trace_cycle", C.c_int32), ("bank_init", C.c_int32), ("reserved", C.c_int32 * 7),
    ]


class TraceEvent(C.Structure):
    _fields_ = [("type", C.c_int32), ("cell", C.c_int32), ("group", C.c_int32),
                ("aux", C.c_int32), ("x", C.c_double), ("mu", C.c_double)]


TRACE_DTYPE = [("type", "<i4"), ("cell", "<i4"), ("group", "<i4"), ("aux", "<i4"),
               ("x", "<f8"), ("mu", "<f8")]


class Result(C.Structure):
    _fields_ = [
        ("flux", _dp), ("assembly_average", _dp), ("fission_source", _dp), ("k", _dp),
        ("k_fund", _dp), ("flux_mean", _dp), ("flux_sigma", _dp), ("tally_last", _dp),
        ("k_ref", _dp), ("sites", _lp),
        ("trace", C.POINTER(TraceEvent)), ("trace_len", _ip), ("draws", _lp), ("kacc", _dp),
        ("k_mean", C.c_double), ("k_sigma", C.c_double), ("histories_run", C.c_int64),
        ("collisions", C.c_int64), ("virtual_collisions", C.c_int64),
        ("crossings", C.c_int64), ("absorptions", C.c_int64), ("leaks", C.c_int64),
        ("capped", C.c_int64), ("mean_flights", C.c_double), ("ms_total", C.c_double),
        ("ms_device", C.c_double), ("ms_transport", C.c_double),
        ("transport_launches", C.c_int64), ("status", C.c_int32),
        ("reserved", C.c_int32 * 7),
    ]


class DeckInfo(C.Structure):
    _fields_ = [
        ("analk", C.c_int32), ("mattypes", C.c_int32), ("energygroups", C.c_int32),
        ("generations", C.c_int64), ("histories", C.c_int64), ("skip", C.c_int64),
        ("numass", C.c_int32), ("numrods", C.c_int32), ("roddia", C.c_double),
        ("rodpitch", C.c_double), ("mpfr", C.c_int64), ("mpwr", C.c_int64),
        ("boundl", C.c_double), ("boundr", C.c_double), ("delta_fuel", C.c_double),
        ("delta_water", C.c_double), ("solution_mc", C.c_int32), ("solver", C.c_int32),
    ]


class ServerOpts(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("precision", C.c_int32), ("budget_bytes", C.c_int64),
        ("max_slots", C.c_int32), ("min_active", C.c_int32),
        ("sigma_k_target", C.c_double), ("use_graph", C.c_int32),
        ("method", C.c_int32), ("reserved", C.c_int32 * 6),
    ]


class DiffusionOpts(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("max_iterations", C.c_int32), ("flux_tol", C.c_double),
        ("k_tol", C.c_double), ("reserved", C.c_int32 * 8),
    ]


class DiffusionResult(C.Structure):
    _fields_ = [
        ("flux", _dp), ("assembly_average", _dp), ("fission_source", _dp),
        ("k", C.c_double), ("delta_flux", C.c_double), ("delta_k", C.c_double),
        ("iterations", C.c_int32), ("status", C.c_int32), ("ms_total", C.c_double),
        ("ms_device", C.c_double), ("reserved", C.c_int32 * 8),
    ]
