"""Host-side problem container: numpy arrays behind an nraps_problem view.

A SlabProblem owns the arrays of one 1D multigroup slab problem in the
reference layout (XS [g][mat] flattened as mat + n_mats*g, scatter matrix
[mat][from][to]) and hands out the ctypes struct that the C ABI takes.
"""
import ctypes as C

import numpy as np

from . import _abi


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class SlabProblem:
    """Arrays of one problem plus its nraps_problem view (kept alive here)."""

    FIELDS_XS = ("sigt", "sigs", "mu", "siga", "sigf", "nut", "chit")

    def __init__(self, *, matid, delta_x, mesh_left, mesh_right, n_groups, n_mats,
                 sigt, sigs, mu, siga, sigf, nut, chit, scat_matrix, boundl, boundr,
                 numass=1, fuel_indices=None, inv_sigtr=None):
        self.matid = _i32(matid)
        self.delta_x = _f64(delta_x)
        self.mesh_left = _f64(mesh_left)
        self.mesh_right = _f64(mesh_right)
        self.n_groups = int(n_groups)
        self.n_mats = int(n_mats)
        for name in self.FIELDS_XS:
            setattr(self, name, _f64(locals()[name]))
        self.scat_matrix = _f64(scat_matrix)
        self.inv_sigtr = None if inv_sigtr is None else _f64(inv_sigtr)
        self.boundl = float(boundl)
        self.boundr = float(boundr)
        self.numass = int(numass)
        self.fuel_indices = None if fuel_indices is None else _i32(fuel_indices)
        self._view = None

    @property
    def n_cells(self):
        return int(self.matid.shape[0])

    def view(self):
        """ctypes nraps_problem pointing into this object's arrays."""
        p = _abi.Problem()
        p.n_cells = self.n_cells
        p.n_groups = self.n_groups
        p.n_mats = self.n_mats
        p.numass = self.numass
        p.boundl = self.boundl
        p.boundr = self.boundr
        p.matid = self.matid.ctypes.data_as(C.POINTER(C.c_int32))
        for name in ("delta_x", "mesh_left", "mesh_right", "scat_matrix") + self.FIELDS_XS:
            setattr(p, name, getattr(self, name).ctypes.data_as(C.POINTER(C.c_double)))
        if self.fuel_indices is not None:
            p.n_fuel = int(self.fuel_indices.shape[0])
            p.fuel_indices = self.fuel_indices.ctypes.data_as(C.POINTER(C.c_int32))
        if self.inv_sigtr is not None:
            p.inv_sigtr = self.inv_sigtr.ctypes.data_as(C.POINTER(C.c_double))
        self._view = p
        return p

    @classmethod
    def from_view(cls, v):
        """Copy a problem out of a C-owned nraps_problem view."""
        nc, G, M = v.n_cells, v.n_groups, v.n_mats

        def arr(ptr, n):
            return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

        kw = dict(matid=arr(v.matid, nc), delta_x=arr(v.delta_x, nc),
                  mesh_left=arr(v.mesh_left, nc), mesh_right=arr(v.mesh_right, nc),
                  n_groups=G, n_mats=M, scat_matrix=arr(v.scat_matrix, M * G * G),
                  boundl=v.boundl, boundr=v.boundr, numass=v.numass)
        for name in cls.FIELDS_XS:
            kw[name] = arr(getattr(v, name), G * M)
        if v.fuel_indices:
            kw["fuel_indices"] = arr(v.fuel_indices, v.n_fuel)
        if v.inv_sigtr:
            kw["inv_sigtr"] = arr(v.inv_sigtr, G * M)
        return cls(**kw)

    @classmethod
    def from_regions(cls, widths, mats, cells_per_region, *, n_groups, n_mats, xs,
                     boundl, boundr, numass=1):
        """Mesh from region widths/materials, uniform cells inside a region.

        Edges are running double sums, so mesh_right[c] == mesh_left[c+1]."""
        matid, dxs, left, right = [], [], [], []
        x = 0.0
        for w, m, n in zip(widths, mats, cells_per_region):
            d = float(w) / int(n)
            for _ in range(int(n)):
                matid.append(m)
                dxs.append(d)
                left.append(x)
                x = x + d
                right.append(x)
        return cls(matid=matid, delta_x=dxs, mesh_left=left, mesh_right=right,
                   n_groups=n_groups, n_mats=n_mats, boundl=boundl, boundr=boundr,
                   numass=numass, **xs)


def run_opts(**kw):
    """nraps_run_opts with the reference defaults, overridden by kw."""
    o = _abi.RunOpts()
    o.histories = 10000
    o.generations = 150
    o.skip = 50
    o.seed = 42
    o.problem_id = 0
    o.precision = _abi.FP64
    o.method = _abi.HISTORY_SURFACE
    o.source = _abi.SOURCE_UNIFORM_FUEL
    o.xs_refresh = _abi.XS_ON_MATERIAL_CHANGE
    o.k_initial = 1.0
    o.max_events = 0
    o.device = 0
    o.use_graph = 1
    o.chunk = 0
    o.trace_histories = 0
    o.trace_events = 0
    o.trace_cycle = 0
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class ResultBuffers:
    """Caller-allocated result arrays for one solve, plus the ctypes view."""

    def __init__(self, problem, opts, *, trace=False):
        G, nc, gens = problem.n_groups, problem.n_cells, int(opts.generations)
        self.flux = np.zeros(G * nc)
        self.assembly_average = np.zeros(G * nc)
        self.fission_source = np.zeros(nc)
        self.k = np.zeros(gens)
        self.k_fund = np.zeros(gens)
        self.flux_mean = np.zeros(G * nc)
        self.flux_sigma = np.zeros(G * nc)
        self.tally_last = np.zeros(G * nc)
        self.k_ref = np.zeros(gens)
        self.sites = np.zeros(gens, dtype=np.int64)
        nt = int(opts.trace_histories) if trace else 0
        ne = int(opts.trace_events) if trace else 0
        self.trace = np.zeros(max(nt * ne, 1), dtype=_abi.TRACE_DTYPE)
        self.trace_len = np.zeros(max(nt, 1), dtype=np.int32)
        self.draws = np.zeros(max(nt, 1), dtype=np.int64)
        self.kacc = np.zeros(max(nt, 1))
        self.traced = trace
        r = _abi.Result()
        dp = C.POINTER(C.c_double)
        for name in ("flux", "assembly_average", "fission_source", "k", "k_fund",
                     "flux_mean", "flux_sigma", "tally_last", "k_ref"):
            setattr(r, name, getattr(self, name).ctypes.data_as(dp))
        r.sites = self.sites.ctypes.data_as(C.POINTER(C.c_int64))
        if trace:
            r.trace = self.trace.ctypes.data_as(C.POINTER(_abi.TraceEvent))
            r.trace_len = self.trace_len.ctypes.data_as(C.POINTER(C.c_int32))
            r.draws = self.draws.ctypes.data_as(C.POINTER(C.c_int64))
            r.kacc = self.kacc.ctypes.data_as(dp)
        self.c = r

    def __getattr__(self, name):
        # scalar outputs live in the ctypes struct
        return getattr(self.__dict__["c"], name)

    def trace_of(self, p, events):
        """Recorded events of history p (numpy structured array)."""
        n = min(int(self.trace_len[p]), events)
        return self.trace[p * events:p * events + n]


def diffusion_opts(**kw):
    """nraps_diffusion_opts with the reference tolerances, updated from kw."""
    o = _abi.DiffusionOpts()
    o.flux_tol = 1e-5
    o.k_tol = 1e-6
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class DiffusionBuffers:
    """Caller-allocated arrays of one diffusion solve, plus the ctypes view."""

    def __init__(self, problem):
        G, nc = problem.n_groups, problem.n_cells
        self.flux = np.zeros(G * nc)
        self.assembly_average = np.zeros(G * nc)
        self.fission_source = np.zeros(nc)
        r = _abi.DiffusionResult()
        dp = C.POINTER(C.c_double)
        for name in ("flux", "assembly_average", "fission_source"):
            setattr(r, name, getattr(self, name).ctypes.data_as(dp))
        self.c = r

    def __getattr__(self, name):
        return getattr(self.__dict__["c"], name)
