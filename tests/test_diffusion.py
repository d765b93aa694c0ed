"""Multigroup diffusion (src/discrete.rs:169-356): the oracle against what the
mathematics and the reference's decks fix, then the device solver and the
diffusion-started fission bank against the oracle.

Oracle pins (no GPU):
- TestCaseA.txt asks for the analytic k (`Analk = 1`): with reflective faces,
  no absorption in group 1 or in water, every neutron is absorbed in fuel
  group 2, so k = nu Sigma_f / Sigma_a = 1.4 * 0.18 / 0.2 = 1.26;
- a reflective homogeneous slab with a symmetric but non-uniform mesh keeps a
  flat flux: k = nuSigf . (diag(Sigma_t - S_gg) - S_off^T)^-1 chi, with
  upscatter so a transposed scatter operand fails;
- a one-group bare slab on a uniform mesh is a tridiagonal Toeplitz
  eigenproblem with corner entries 3 - 2 beta: k = nuSigf dx / (dx Sigma_r +
  (D/dx) mu_min), mu_min from scipy's tridiagonal eigensolver;
- a mirror-symmetric slab gives a mirror-symmetric flux;
- on a uniform mesh the power normalisation makes
  sum nuSigf phi dx = 3565e6 / (1.6022e-13 * 200);
- the bank start puts source p at the quantile (p + 1/2)/N of hand-computed
  weights.
Transcribed without an independent pin: the beta expression for
0 < bound < 1 (src/discrete.rs:26-33, :81-88) and the last row's coupling
of cells 0 and 1 on meshes whose ends differ (:93-99).
"""
import numpy as np
import pytest
from scipy import linalg

import nraps_amd as nr
from nraps_amd import _abi
from nraps_amd.problem import SlabProblem, diffusion_opts
from test_oracle_physics import homogeneous, k_infinite

REF = "tests/golden/reference"
TIGHT = dict(flux_tol=1e-13, k_tol=1e-14)


def _deck(name):
    import os
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    return nr.load_deck(os.path.join(here, REF, name + ".txt"))[0]


def k_diffusion_infinite(x):
    """Flat-flux k of the diffusion equations: removal Sigma_t - S_gg,
    in-scatter S[e->g] for e != g."""
    S = x["scat"]
    A = np.diag(x["sigt"] - np.diag(S)) - (S - np.diag(np.diag(S))).T
    return float(x["nsf"] @ np.linalg.solve(A, x["chi"]))


def symmetric_mesh_problem(G=4, upscatter=True, bl=1.0, br=1.0):
    """Homogeneous material on a palindromic non-uniform mesh."""
    p, x = homogeneous(G=G, upscatter=upscatter, cells=10, bl=bl, br=br)
    half = np.array([0.31, 0.12, 0.27, 0.05, 0.44, 0.18, 0.09, 0.36, 0.21, 0.14])
    dx = np.concatenate([half, half[::-1]])
    left = np.concatenate([[0.0], np.cumsum(dx)[:-1]])
    right = left + dx
    right[:-1] = left[1:]
    q = SlabProblem(matid=np.zeros(len(dx), np.int32), delta_x=dx, mesh_left=left,
                    mesh_right=right, n_groups=G, n_mats=1, sigt=p.sigt, sigs=p.sigs, mu=p.mu,
                    siga=p.siga, sigf=p.sigf, nut=p.nut, chit=p.chit, scat_matrix=p.scat_matrix,
                    boundl=bl, boundr=br, numass=2)
    return q, x


def one_group_slab(n=60, dx=0.25, bound=0.0, sigt=0.8, c=0.7, nsf=0.3):
    xs = dict(sigt=[sigt], sigs=[sigt * c], mu=[0.0], siga=[sigt * (1 - c)], sigf=[nsf / 2.5],
              nut=[2.5], chit=[1.0], scat_matrix=[sigt * c])
    return SlabProblem.from_regions([n * dx], [0], [n], n_groups=1, n_mats=1, xs=xs,
                                    boundl=bound, boundr=bound)


def random_slab(seed, G=4, regions=9, cells=7, bl=1.0, br=0.0, symmetric=False):
    rng = np.random.default_rng(seed)
    M = 3
    sigt = rng.uniform(0.3, 1.2, (G, M))
    c = rng.uniform(0.5, 0.95, (G, M))
    sigs = sigt * c
    scat = np.zeros((M, G, G))
    for m in range(M):
        for g in range(G):
            w = rng.uniform(0.1, 1.0, G)
            w[: max(g - 1, 0)] = 0.0
            scat[m, g] = sigs[g, m] * w / w.sum()
    nsf = np.zeros((G, M))
    nsf[G // 2:, :2] = rng.uniform(0.02, 0.2, (G - G // 2, 2))
    chi = np.zeros((G, M))
    chi[0, :] = 0.7
    chi[1 % G, :] += 0.3
    mats = [0 if r % 2 == 0 else 2 for r in range(regions)]
    widths = rng.uniform(0.5, 1.5, regions)
    if symmetric:
        widths = (widths + widths[::-1]) / 2
    xs = dict(sigt=sigt.ravel(), sigs=sigs.ravel(), mu=np.zeros(G * M), siga=(sigt - sigs).ravel(),
              sigf=(nsf / 2.5).ravel(), nut=np.full(G * M, 2.5), chit=chi.ravel(),
              scat_matrix=scat.ravel())
    return SlabProblem.from_regions(widths, mats, [cells] * regions, n_groups=G, n_mats=M, xs=xs,
                                    boundl=bl, boundr=br, numass=3)


# ---------------------------------------------------------------------------
# oracle pins (CPU)
# ---------------------------------------------------------------------------

def test_testcasea_analytic_k(oracle):
    r = oracle.diffusion(_deck("TestCaseA"))
    assert r.status_code == 0
    assert r.k == pytest.approx(1.4 * 0.18 / 0.2, rel=1e-9)


@pytest.mark.parametrize("upscatter", [False, True])
def test_reflective_homogeneous_closed_form(oracle, upscatter):
    p, x = symmetric_mesh_problem(upscatter=upscatter)
    r = oracle.diffusion(p, diffusion_opts(**TIGHT))
    assert r.status_code == 0
    assert r.k == pytest.approx(k_diffusion_infinite(x), rel=1e-11)
    # with mu = 0 and a consistent deck this is the transport k_infinite too
    assert r.k == pytest.approx(k_infinite(x), rel=1e-11)
    G, n = p.n_groups, p.n_cells
    flux = r.flux.reshape(G, n)
    assert np.allclose(flux, flux[:, :1], rtol=1e-10)


@pytest.mark.parametrize("bound,beta", [(0.0, 0.25), (1.0, 1.0)])
def test_one_group_bare_slab_eigenvalue(oracle, bound, beta):
    n, dx, sigt, c, nsf = 60, 0.25, 0.8, 0.7, 0.3
    p = one_group_slab(n, dx, bound, sigt, c, nsf)
    r = oracle.diffusion(p, diffusion_opts(**TIGHT))
    assert r.status_code == 0
    D = (1.0 / 3.0) / sigt / dx
    diag = np.full(n, 2.0)
    diag[0] = diag[-1] = 3.0 - 2.0 * beta
    mu_min = linalg.eigh_tridiagonal(diag, np.full(n - 1, -1.0), eigvals_only=True)[0]
    sig_r = sigt * (1 - c)
    k = nsf * dx / (dx * sig_r + D * mu_min)
    assert r.k == pytest.approx(k, rel=1e-10)


def test_oracle_matrix_rows(oracle):
    """Interior rows conserve: -sub - sup + removal = diag."""
    p = random_slab(3)
    for g in range(p.n_groups):
        sub, diag, sup = oracle.diffusion_matrix(p, g)
        M, G = p.n_mats, p.n_groups
        m = p.matid
        rem = p.delta_x * (p.sigt[m + M * g] - p.scat_matrix[G * G * m + (G + 1) * g])
        inner = slice(1, p.n_cells - 1)
        assert np.allclose(-sub[inner] - sup[inner] + rem[inner], diag[inner], rtol=1e-13)
        assert sub[0] == 0.0 and sup[-1] == 0.0
        assert np.all(sub[1:] < 0) and np.all(sup[:-1] < 0)
        assert np.allclose(sub[1:], sup[:-1], rtol=1e-14)  # symmetric coupling


def test_mirror_symmetric_slab(oracle):
    p = random_slab(11, bl=0.0, br=0.0, symmetric=True)
    r = oracle.diffusion(p, diffusion_opts(**TIGHT))
    assert r.status_code == 0
    flux = r.flux.reshape(p.n_groups, p.n_cells)
    assert np.allclose(flux, flux[:, ::-1], rtol=1e-9)


def test_power_normalisation_uniform_mesh(oracle):
    p = random_slab(5, regions=5, cells=8)
    p.delta_x[:] = p.delta_x[0]
    edges = np.arange(p.n_cells + 1) * p.delta_x[0]
    p.mesh_left[:] = edges[:-1]
    p.mesh_right[:] = edges[1:]
    r = oracle.diffusion(p)
    G, M, n = p.n_groups, p.n_mats, p.n_cells
    nsf = (p.nut * p.sigf)[p.matid[None, :] + M * np.arange(G)[:, None]]
    total = float((nsf * r.flux.reshape(G, n)).sum() * p.delta_x[0])
    assert total == pytest.approx(3565e6 / (1.6022e-13 * 200.0), rel=1e-12)
    assert np.allclose(r.fission_source, (nsf * r.flux.reshape(G, n)).sum(0), rtol=1e-14)


def test_assembly_average_blocks(oracle):
    p = random_slab(2, regions=9, cells=7)
    p.numass = 4
    r = oracle.diffusion(p)
    G, n = p.n_groups, p.n_cells
    per = n // 4
    f, a = r.flux.reshape(G, n), r.assembly_average.reshape(G, n)
    for b in range(4):
        blk = slice(b * per, (b + 1) * per)
        assert np.allclose(a[:, blk], f[:, blk].mean(axis=1, keepdims=True), rtol=1e-13)
    assert np.all(a[:, 4 * per:] == 0.0)


def test_iteration_limit_reports_not_converged(oracle):
    r = oracle.diffusion(random_slab(4), diffusion_opts(max_iterations=2, **TIGHT))
    assert r.status_code == _abi.NRAPS_E_NOT_CONVERGED and r.iterations == 2


def test_bank_start_quantiles(oracle):
    p = one_group_slab(n=3, dx=2.0)
    fs = np.array([1.0, 0.0, 3.0])
    st, x, c = oracle.diffusion_bank(p, fs, 8)
    assert st == 0
    t = (np.arange(8) + 0.5) / 8 * 8.0          # weights 2, 0, 6: total 8
    want_c = np.where(t < 2.0, 0, 2)
    want_x = np.where(t < 2.0, 0.0 + t / 2.0 * 2.0, 4.0 + (t - 2.0) / 6.0 * 2.0)
    assert np.array_equal(c, want_c)
    assert np.allclose(x, want_x, rtol=0, atol=1e-15)
    st, _, _ = oracle.diffusion_bank(p, np.zeros(3), 4)
    assert st == _abi.NRAPS_E_NO_FISSION


def test_diffusion_start_runs_on_the_oracle(oracle):
    """A diffusion-started bank gives the same k as the uniform start
    within statistics, from generation 0 on."""
    p, o = nr.generate(2)
    o.histories, o.generations, o.skip = 20000, 12, 4
    o.source = _abi.SOURCE_FISSION_BANK
    d = oracle.diffusion(p)
    st, x, c = oracle.diffusion_bank(p, d.fission_source, o.histories)
    assert st == 0
    o.bank_init = _abi.BANK_INIT_DIFFUSION
    r = oracle.solve_init(p, o, x, c, d.k)
    assert r.status_code == 0 and r.sites[0] > 0
    o2 = nr.run_opts(histories=20000, generations=12, skip=4, source=_abi.SOURCE_FISSION_BANK)
    r2 = oracle.solve(p, o2)
    assert abs(r.k_mean - r2.k_mean) < 4 * np.hypot(r.k_sigma, r2.k_sigma)


# ---------------------------------------------------------------------------
# device solver against the oracle
# ---------------------------------------------------------------------------

def _workloads():
    cases = [("9-region 4-group slab", lambda: nr.generate(0)[0]),
             ("201-region 8-group slab", lambda: nr.generate(2)[0])]
    cases += [(f"uneven batch problem {i}", (lambda i=i: nr.generate(3, i)[0])) for i in range(4)]
    cases += [(name, (lambda name=name: _deck(name))) for name in
              ("TestCaseA", "TestCaseB", "TestCaseC")]
    cases += [("random vacuum slab", lambda: random_slab(7, bl=0.0, br=0.0)),
              ("random albedo slab", lambda: random_slab(8, bl=0.5, br=0.8)),
              ("two cells", lambda: one_group_slab(n=2, bound=0.0)),
              ("three cells", lambda: one_group_slab(n=3, bound=0.3))]
    return cases


def _compare(d, o, p):
    assert d.status_code == o.status_code
    assert d.iterations == o.iterations
    assert d.k == pytest.approx(o.k, rel=1e-12)
    scale = np.abs(o.flux).max()
    assert np.abs(d.flux - o.flux).max() <= 1e-10 * scale
    assert np.abs(d.assembly_average - o.assembly_average).max() <= 1e-10 * scale
    fs = np.abs(o.fission_source).max()
    assert np.abs(d.fission_source - o.fission_source).max() <= 1e-10 * fs


@pytest.mark.gpu
@pytest.mark.parametrize("name,make", _workloads(), ids=[c[0] for c in _workloads()])
def test_device_diffusion_matches_oracle(gpu, oracle, name, make):
    p = make()
    print(f"diffusion parity: {name}, {p.n_cells} cells, {p.n_groups} groups", flush=True)
    o = oracle.diffusion(p)
    d = nr.diffusion_solve(p, check=False)
    _compare(d, o, p)


@pytest.mark.gpu
@pytest.mark.parametrize("G,regions,cells", [(8, 41, 60), (10, 61, 70)])
def test_device_diffusion_global_memory_form(gpu, oracle, monkeypatch, G, regions, cells):
    """Problems too large for LDS run with the flux in global memory: a
    2460-cell 8-group slab on the LDS/global kernel (forced) and a
    4270-cell 10-group slab, which takes that kernel by itself (more than 8
    groups, 17 cells per thread)."""
    p = random_slab(9, G=G, regions=regions, cells=cells, bl=0.0, br=1.0)
    assert (p.n_groups + 1) * p.n_cells * 8 > 160 * 1024
    if G <= 8:
        monkeypatch.setenv("NRAPS_DIFFUSION_GENERIC", "1")
    print(f"diffusion parity: global form, {p.n_cells} cells, {G} groups", flush=True)
    _compare(nr.diffusion_solve(p, check=False), oracle.diffusion(p), p)


@pytest.mark.gpu
def test_device_diffusion_batch_equals_single(gpu):
    probs = [nr.generate(3, i)[0] for i in range(6)] + [_deck("TestCaseB")]
    batch = nr.diffusion_solve_batch(probs)
    for p, b in zip(probs, batch):
        s = nr.diffusion_solve(p)
        assert b.k == s.k and b.iterations == s.iterations
        assert np.array_equal(b.flux, s.flux)


@pytest.mark.gpu
def test_register_and_lds_forms_identical(gpu, monkeypatch):
    """The register-resident kernel (G, cells per thread <= 8) and the
    LDS/global-memory kernel do the same arithmetic in the same order."""
    probs = [nr.generate(0)[0], nr.generate(2)[0], _deck("TestCaseC"),
             random_slab(8, bl=0.5, br=0.8), one_group_slab(n=3, bound=0.3)]
    monkeypatch.setenv("NRAPS_DIFFUSION_GENERIC", "1")
    lds = nr.diffusion_solve_batch(probs)
    monkeypatch.delenv("NRAPS_DIFFUSION_GENERIC")
    regs = nr.diffusion_solve_batch(probs)
    for a, b in zip(lds, regs):
        assert a.k == b.k and a.iterations == b.iterations and a.delta_flux == b.delta_flux
        assert np.array_equal(a.flux, b.flux) and np.array_equal(a.fission_source, b.fission_source)


@pytest.mark.gpu
def test_device_diffusion_iteration_limit(gpu, oracle):
    p = random_slab(4)
    o = diffusion_opts(max_iterations=3, **TIGHT)
    d = nr.diffusion_solve(p, o, check=False)
    assert d.status_code == _abi.NRAPS_E_NOT_CONVERGED and d.iterations == 3
    _compare(d, oracle.diffusion(p, o), p)


def test_diffusion_rejects_bad_input():
    p = one_group_slab(n=1)
    with pytest.raises(nr.NrapsError) as e:
        nr.diffusion_solve(p)
    assert e.value.status == _abi.NRAPS_E_INVALID


@pytest.mark.gpu
@pytest.mark.parametrize("method", [_abi.HISTORY_SURFACE, _abi.EVENT_WOODCOCK,
                                    _abi.EVENT_REGION])
def test_diffusion_started_bank_bit_exact(gpu, oracle, method):
    """FP64 Monte Carlo from a diffusion-started bank, against the oracle run
    from the same diffusion source: every generation's k (a canonical-order
    sum) and site count bit for bit, the event counts exactly, and the
    tallies (summed in atomic order on the device) to rounding."""
    p, o = nr.generate(2)
    o.histories, o.generations, o.skip = 4000, 5, 2
    o.precision, o.method = _abi.FP64, method
    o.source, o.bank_init = _abi.SOURCE_FISSION_BANK, _abi.BANK_INIT_DIFFUSION
    print(f"diffusion-started bank: method {method}", flush=True)
    d = nr.diffusion_solve(p)
    st, x, c = oracle.diffusion_bank(p, d.fission_source, o.histories)
    assert st == 0
    want = oracle.solve_init(p, o, x, c, d.k)
    got = nr.solve(p, o)
    assert np.array_equal(got.k.view(np.uint64), want.k.view(np.uint64))
    assert np.array_equal(got.sites, want.sites)
    for f in ("collisions", "absorptions", "leaks", "capped"):
        assert getattr(got, f) == getattr(want, f), f
    np.testing.assert_allclose(got.tally_last, want.tally_last, rtol=1e-10, atol=1e-10)
    # (the event path's difference-array tally leaves ~1e-15-relative dust in
    # cells whose oracle flux is exactly zero)
    np.testing.assert_allclose(got.flux, want.flux, rtol=1e-10, atol=1e-12 * np.abs(want.flux).max())


@pytest.mark.gpu
def test_diffusion_start_in_batch_and_server(gpu):
    """Diffusion-started FP32 requests give the same k, bit for bit, solved
    alone, in one batch and through the continuous-batching server."""
    probs = []
    for i in range(4):
        p, o = nr.generate(3, i)
        o.histories, o.generations, o.skip = min(int(o.histories), 15000), 8, 2
        o.bank_init = _abi.BANK_INIT_DIFFUSION
        probs.append((p, o))
    print("diffusion start: solo, batch, server", flush=True)
    solo = [nr.solve(p, o) for p, o in probs]
    batch = nr.solve_batch([p for p, _ in probs], [o for _, o in probs])
    with nr.Server(precision=_abi.FP32, budget_bytes=8 * 10**9, max_slots=2) as srv:
        tickets = [srv.submit(p, o) for p, o in probs]
        served = [srv.wait(t, timeout_ms=120000) for t in tickets]
    for s, b, v in zip(solo, batch, served):
        assert s.status_code == 0 and v.status_code == 0
        assert s.sites[0] > 0  # generation 0 banks sites
        assert np.array_equal(s.k.view(np.uint64), b.k.view(np.uint64))
        assert np.array_equal(s.k.view(np.uint64), v.k.view(np.uint64))
