"""Host-side pieces of the library (no GPU): the input-deck reader and mesh
expansion (src/process_input.rs, src/main.rs:85-143), the CSV writers
(src/plot_solution.rs) checked byte for byte against the reference's shipped
output files, the seeded workload generators, validation, and the ABI.
"""
import ctypes as C
import json
import math
import os
import re

import numpy as np
import pytest

import nraps_amd as nr
from nraps_amd import _abi
from conftest import GOLDEN, REFDATA, ROOT


def deck(name):
    return nr.load_deck(os.path.join(REFDATA, name))


# ---- input deck + mesh_gen --------------------------------------------------
def test_testcase_c_deck_values():
    p, d = deck("TestCaseC.txt")
    assert (d.mattypes, d.energygroups, d.generations, d.histories, d.skip) == (4, 4, 100, 10**6, 1)
    assert (d.numass, d.numrods, d.mpfr, d.mpwr) == (2, 17, 8, 4)
    assert d.roddia == 0.94 and abs(d.rodpitch - 0.322) < 1e-15
    assert (d.boundl, d.boundr, d.solution_mc) == (1.0, 1.0, 1)
    # XS in the reference layout mat + M*g: both SigT lines appended
    assert list(p.sigt[:8]) == [0.207642, 0.208183, 0.176541, 0.1, 0.599466, 0.591052, 0.74598, 1.1]
    assert list(p.sigt[8:]) == [0.833793, 0.765393, 1.078194, 0.1, 1.330174, 1.46294, 1.946933, 1.1]
    assert p.scat_matrix.shape == (64,) and p.scat_matrix[16 * 1 + 4 * 2 + 3] == 0.004219
    assert list(p.chit[:4]) == [0.698645, 0.709543, 0.0, 0.0]


def test_mesh_gen_matches_shipped_vars_csv():
    p, d = deck("TestCaseC.txt")
    with open(os.path.join(REFDATA, "vars.csv")) as f:
        lines = f.read().split("\n")
    assert p.n_cells == int(lines[1]) == 408
    # running f64 sum of delta_x reproduces the shipped length to the last digit
    assert repr(float(p.mesh_right[-1])) == lines[0] == "42.90800000000001"
    assert np.all(p.mesh_right[:-1] == p.mesh_left[1:])
    # MPWR/2 = 2 water cells are drained at each end (src/main.rs:105-108)
    assert p.matid[:3].tolist() == [2, 2, 0] and p.matid[-3:].tolist() == [1, 2, 2]
    # fuel cells are matid 0/1; the two assemblies split at the middle
    fuel = p.fuel_indices
    assert np.all(np.isin(p.matid[fuel], [0, 1]))
    assert np.all(p.matid[:204][np.isin(p.matid[:204], [0, 1])] == 0)
    assert np.all(p.matid[204:][np.isin(p.matid[204:], [0, 1])] == 1)


@pytest.mark.parametrize("name,G,nc", [("TestCaseA.txt", 2, 408),
                                       # two control-rod positions (matid 3) expand
                                       # to MPWR cells instead of MPFR
                                       ("TestCaseB.txt", 2, 400)])
def test_two_group_decks(name, G, nc):
    p, d = deck(name)
    assert p.n_groups == G and p.n_cells == nc and p.n_mats == 4
    assert nr.validate(p, nr.run_opts()) == 0


def test_sample_deck_old_keys_are_dropped():
    # SigTR/SigIS/SigDS keys hash to the catch-all bucket (src/process_input.rs:40):
    # the reader reports what is missing instead of guessing
    with pytest.raises(nr.NrapsError):
        deck("SampleInputFile.txt")


def test_missing_deck_is_an_io_error(tmp_path):
    with pytest.raises(nr.NrapsError) as e:
        nr.load_deck(str(tmp_path / "nope.txt"))
    assert e.value.status == _abi.NRAPS_E_IO


# ---- output writers ---------------------------------------------------------
def _rows(path):
    with open(path) as f:
        return [np.array([float(v) for v in line.split(",")]) for line in f.read().splitlines()]


def test_writers_reproduce_shipped_files_byte_for_byte(tmp_path):
    p, d = deck("TestCaseC.txt")
    o = nr.run_opts(generations=100, skip=1)
    rb = nr.ResultBuffers(p, o)
    iface = _rows(os.path.join(REFDATA, "interface.csv"))
    keff = _rows(os.path.join(REFDATA, "k_eff.csv"))
    G, nc = 4, p.n_cells
    rb.flux[:] = np.concatenate(iface[:G])
    rb.assembly_average[:] = np.concatenate(iface[G:2 * G])
    rb.fission_source[:] = iface[2 * G]
    rb.k[:] = keff[0]
    rb.k_fund[:] = keff[1]
    nr.write_outputs(str(tmp_path), p, o, rb)
    for name in ("vars.csv", "interface.csv", "k_eff.csv"):
        with open(tmp_path / name, "rb") as a, open(os.path.join(REFDATA, name), "rb") as b:
            assert a.read() == b.read(), name


# ---- generators -------------------------------------------------------------
def _pinned():
    with open(os.path.join(GOLDEN, "workloads.json")) as f:
        return json.load(f)["workloads"]


@pytest.mark.parametrize("key,config,index", [("slab9_g4", 0, 0), ("slab201_g8", 2, 0),
                                              ("batch_0", 3, 0), ("batch_1", 3, 1),
                                              ("batch_2", 3, 2), ("batch_3", 3, 3)])
def test_generators_match_pinned_workloads(key, config, index):
    w = _pinned()[key]
    p, o = nr.generate(config, index)
    assert (p.n_cells, p.n_groups, int(o.histories)) == (w["n_cells"], w["n_groups"], w["histories"])
    for name in ("matid", "mesh_left", "mesh_right", "delta_x", "sigt", "sigs", "siga", "sigf",
                 "nut", "chit", "scat_matrix"):
        assert np.array_equal(getattr(p, name), np.array(w[name])), name


def _check_xs(p, upscatter):
    G, M = p.n_groups, p.n_mats
    sigt, sigs = p.sigt.reshape(G, M), p.sigs.reshape(G, M)
    scat = p.scat_matrix.reshape(M, G, G)
    nsf = (p.nut * p.sigf).reshape(G, M)
    assert np.all((sigt[:, 0] >= 0.2) & (sigt[:, 0] <= 0.6))
    assert np.all((sigs[:, 0] / sigt[:, 0] >= 0.5) & (sigs[:, 0] / sigt[:, 0] <= 0.9))
    assert np.all((sigt[:, 2] >= 0.3) & (sigt[:, 2] <= 2.0))
    assert np.all((sigs[:, 2] / sigt[:, 2] >= 0.95) & (sigs[:, 2] / sigt[:, 2] <= 0.99))
    assert np.all(nsf[:G // 2, 0] == 0) and np.all((nsf[G // 2:, 0] >= 0.005) & (nsf[G // 2:, 0] <= 0.1))
    assert np.all(nsf[:, 2] == 0)
    assert np.array_equal(sigt[:, 0], sigt[:, 1])  # material 1 copies the fuel
    np.testing.assert_allclose(scat.sum(2), sigs.T, rtol=1e-14)
    for m in range(M):
        for g in range(G):
            row = scat[m, g]
            if upscatter and g >= max(1, G - 2):
                assert row[g - 1] <= 0.05 * sigs[g, m] + 1e-15
                assert np.all(row[:g - 1] == 0)
            else:
                assert np.all(row[:g] == 0)
    assert list(p.chit.reshape(G, M)[:2, 0]) == [0.75, 0.25]


def test_config0_geometry_and_xs():
    p, o = nr.generate(0)
    assert p.n_cells == 180 and p.n_groups == 4 and (p.boundl, p.boundr) == (1.0, 0.0)
    assert abs(p.mesh_right[-1] - 7.0) < 1e-12
    _check_xs(p, upscatter=False)
    _, o1 = nr.generate(1)
    assert (o1.histories, o1.generations, o1.skip) == (10**7, 120, 20)


def test_config2_geometry_and_xs():
    p, o = nr.generate(2)
    regions = 1 + int(np.sum(np.diff(p.matid) != 0))
    assert regions == 201 and p.n_cells == 2000 and p.n_groups == 8
    w = np.array([p.mesh_right[p.matid == m].size for m in (0, 2)])
    assert w.sum() == 2000
    edges = np.flatnonzero(np.diff(p.matid) != 0)
    widths = np.diff(np.concatenate([[0.0], p.mesh_right[edges], [p.mesh_right[-1]]]))
    assert np.all((widths >= 0.4 - 1e-12) & (widths <= 1.2 + 1e-12))
    _check_xs(p, upscatter=True)
    assert o.source == _abi.SOURCE_FISSION_BANK and o.precision == _abi.FP32


def test_config3_batch_ranges():
    for i in range(12):
        p, o = nr.generate(3, i)
        regions = 1 + int(np.sum(np.diff(p.matid) != 0))
        assert 3 <= regions <= 201 and p.n_cells == 10 * regions
        assert p.n_groups in (2, 4, 8)
        assert 10**4 <= o.histories <= 10**6 and o.problem_id == i
        _check_xs(p, upscatter=p.n_groups >= 4)


# ---- validation -------------------------------------------------------------
@pytest.mark.parametrize("change,msg", [
    (dict(skip=0), "skip"), (dict(skip=150), "skip"), (dict(histories=0), "histories"),
    (dict(precision=_abi.FP32, method=_abi.HISTORY_SURFACE), "FP64"),
    (dict(source=_abi.SOURCE_FISSION_BANK, skip=1), "skip"),
    (dict(k_initial=0.0), "k_initial"), (dict(problem_id=1 << 19), "problem_id"),
])
def test_invalid_options_are_rejected(change, msg):
    p, _ = nr.generate(0)
    o = nr.run_opts(**change)
    assert nr.validate(p, o) == _abi.NRAPS_E_INVALID
    assert msg in nr.last_error()


def test_invalid_meshes_are_rejected():
    p, _ = nr.generate(0)
    p.matid = p.matid.copy()
    p.matid[5] = 7
    assert nr.validate(p, nr.run_opts()) == _abi.NRAPS_E_INVALID
    p, _ = nr.generate(0)
    p.mesh_left = p.mesh_left.copy()
    p.mesh_left[10] += 1e-9
    assert nr.validate(p, nr.run_opts()) == _abi.NRAPS_E_INVALID
    p, _ = nr.generate(0)
    p.boundr = 1.5
    assert nr.validate(p, nr.run_opts()) == _abi.NRAPS_E_INVALID


# ---- ABI --------------------------------------------------------------------
def test_library_exports_every_declared_entry_point():
    with open(os.path.join(ROOT, "include", "nraps.h")) as f:
        text = f.read()
    names = set(re.findall(r"\b(nraps_[a-z0-9_]+)\s*\(", text))
    assert len(names) >= 19
    L = nr.lib()
    for n in sorted(names):
        assert hasattr(L, n), n


def test_no_gpu_means_an_error_not_a_fallback():
    from conftest import has_gpu
    if has_gpu():
        pytest.skip("a GPU is present")
    p, o = nr.generate(0)
    o.histories = 10
    rb = nr.solve(p, o, check=False)
    assert rb.status_code == _abi.NRAPS_E_DEVICE
    assert "device" in nr.last_error()


def test_command_line_driver_usage_and_no_gpu_error():
    """nraps_run (deck -> solve -> CSV) is built beside the library; bad
    arguments print the usage (status 2) and without a device the solve
    fails loudly (status 1), with no fallback."""
    import subprocess
    cli = os.path.join(ROOT, "nraps_amd", "nraps_run")
    assert os.access(cli, os.X_OK)
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([cli, os.path.join(REFDATA, "TestCaseA.txt"), "--bogus"], capture_output=True,
                       text=True)
    assert r.returncode == 2
    from conftest import has_gpu
    if not has_gpu():
        r = subprocess.run([cli, os.path.join(REFDATA, "TestCaseA.txt"), "/tmp", "--histories",
                            "100"], capture_output=True, text=True)
        assert r.returncode == 1 and "no HIP device" in r.stderr


def test_serving_arrivals_use_pcg32():
    """The serving benchmark's arrival generator is PCG32 (pcg-c-basic KAT
    srandom(42, 54)) and gives Poisson arrivals at the requested rate."""
    from nraps_amd.serve_bench import Pcg32, arrival_times, percentile
    g = Pcg32(42, 54)
    assert [g.next() for _ in range(6)] == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293,
                                            0xbfa4784b, 0xcbed606e]
    t = arrival_times(200.0, 60.0)
    assert abs(len(t) - 12000) < 5 * math.sqrt(12000)
    assert all(0 <= a < b < 60.0 for a, b in zip(t, t[1:]))
    assert percentile([3.0, 1.0, 2.0], 0.5) == 2.0
