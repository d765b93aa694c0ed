"""Output plots (plot.py of the reference, generalised to G groups): the
parser reads the reference's own shipped 4-group TestCaseC outputs, and
files written by nraps_write_outputs / nraps_write_diffusion_outputs round
trip through it for 2 and 4 groups."""
import os

import numpy as np
import pytest

import nraps_amd as nr
from nraps_amd import plot
from nraps_amd.problem import DiffusionBuffers

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference")


def test_reads_the_shipped_four_group_outputs():
    d = plot.read_outputs(REF)
    assert d["groups"] == 4 and d["meshed"] == 408 and d["generations"] == 100
    assert d["flux"].shape == (4, 408) and d["average"].shape == (4, 408)
    assert d["fission"].shape == (408,)
    assert d["k"].size == 100 and d["k_fund"].size == 100
    assert d["length"] == pytest.approx(42.908)


@pytest.mark.parametrize("deck,G", [("TestCaseA", 2), ("TestCaseC", 4)])
def test_written_outputs_round_trip(tmp_path, deck, G):
    p, info = nr.load_deck(os.path.join(REF, deck + ".txt"))
    assert p.n_groups == G
    o = nr.run_opts(generations=7, skip=2)
    rb = nr.ResultBuffers(p, o)
    rng = np.random.default_rng(G)
    for name in ("flux", "assembly_average", "fission_source", "k", "k_fund"):
        getattr(rb, name)[:] = rng.uniform(0.1, 3.0, getattr(rb, name).shape) * 10.0 ** rng.integers(-3, 18)
    nr.write_outputs(str(tmp_path), p, o, rb)
    d = plot.read_outputs(str(tmp_path))
    assert d["groups"] == G and d["generations"] == 7
    assert np.array_equal(d["flux"].ravel(), rb.flux)
    assert np.array_equal(d["average"].ravel(), rb.assembly_average)
    assert np.array_equal(d["fission"], rb.fission_source)
    assert np.array_equal(d["k"], rb.k) and np.array_equal(d["k_fund"], rb.k_fund)
    paths = plot.plot(str(tmp_path))
    names = {os.path.basename(x) for x in paths}
    assert names == {"k_eff.png", "fission.png", *plot.flux_figure_names(G)}
    assert all(os.path.getsize(x) > 0 for x in paths)


def test_diffusion_outputs_round_trip(tmp_path):
    p, info = nr.load_deck(os.path.join(REF, "TestCaseB.txt"))
    db = DiffusionBuffers(p)
    db.flux[:] = np.linspace(1.0, 2.0, db.flux.size) * 1e14
    db.assembly_average[:] = 3.0e13
    db.fission_source[:] = np.linspace(0.5, 0.7, db.fission_source.size)
    db.c.k = 1.1972997251893478
    nr.write_diffusion_outputs(str(tmp_path), p, info.generations, db)
    d = plot.read_outputs(str(tmp_path))
    assert d["groups"] == 2 and d["k"].tolist() == [1.1972997251893478] and d["k_fund"].size == 0
    assert np.array_equal(d["flux"].ravel(), db.flux)
    assert np.array_equal(d["fission"], db.fission_source)
    assert {os.path.basename(x) for x in plot.plot(str(tmp_path))} == {
        "k_eff.png", "fission.png", "fast_flux.png", "thermal_flux.png"}
