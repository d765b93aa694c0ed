"""Plots of a run's CSV outputs (vars.csv, k_eff.csv, interface.csv).

    python -m nraps_amd.plot [DIR]

The reference's plot.py (run by plot_solution, src/plot_solution.rs:60)
reads the same three files but parses interface.csv as exactly two flux rows,
two assembly-average rows and the fission row (plot.py:26-37).  Here the
group count comes from the file: interface.csv holds G flux rows, G
assembly-average rows and one fission-source row, so G = (rows - 1) / 2.
The figures are the reference's: k_eff.png (k per generation and the running
active mean), the flux with its assembly average per group (fast_flux.png and
thermal_flux.png for two groups, as the reference names them, else
group_<g>_flux.png), and fission.png.
"""
import os
import sys

import numpy as np


def _row(line):
    line = line.strip()
    return np.array(line.split(","), dtype=float) if line else np.zeros(0)


def read_outputs(directory="."):
    """Parse the three CSV files of one run (any number of groups)."""
    with open(os.path.join(directory, "vars.csv")) as f:
        v = [x.strip() for x in f.read().splitlines()]
    length, meshed, generations = float(v[0]), int(v[1]), int(v[2])
    with open(os.path.join(directory, "k_eff.csv")) as f:
        krows = f.read().splitlines()
    k = _row(krows[0]) if krows else np.zeros(0)
    k_fund = _row(krows[1]) if len(krows) > 1 else np.zeros(0)
    with open(os.path.join(directory, "interface.csv")) as f:
        rows = [_row(x) for x in f.read().splitlines()]
    if len(rows) < 3 or len(rows) % 2 == 0:
        raise ValueError(f"interface.csv has {len(rows)} rows; expected 2G + 1")
    G = (len(rows) - 1) // 2
    flux = np.vstack(rows[:G])
    average = np.vstack(rows[G:2 * G])
    fission = rows[2 * G]
    if flux.shape[1] != meshed:
        raise ValueError(f"interface.csv rows have {flux.shape[1]} values; vars.csv says {meshed}")
    return dict(length=length, meshed=meshed, generations=generations, k=k, k_fund=k_fund,
                groups=G, flux=flux, average=average, fission=fission)


def flux_figure_names(G):
    if G == 2:
        return ["fast_flux.png", "thermal_flux.png"]
    return [f"group_{g + 1}_flux.png" for g in range(G)]


def plot(directory=".", out=None):
    """Write the figures next to the CSV files (or into `out`); returns their paths."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    d = read_outputs(directory)
    out = out or directory
    os.makedirs(out, exist_ok=True)
    x = np.linspace(0.0, d["length"], num=d["meshed"])
    paths = []

    def save(name):
        p = os.path.join(out, name)
        plt.savefig(p, bbox_inches="tight")
        plt.clf()
        paths.append(p)

    plt.plot(d["k"], "o", markerfacecolor="none", label="multiplication factor")
    if d["k_fund"].size:
        plt.plot(d["k_fund"], color="orange")
    plt.title("Multiplication Factor")
    plt.xlim(0, max(d["generations"], 1))
    plt.ylim(0, 2)
    save("k_eff.png")
    titles = ["Fast Flux", "Thermal Flux"] if d["groups"] == 2 else \
        [f"Group {g + 1} Flux" for g in range(d["groups"])]
    for g, (name, title) in enumerate(zip(flux_figure_names(d["groups"]), titles)):
        plt.plot(x, d["flux"][g], label=title.lower())
        plt.plot(x, d["average"][g], color="orange")
        plt.title(title)
        plt.xlim(0, d["length"])
        save(name)
    plt.plot(x, d["fission"], label="fission")
    plt.title("Fission Density")
    plt.xlim(0, d["length"])
    save("fission.png")
    plt.close("all")
    return paths


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    directory = argv[0] if argv else "."
    for p in plot(directory):
        print(p)


if __name__ == "__main__":
    main()
