"""Writes the golden fixtures under tests/golden/ from the reference's own
regression data (higj/pimd-b tests/cases/<case>/).

The fixtures are data only: the case .ini inputs, the expected per-bead
trajectories (positions in angstrom, velocities in angstrom/ps, forces in
eV/angstrom, as the reference prints them, src/states/{position,velocity,
force}.cpp) and output/simulation.out (src/observables/observable.cpp:61-116).
The first NFRAMES frames and NROWS rows are kept in float64 (compared at 1e-9
and below); every stored frame is also kept in float32 (`*_all`, compared at
the reference harness's rtol 1e-5, tests/main.py:77-86, which float32's 6e-8
rounding leaves intact), and every simulation.out row in float64
(`sim_out_all`).  Run once in a container that has the reference tree:

    python tests/golden/make_fixtures.py [/root/reference]
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NFRAMES = 21          # frame 0 .. 20 of every trajectory
NROWS = 21            # simulation.out rows kept

DYNAMICS = [
    "bosonic_quadratic_harmonic_dynamics",
    "bosonic_quadratic_harmonic_nmthermostat_dynamics",
    "dist_harmonic_dynamics",
    "dist_harmonic_nm_propagation_dynamics",
    "bosonic_factorial_harmonic_dynamics",
    "bosonic_quadratic_harmonic_nh_dynamics",
    "bosonic_quadratic_harmonic_nh_np_dynamics",
    "bosonic_quadratic_harmonic_nh_np_dim_dynamics",
]
OUT_ONLY = ["bosonic_quadratic_harmonic", "dist_harmonic", "bosonic_factorial_harmonic"]


def read_frames(path, nframes):
    """Frames of an xyz/dat state file: the last three numeric tokens of each
    particle row (the reference harness reads them the same way,
    tests/main.py:145-161)."""
    frames = []
    with open(path) as fh:
        lines = fh.read().split("\n")
    i = 0
    while i < len(lines) and lines[i].strip() and len(frames) < nframes:
        n = int(lines[i])
        rows = [[float(t) for t in lines[i + 2 + r].split()[-3:]] for r in range(n)]
        frames.append(rows)
        i += 2 + n
    return np.array(frames, dtype=np.float64)


def read_sim_out(path, nrows):
    with open(path) as fh:
        lines = [l for l in fh.read().split("\n") if l.strip()]
    header = lines[0].split()
    rows = np.array([[float(t) for t in l.split()] for l in lines[1:1 + nrows]])
    return header, rows


def main(ref):
    cases = os.path.join(ref, "tests", "cases")
    os.makedirs(os.path.join(HERE, "cases"), exist_ok=True)
    for case in DYNAMICS + OUT_ONLY:
        src = os.path.join(cases, case)
        shutil.copyfile(os.path.join(src, case + ".ini"), os.path.join(HERE, "cases", case + ".ini"))
        header, rows = read_sim_out(os.path.join(src, "simulation.out"), NROWS)
        _, all_rows = read_sim_out(os.path.join(src, "simulation.out"), 10 ** 9)
        data = dict(sim_out_header=np.array(header), sim_out=rows, sim_out_all=all_rows)
        if case in DYNAMICS:
            nbeads = len([f for f in os.listdir(src) if f.startswith("position_")])
            for name, stem in (("positions", "position_%d.xyz"), ("velocities", "velocity_%d.dat"),
                               ("forces", "force_%d.dat")):
                full = np.stack([read_frames(os.path.join(src, stem % b), 10 ** 9)
                                 for b in range(nbeads)], axis=1)
                data[name] = full[:NFRAMES].copy()
                data[name + "_all"] = full.astype(np.float32)
        np.savez_compressed(os.path.join(HERE, case + ".npz"), **data)
        print("wrote", case, {k: v.shape for k, v in data.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
