"""The benchmark configurations as parity cases.

Config 1 (the reference-runnable case): N=4 bosons, P=32, He-4, T=1 K, free
pairs, harmonic trap 3 meV, no PBC, positions from the reference's `random`
initialisation in a 1 nm box (mt19937(seed + bead), simulation.cpp:708-728),
Maxwell-Boltzmann momenta, dt = 1 fs, Cartesian Langevin: 10^4 steps with the
reference's RanMars noise fed from the host, GPU against the oracle.

Config 2: N=64 bosons, P=64, T=2 K, trap 3 meV plus harmonic pairs 1 meV (all
pairs), uniform positions in a 2 nm box: 100 host-noise steps.

Config 3: N=256 He-4, P=64, Aziz, PBC, cutoff L/2, dt = 0.5 fs, normal-mode
Langevin, fixcom (the benchmark's own initial state): 100 host-noise steps.

Config 4: the 512 uneven two-dimensional dipole replicas of the benchmark in
one batch: every replica's forces against the oracle, and 100 device-noise
steps of every replica against a single-system context with its seed.

Reference: Simulation::run (src/simulation.cpp:222-290), langevin.cpp:15-27,
random_mars.cpp:83-103, updateForces (simulation.cpp:353-455).
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
import pimdb_amd as G
from helpers import AMU, ANGSTROM, FS, KELVIN, MEV, per_bead_err, rel_err

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def host_noise_run(kw, nsteps, x=None, p=None):
    """Oracle and GPU from the same state (the oracle's reference-style
    initialisation when x is None) with the same RanMars noise."""
    sim = O.Sim(O.make_params(**kw))
    if x is None:
        sim.init_random()
    else:
        sim.set_state(x, p)
        sim.forces()
    x0, p0, f0 = sim.state()
    ctx = G.Context(G.make_params(**kw))
    ctx.set_state(x0, p0)
    if x is not None:
        ctx.forces()
    noise = np.stack([sim.draw_noise() for _ in range(nsteps)])
    for s in range(nsteps):
        sim.step(noise[s])
    ctx.step(nsteps, noise)
    g = ctx.state()
    ctx.close()
    return g, sim.state()


def config1():
    return dict(natoms=4, nbeads=32, ndim=3, mass=4.0026 * AMU, temperature=1.0 * KELVIN, dt=1.0 * FS,
                gamma=1.0 / (100.0 * FS), bosonic=1, pbc=0, size=10.0 * ANGSTROM, ext_kind="harmonic",
                ext_omega=3.0 * MEV, int_kind="free", cutoff=-1.0, propagator="cartesian",
                thermostat="langevin", nmthermostat=0, fixcom=1, seed=12345)


def test_config1_ten_thousand_host_noise_steps():
    kw = config1()
    (xg, pg, fg), (xo, po, fo) = host_noise_run(kw, 10000)
    ex, ep, ef = rel_err(xg, xo), rel_err(pg, po), per_bead_err(fg, fo)
    print("config 1 after 1e4 steps: x %.2e p %.2e f %.2e" % (ex, ep, ef))
    assert ex < 1e-9 and ep < 1e-9 and ef < 1e-9


def test_config2_hundred_host_noise_steps():
    L = 20.0 * ANGSTROM
    rng = np.random.default_rng(202)
    N, P = 64, 64
    kw = dict(natoms=N, nbeads=P, ndim=3, mass=4.0026 * AMU, temperature=2.0 * KELVIN, dt=1.0 * FS,
              gamma=1.0 / (100.0 * FS), bosonic=1, pbc=0, size=L, ext_kind="harmonic", ext_omega=3.0 * MEV,
              int_kind="harmonic", int_omega=1.0 * MEV, cutoff=-1.0, propagator="cartesian",
              thermostat="langevin", nmthermostat=0, fixcom=1, seed=2002)
    x = rng.uniform(-0.5 * L, 0.5 * L, size=(P, N, 3))
    p = rng.normal(0.0, 0.2, size=x.shape)
    (xg, pg, fg), (xo, po, fo) = host_noise_run(kw, 100, x, p)
    ex, ep, ef = rel_err(xg, xo), rel_err(pg, po), per_bead_err(fg, fo)
    print("config 2 after 100 steps: x %.2e p %.2e f %.2e" % (ex, ep, ef))
    assert ex < 1e-10 and ep < 1e-10 and ef < 1e-10


def test_config3_hundred_host_noise_steps():
    kw, x, p = bench.workload()
    (xg, pg, fg), (xo, po, fo) = host_noise_run(kw, 100, x, p)
    ex, ep, ef = rel_err(xg, xo), rel_err(pg, po), per_bead_err(fg, fo)
    print("config 3 after 100 steps: x %.2e p %.2e f %.2e" % (ex, ep, ef))
    assert ex < 1e-10 and ep < 1e-9 and ef < 1e-9


def test_config4_all_512_replicas():
    reps = bench.uneven_dipole_replicas()
    b = G.Batch([G.make_params(**kw) for kw, _, _ in reps])
    for r, (kw, x, p) in enumerate(reps):
        b.set_state(r, x, p)
    threads, resident = b.autotune(trial_steps=2)
    print("config 4 launch shape: %d threads, %d workgroups" % (threads, resident))
    b.forces()
    worst = 0.0
    for r, (kw, x, p) in enumerate(reps):
        _, _, f = b.state(r)
        worst = max(worst, per_bead_err(f, O.update_forces(O.make_params(**kw), x)))
    print("config 4 worst per-replica force error %.2e" % worst)
    assert worst < 1e-10
    b.step(100)
    wx = wp = wf = 0.0
    for r, (kw, x, p) in enumerate(reps):
        c = G.Context(G.make_params(**kw))
        c.set_state(x, p)
        c.forces()
        c.step(100)
        xc, pc, fc = c.state()
        c.close()
        xb, pb, fb = b.state(r)
        wx, wp, wf = max(wx, rel_err(xb, xc)), max(wp, rel_err(pb, pc)), max(wf, per_bead_err(fb, fc))
    print("config 4 after 100 steps, worst replica: x %.2e p %.2e f %.2e" % (wx, wp, wf))
    b.close()
    assert wx < 1e-10 and wp < 1e-9 and wf < 1e-9
